"""The uneven batch packed onto the device (several contexts on their own
streams, driven from host threads) against the one-at-a-time batch: the
same roots, bit for bit, for draws of the 48-draw batch at reduced grid
size; a budget that holds only about one draw at a time makes workers
release and recreate their contexts."""
import numpy as np
import pytest

from kalnajs_log_spirals_amd.UnevenBatch import config_for, draw_batch, footprint_bytes, run_batch

pytestmark = pytest.mark.gpu

SCALE = 0.15


def same(a, b):
    assert a.draw.index == b.draw.index
    assert a.converged == b.converged
    assert (a.Omega_p, a.gamma) == (b.Omega_p, b.gamma), (a.draw.index, a.Omega_p, b.Omega_p)
    assert a.n_fp64_builds == b.n_fp64_builds and a.n_lp_builds == b.n_lp_builds
    if a.converged and not np.isnan(a.polished):
        assert a.polished == b.polished


@pytest.mark.parametrize("workers", [2, 4])
def test_packed_batch_equals_one_at_a_time(workers):
    draws = draw_batch(48, 0)[:10]
    seq = run_batch(draws, precision="mixed", scale=SCALE, orbits="device", polish=1)
    packed = run_batch(draws, precision="mixed", scale=SCALE, orbits="device", polish=1,
                       workers=workers)
    assert len(packed) == len(seq)
    for a, b in zip(seq, packed):
        same(a, b)


def test_packed_batch_under_a_tight_budget():
    draws = draw_batch(48, 0)[:6]
    cfgs = [config_for(d, SCALE) for d in draws]
    big = max(sum(footprint_bytes(c, "fp64")) for c in cfgs)
    seq = run_batch(draws, precision="fp64", scale=SCALE, orbits="device")
    packed = run_batch(draws, precision="fp64", scale=SCALE, orbits="device", workers=3,
                       budget=int(1.3 * big))
    for a, b in zip(seq, packed):
        same(a, b)
