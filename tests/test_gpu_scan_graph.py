"""klsp_grid_scan as one captured HIP graph, against the same scan launched
kernel by kernel (klsp_set_scan_graph(ctx, 0)) on the published small grid.

The graph reads its omegas from device memory at launch, so a later scan of
the same size on the same context launches the same graph with new omegas.
Every graph launch below is checked against the kernel-by-kernel result of
the same window, with eager dets and synchronisations of the program's own
between the launches; klsp_scan_graph_stats shows which launches reused an
instantiated graph.  The kernels and their order are the same in both forms,
so the dets must be bit-identical.
"""
import numpy as np
import pytest

from kalnajs_log_spirals_amd import _ffi
from kalnajs_log_spirals_amd.BasisFunctions import from_arrays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_pre(oracle_small):
    o = oracle_small
    pre = from_arrays(o["m"], o["G"], o["W"], o["DJ"], o["F0l"], o["Omega_res"], o["N_kernel"])
    yield pre
    pre.ctx.close()


def eager_scan(ctx, *args, **kw):
    ctx.set_scan_graph(False)
    try:
        return ctx.grid_scan(*args, **kw)
    finally:
        ctx.set_scan_graph(True)


@pytest.mark.parametrize("prec", [_ffi.PREC_FP64, _ffi.PREC_BF16X3])
def test_graph_scan_equals_kernel_by_kernel_on_every_launch(small_pre, prec):
    ctx = small_pre.ctx
    ctx.set_scan_graph(True)
    # 7 x 5 = 35 omegas: BF16x3 batches of three leave a remainder of two
    win_a = (7, 5, 0.40, 0.48, 0.08, 0.16)
    win_b = (7, 5, 0.30, 0.60, 0.02, 0.25)
    ref_a = eager_scan(ctx, *win_a, precision=prec)
    ref_b = eager_scan(ctx, *win_b, precision=prec)
    b0, l0 = ctx.scan_graph_stats()
    seq = [win_a, win_b, win_a, win_b]
    for k, win in enumerate(seq):
        # the program's own eager work between graph launches
        ctx.det([0.88 + 0.12j, 0.7 + 0.05j], prec)
        dg, op0, g0 = ctx.grid_scan(*win, precision=prec)
        ref = ref_a if win is win_a else ref_b
        assert np.array_equal(dg, ref[0]), (k, np.max(np.abs(dg - ref[0])))
        assert (op0, g0) == (ref[1], ref[2])
    b1, l1 = ctx.scan_graph_stats()
    assert l1 - l0 == len(seq)
    assert b1 - b0 == 1, "scans of one size and precision reuse one instantiated graph"


def test_graph_scan_new_size_instantiates_again(small_pre):
    ctx = small_pre.ctx
    b0, _ = ctx.scan_graph_stats()
    for n_op, n_g in [(1, 1), (2, 1), (4, 4)]:
        ref = eager_scan(ctx, n_op, n_g, 0.41, 0.47, 0.10, 0.14, precision=_ffi.PREC_BF16X3)
        dg = ctx.grid_scan(n_op, n_g, 0.41, 0.47, 0.10, 0.14, precision=_ffi.PREC_BF16X3)
        assert np.array_equal(dg[0], ref[0])
    b1, _ = ctx.scan_graph_stats()
    assert b1 - b0 == 3


def test_graph_scan_follows_a_new_problem(oracle_small):
    """A context given a different problem of another size instantiates a
    new graph (sizes and buffers are part of the key) and matches the
    kernel-by-kernel scan of the new problem."""
    o = oracle_small
    pre = from_arrays(o["m"], o["G"], o["W"], o["DJ"], o["F0l"], o["Omega_res"], o["N_kernel"])
    ctx = pre.ctx
    try:
        win = (5, 4, 0.40, 0.48, 0.08, 0.16)
        first = ctx.grid_scan(*win, precision=_ffi.PREC_FP64)
        assert np.array_equal(first[0], eager_scan(ctx, *win, precision=_ffi.PREC_FP64)[0])
        # a problem with fewer alpha columns (an asymmetric alpha window, so
        # the builds take the complex kernels) uploaded into the same context
        NPh, N, n_l = o["W"].shape
        N2 = N - 8
        W2 = np.ascontiguousarray(o["W"][:, 3:3 + N2, :])
        ctx.upload_problem(NPh, N2, n_l, o["m"], o["G"], o["DJ"], o["F0l"], o["Omega_res"],
                           o["N_kernel"][3:3 + N2])
        ctx.upload_W(W2)
        assert not ctx.symmetric_basis()
        b0, _ = ctx.scan_graph_stats()
        dg = ctx.grid_scan(*win, precision=_ffi.PREC_FP64)
        ref = eager_scan(ctx, *win, precision=_ffi.PREC_FP64)
        assert np.array_equal(dg[0], ref[0])
        assert ctx.scan_graph_stats()[0] == b0 + 1
    finally:
        ctx.close()


@pytest.mark.parametrize("prec", [_ffi.PREC_FP64, _ffi.PREC_BF16X3])
def test_graph_scan_over_several_lu_launches(small_pre, prec):
    """132 omegas: several LU launches of whole three-omega batches (the
    grouping both forms share) and a short last launch."""
    ctx = small_pre.ctx
    win = (12, 11, 0.30, 0.60, 0.02, 0.25)
    ref = eager_scan(ctx, *win, precision=prec)
    for _ in range(2):
        dg = ctx.grid_scan(*win, precision=prec)
        assert np.array_equal(dg[0], ref[0])
        ctx.det([0.9 + 0.1j], prec)
