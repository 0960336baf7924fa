"""The command-line runner end to end on the device, on the published small
grid (convergence_table.txt:13: Omega_p = 0.4397142267, gamma = 0.1230663944,
printed to 10 decimals): the reference's own run (Newton from the model
file's reference root, FP64 builds) and the scan-seeded mixed-precision run
must both land on the published root, and their log files must carry the
reference runner's result block (run_kalnajs_gpu.jl:310-321)."""
import os
import re

import pytest

from kalnajs_log_spirals_amd import run_kalnajs

from conftest import CONFIGS

pytestmark = pytest.mark.gpu

KAT = (0.4397142267, 0.1230663944)


def _run(tmp_path, *opts):
    text = open(os.path.join(CONFIGS, "small.toml")).read()
    out = tmp_path / "results"
    cfg = tmp_path / "small.toml"
    cfg.write_text(text.replace('output_path = "results"', f'output_path = "{out}"'))
    print("runner", opts, flush=True)
    assert run_kalnajs.main([str(cfg), "kalnajs_zang4.toml", "--gpu=0", *opts]) == 0
    (name,) = os.listdir(out)
    assert re.fullmatch(r"kalnajs_m2_n4_\d{8}_\d{6}\.log", name)
    log = open(os.path.join(out, name)).read()
    Op = float(re.search(r"^Ω_p = ([0-9.]+)$", log, re.M).group(1))
    g = float(re.search(r"^γ   = ([0-9.]+)$", log, re.M).group(1))
    assert re.search(r"^\|det\| = \d\.\d\de[-+]\d+$", log, re.M)
    assert re.search(r"^Converged: true \(\d+ iterations\)$", log, re.M)
    assert re.search(r"^Runtime: [0-9.]+s \(\d+ min [0-9.]+s\)$", log, re.M)
    return log, Op, g


def test_runner_reference_run(tmp_path):
    log, Op, g = _run(tmp_path)
    assert "PHASE 2: GRID SCAN (SKIPPED)" in log
    assert "Using reference initial guess: Ω_p = 0.43971423, γ = 0.12306639" in log
    assert re.search(r"^        Device 0: .+ \([0-9.]+ GB\) \[ROCm\]$", log, re.M)
    assert abs(Op - KAT[0]) <= 1e-10 and abs(g - KAT[1]) <= 1e-10


def test_runner_scan_seeded_mixed(tmp_path):
    log, Op, g = _run(tmp_path, "--scan", "--newton=mixed", "--orbits=device")
    assert "PHASE 2: GRID SCAN\n" in log
    assert "Scan: 820 dets" in log
    assert "Using grid-scan initial guess" in log
    assert abs(Op - KAT[0]) <= 1e-10 and abs(g - KAT[1]) <= 1e-10
