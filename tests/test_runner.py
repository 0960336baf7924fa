"""The command-line runner (run_kalnajs_gpu.jl:84-353) on the CPU: argument
parsing, path resolution, the log file's name and header, and the error
paths that end before any device work (CPU mode, devices that do not exist).
The end-to-end run is in test_gpu_runner.py."""
import os
import re

import pytest

from kalnajs_log_spirals_amd import _ffi, run_kalnajs

from conftest import CONFIGS, MODELS


def _config_in(tmp_path, name="small.toml"):
    """The package config with io.output_path pointing into tmp_path."""
    text = open(os.path.join(CONFIGS, name)).read()
    out = tmp_path / "results"
    text = text.replace('output_path = "results"', f'output_path = "{out}"')
    p = tmp_path / name
    p.write_text(text)
    return str(p), out


def _only_log(out):
    logs = sorted(os.listdir(out))
    assert len(logs) == 1, logs
    assert re.fullmatch(r"kalnajs_m2_n4_\d{8}_\d{6}\.log", logs[0]), logs[0]
    return open(os.path.join(out, logs[0])).read()


def test_parse_arguments():
    assert run_kalnajs.parse_arguments(["--help"]) == 0
    assert run_kalnajs.parse_arguments(["only_one.toml"]) == 1
    cfg, mdl, gpu, thr, opts = run_kalnajs.parse_arguments(
        ["c.toml", "m.toml", "--gpu=01", "--threads=3", "--scan", "--newton=mixed",
         "--orbits=device"])
    assert (cfg, mdl, gpu, thr) == ("c.toml", "m.toml", "01", 3)
    assert opts == {"scan": True, "newton": "mixed", "orbits": "device"}
    cfg, mdl, gpu, thr, opts = run_kalnajs.parse_arguments(["c.toml", "m.toml"])
    assert (gpu, thr, opts) == ("auto", None, {"scan": False, "newton": None, "orbits": "host"})
    assert run_kalnajs.parse_arguments(["c.toml", "m.toml", "--newton=fp16"]) == 1
    assert run_kalnajs.parse_arguments(["c.toml", "m.toml", "--orbits=gpu"]) == 1
    assert run_kalnajs.parse_arguments(["c.toml", "m.toml", "--threads=x"]) == 1
    assert run_kalnajs.parse_arguments(["c.toml", "m.toml", "--bogus"]) == 1


def test_missing_files(tmp_path):
    assert run_kalnajs.main([str(tmp_path / "none.toml"), "kalnajs_zang4.toml"]) == 1
    cfg, _ = _config_in(tmp_path)
    assert run_kalnajs.main([cfg, str(tmp_path / "none.toml")]) == 1


def test_package_paths_resolve():
    """configs/<name> and models/<name> given relative to another directory
    fall back to the package's files."""
    assert run_kalnajs._resolve("configs/default.toml", "configs") == os.path.join(
        CONFIGS, "default.toml")
    assert run_kalnajs._resolve("models/kalnajs_zang4.toml", "models") == os.path.join(
        MODELS, "kalnajs_zang4.toml")


def test_cpu_mode_is_an_error_and_is_logged(tmp_path):
    cfg, out = _config_in(tmp_path)
    assert run_kalnajs.main([cfg, "kalnajs_zang4.toml", "--gpu=CPU"]) == 1
    log = _only_log(out)
    assert "KALNAJS LOG-SPIRAL EIGENVALUE SOLVER - GPU VERSION" in log
    assert f"Config: {cfg}" in log
    assert "GPU:    CPU mode (forced)" in log
    assert "Error: CPU mode is not available" in log
    assert "PHASE 1" not in log


@pytest.mark.skipif(_ffi.device_count() > 0, reason="needs a host without HIP devices")
def test_no_device_paths(tmp_path):
    cfg, out = _config_in(tmp_path)
    assert run_kalnajs.main([cfg, "kalnajs_zang4.toml", "--gpu=7"]) == 1
    log = _only_log(out)
    assert "GPU:    Device(s) [7]" in log
    assert "Error: GPU 7 not available (have 0 GPUs)" in log
    for f in os.listdir(out):
        os.remove(os.path.join(out, f))
    assert run_kalnajs.main([cfg, "kalnajs_zang4.toml"]) == 1
    assert "Error: no HIP device found" in _only_log(out)


def test_julia_value_formatting():
    j = run_kalnajs._jl
    assert (j(1e-12), j(100.0), j(0.5), j(2.5e20), j(True), j(False), j(7)) == (
        "1.0e-12", "100.0", "0.5", "2.5e20", "true", "false", "7")


def test_find_minimum_column_major_and_nan():
    """find_minimum (GridScan.jl:234-246): argmin |det| over det_grid[n_gamma,
    n_Op] in column-major order, first occurrence on ties, and NaN counted as
    the minimum, as Julia's argmin does ("NaN is treated as less than all
    other values except missing")."""
    import numpy as np
    from kalnajs_log_spirals_amd.GridScan import find_minimum
    Op = np.array([0.2, 0.3, 0.4])
    g = np.array([0.01, 0.02])
    d = np.array([[3.0, 1.0, 1.0], [2.0, 5.0, 0.5]], dtype=complex)
    assert find_minimum(d, Op, g) == (0.4, 0.02)
    d[1, 2] = 1.0                      # tie at (0, 1), (0, 2), (1, 2): first column-major
    assert find_minimum(d, Op, g) == (0.3, 0.01)
    d[1, 1] = complex(np.nan, 0.0)     # NaN wins over the smaller finite values
    d[0, 2] = 0.1
    assert find_minimum(d, Op, g) == (0.3, 0.02)
