"""run_kalnajs -- the command-line runner of run_kalnajs_gpu.jl.

    python -m kalnajs_log_spirals_amd.run_kalnajs config.toml model.toml [options]

Same arguments (``--gpu=auto|0|1|01|CPU``, ``--threads=N``), the same log file
(``<io.output_path>/kalnajs_m<m>_n<n_zang>_<yyyymmdd_HHMMSS>.log``, every line
also printed to stdout) with the same header, configuration summary, phase
banners, final-result block and error block (run_kalnajs_gpu.jl:84-148,
:151-353).  Config and model paths that do not exist as given are also looked
up in the package (``configs/``, ``models/``).

Where it differs from the reference runner:

* there is no CPU path: ``--gpu=CPU``, or ``auto`` with no HIP device, ends
  with an error (exit status 1) after the header is logged;
* one device per process: of a multi-digit ``--gpu`` list the first device
  is used (the reference uses the rest only in its grid scan);
* ``--threads`` sets the host threads of the orbit integration (the
  reference sets its BLAS threads; this build does no host BLAS);
* options the reference lacks: ``--scan`` runs the grid scan of phase 2 and
  seeds Newton from its minimum (the reference skips the scan and starts from
  the model file's reference root, run_kalnajs_gpu.jl:289-296);
  ``--newton=fp64|fp32|bf16x3|mixed|mixed_fp32`` sets the build precision of
  the search (default: the configuration's GPU precision);
  ``--orbits=host|device`` selects where the orbits are integrated.
"""
from __future__ import annotations

import datetime
import os
import sys
import time
from typing import List, Optional

import numpy as np

from . import GPUBackend, _ffi
from .BasisFunctions import precompute_all, to_device
from .Configuration import get_float_type, load_config, load_model
from .GridScan import grid_scan
from .Models import create_toomre_zang_model
from .NewtonSolver import newton_search

__all__ = ["main", "parse_arguments", "HELP_TEXT"]

HELP_TEXT = """\
Kalnajs log-spiral mode solver on a gfx950 device (HIP)

Usage: python -m kalnajs_log_spirals_amd.run_kalnajs config.toml model.toml [options]

Arguments:
  config.toml    configuration file (grid, Newton, precision, io)
  model.toml     model file (physics, model, reference root)

Options:
  --gpu=IDS        device: auto, 0, 1, 01 (first listed is used); CPU is not available
  --threads=N      host threads of the orbit integration (default: [cpu] max_threads)
  --scan           run the (Omega_p, gamma) grid scan and start Newton at its minimum
  --newton=P       build precision of the Newton search: fp64, fp32, bf16x3, mixed, mixed_fp32
  --orbits=WHERE   orbit integration on the host library (host) or the device (device)

Examples:
  python -m kalnajs_log_spirals_amd.run_kalnajs configs/default.toml models/kalnajs_zang4.toml --gpu=auto
  python -m kalnajs_log_spirals_amd.run_kalnajs configs/small.toml models/kalnajs_zang4.toml --gpu=0 --scan --newton=mixed
"""

_PKG = os.path.dirname(os.path.abspath(__file__))
_RULE = "=" * 60


class _Log:
    """log_print / log_println of run_kalnajs_gpu.jl:84-100: stdout and the
    log file, flushed per line."""

    def __init__(self):
        self.f = None
        self.path = ""

    def start(self, config) -> None:
        """run_kalnajs_gpu.jl:102-110."""
        d = config.io.output_path
        os.makedirs(d, exist_ok=True)
        stamp = datetime.datetime.now().strftime("%Y%m%d_%H%M%S")
        self.path = os.path.join(d, "kalnajs_m%d_n%d_%s.log"
                                 % (config.physics.m, config.model.n_zang, stamp))
        self.f = open(self.path, "w")

    def stop(self) -> None:
        if self.f is not None:
            self.f.flush()
            self.f.close()
            self.f = None

    def __call__(self, *parts) -> None:
        line = "".join(str(p) for p in parts)
        print(line, flush=True)
        if self.f is not None:
            self.f.write(line + "\n")
            self.f.flush()


def _jl(x) -> str:
    """A value as Julia's print shows it (true/false, 1.0e-12, 100.0)."""
    if isinstance(x, (bool, np.bool_)):
        return "true" if x else "false"
    if isinstance(x, (float, np.floating)):
        r = repr(float(x))
        if "e" in r:
            mant, ex = r.split("e")
            if "." not in mant:
                mant += ".0"
            return f"{mant}e{int(ex)}"
        return r
    return str(x)


def _jl_type(T) -> str:
    return "Float64" if np.dtype(T) == np.float64 else "Float32"


def _resolve(path: str, sub: str) -> str:
    if os.path.isfile(path):
        return path
    alt = os.path.join(_PKG, sub, os.path.basename(path))
    return alt if os.path.isfile(alt) else path


def parse_arguments(argv: List[str]):
    """run_kalnajs_gpu.jl:121-148, plus --scan, --newton=, --orbits=.
    Returns (config_file, model_file, gpu_spec, threads, opts) or an int exit
    status after printing help or an error."""
    if argv and argv[0] in ("--help", "-h"):
        print(HELP_TEXT)
        return 0
    if len(argv) < 2:
        print("Error: Insufficient arguments")
        print(HELP_TEXT)
        return 1
    config_file, model_file = argv[0], argv[1]
    gpu_spec, threads = "auto", None
    opts = {"scan": False, "newton": None, "orbits": "host"}
    for arg in argv[2:]:
        if arg.startswith("--gpu="):
            gpu_spec = arg.split("=", 1)[1]
        elif arg.startswith("--threads="):
            try:
                threads = int(arg.split("=", 1)[1])
            except ValueError:
                print(f"Error: Invalid thread count: {arg}")
                return 1
        elif arg == "--scan":
            opts["scan"] = True
        elif arg.startswith("--newton="):
            opts["newton"] = arg.split("=", 1)[1]
        elif arg.startswith("--orbits="):
            opts["orbits"] = arg.split("=", 1)[1]
        elif arg in ("--help", "-h"):
            print(HELP_TEXT)
            return 0
        else:
            print(f"Error: Unknown option: {arg}")
            print(HELP_TEXT)
            return 1
    if opts["newton"] not in (None, "fp64", "fp32", "bf16x3", "mixed", "mixed_fp32"):
        print(f"Error: Invalid Newton precision: {opts['newton']}")
        return 1
    if opts["orbits"] not in ("host", "device"):
        print(f"Error: Invalid orbit integration site: {opts['orbits']}")
        return 1
    return config_file, model_file, gpu_spec, threads, opts


def _select_devices(gpu_spec: str, log: _Log):
    """run_kalnajs_gpu.jl:173-232.  Returns the device list, or an error
    message (the reference's CPU modes have no counterpart here)."""
    if gpu_spec == "auto":
        n = GPUBackend.gpu_ndevices() if GPUBackend.detect_gpu_type() != GPUBackend.NONE else 0
        if n == 0:
            log("GPU:    None")
            return "no HIP device found; this build evaluates det(I - M) on a gfx950 device only"
        log(f"GPU:    Auto-detected {n} device(s)")
        devices = list(range(n))
    elif gpu_spec in ("CPU", "cpu"):
        log("GPU:    CPU mode (forced)")
        return "CPU mode is not available: this build evaluates det(I - M) on a gfx950 device only"
    elif gpu_spec.isdigit():
        devices = [int(gpu_spec)] if len(gpu_spec) == 1 else [int(c) for c in gpu_spec]
        log(f"GPU:    Device(s) {devices}")
    else:
        return f"Invalid GPU specification: {gpu_spec}"
    n_available = GPUBackend.gpu_ndevices()
    for d in devices:
        if d >= n_available:
            return f"GPU {d} not available (have {n_available} GPUs)"
    return devices


def main(argv: Optional[List[str]] = None) -> int:
    """run_kalnajs_gpu.jl:150-353: phases 1 (precompute), 2 (grid scan, run
    with --scan) and 3 (Newton), logged; errors are logged and re-raised."""
    parsed = parse_arguments(sys.argv[1:] if argv is None else list(argv))
    if isinstance(parsed, int):
        return parsed
    config_file, model_file, gpu_spec, threads_arg, opts = parsed
    config_file = _resolve(config_file, "configs")
    model_file = _resolve(model_file, "models")
    if not os.path.isfile(config_file):
        print(f"Error: Configuration file not found: {config_file}")
        return 1
    if not os.path.isfile(model_file):
        print(f"Error: Model file not found: {model_file}")
        return 1

    config = load_model(load_config(config_file), model_file)
    log = _Log()
    log.start(config)
    try:
        return _run(config, config_file, model_file, gpu_spec, threads_arg, opts, log)
    finally:
        log.stop()


def _run(config, config_file, model_file, gpu_spec, threads_arg, opts, log: _Log) -> int:
    log()
    log(_RULE)
    log("KALNAJS LOG-SPIRAL EIGENVALUE SOLVER - GPU VERSION")
    log(_RULE)
    log("Started: ", datetime.datetime.now().strftime("%Y-%m-%d %H:%M:%S"))
    log(f"Config: {config_file}")
    log(f"Model:  {model_file}")
    log()

    devices = _select_devices(gpu_spec, log)
    if isinstance(devices, str):
        log(f"Error: {devices}")
        log()
        log("Log file: ", log.path)
        return 1
    gpu_id = devices[0]
    GPUBackend.gpu_device(gpu_id)
    name, mem, _ = _ffi.device_info(gpu_id)
    log("        Device %d: %s (%.1f GB) [%s]" % (gpu_id, name, mem / 1024**3, "ROCm"))
    if len(devices) > 1:
        log(f"        One device per process: device {gpu_id} is used, {devices[1:]} are not")

    # host threads of the orbit integration (the reference's BLAS threads)
    threads = threads_arg if threads_arg is not None else config.cpu.max_threads
    threads = max(1, min(threads, config.cpu.max_threads, os.cpu_count() or 1))
    os.environ["KLSP_HOST_THREADS"] = str(threads)
    log(f"Threads: {threads} host threads (orbit integration)")
    log()

    Tgpu = get_float_type(config, True)
    Tcpu = get_float_type(config, False)
    log("Precision:")
    log("  GPU: ", _jl_type(Tgpu))
    log("  CPU: ", _jl_type(Tcpu))
    if opts["newton"] is not None:
        log("  Newton builds: ", opts["newton"])
    log()

    log(_RULE)
    log("CONFIGURATION")
    log(_RULE)
    log("Physics: m = ", config.physics.m)
    log("Model: n_zang = ", config.model.n_zang, ", q1 = ", config.model.q1)
    log("Grid: NR×Ne = ", config.grid.NR, "×", config.grid.Ne, ", N_alpha = ", config.grid.N_alpha)
    log("Newton: max_iter = ", config.newton.max_iter, ", tol = ", _jl(config.newton.tol))
    log()
    log("Reference: Ω_p = %.10f, γ = %.10f" % (config.reference.Omega_p, config.reference.gamma))
    log()

    start = time.time()
    pre = None
    try:
        log(_RULE)
        log(f"PHASE 1: PRECOMPUTATION (Float64, {opts['orbits']} orbits, device W)")
        log(_RULE)
        model = create_toomre_zang_model(config)
        pre = precompute_all(config, model, GPUBackend.AMD, verbose=True, orbits=opts["orbits"])
        log()
        log(f"Transferring arrays to GPU in precision {_jl_type(Tgpu)}...")
        dev = to_device(pre, GPUBackend.AMD, Tgpu)
        log("✓ GPU transfer complete")
        log()

        if opts["scan"]:
            s = config.scan
            log(_RULE)
            log("PHASE 2: GRID SCAN")
            log(_RULE)
            log("Grid: n_Op×n_gamma = %d×%d, Ω_p ∈ [%g, %g], γ ∈ [%g, %g]"
                % (s.n_Op, s.n_gamma, s.Omega_p_min, s.Omega_p_max, s.gamma_min, s.gamma_max))
            t0 = time.time()
            det_grid, Omega_p_init, gamma_init = grid_scan(config, pre, dev, GPUBackend.AMD, Tcpu)
            log("Scan: %d dets in %.1fs; minimum |det| = %.2e"
                % (det_grid.size, time.time() - t0, float(np.abs(det_grid).min())))
            log()
            log("Using grid-scan initial guess: Ω_p = %.8f, γ = %.8f" % (Omega_p_init, gamma_init))
        else:
            log(_RULE)
            log("PHASE 2: GRID SCAN (SKIPPED)")
            log(_RULE)
            Omega_p_init = float(config.reference.Omega_p)
            gamma_init = float(config.reference.gamma)
            log()
            log("Using reference initial guess: Ω_p = %.8f, γ = %.8f" % (Omega_p_init, gamma_init))
        log()

        log(_RULE)
        log("PHASE 3: NEWTON REFINEMENT")
        log(_RULE)
        result = newton_search(config, pre, dev, GPUBackend.AMD, Tcpu, Omega_p_init, gamma_init,
                               verbose=True, precision=opts["newton"])

        elapsed = time.time() - start
        minutes = int(elapsed // 60)
        log()
        log(_RULE)
        log("FINAL RESULT")
        log(_RULE)
        log("Ω_p = %.12f" % result.Omega_p)
        log("γ   = %.12f" % result.gamma)
        log("|det| = %.2e" % abs(result.det_val))
        log("Converged: ", _jl(result.converged), " (", result.iterations, " iterations)")
        log()
        log("Runtime: %.1fs (%d min %.1fs)" % (elapsed, minutes, elapsed - 60 * minutes))
        log(_RULE)
    except Exception as e:
        minutes = int((time.time() - start) // 60)
        log()
        log(_RULE)
        log("ERROR")
        log(_RULE)
        log(f"Error after {minutes} minutes:")
        log("Type: ", type(e).__name__)
        log("Message: ", e)
        log()
        log("Log file: ", log.path)
        log(_RULE)
        raise
    finally:
        if pre is not None and pre.ctx is not None:
            pre.ctx.close()
    print("Log saved to: ", log.path, sep="")
    return 0


if __name__ == "__main__":
    sys.exit(main())
