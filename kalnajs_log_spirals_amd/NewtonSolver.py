"""NewtonSolver -- mirror of src/KalnajsLogSpiral/NewtonSolver.jl.

The iteration (forward-difference Jacobian, optional halving line search,
bounds gamma >= 1e-4 and Omega_p in [0, 1], final re-evaluation) runs in
libklsp's host driver, with every det(I - M) on the device and the
difference pair of each step evaluated as one two-omega batch.

``precision="mixed"`` (no counterpart in the reference) selects libklsp's
mixed-precision search: BF16x3 passes ("mixed_fp32": FP32 passes) that each
build the first four (FP32: three) Taylor terms of M around a centre and run Newton on
that model, then an FP64 build that anchors the model for the FP64 finish
(a second one where the estimated error of the anchored model's zero exceeds
5e-11); it returns an FP64 root to |d omega| <= 1e-10 by that calibrated
estimate, which from far starts may differ from the root the FP64 iteration
reaches.  In
that mode ``det_val`` is the anchored model's det at the returned omega (the
last FP64 build plus the low-precision Taylor corrections, which enter
multiplied by the distance from that anchor), not a fresh FP64 build; the
reference recomputes it with a full evaluation (NewtonSolver.jl:157-158).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _ffi
from .BasisFunctions import DevicePrecomputed, PrecomputedData
from .Configuration import KalnajsConfig
from .MatrixBuilder import _device

__all__ = ["NewtonResult", "newton_search"]


@dataclass
class NewtonResult:
    """NewtonSolver.jl:19-27."""
    Omega_p: float
    gamma: float
    omega: complex
    det_val: complex
    iterations: int
    converged: bool
    history: np.ndarray          # [iterations, 3]: (Omega_p, gamma, |det|)
    n_det_evals: int = 0         # M(omega) builds
    n_fp64_builds: int = 0
    n_lp_builds: int = 0         # FP32 builds (Taylor passes in mixed mode)


_PRECISIONS = {"fp64": _ffi.PREC_FP64, "fp32": _ffi.PREC_FP32, "bf16x3": _ffi.PREC_BF16X3,
               "mixed": _ffi.PREC_MIXED, "mixed_fp32": _ffi.PREC_MIXED_FP32}


def newton_search(config: KalnajsConfig, precomputed: PrecomputedData,
                  devprecomp: DevicePrecomputed, gpu_type, Tcpu, Omega_p_init: float,
                  gamma_init: float, verbose: bool = False,
                  precision: str | None = None) -> NewtonResult:
    """NewtonSolver.jl:36-172.  ``precision`` None uses the device bundle's
    build precision; "fp64", "fp32" or "mixed" override it."""
    dev = _device(devprecomp, gpu_type, "newton_search")
    if precision is None:
        prec = dev.precision
    elif precision in _PRECISIONS:
        prec = _PRECISIONS[precision]
    else:
        raise ValueError(f"newton_search: precision must be one of {sorted(_PRECISIONS)}")
    nc = config.newton
    res, H = dev.ctx.newton(Omega_p_init, gamma_init, max_iter=nc.max_iter, tol=nc.tol,
                            delta=nc.delta, use_line_search=nc.use_line_search,
                            max_line_search=nc.max_line_search, precision=prec)
    ct = np.complex128 if np.dtype(Tcpu) == np.float64 else np.complex64
    if verbose:
        for i, (op, g, a) in enumerate(H, 1):
            print(f"Iter {i:2d}: Omega_p={op:.10f}  gamma={g:.10f}  |det|={a:.2e}")
        print("Converged!" if res.converged else "Not converged")
    return NewtonResult(res.Omega_p, res.gamma, complex(res.omega[0], res.omega[1]),
                        ct(complex(res.det[0], res.det[1])), res.iterations,
                        bool(res.converged), H, res.n_det_evals, res.n_fp64_builds,
                        res.n_lp_builds)
