"""NewtonSolver -- mirror of src/KalnajsLogSpiral/NewtonSolver.jl.

The iteration (forward-difference Jacobian, optional halving line search,
bounds gamma >= 1e-4 and Omega_p in [0, 1], final re-evaluation) runs in
libklsp's host driver, with every det(I - M) on the device and the
difference pair of each step evaluated as one two-omega batch.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

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
    n_det_evals: int = 0


def newton_search(config: KalnajsConfig, precomputed: PrecomputedData,
                  devprecomp: DevicePrecomputed, gpu_type, Tcpu, Omega_p_init: float,
                  gamma_init: float, verbose: bool = False) -> NewtonResult:
    """NewtonSolver.jl:36-172."""
    dev = _device(devprecomp, gpu_type, "newton_search")
    nc = config.newton
    res, H = dev.ctx.newton(Omega_p_init, gamma_init, max_iter=nc.max_iter, tol=nc.tol,
                            delta=nc.delta, use_line_search=nc.use_line_search,
                            max_line_search=nc.max_line_search, precision=dev.precision)
    ct = np.complex128 if np.dtype(Tcpu) == np.float64 else np.complex64
    if verbose:
        for i, (op, g, a) in enumerate(H, 1):
            print(f"Iter {i:2d}: Omega_p={op:.10f}  gamma={g:.10f}  |det|={a:.2e}")
        print("Converged!" if res.converged else "Not converged")
    return NewtonResult(res.Omega_p, res.gamma, complex(res.omega[0], res.omega[1]),
                        ct(complex(res.det[0], res.det[1])), res.iterations,
                        bool(res.converged), H, res.n_det_evals)
