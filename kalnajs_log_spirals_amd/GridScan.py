"""GridScan -- mirror of src/KalnajsLogSpiral/GridScan.jl.

grid_scan evaluates det(I - M) on the (Omega_p, gamma) grid in one device
call (omegas batched inside the library) and returns det_grid[n_gamma, n_Op]
plus the argmin |det| starting point, as the reference does.
"""
from __future__ import annotations

import numpy as np

from .BasisFunctions import DevicePrecomputed, PrecomputedData
from .Configuration import KalnajsConfig
from .MatrixBuilder import _device

__all__ = ["grid_scan", "find_minimum"]


def grid_scan(config: KalnajsConfig, precomputed: PrecomputedData,
              devprecomp: DevicePrecomputed, gpu_type, Tcpu=np.float64, verbose: bool = False,
              gpu_devices=()):
    """GridScan.jl:25-82.  One device only: gpu_devices beyond the first are
    ignored (the reference splits Omega_p columns over devices)."""
    dev = _device(devprecomp, gpu_type, "grid_scan")
    s = config.scan
    det_grid, Op0, g0 = dev.ctx.grid_scan(s.n_Op, s.n_gamma, s.Omega_p_min, s.Omega_p_max,
                                          s.gamma_min, s.gamma_max, dev.precision)
    ct = np.complex128 if np.dtype(Tcpu) == np.float64 else np.complex64
    if verbose:
        print(f"Minimum |det| = {np.abs(det_grid).min():.2e} at Omega_p = {Op0:.6f}, gamma = {g0:.6f}")
    return det_grid.astype(ct), Op0, g0


def find_minimum(det_grid: np.ndarray, Omega_p_grid, gamma_grid):
    """GridScan.jl:234-246: argmin |det| (first in column-major order)."""
    a = np.abs(np.asarray(det_grid))
    flat = a.reshape(-1, order="F")
    idx = int(np.argmin(flat))
    n_gamma = a.shape[0]
    return float(np.asarray(Omega_p_grid)[idx // n_gamma]), float(np.asarray(gamma_grid)[idx % n_gamma])
