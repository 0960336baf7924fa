"""OrbitIntegration -- mirror of src/KalnajsLogSpiral/OrbitIntegration.jl.

compute_grids / compute_orbits run in libklsp's host C++ (FP64; the orbits are
spread over host threads).  Array shapes follow the reference: [NR, Ne] grids
and [nwa, NR, Ne] orbit samples, returned as Fortran-ordered numpy arrays.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _ffi
from .Configuration import KalnajsConfig

__all__ = ["GridData", "OrbitData", "compute_grids", "compute_orbits", "trapezoid_coef",
           "simpson_coef", "params_from_config", "jacobian_df", "orbit_phase_grid"]


@dataclass
class GridData:
    """OrbitIntegration.jl:26-38."""
    Rc: np.ndarray
    e: np.ndarray
    v: np.ndarray
    R1: np.ndarray
    R2: np.ndarray
    E: np.ndarray
    L_m: np.ndarray
    L2_m: np.ndarray
    SGNL: np.ndarray
    S_RC: np.ndarray
    S_e: np.ndarray


@dataclass
class OrbitData:
    """OrbitIntegration.jl:43-54."""
    grids: GridData
    Omega_1: np.ndarray
    Omega_2: np.ndarray
    w1: np.ndarray
    ra: np.ndarray
    pha: np.ndarray
    jacobian: np.ndarray
    F0: np.ndarray
    FE: np.ndarray
    FL: np.ndarray


def params_from_config(config: KalnajsConfig) -> _ffi.klsp_params:
    g = config.grid
    return _ffi.klsp_params(int(config.physics.m), float(config.physics.G),
                            float(config.model.L0), int(config.model.n_zang),
                            int(config.model.q1), int(g.NR), int(g.Ne), float(g.Rc_min),
                            float(g.Rc_max), float(g.v_min), int(g.N_alpha),
                            float(g.alpha_max), int(g.l_min), int(g.l_max), int(g.nw),
                            int(g.nwa))


def trapezoid_coef(x) -> np.ndarray:
    """OrbitIntegration.jl:63-74."""
    x = _ffi.f64c(x)
    S = np.empty_like(x)
    _ffi.check(_ffi.lib().klsp_trapezoid_coef(_ffi.dptr(x), x.size, _ffi.dptr(S)),
               what="trapezoid_coef")
    return S


def simpson_coef(x) -> np.ndarray:
    """OrbitIntegration.jl:79-95."""
    x = _ffi.f64c(x)
    S = np.empty_like(x)
    _ffi.check(_ffi.lib().klsp_simpson_coef(_ffi.dptr(x), x.size, _ffi.dptr(S)),
               what="simpson_coef")
    return S


def _grid(flat: np.ndarray, NR: int, Ne: int) -> np.ndarray:
    return flat.reshape((NR, Ne), order="F")


def compute_grids(config: KalnajsConfig, model=None) -> GridData:
    """OrbitIntegration.jl:102-170."""
    p = params_from_config(config)
    NR, Ne = p.NR, p.Ne
    names = ("Rc", "e", "v", "R1", "R2", "E", "L", "L2", "SGNL")
    a = {k: np.empty(NR * Ne) for k in names}
    S_RC, S_e = np.empty(NR), np.empty(NR * Ne)
    import ctypes as ct
    _ffi.check(_ffi.lib().klsp_compute_grids(ct.byref(p), *[_ffi.dptr(a[k]) for k in names],
                                             _ffi.dptr(S_RC), _ffi.dptr(S_e)),
               what="compute_grids: invalid grid parameters")
    return GridData(*[_grid(a[k], NR, Ne) for k in names], S_RC, _grid(S_e, NR, Ne))


def compute_orbits(config: KalnajsConfig, model, grids: GridData, verbose: bool = False,
                   n_threads: int | None = None) -> OrbitData:
    """OrbitIntegration.jl:181-336 (+ compute_jacobian! :341-382)."""
    import ctypes as ct
    p = params_from_config(config)
    NR, Ne, nwa = p.NR, p.Ne, p.nwa
    NPh = NR * Ne
    ins = [np.ascontiguousarray(getattr(grids, k).reshape(-1, order="F"))
           for k in ("Rc", "e", "R1", "R2", "E", "L_m", "L2_m", "SGNL")]
    o = {k: np.empty(NPh) for k in ("Om1", "Om2", "jac", "F0", "FE", "FL")}
    w1, ra, pha = (np.empty(NPh * nwa) for _ in range(3))
    nt = n_threads if n_threads else _ffi.host_threads()
    _ffi.check(_ffi.lib().klsp_compute_orbits(
        ct.byref(p), *[_ffi.dptr(x) for x in ins], _ffi.dptr(o["Om1"]), _ffi.dptr(o["Om2"]),
        _ffi.dptr(w1), _ffi.dptr(ra), _ffi.dptr(pha), _ffi.dptr(o["jac"]), _ffi.dptr(o["F0"]),
        _ffi.dptr(o["FE"]), _ffi.dptr(o["FL"]), int(nt)),
        what="compute_orbits: (nw-1) must be divisible by (nwa-1)")
    shp3 = (nwa, NR, Ne)
    return OrbitData(grids, _grid(o["Om1"], NR, Ne), _grid(o["Om2"], NR, Ne),
                     w1.reshape(shp3, order="F"), ra.reshape(shp3, order="F"),
                     pha.reshape(shp3, order="F"), _grid(o["jac"], NR, Ne),
                     _grid(o["F0"], NR, Ne), _grid(o["FE"], NR, Ne), _grid(o["FL"], NR, Ne))


def orbit_phase_grid(config: KalnajsConfig):
    """The radial-phase samples w, cos w, sin w (OrbitIntegration.jl:190-193)."""
    import ctypes as ct
    p = params_from_config(config)
    w, cw, sw = np.empty(p.nw), np.empty(p.nw), np.empty(p.nw)
    _ffi.check(_ffi.lib().klsp_orbit_phase_grid(ct.byref(p), _ffi.dptr(w), _ffi.dptr(cw),
                                                _ffi.dptr(sw)), what="orbit_phase_grid")
    return w, cw, sw


def jacobian_df(config: KalnajsConfig, grids: GridData, Omega_1):
    """compute_jacobian! (OrbitIntegration.jl:341-382) and the DF on the grid
    (:317-333) for given radial frequencies: (jacobian, F0, FE, FL), each
    (NR, Ne)."""
    import ctypes as ct
    p = params_from_config(config)
    NR, Ne = p.NR, p.Ne
    flat = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, order="F"))  # noqa: E731
    ins = [flat(getattr(grids, k)) for k in ("Rc", "R1", "R2", "E", "L_m")] + [flat(Omega_1)]
    if any(x.size != NR * Ne for x in ins):
        raise ValueError("jacobian_df: arrays must have NR * Ne entries")
    out = [np.empty(NR * Ne) for _ in range(4)]
    _ffi.check(_ffi.lib().klsp_compute_jacobian_df(ct.byref(p), *[_ffi.dptr(x) for x in ins],
                                                   *[_ffi.dptr(x) for x in out]),
               what="jacobian_df: invalid parameters")
    return tuple(_grid(x, NR, Ne) for x in out)
