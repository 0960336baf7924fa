"""BasisFunctions -- mirror of src/KalnajsLogSpiral/BasisFunctions.jl.

precompute_all computes the orbits on the host (libklsp C++) and W_l(alpha)
on the device (libklsp HIP, FP64 MFMA), and keeps W device-resident: the
24.7 GB default-grid W never crosses PCIe unless ``W_l_mat`` is read.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _ffi
from . import GPUBackend
from .Configuration import KalnajsConfig
from .OrbitIntegration import (OrbitData, compute_grids, compute_orbits, jacobian_df,
                               params_from_config)

__all__ = ["PrecomputedData", "precompute_all", "precompute_device_orbits", "trapezoid_coef",
           "compute_W_l", "compute_N_kernel", "gamma_complex", "DevicePrecomputed", "to_device"]


def _require_device(backend, who: str):
    if backend == GPUBackend.GPUType.NONE:
        raise RuntimeError(f"{who}: this build has no CPU path; a HIP (gfx950) device is required")


@dataclass
class PrecomputedData:
    """BasisFunctions.jl:30-60.  W_l_mat is held on the device by ``ctx``;
    reading the attribute copies it back (layout [NPh, N_alpha, n_l])."""
    m: int
    G: float
    N_alpha: int
    n_l: int
    l_arr: np.ndarray
    NPh: int
    alpha_arr: np.ndarray
    d_alpha: np.ndarray
    orbit_data: Optional[OrbitData]
    DJ_vec: np.ndarray
    F0l_all: np.ndarray
    Omega_res: np.ndarray
    N_kernel: np.ndarray
    ctx: Optional[_ffi.Context] = None
    _W_host: Optional[np.ndarray] = field(default=None, repr=False)

    @property
    def W_l_mat(self) -> np.ndarray:
        if self._W_host is None:
            if self.ctx is None:
                raise RuntimeError("PrecomputedData has no W")
            self._W_host = self.ctx.download_W()
        return self._W_host

    @W_l_mat.setter
    def W_l_mat(self, W) -> None:
        self._W_host = np.asarray(W, dtype=np.complex128)
        if self.ctx is not None:
            self.ctx.upload_W(self._W_host)


def trapezoid_coef(x) -> np.ndarray:
    """BasisFunctions.jl:174-188."""
    x = _ffi.f64c(x)
    S = np.empty_like(x)
    _ffi.check(_ffi.lib().klsp_trapezoid_coef(_ffi.dptr(x), x.size, _ffi.dptr(S)),
               what="trapezoid_coef")
    return S


def gamma_complex(z) -> complex:
    """BasisFunctions.jl:309-366 (Godfrey's Lanczos approximation)."""
    if np.ndim(z) > 0:
        return np.array([gamma_complex(v) for v in np.ravel(z)]).reshape(np.shape(z))
    out = np.empty(2)
    _ffi.check(_ffi.lib().klsp_gamma_complex(float(np.real(z)), float(np.imag(z)),
                                             _ffi.dptr(out)), what="gamma_complex")
    return complex(out[0], out[1])


def compute_N_kernel(m: int, alpha_arr, d_alpha) -> np.ndarray:
    """BasisFunctions.jl:281-304: N[a] = Re[G(z1)G(z2)/(G(z3)G(z4))] pi d_alpha[a]."""
    al, da = _ffi.f64c(alpha_arr), _ffi.f64c(d_alpha)
    if al.size != da.size:
        raise ValueError("compute_N_kernel: alpha_arr and d_alpha differ in length")
    out = np.empty_like(al)
    _ffi.check(_ffi.lib().klsp_compute_N_kernel(int(m), _ffi.dptr(al), _ffi.dptr(da), al.size,
                                                _ffi.dptr(out)), what="compute_N_kernel")
    return out


def _linrange(a: float, b: float, n: int) -> np.ndarray:
    out = np.empty(n)
    _ffi.check(_ffi.lib().klsp_linrange(float(a), float(b), int(n), _ffi.dptr(out)),
               what="range")
    return out


def _omega_independent(config: KalnajsConfig, od: OrbitData):
    """BasisFunctions.jl:127-153 (DJ, F0l_all, Omega_res)."""
    p = params_from_config(config)
    NPh = p.NR * p.Ne
    n_l = p.l_max - p.l_min + 1
    DJ, F0l, Ores = np.empty(NPh), np.empty(NPh * n_l), np.empty(NPh * n_l)
    flat = [np.ascontiguousarray(x.reshape(-1, order="F"))
            for x in (od.grids.S_e, od.jacobian, od.Omega_1, od.Omega_2, od.FE, od.FL)]
    _ffi.check(_ffi.lib().klsp_precompute_arrays(
        ct.byref(p), _ffi.dptr(_ffi.f64c(od.grids.S_RC)), *[_ffi.dptr(x) for x in flat],
        _ffi.dptr(DJ), _ffi.dptr(F0l), _ffi.dptr(Ores)), what="precompute arrays")
    return DJ, F0l.reshape((NPh, n_l), order="F"), Ores.reshape((NPh, n_l), order="F")


def _orbit_inputs(od: OrbitData):
    NR, Ne = od.Omega_1.shape
    nwa = od.w1.shape[0]
    flat = lambda a: np.ascontiguousarray(a.reshape(-1, order="F"))  # noqa: E731
    return (flat(od.w1), flat(od.ra), flat(od.pha), flat(od.Omega_1), flat(od.Omega_2),
            NR * Ne, nwa)


def precompute_host(config: KalnajsConfig, model, verbose: bool = False) -> dict:
    """Host half of precompute_all (BasisFunctions.jl:73-153 without W_l):
    grids, orbits, alpha grid, N_m(alpha) and DJ / F0l / Omega_res."""
    m = int(config.physics.m)
    N = int(config.grid.N_alpha)
    if verbose:
        print("Computing grids...")
    grids = compute_grids(config, model)
    if verbose:
        print("Computing orbits...")
    od = compute_orbits(config, model, grids, verbose=verbose)
    alpha = _linrange(-config.grid.alpha_max, config.grid.alpha_max, N)
    d_alpha = trapezoid_coef(alpha)
    Nk = compute_N_kernel(m, alpha, d_alpha)
    DJ, F0l, Ores = _omega_independent(config, od)
    return dict(od=od, alpha=alpha, d_alpha=d_alpha, Nk=Nk, DJ=DJ, F0l=F0l, Ores=Ores)


def precompute_device(config: KalnajsConfig, host: dict, verbose: bool = False,
                      ctx: _ffi.Context | None = None) -> PrecomputedData:
    """Device half of precompute_all: a context with the problem arrays and
    W_l (BasisFunctions.jl:155-165) computed on the device.  An existing
    context passed as ctx is reused (its previous problem is replaced and its
    W buffers are kept when the new W fits)."""
    m, G = int(config.physics.m), float(config.physics.G)
    N = int(config.grid.N_alpha)
    l_arr = np.arange(config.grid.l_min, config.grid.l_max + 1, dtype=np.int32)
    NPh = config.grid.NR * config.grid.Ne
    if verbose:
        print("Computing W_l basis functions (complex, on device)...")
    if ctx is None:
        ctx = _ffi.Context(GPUBackend.current_device())
    ctx.upload_problem(NPh, N, l_arr.size, m, G, host["DJ"], host["F0l"], host["Ores"],
                       host["Nk"])
    w1, ra, pha, o1, o2, _, _ = _orbit_inputs(host["od"])
    ctx.compute_W(w1, ra, pha, o1, o2, host["alpha"], l_arr)
    if verbose:
        print("Precomputation complete.")
    return PrecomputedData(m, G, N, l_arr.size, l_arr.astype(np.int64), NPh, host["alpha"],
                           host["d_alpha"], host["od"], host["DJ"], host["F0l"], host["Ores"],
                           host["Nk"], ctx=ctx)


def precompute_device_orbits(config: KalnajsConfig, model, verbose: bool = False,
                             download_orbits: bool = True,
                             ctx: _ffi.Context | None = None) -> PrecomputedData:
    """precompute_all with the orbits integrated on the device
    (OrbitIntegration.jl:181-302 as one workgroup per orbit) and fed to the W
    quadrature without leaving it.  Grids, N_m(alpha), the Jacobian / DF and
    DJ / F0l / Omega_res stay on the host (O(NPh) work).  The orbit samples
    are copied back into orbit_data unless download_orbits is False (then
    w1 / ra / pha are None).  ctx: an existing context to reuse, as in
    precompute_device."""
    m, G = int(config.physics.m), float(config.physics.G)
    N = int(config.grid.N_alpha)
    l_arr = np.arange(config.grid.l_min, config.grid.l_max + 1, dtype=np.int32)
    p = params_from_config(config)
    NPh = p.NR * p.Ne
    if verbose:
        print("Computing grids...")
    grids = compute_grids(config, model)
    alpha = _linrange(-config.grid.alpha_max, config.grid.alpha_max, N)
    d_alpha = trapezoid_coef(alpha)
    Nk = compute_N_kernel(m, alpha, d_alpha)
    if verbose:
        print("Computing orbits and W_l basis functions (on device)...")
    own = ctx is None
    if own:
        ctx = _ffi.Context(GPUBackend.current_device())
    try:
        ctx.upload_problem(NPh, N, l_arr.size, m, G, None, None, None, Nk)
        g = dict(Rc=grids.Rc, e=grids.e, R1=grids.R1, R2=grids.R2, E=grids.E, L=grids.L_m,
                 L2=grids.L2_m, SGNL=grids.SGNL)
        o1, o2, w1, ra, pha = ctx.compute_orbits_W(p, g, alpha, l_arr, download_orbits)
        shp = (p.NR, p.Ne)
        Om1, Om2 = o1.reshape(shp, order="F"), o2.reshape(shp, order="F")
        jac, F0, FE, FL = jacobian_df(config, grids, Om1)
        shp3 = (p.nwa, p.NR, p.Ne)
        orb = [x.reshape(shp3, order="F") if x is not None else None for x in (w1, ra, pha)]
        od = OrbitData(grids, Om1, Om2, *orb, jac, F0, FE, FL)
        DJ, F0l, Ores = _omega_independent(config, od)
        ctx.set_problem_arrays(DJ, F0l, Ores)
    except BaseException:
        if own:
            ctx.close()
        raise
    if verbose:
        print("Precomputation complete.")
    return PrecomputedData(m, G, N, l_arr.size, l_arr.astype(np.int64), NPh, alpha, d_alpha,
                           od, DJ, F0l, Ores, Nk, ctx=ctx)


def precompute_all(config: KalnajsConfig, model, backend=GPUBackend.GPUType.AMD,
                   verbose: bool = False, orbits: str = "host") -> PrecomputedData:
    """BasisFunctions.jl:73-165: grids, orbits, alpha grid, W_l (device),
    N_m(alpha), DJ / F0l / Omega_res.  orbits = "host" integrates the orbits
    with the host library (threads), "device" on the GPU
    (precompute_device_orbits)."""
    _require_device(backend, "precompute_all")
    if orbits == "device":
        return precompute_device_orbits(config, model, verbose)
    if orbits != "host":
        raise ValueError(f"precompute_all: orbits must be 'host' or 'device', not {orbits!r}")
    return precompute_device(config, precompute_host(config, model, verbose), verbose)


def compute_W_l(config: KalnajsConfig, orbit_data: OrbitData, alpha_arr, l_arr,
                backend=GPUBackend.GPUType.AMD, verbose: bool = False) -> np.ndarray:
    """BasisFunctions.jl:199-266: W_l [NPh, N_alpha, n_l] complex128.  Computed
    on the device and copied back (the reference returns a host array)."""
    _require_device(backend, "compute_W_l")
    alpha = _ffi.f64c(alpha_arr)
    la = np.asarray(l_arr, dtype=np.int32)
    w1, ra, pha, o1, o2, NPh, _ = _orbit_inputs(orbit_data)
    N, n_l = alpha.size, la.size
    ctx = _ffi.Context(GPUBackend.current_device())
    try:
        ctx.upload_problem(NPh, N, n_l, int(config.physics.m), 1.0, np.zeros(NPh),
                           np.zeros((NPh, n_l)), np.zeros((NPh, n_l)), np.zeros(N))
        return ctx.compute_W(w1, ra, pha, o1, o2, alpha, la, download=True)
    finally:
        ctx.close()


@dataclass
class DevicePrecomputed:
    """BasisFunctions.jl:381-395: the device-resident view used by the GPU
    determinant.  ``precision`` is the arithmetic of M builds (FP64 for
    Tgpu = float64, the complex64/FP32-MFMA path for Tgpu = float32)."""
    m: int
    G: float
    N_alpha: int
    n_l: int
    NPh: int
    ctx: _ffi.Context
    precision: int
    DJ_vec: np.ndarray
    F0l_all: np.ndarray
    Omega_res: np.ndarray
    N_kernel: np.ndarray
    d_alpha: np.ndarray

    @property
    def W_l_mat(self) -> _ffi.Context:
        return self.ctx


def to_device(precomp: PrecomputedData, gpu_type, Tgpu=np.float64) -> DevicePrecomputed:
    """BasisFunctions.jl:397-437.  With gpu_type NONE the reference builds a host
    copy; this build has no host path and raises."""
    _require_device(gpu_type, "to_device")
    prec = _ffi.PREC_FP64 if np.dtype(Tgpu) == np.float64 else _ffi.PREC_FP32
    if precomp.ctx is None:
        ctx = _ffi.Context(GPUBackend.current_device())
        ctx.upload_problem(precomp.NPh, precomp.N_alpha, precomp.n_l, precomp.m, precomp.G,
                           precomp.DJ_vec, precomp.F0l_all, precomp.Omega_res, precomp.N_kernel)
        ctx.upload_W(precomp.W_l_mat)
        precomp.ctx = ctx
    precomp.ctx.prepare(prec)
    return DevicePrecomputed(precomp.m, precomp.G, precomp.N_alpha, precomp.n_l, precomp.NPh,
                             precomp.ctx, prec, precomp.DJ_vec, precomp.F0l_all,
                             precomp.Omega_res, precomp.N_kernel, precomp.d_alpha)


def from_arrays(m: int, G: float, W, DJ, F0l, Omega_res, N_kernel, alpha_arr=None,
                d_alpha=None) -> PrecomputedData:
    """PrecomputedData from caller-supplied arrays (W [NPh, N, n_l] complex,
    DJ [NPh], F0l / Omega_res [NPh, n_l], N_kernel [N]); W is uploaded."""
    W = np.asarray(W, dtype=np.complex128)
    NPh, N, n_l = W.shape
    ctx = _ffi.Context(GPUBackend.current_device())
    ctx.upload_problem(NPh, N, n_l, m, G, DJ, F0l, Omega_res, N_kernel)
    ctx.upload_W(W)
    alpha = np.zeros(N) if alpha_arr is None else _ffi.f64c(alpha_arr)
    da = np.zeros(N) if d_alpha is None else _ffi.f64c(d_alpha)
    return PrecomputedData(m, G, N, n_l, np.arange(n_l), NPh, alpha, da, None,
                           _ffi.f64c(DJ), np.asfortranarray(F0l, dtype=np.float64),
                           np.asfortranarray(Omega_res, dtype=np.float64), _ffi.f64c(N_kernel),
                           ctx=ctx, _W_host=W)
