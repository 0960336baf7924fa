"""ctypes front end of the C oracle (test infrastructure only).

Array layouts follow the reference (Julia column-major):
  grids/orbit scalars  [NR, Ne]       -> flat index k = iR + NR*ie
  w1, ra, pha          [nwa, NR, Ne]  -> numpy shape (NPh, nwa) C-order
  W_l_mat              [NPh, N, n_l]  -> numpy complex128, Fortran order
  F0l_all, Omega_res   [NPh, n_l]     -> numpy float64, Fortran order
  M                    [N, N]         -> numpy complex128, Fortran order
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "Params", "lib", "build", "linrange", "trapezoid_coef", "simpson_coef",
    "gamma_complex", "N_kernel", "df_const", "df_eval", "compute_grids",
    "compute_orbits", "compute_W", "precompute", "build_M", "det_matrix",
    "compute_det", "newton", "newton_fn", "grid_scan", "load_toml_params",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libklsp_oracle.so")
_lib = None

_D = ct.POINTER(ct.c_double)
_I = ct.POINTER(ct.c_int)


def build() -> str:
    """Compile the C oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ct.CDLL(_LIB_PATH)
        L.or_linrange.argtypes = [ct.c_double, ct.c_double, ct.c_long, _D]
        L.or_trapezoid_coef.argtypes = [_D, ct.c_long, _D]
        L.or_simpson_coef.argtypes = [_D, ct.c_long, _D]
        L.or_gamma_complex.argtypes = [ct.c_double, ct.c_double, _D, _D]
        L.or_N_kernel.argtypes = [ct.c_int, _D, _D, ct.c_long, _D]
        L.or_df_const.argtypes = [ct.c_double, ct.c_int, ct.c_int, ct.c_double]
        L.or_df_const.restype = ct.c_double
        L.or_df_eval.argtypes = [ct.c_double, ct.c_int, ct.c_int, ct.c_double,
                                 ct.c_double, ct.c_double, _D, _D, _D]
        L.or_compute_grids.argtypes = [ct.c_int, ct.c_int, ct.c_double, ct.c_double,
                                       ct.c_double] + [_D] * 11
        L.or_compute_orbits.argtypes = ([ct.c_int] * 4 + [ct.c_double, ct.c_int, ct.c_int,
                                                         ct.c_double] + [_D] * 17)
        L.or_compute_W.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, _D, _D, _D, _D,
                                   _D, _D, ct.c_long, _I, ct.c_long, _D]
        L.or_precompute_arrays.argtypes = [ct.c_int, ct.c_int, ct.c_int, _I, ct.c_long] + [_D] * 10
        L.or_build_M.argtypes = [ct.c_long] * 3 + [_D] * 5 + [ct.c_double] * 3 + [_D]
        L.or_det_inplace.argtypes = [ct.c_long, _D, _D, _D]
        L.or_compute_det.argtypes = [ct.c_long] * 3 + [_D] * 5 + [ct.c_double] * 3 + [_D, _D]
        L.or_newton_problem.argtypes = ([ct.c_long] * 3 + [_D] * 5 + [ct.c_double, ct.c_int,
                                        ct.c_int, ct.c_double, ct.c_double, ct.c_int, ct.c_int,
                                        ct.c_double, ct.c_double, _D, _I, _I, _D])
        L.or_grid_scan.argtypes = ([ct.c_long] * 3 + [_D] * 5 + [ct.c_double, ct.c_int,
                                   ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_double,
                                   ct.c_double, _D, _D, _D])
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_D)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


# ---------------------------------------------------------------------------
# Parameters (Configuration.jl:31-149 defaults; the TOML keys are the same)
# ---------------------------------------------------------------------------
@dataclass
class Params:
    m: int = 2
    G: float = 1.0
    L0: float = 1.0
    n_zang: int = 4
    q1: int = 7
    NR: int = 201
    Ne: int = 21
    Rc_min: float = 1e-3
    Rc_max: float = 100.0
    v_min: float = 1e-2
    N_alpha: int = 101
    alpha_max: float = 10.0
    l_min: int = -25
    l_max: int = 25
    nw: int = 10001
    nwa: int = 201
    extra: dict = field(default_factory=dict)

    @property
    def NPh(self) -> int:
        return self.NR * self.Ne

    @property
    def n_l(self) -> int:
        return self.l_max - self.l_min + 1


def load_toml_params(config_path: str, model_path: str | None = None) -> Params:
    """Read the reference TOML schema (Configuration.jl:160-300)."""
    import tomli

    p = Params()
    with open(config_path, "rb") as f:
        cfg = tomli.load(f)
    secs = [cfg]
    if model_path is not None:
        with open(model_path, "rb") as f:
            secs.append(tomli.load(f))
    for d in secs:
        for key in ("m", "G"):
            if key in d.get("physics", {}):
                setattr(p, key, d["physics"][key])
        for key in ("L0", "n_zang", "q1", "G"):
            if key in d.get("model", {}):
                setattr(p, key, d["model"][key])
        for key in ("NR", "Ne", "Rc_min", "Rc_max", "v_min", "N_alpha", "alpha_max",
                    "l_min", "l_max", "nw", "nwa"):
            if key in d.get("grid", {}):
                setattr(p, key, d["grid"][key])
        for sec in ("newton", "scan", "reference"):
            if sec in d:
                p.extra.setdefault(sec, {}).update(d[sec])
    return p


# ---------------------------------------------------------------------------
# Small pieces
# ---------------------------------------------------------------------------
def linrange(a: float, b: float, n: int) -> np.ndarray:
    out = np.empty(n)
    lib().or_linrange(a, b, n, _p(out))
    return out


def trapezoid_coef(x) -> np.ndarray:
    x = _f64(x)
    S = np.empty_like(x)
    lib().or_trapezoid_coef(_p(x), x.size, _p(S))
    return S


def simpson_coef(x) -> np.ndarray:
    x = _f64(x)
    S = np.empty_like(x)
    lib().or_simpson_coef(_p(x), x.size, _p(S))
    return S


def gamma_complex(z: complex) -> complex:
    re, im = ct.c_double(), ct.c_double()
    lib().or_gamma_complex(float(np.real(z)), float(np.imag(z)), ct.byref(re), ct.byref(im))
    return complex(re.value, im.value)


def N_kernel(m: int, alpha, d_alpha) -> np.ndarray:
    alpha, d_alpha = _f64(alpha), _f64(d_alpha)
    out = np.empty_like(alpha)
    lib().or_N_kernel(m, _p(alpha), _p(d_alpha), alpha.size, _p(out))
    return out


def df_const(p: Params) -> float:
    return lib().or_df_const(p.L0, p.n_zang, p.q1, p.G)


def df_eval(p: Params, E: float, L: float):
    F, FE, FL = ct.c_double(), ct.c_double(), ct.c_double()
    lib().or_df_eval(p.L0, p.n_zang, p.q1, p.G, E, L, ct.byref(F), ct.byref(FE), ct.byref(FL))
    return F.value, FE.value, FL.value


# ---------------------------------------------------------------------------
# Precompute chain (BasisFunctions.precompute_all, BasisFunctions.jl:73-165)
# ---------------------------------------------------------------------------
def compute_grids(p: Params) -> dict:
    NPh = p.NPh
    g = {k: np.empty(NPh) for k in ("Rc", "e", "v", "R1", "R2", "E", "L", "L2", "SGNL", "S_e")}
    g["S_RC"] = np.empty(p.NR)
    rc = lib().or_compute_grids(p.NR, p.Ne, p.Rc_min, p.Rc_max, p.v_min,
                                *[_p(g[k]) for k in ("Rc", "e", "v", "R1", "R2", "E", "L",
                                                     "L2", "SGNL", "S_RC", "S_e")])
    if rc != 0:
        raise ValueError("compute_grids: bad grid sizes")
    return g


def compute_orbits(p: Params, g: dict) -> dict:
    NPh = p.NPh
    o = {k: np.empty(NPh) for k in ("Om1", "Om2", "jac", "F0", "FE", "FL")}
    for k in ("w1", "ra", "pha"):
        o[k] = np.empty((NPh, p.nwa))
    rc = lib().or_compute_orbits(
        p.NR, p.Ne, p.nw, p.nwa, p.L0, p.n_zang, p.q1, p.G,
        *[_p(g[k]) for k in ("Rc", "e", "R1", "R2", "E", "L", "L2", "SGNL")],
        *[_p(o[k]) for k in ("Om1", "Om2", "w1", "ra", "pha", "jac", "F0", "FE", "FL")])
    if rc != 0:
        raise ValueError("compute_orbits: (nw-1) must be divisible by (nwa-1)")
    return o


def compute_W(p: Params, o: dict, alpha: np.ndarray, l_arr: np.ndarray) -> np.ndarray:
    """W_l_mat [NPh, N, n_l] complex128 (Fortran order)."""
    alpha = _f64(alpha)
    l_arr = np.ascontiguousarray(l_arr, dtype=np.int32)
    N, n_l = alpha.size, l_arr.size
    W = np.empty(2 * p.NPh * N * n_l)
    lib().or_compute_W(p.NR, p.Ne, p.nwa, p.m, _p(o["w1"]), _p(o["ra"]), _p(o["pha"]),
                       _p(o["Om1"]), _p(o["Om2"]), _p(alpha), N,
                       l_arr.ctypes.data_as(_I), n_l, _p(W))
    return W.view(np.complex128).reshape((p.NPh, N, n_l), order="F")


def precompute(p: Params, with_W: bool = True) -> dict:
    g = compute_grids(p)
    o = compute_orbits(p, g)
    alpha = linrange(-p.alpha_max, p.alpha_max, p.N_alpha)
    d_alpha = trapezoid_coef(alpha)
    l_arr = np.arange(p.l_min, p.l_max + 1, dtype=np.int32)
    Nk = N_kernel(p.m, alpha, d_alpha)
    NPh, n_l = p.NPh, l_arr.size
    DJ = np.empty(NPh)
    F0l = np.empty(NPh * n_l)
    Ores = np.empty(NPh * n_l)
    lib().or_precompute_arrays(p.NR, p.Ne, p.m, l_arr.ctypes.data_as(_I), n_l, _p(g["S_RC"]),
                               _p(g["S_e"]), _p(o["jac"]), _p(o["Om1"]), _p(o["Om2"]),
                               _p(o["FE"]), _p(o["FL"]), _p(DJ), _p(F0l), _p(Ores))
    out = dict(params=p, grids=g, orbits=o, alpha=alpha, d_alpha=d_alpha, l_arr=l_arr,
               N_kernel=Nk, DJ=DJ, F0l=F0l.reshape((NPh, n_l), order="F"),
               Omega_res=Ores.reshape((NPh, n_l), order="F"), G=float(p.G), m=int(p.m))
    if with_W:
        out["W"] = compute_W(p, o, alpha, l_arr)
    return out


# ---------------------------------------------------------------------------
# M, det, Newton, grid scan (MatrixBuilder.jl, NewtonSolver.jl, GridScan.jl)
# ---------------------------------------------------------------------------
def _problem_args(pre: dict):
    W = pre["W"]
    NPh, N, n_l = W.shape
    Wf = np.asfortranarray(W).reshape(-1, order="F").view(np.float64)
    F0l = np.asfortranarray(pre["F0l"]).reshape(-1, order="F")
    Ores = np.asfortranarray(pre["Omega_res"]).reshape(-1, order="F")
    keep = (Wf, _f64(pre["DJ"]), _f64(F0l), _f64(Ores), _f64(pre["N_kernel"]))
    return (NPh, N, n_l), keep


def build_M(pre: dict, omega: complex) -> np.ndarray:
    (NPh, N, n_l), keep = _problem_args(pre)
    M = np.empty(2 * N * N)
    lib().or_build_M(NPh, N, n_l, *[_p(a) for a in keep], pre["G"], omega.real, omega.imag, _p(M))
    return M.view(np.complex128).reshape((N, N), order="F")


def det_matrix(A: np.ndarray) -> complex:
    """det by LU with partial pivoting (zgetf2 semantics)."""
    A = np.asfortranarray(np.array(A, dtype=np.complex128))
    flat = A.reshape(-1, order="F").view(np.float64).copy()
    re, im = ct.c_double(), ct.c_double()
    lib().or_det_inplace(A.shape[0], _p(flat), ct.byref(re), ct.byref(im))
    return complex(re.value, im.value)


def compute_det(pre: dict, omega: complex) -> complex:
    (NPh, N, n_l), keep = _problem_args(pre)
    re, im = ct.c_double(), ct.c_double()
    lib().or_compute_det(NPh, N, n_l, *[_p(a) for a in keep], pre["G"], omega.real,
                         omega.imag, ct.byref(re), ct.byref(im))
    return complex(re.value, im.value)


def newton(pre: dict, Op0: float, g0: float, max_iter: int = 50, tol: float = 1e-12,
           delta: float = 1e-6, use_line_search: bool = False, max_line_search: int = 10) -> dict:
    (NPh, N, n_l), keep = _problem_args(pre)
    out = np.empty(4)
    hist = np.zeros(3 * max(max_iter, 1))
    it, cv = ct.c_int(), ct.c_int()
    lib().or_newton_problem(NPh, N, n_l, *[_p(a) for a in keep], pre["G"], pre["m"], max_iter,
                            tol, delta, int(use_line_search), max_line_search, Op0, g0,
                            _p(out), ct.byref(it), ct.byref(cv), _p(hist))
    n = it.value
    H = hist.reshape((max(max_iter, 1), 3), order="F")[:n].copy()
    return dict(Omega_p=out[0], gamma=out[1], det=complex(out[2], out[3]),
                omega=complex(pre["m"] * out[0], out[1]), iterations=n,
                converged=bool(cv.value), history=H)


_DETFN = ct.CFUNCTYPE(None, ct.c_double, ct.c_double, _D, _D, ct.c_void_p)


def newton_fn(fn, m: int, Op0: float, g0: float, max_iter: int = 50, tol: float = 1e-12,
              delta: float = 1e-6, use_line_search: bool = False,
              max_line_search: int = 10) -> dict:
    """Run the oracle Newton driver on an arbitrary det(omega) callable."""
    L = lib()
    if not hasattr(L, "_newton_raw"):
        L.or_newton.argtypes = [_DETFN, ct.c_void_p, ct.c_int, ct.c_int, ct.c_double,
                                ct.c_double, ct.c_int, ct.c_int, ct.c_double, ct.c_double,
                                _D, _I, _I, _D]
        L._newton_raw = True

    def cb(wr, wi, dr, di, _user):
        v = complex(fn(complex(wr, wi)))
        dr[0] = v.real
        di[0] = v.imag

    cfn = _DETFN(cb)
    out = np.empty(4)
    hist = np.zeros(3 * max(max_iter, 1))
    it, cv = ct.c_int(), ct.c_int()
    L.or_newton(cfn, None, m, max_iter, tol, delta, int(use_line_search), max_line_search,
                Op0, g0, _p(out), ct.byref(it), ct.byref(cv), _p(hist))
    n = it.value
    H = hist.reshape((max(max_iter, 1), 3), order="F")[:n].copy()
    return dict(Omega_p=out[0], gamma=out[1], det=complex(out[2], out[3]), iterations=n,
                converged=bool(cv.value), history=H)


def grid_scan(pre: dict, n_Op: int, n_gamma: int, Op_min: float, Op_max: float,
              g_min: float, g_max: float):
    (NPh, N, n_l), keep = _problem_args(pre)
    dg = np.empty(2 * n_Op * n_gamma)
    o1, o2 = ct.c_double(), ct.c_double()
    lib().or_grid_scan(NPh, N, n_l, *[_p(a) for a in keep], pre["G"], pre["m"], n_Op, n_gamma,
                       Op_min, Op_max, g_min, g_max, _p(dg), ct.byref(o1), ct.byref(o2))
    det_grid = dg.view(np.complex128).reshape((n_gamma, n_Op), order="F")
    return det_grid, o1.value, o2.value
