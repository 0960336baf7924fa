"""ctypes binding of libklsp.so (the C ABI declared in include/klsp.h).

There is no fallback: if the library is missing the import of any compute
function raises, telling the user to build it.
"""
from __future__ import annotations

import ctypes as ct
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libklsp.so")

KLSP_OK = 0
KLSP_ERR_INVALID_ARG = 1
KLSP_ERR_OOM = 2
KLSP_ERR_HIP = 3
KLSP_ERR_NOT_READY = 4
KLSP_ERR_NO_DEVICE = 5
KLSP_ERR_HOST = 6

PREC_FP64 = 0
PREC_FP32 = 1
PREC_MIXED = 2
PREC_BF16X3 = 3
PREC_MIXED_FP32 = 4

_STATUS_NAMES = {
    KLSP_ERR_INVALID_ARG: "invalid argument",
    KLSP_ERR_OOM: "out of memory",
    KLSP_ERR_HIP: "HIP runtime error",
    KLSP_ERR_NOT_READY: "not ready",
    KLSP_ERR_NO_DEVICE: "no gfx950 device",
    KLSP_ERR_HOST: "host failure",
}

# Every symbol include/klsp.h declares (checked by tests/test_abi.py).
EXPORTED = [
    "klsp_version", "klsp_device_count", "klsp_device_info", "klsp_create", "klsp_destroy", "klsp_last_error",
    "klsp_trapezoid_coef", "klsp_simpson_coef", "klsp_linrange", "klsp_gamma_complex",
    "klsp_compute_N_kernel", "klsp_compute_grids", "klsp_compute_orbits",
    "klsp_compute_jacobian_df", "klsp_orbit_phase_grid", "klsp_precompute_arrays",
    "klsp_upload_problem", "klsp_set_problem_arrays", "klsp_compute_orbits_W",
    "klsp_compute_W", "klsp_upload_W", "klsp_download_W", "klsp_prepare", "klsp_build_M",
    "klsp_build_taylor", "klsp_build_pencil", "klsp_det", "klsp_det_noalloc", "klsp_newton", "klsp_grid_scan",
    "klsp_set_symmetric_basis", "klsp_symmetric_basis", "klsp_autotune", "klsp_variant_name",
    "klsp_set_variant", "klsp_set_scan_graph", "klsp_set_lu_form", "klsp_set_bf16_form",
    "klsp_scan_graph_stats", "klsp_device_bytes",
    "klsp_set_profiling",
    "klsp_last_timing", "klsp_last_W_ms", "klsp_last_orbit_ms",
]


class KlspError(RuntimeError):
    """A non-zero status returned by libklsp."""

    def __init__(self, status: int, message: str):
        super().__init__(f"klsp: {_STATUS_NAMES.get(status, 'error')} ({status}): {message}")
        self.status = status


class klsp_params(ct.Structure):
    _fields_ = [("m", ct.c_int), ("G", ct.c_double), ("L0", ct.c_double),
                ("n_zang", ct.c_int), ("q1", ct.c_int), ("NR", ct.c_int), ("Ne", ct.c_int),
                ("Rc_min", ct.c_double), ("Rc_max", ct.c_double), ("v_min", ct.c_double),
                ("N_alpha", ct.c_int), ("alpha_max", ct.c_double), ("l_min", ct.c_int),
                ("l_max", ct.c_int), ("nw", ct.c_int), ("nwa", ct.c_int)]


_D = ct.POINTER(ct.c_double)


class klsp_problem(ct.Structure):
    _fields_ = [("NPh", ct.c_int64), ("N_alpha", ct.c_int64), ("n_l", ct.c_int64),
                ("m", ct.c_int), ("G", ct.c_double), ("DJ", _D), ("F0l", _D),
                ("Omega_res", _D), ("N_kernel", _D)]


class klsp_newton_params(ct.Structure):
    _fields_ = [("max_iter", ct.c_int), ("tol", ct.c_double), ("delta", ct.c_double),
                ("use_line_search", ct.c_int), ("max_line_search", ct.c_int),
                ("Omega_p_init", ct.c_double), ("gamma_init", ct.c_double),
                ("precision", ct.c_int)]


class klsp_newton_result(ct.Structure):
    _fields_ = [("Omega_p", ct.c_double), ("gamma", ct.c_double),
                ("omega", ct.c_double * 2), ("det", ct.c_double * 2),
                ("iterations", ct.c_int), ("converged", ct.c_int), ("n_det_evals", ct.c_int),
                ("n_fp64_builds", ct.c_int), ("n_lp_builds", ct.c_int)]


class klsp_scan_params(ct.Structure):
    _fields_ = [("n_Op", ct.c_int), ("n_gamma", ct.c_int), ("Omega_p_min", ct.c_double),
                ("Omega_p_max", ct.c_double), ("gamma_min", ct.c_double),
                ("gamma_max", ct.c_double), ("precision", ct.c_int)]


class klsp_timing(ct.Structure):
    _fields_ = [("weights_ms", ct.c_double), ("contract_ms", ct.c_double),
                ("reduce_ms", ct.c_double), ("lu_ms", ct.c_double),
                ("n_omega", ct.c_int64), ("contract_launches", ct.c_int64)]


_lib = None


def lib() -> ct.CDLL:
    """Load libklsp.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libklsp.so not found at {LIB_PATH}; build it with "
            "`python -m kalnajs_log_spirals_amd.build_lib` (this package has no CPU fallback)")
    L = ct.CDLL(LIB_PATH)
    P = ct.c_void_p
    i64 = ct.c_int64
    L.klsp_version.restype = ct.c_char_p
    L.klsp_version.argtypes = []
    L.klsp_device_count.argtypes = [ct.POINTER(ct.c_int)]
    L.klsp_device_info.argtypes = [ct.c_int, ct.c_char_p, ct.c_int, ct.POINTER(i64),
                                   ct.POINTER(ct.c_int)]
    L.klsp_create.argtypes = [ct.c_int, ct.POINTER(P)]
    L.klsp_destroy.argtypes = [P]
    L.klsp_last_error.restype = ct.c_char_p
    L.klsp_last_error.argtypes = [P]
    L.klsp_trapezoid_coef.argtypes = [_D, i64, _D]
    L.klsp_simpson_coef.argtypes = [_D, i64, _D]
    L.klsp_linrange.argtypes = [ct.c_double, ct.c_double, i64, _D]
    L.klsp_gamma_complex.argtypes = [ct.c_double, ct.c_double, _D]
    L.klsp_compute_N_kernel.argtypes = [ct.c_int, _D, _D, i64, _D]
    L.klsp_compute_grids.argtypes = [ct.POINTER(klsp_params)] + [_D] * 11
    L.klsp_compute_orbits.argtypes = [ct.POINTER(klsp_params)] + [_D] * 17 + [ct.c_int]
    L.klsp_compute_jacobian_df.argtypes = [ct.POINTER(klsp_params)] + [_D] * 10
    L.klsp_orbit_phase_grid.argtypes = [ct.POINTER(klsp_params)] + [_D] * 3
    L.klsp_precompute_arrays.argtypes = [ct.POINTER(klsp_params)] + [_D] * 10
    L.klsp_upload_problem.argtypes = [P, ct.POINTER(klsp_problem)]
    L.klsp_set_problem_arrays.argtypes = [P, _D, _D, _D]
    L.klsp_compute_orbits_W.argtypes = ([P, ct.POINTER(klsp_params)] + [_D] * 9
                                        + [ct.POINTER(ct.c_int32)] + [_D] * 5)
    L.klsp_compute_W.argtypes = [P, i64, _D, _D, _D, _D, _D, _D, ct.POINTER(ct.c_int32), _D]
    L.klsp_upload_W.argtypes = [P, _D]
    L.klsp_download_W.argtypes = [P, _D]
    L.klsp_prepare.argtypes = [P, ct.c_int]
    L.klsp_build_M.argtypes = [P, ct.c_double, ct.c_double, ct.c_int, _D]
    L.klsp_build_taylor.argtypes = [P, ct.c_double, ct.c_double, ct.c_int, ct.c_int, _D]
    L.klsp_det.argtypes = [P, _D, i64, ct.c_int, _D]
    L.klsp_build_pencil.argtypes = [P, _D, _D, _D]
    L.klsp_det_noalloc.argtypes = [P, _D, i64, _D]
    L.klsp_newton.argtypes = [P, ct.POINTER(klsp_newton_params),
                              ct.POINTER(klsp_newton_result), _D]
    L.klsp_grid_scan.argtypes = [P, ct.POINTER(klsp_scan_params), _D, _D, _D]
    L.klsp_autotune.argtypes = [P, ct.c_int, ct.c_int, ct.POINTER(ct.c_int), _D, ct.c_int,
                                ct.POINTER(ct.c_int)]
    L.klsp_set_variant.argtypes = [P, ct.c_int, ct.c_int]
    L.klsp_variant_name.restype = ct.c_char_p
    L.klsp_variant_name.argtypes = [ct.c_int, ct.c_int]
    L.klsp_set_symmetric_basis.argtypes = [P, ct.c_int]
    L.klsp_set_scan_graph.argtypes = [P, ct.c_int]
    L.klsp_set_lu_form.argtypes = [P, ct.c_int]
    L.klsp_set_bf16_form.argtypes = [P, ct.c_int]
    L.klsp_device_bytes.argtypes = [P, ct.POINTER(ct.c_int64)]
    L.klsp_scan_graph_stats.argtypes = [P, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64)]
    L.klsp_symmetric_basis.argtypes = [P, ct.POINTER(ct.c_int)]
    L.klsp_set_profiling.argtypes = [P, ct.c_int]
    L.klsp_last_timing.argtypes = [P, ct.POINTER(klsp_timing)]
    L.klsp_last_W_ms.argtypes = [P, _D]
    L.klsp_last_orbit_ms.argtypes = [P, _D]
    for name in EXPORTED:
        if name not in ("klsp_version", "klsp_last_error", "klsp_variant_name"):
            getattr(L, name).restype = ct.c_int
    _lib = L
    return L


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_D)


def f64c(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def check(status: int, ctx_handle=None, what: str = "") -> None:
    if status != KLSP_OK:
        msg = lib().klsp_last_error(ctx_handle).decode() if ctx_handle else what
        raise KlspError(status, msg or what)


def version() -> str:
    return lib().klsp_version().decode()


def device_count() -> int:
    n = ct.c_int(0)
    lib().klsp_device_count(ct.byref(n))
    return n.value


def device_info(device: int = 0):
    """(name, total memory in bytes, compute units) of HIP device `device`."""
    buf = ct.create_string_buffer(256)
    mem, cus = ct.c_int64(0), ct.c_int(0)
    check(lib().klsp_device_info(int(device), buf, len(buf), ct.byref(mem), ct.byref(cus)),
          what=f"device_info: no HIP device {device}")
    return buf.value.decode(errors="replace"), int(mem.value), int(cus.value)


def host_threads() -> int:
    env = os.environ.get("KLSP_HOST_THREADS")
    if env:
        return max(1, int(env))
    return max(1, min(os.cpu_count() or 1, 16))


def _destroy(handle):
    if handle:
        lib().klsp_destroy(handle)


class Context:
    """One device context (stream + device-resident problem and W)."""

    def __init__(self, device: int = 0):
        h = ct.c_void_p()
        st = lib().klsp_create(device, ct.byref(h))
        if st != KLSP_OK:
            raise KlspError(st, f"cannot create a context on device {device}")
        self._handle = h
        self.device = device
        self.dims = None
        self._fin = weakref.finalize(self, _destroy, h)

    @property
    def handle(self):
        """The native context; a closed context raises instead of handing the
        library a pointer to freed memory."""
        if not self._fin.alive:
            raise KlspError(KLSP_ERR_INVALID_ARG, "context is closed")
        return self._handle

    @property
    def closed(self) -> bool:
        return not self._fin.alive

    def close(self):
        self._fin()

    def _chk(self, st, what=""):
        check(st, self.handle, what)

    # -- residency -----------------------------------------------------------
    @staticmethod
    def _problem_arrays(NPh, n_l, DJ, F0l, Omega_res):
        DJ = f64c(DJ)
        F0l = f64c(np.asarray(F0l, dtype=np.float64).reshape(-1, order="F"))
        Ores = f64c(np.asarray(Omega_res, dtype=np.float64).reshape(-1, order="F"))
        if DJ.size != NPh or F0l.size != NPh * n_l or Ores.size != NPh * n_l:
            raise ValueError("problem arrays: sizes do not match (NPh, n_l)")
        return DJ, F0l, Ores

    def upload_problem(self, NPh: int, N: int, n_l: int, m: int, G: float, DJ, F0l, Omega_res,
                       N_kernel):
        """DJ, F0l and Omega_res may all be None (set them later with
        set_problem_arrays, e.g. after compute_orbits_W)."""
        Nk = f64c(N_kernel)
        if Nk.size != N:
            raise ValueError("upload_problem: N_kernel must have N entries")
        if DJ is None and F0l is None and Omega_res is None:
            arrs = (None, None, None)
            ptrs = (None, None, None)
        else:
            arrs = self._problem_arrays(NPh, n_l, DJ, F0l, Omega_res)
            ptrs = tuple(dptr(a) for a in arrs)
        pr = klsp_problem(NPh, N, n_l, int(m), float(G), *ptrs, dptr(Nk))
        self._chk(lib().klsp_upload_problem(self.handle, ct.byref(pr)))
        self.dims = (int(NPh), int(N), int(n_l))

    def set_problem_arrays(self, DJ, F0l, Omega_res):
        NPh, _, n_l = self.dims
        DJ, F0l, Ores = self._problem_arrays(NPh, n_l, DJ, F0l, Omega_res)
        self._chk(lib().klsp_set_problem_arrays(self.handle, dptr(DJ), dptr(F0l), dptr(Ores)))

    def compute_orbits_W(self, params: klsp_params, grids: dict, alpha, l_arr,
                         download_orbits: bool = False):
        """Orbits and W on the device (klsp_compute_orbits_W).  grids holds the
        [NPh] arrays Rc, e, R1, R2, E, L, L2, SGNL (column-major order of the
        (NR, Ne) grid).  Returns (Omega_1, Omega_2, w1, ra, pha) with the
        orbit samples [nwa, NPh] (None unless download_orbits)."""
        NPh, N, n_l = self.dims
        ins = [f64c(np.asarray(grids[k]).reshape(-1, order="F"))
               for k in ("Rc", "e", "R1", "R2", "E", "L", "L2", "SGNL")]
        if any(x.size != NPh for x in ins):
            raise ValueError("compute_orbits_W: grid arrays must have NPh entries")
        al = f64c(alpha)
        la = np.ascontiguousarray(l_arr, dtype=np.int32)
        if al.size != N or la.size != n_l:
            raise ValueError("compute_orbits_W: sizes do not match the uploaded problem")
        o1, o2 = np.empty(NPh), np.empty(NPh)
        nwa = int(params.nwa)
        orb = [np.empty(NPh * nwa) for _ in range(3)] if download_orbits else [None] * 3
        self._chk(lib().klsp_compute_orbits_W(
            self.handle, ct.byref(params), *[dptr(x) for x in ins], dptr(al),
            la.ctypes.data_as(ct.POINTER(ct.c_int32)), dptr(o1), dptr(o2),
            *[dptr(x) if x is not None else None for x in orb]))
        return (o1, o2, *orb)

    def compute_W(self, w1, ra, pha, Omega_1, Omega_2, alpha, l_arr, download: bool = False):
        NPh, N, n_l = self.dims
        w1, ra, pha = f64c(w1), f64c(ra), f64c(pha)
        nwa = w1.size // NPh
        if w1.size != NPh * nwa or ra.size != w1.size or pha.size != w1.size:
            raise ValueError("compute_W: orbit arrays must be [nwa, NPh]")
        o1, o2, al = f64c(Omega_1), f64c(Omega_2), f64c(alpha)
        la = np.ascontiguousarray(l_arr, dtype=np.int32)
        if o1.size != NPh or o2.size != NPh or al.size != N or la.size != n_l:
            raise ValueError("compute_W: sizes do not match the uploaded problem")
        out = np.empty(2 * NPh * N * n_l) if download else None
        self._chk(lib().klsp_compute_W(self.handle, nwa, dptr(w1), dptr(ra), dptr(pha), dptr(o1),
                                       dptr(o2), dptr(al), la.ctypes.data_as(ct.POINTER(ct.c_int32)),
                                       dptr(out) if download else None))
        if download:
            return out.view(np.complex128).reshape((NPh, N, n_l), order="F")
        return None

    def upload_W(self, W):
        NPh, N, n_l = self.dims
        W = np.asarray(W, dtype=np.complex128)
        if W.shape != (NPh, N, n_l):
            raise ValueError(f"upload_W: expected shape {(NPh, N, n_l)}, got {W.shape}")
        flat = np.ascontiguousarray(W.reshape(-1, order="F")).view(np.float64)
        self._chk(lib().klsp_upload_W(self.handle, dptr(flat)))

    def download_W(self) -> np.ndarray:
        NPh, N, n_l = self.dims
        out = np.empty(2 * NPh * N * n_l)
        self._chk(lib().klsp_download_W(self.handle, dptr(out)))
        return out.view(np.complex128).reshape((NPh, N, n_l), order="F")

    def prepare(self, precision: int):
        self._chk(lib().klsp_prepare(self.handle, precision))

    # -- hot path ------------------------------------------------------------
    def build_M(self, omega: complex, precision: int = PREC_FP64) -> np.ndarray:
        N = self.dims[1]
        out = np.empty(2 * N * N)
        self._chk(lib().klsp_build_M(self.handle, float(np.real(omega)), float(np.imag(omega)),
                                     precision, dptr(out)))
        return out.view(np.complex128).reshape((N, N), order="F")

    def build_taylor(self, omega: complex, n_terms: int = 3,
                     precision: int = PREC_FP64) -> np.ndarray:
        """[n_terms, N, N] Taylor terms T_j of M around omega (klsp_build_taylor)."""
        N = self.dims[1]
        out = np.empty(2 * n_terms * N * N)
        self._chk(lib().klsp_build_taylor(self.handle, float(np.real(omega)),
                                          float(np.imag(omega)), int(n_terms), precision,
                                          dptr(out)))
        return np.stack([out[2 * j * N * N:2 * (j + 1) * N * N].view(np.complex128)
                         .reshape((N, N), order="F") for j in range(n_terms)])

    def build_pencil(self):
        """(D_l, E_l, F_l), each [N, N, n_l] complex (klsp_build_pencil)."""
        NPh, N, n_l = self.dims
        outs = [np.empty(2 * N * N * n_l) for _ in range(3)]
        self._chk(lib().klsp_build_pencil(self.handle, *[dptr(o) for o in outs]))
        return tuple(o.view(np.complex128).reshape((N, N, n_l), order="F") for o in outs)

    def det(self, omegas, precision: int = PREC_FP64) -> np.ndarray:
        om = np.ascontiguousarray(np.atleast_1d(np.asarray(omegas, dtype=np.complex128)))
        out = np.empty(om.size, dtype=np.complex128)
        self._chk(lib().klsp_det(self.handle, dptr(om.view(np.float64)), om.size, precision,
                                 dptr(out.view(np.float64))))
        return out

    def newton(self, Omega_p_init, gamma_init, max_iter=50, tol=1e-12, delta=1e-6,
               use_line_search=False, max_line_search=10, precision=PREC_FP64):
        npar = klsp_newton_params(int(max_iter), float(tol), float(delta), int(bool(use_line_search)),
                                  int(max_line_search), float(Omega_p_init), float(gamma_init),
                                  int(precision))
        res = klsp_newton_result()
        hist = np.zeros(3 * max(int(max_iter), 1))
        self._chk(lib().klsp_newton(self.handle, ct.byref(npar), ct.byref(res), dptr(hist)))
        n = min(res.iterations, max(int(max_iter), 1))
        H = hist.reshape((max(int(max_iter), 1), 3), order="F")[:n].copy()
        return res, H

    def grid_scan(self, n_Op, n_gamma, Op_min, Op_max, g_min, g_max, precision=PREC_FP64):
        sp = klsp_scan_params(int(n_Op), int(n_gamma), float(Op_min), float(Op_max),
                              float(g_min), float(g_max), int(precision))
        dg = np.empty(2 * n_Op * n_gamma)
        o1, o2 = ct.c_double(), ct.c_double()
        self._chk(lib().klsp_grid_scan(self.handle, ct.byref(sp), dptr(dg), ct.byref(o1),
                                       ct.byref(o2)))
        return dg.view(np.complex128).reshape((n_gamma, n_Op), order="F"), o1.value, o2.value

    def device_bytes(self) -> int:
        """Device memory this context holds between calls (klsp_device_bytes)."""
        v = ct.c_int64(0)
        self._chk(lib().klsp_device_bytes(self.handle, ct.byref(v)))
        return v.value

    def set_lu_form(self, form: int):
        """klsp_set_lu_form: 0 automatic, 1 one workgroup per matrix, 2 step-wise."""
        self._chk(lib().klsp_set_lu_form(self.handle, int(form)))

    def set_bf16_form(self, form: int):
        """klsp_set_bf16_form: 0 automatic, 1 folded column pairs, 2 folded strips."""
        self._chk(lib().klsp_set_bf16_form(self.handle, int(form)))

    def set_scan_graph(self, on: bool):
        """grid_scan as one HIP graph launch (default on) or kernel by kernel."""
        self._chk(lib().klsp_set_scan_graph(self.handle, int(bool(on))))

    def scan_graph_stats(self):
        """(graphs instantiated, graphs launched) by grid_scan on this context."""
        b, l = ct.c_int64(0), ct.c_int64(0)
        self._chk(lib().klsp_scan_graph_stats(self.handle, ct.byref(b), ct.byref(l)))
        return b.value, l.value

    def set_symmetric_basis(self, on: bool):
        """Real-basis FP64 builds where W is alpha-mirror symmetric (default on)."""
        self._chk(lib().klsp_set_symmetric_basis(self.handle, int(bool(on))))

    def autotune(self, precision: int = PREC_FP64, reps: int = 2):
        """Times the contraction variants of one-omega `precision` builds and
        keeps the fastest; returns ({name: mean ms}, chosen name)."""
        n = ct.c_int(0)
        ch = ct.c_int(0)
        ms = np.zeros(16)
        self._chk(lib().klsp_autotune(self.handle, int(precision), int(reps), ct.byref(n),
                                      dptr(ms), ms.size, ct.byref(ch)))
        names = [lib().klsp_variant_name(int(precision), v).decode() or "default"
                 for v in range(n.value)]
        return {names[v]: float(ms[v]) for v in range(n.value)}, names[ch.value]

    def set_variant(self, variant: int, precision: int = PREC_FP64):
        self._chk(lib().klsp_set_variant(self.handle, int(precision), int(variant)))

    def symmetric_basis(self) -> bool:
        """True if one-omega FP64 builds of the current W use the real basis."""
        v = ct.c_int(0)
        self._chk(lib().klsp_symmetric_basis(self.handle, ct.byref(v)))
        return bool(v.value)

    # -- measurement ---------------------------------------------------------
    def set_profiling(self, on: bool):
        self._chk(lib().klsp_set_profiling(self.handle, int(bool(on))))

    def last_timing(self) -> dict:
        t = klsp_timing()
        self._chk(lib().klsp_last_timing(self.handle, ct.byref(t)))
        return {k: getattr(t, k) for k, _ in klsp_timing._fields_}

    def last_W_ms(self) -> float:
        v = ct.c_double()
        self._chk(lib().klsp_last_W_ms(self.handle, ct.byref(v)))
        return v.value

    def last_orbit_ms(self) -> float:
        v = ct.c_double()
        self._chk(lib().klsp_last_orbit_ms(self.handle, ct.byref(v)))
        return v.value
