"""The C-ABI library without a GPU: it loads, exports every symbol the header
declares, its host (CPU) entry points match the oracle, its argument checks
reject bad input, and the compute entry points refuse to run without a device
(no CPU fallback)."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

import oracle
import kalnajs_log_spirals_amd as K
from kalnajs_log_spirals_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "klsp.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(klsp_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _ffi.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(_ffi.EXPORTED) == syms
    assert _ffi.version().startswith("klsp ") and "gfx950" in _ffi.version()


def test_no_device_means_no_compute():
    if _ffi.device_count() > 0:
        pytest.skip("a GPU is visible")
    h = ct.c_void_p()
    assert _ffi.lib().klsp_create(0, ct.byref(h)) == _ffi.KLSP_ERR_NO_DEVICE
    assert K.detect_gpu_type() == K.GPUType.NONE
    cfg = K.KalnajsConfig()
    with pytest.raises(RuntimeError, match="no CPU path"):
        K.precompute_all(cfg, K.create_toomre_zang_model(cfg), K.GPUType.NONE)
    with pytest.raises(RuntimeError, match="CPU route"):
        K.compute_determinant(0.9 + 0.1j, None, None, K.GPUType.NONE, np.float64)


def test_null_context_and_bad_arguments():
    L = _ffi.lib()
    assert L.klsp_det(None, None, 1, 0, None) == _ffi.KLSP_ERR_INVALID_ARG
    assert L.klsp_destroy(None) == _ffi.KLSP_OK
    assert L.klsp_last_error(None) == b"klsp: null context"
    p = _ffi.klsp_params(2, 1.0, 1.0, 4, 7, 5, 5, 0.1, 10.0, 0.01, 11, 5.0, -2, 2, 100, 21)
    buf = [np.empty(25) for _ in range(11)]
    # (nw - 1) not divisible by (nwa - 1)
    assert L.klsp_compute_grids(ct.byref(p), *[_ffi.dptr(b) for b in buf]) == \
        _ffi.KLSP_ERR_INVALID_ARG
    assert L.klsp_trapezoid_coef(None, 5, None) == _ffi.KLSP_ERR_INVALID_ARG


def test_host_precompute_matches_oracle(small_cfg_paths):
    cfg = K.load_model(K.load_config(small_cfg_paths[0]), small_cfg_paths[1])
    p = oracle.load_toml_params(*small_cfg_paths)
    g = oracle.compute_grids(p)
    o = oracle.compute_orbits(p, g)
    G = K.compute_grids(cfg)
    od = K.compute_orbits(cfg, None, G)
    for a, b in ((G.Rc, g["Rc"]), (G.E, g["E"]), (G.L_m, g["L"]), (G.S_e, g["S_e"])):
        assert np.array_equal(a.reshape(-1, order="F"), b)
    assert np.array_equal(G.S_RC, g["S_RC"])
    for a, b in ((od.Omega_1, o["Om1"]), (od.Omega_2, o["Om2"]), (od.jacobian, o["jac"]),
                 (od.FE, o["FE"]), (od.FL, o["FL"])):
        np.testing.assert_allclose(a.reshape(-1, order="F"), b, rtol=1e-15, atol=0)
    for k in ("w1", "ra", "pha"):
        np.testing.assert_allclose(getattr(od, k).reshape(-1, order="F"), o[k].reshape(-1),
                                   rtol=1e-15, atol=1e-300)
    # thread count does not change the result
    od1 = K.compute_orbits(cfg, None, G, n_threads=1)
    assert np.array_equal(od1.w1, od.w1) and np.array_equal(od1.Omega_2, od.Omega_2)
    alpha = oracle.linrange(-10, 10, 101)
    da = oracle.trapezoid_coef(alpha)
    assert np.array_equal(K.compute_N_kernel(2, alpha, da), oracle.N_kernel(2, alpha, da))
    for z in (0.3 + 2j, -1.3 + 0.2j, 1.0 + 0j, 2.0 + 0j):
        assert K.gamma_complex(z) == oracle.gamma_complex(z)


_OOM_CHILD = r"""
import ctypes as ct, resource, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from kalnajs_log_spirals_amd import _ffi
L = _ffi.lib()
D = ct.POINTER(ct.c_double)
ptr = lambda a: a.ctypes.data_as(D)
mode = sys.argv[2]
if mode == "grids":
    # the grid routine's first host vector (NR doubles = 1.6 GB) cannot be
    # allocated; it must fail before writing any output
    p = _ffi.klsp_params(m=2, G=1.0, L0=1.0, n_zang=4, q1=7, NR=200_000_000, Ne=3, Rc_min=0.1,
                         Rc_max=10.0, v_min=0.05, N_alpha=5, alpha_max=5.0, l_min=-1, l_max=1,
                         nw=401, nwa=41)
else:
    # one orbit with nw = 3e7 phase samples: the shared phase tables (0.72 GB)
    # fit under the cap, each worker's scratch (0.96 GB) does not
    p = _ffi.klsp_params(m=2, G=1.0, L0=1.0, n_zang=4, q1=7, NR=1, Ne=1, Rc_min=1.0,
                         Rc_max=1.0, v_min=0.5, N_alpha=5, alpha_max=5.0, l_min=-1, l_max=1,
                         nw=30_000_001, nwa=2)
    g = [np.zeros(1) for _ in range(11)]
    assert L.klsp_compute_grids(ct.byref(p), *[ptr(a) for a in g]) == 0
outs = [np.zeros(2) for _ in range(11)]
# address space capped 1 GiB above what the process maps now
with open("/proc/self/statm") as fh:
    mapped = int(fh.read().split()[0]) * resource.getpagesize()
resource.setrlimit(resource.RLIMIT_AS, (mapped + (1 << 30), resource.RLIM_INFINITY))
if mode == "grids":
    rc = L.klsp_compute_grids(ct.byref(p), *[ptr(o) for o in outs])
else:
    rc = L.klsp_compute_orbits(ct.byref(p), *[ptr(a) for a in (g[0], g[1], g[3], g[4], g[5],
                               g[6], g[7], g[8])], *[ptr(o) for o in outs[:9]], 2)
print("status", rc)
sys.exit(0 if rc == _ffi.KLSP_ERR_OOM else 1)
"""


@pytest.mark.parametrize("mode", ["grids", "orbit_workers"])
def test_host_allocation_failure_is_a_status(mode):
    """A failed host allocation inside an entry point comes back as
    KLSP_ERR_OOM instead of a C++ exception through the C ABI (which would
    terminate the caller): in the entry point's own thread (grids) and in an
    orbit worker thread (orbit_workers).  Run in a child process with a
    capped address space."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, "-c", _OOM_CHILD, ROOT, mode], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr[-2000:])
    assert "status 2" in r.stdout
