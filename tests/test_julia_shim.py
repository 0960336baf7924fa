"""The Julia shim (julia/KLSP.jl) against the C ABI.

Julia is not installed here, so the shim is checked in two ways: statically
(every ccall names a symbol klsp.h declares, with as many argument types as
the C prototype has parameters, and its structs mirror klsp.h's field by
field), and on the device by driving exactly the ABI call sequence the shim
makes -- to_device (create, upload_problem, upload_W, prepare), then
compute_determinant, compute_determinant_batched, build_M_matrix,
grid_scan and newton_search -- through ctypes on the published small grid."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "julia", "KLSP.jl")
HEADER = os.path.join(ROOT, "include", "klsp.h")


def c_prototypes():
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    protos = {}
    for m in re.finditer(r"\b(?:int|const char\*)\s+(klsp_\w+)\s*\(([^)]*)\)\s*;", txt):
        args = [a for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        protos[m.group(1)] = len(args)
    return protos


def c_struct_fields(name):
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), txt, flags=re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        names = decl.split(None, 1)[1] if not decl.startswith("const") else decl.split(None, 2)[2]
        for n in names.split(","):
            n = n.strip().lstrip("*")
            fields.append(re.sub(r"\[.*\]", "", n))
    return fields


def julia_struct_fields(name):
    src = open(SHIM).read()
    body = re.search(r"(?:mutable )?struct %s\b.*?\n(.*?)\nend" % name, src, flags=re.S).group(1)
    return [ln.split("::")[0].strip() for ln in body.splitlines()
            if "::" in ln and not ln.strip().startswith("#") and "=" not in ln.split("::")[0]]


def test_shim_ccalls_match_the_header():
    protos = c_prototypes()
    src = open(SHIM).read()
    calls = re.findall(r"ccall\(\(:(klsp_\w+), lib\),\s*(\w+(?:\{\w+\})?),\s*\(([^()]*(?:\([^()]*\))?[^()]*)\)",
                       src)
    names = {c[0] for c in calls}
    assert {"klsp_create", "klsp_destroy", "klsp_last_error", "klsp_upload_problem",
            "klsp_upload_W", "klsp_prepare", "klsp_det", "klsp_build_M", "klsp_grid_scan",
            "klsp_newton"} <= names
    for name, ret, types in calls:
        assert name in protos, name
        ntypes = len([t for t in types.split(",") if t.strip()])
        assert ntypes == protos[name], (name, ntypes, protos[name])
        assert ret in ("Cint", "Cstring"), (name, ret)


@pytest.mark.parametrize("jl,c", [("Problem", "klsp_problem"), ("NewtonParams", "klsp_newton_params"),
                                  ("CNewtonResult", "klsp_newton_result"),
                                  ("ScanParams", "klsp_scan_params")])
def test_shim_structs_mirror_the_header(jl, c):
    assert julia_struct_fields(jl) == c_struct_fields(c)


def test_shim_keeps_the_reference_signatures():
    """The public functions keep the reference's names and positional
    arguments (BasisFunctions.jl:397, MatrixBuilder.jl:197, :215, :242,
    GridScan.jl:25, :234, NewtonSolver.jl:36) and the NewtonResult fields
    (NewtonSolver.jl:19-27); DevicePrecomputed keeps its name and fields."""
    src = open(SHIM).read()
    sigs = {
        "to_device": ["precomp", "gpu_type", "Tgpu"],
        "compute_determinant": ["omega", "precomputed", "devprecomp", "gpu_type", "Tcpu"],
        "compute_determinant_batched": ["omegas", "precomputed", "devprecomp", "gpu_type", "Tcpu"],
        "grid_scan": ["config", "precomputed", "devprecomp", "gpu_type", "Tcpu"],
        "newton_search": ["config", "precomputed", "devprecomp", "gpu_type", "Tcpu",
                          "Omega_p_init", "gamma_init"],
    }
    for fn, args in sigs.items():
        m = re.search(r"\nfunction %s\((.*?)\)\s*(?:where T)?\n" % fn, src, flags=re.S)
        assert m, fn
        pos = m.group(1).split(";")[0]
        while re.search(r"\{[^{}]*\}", pos):
            pos = re.sub(r"\{[^{}]*\}", "", pos)
        got = [a.split("::")[0].strip() for a in pos.split(",") if a.strip()]
        assert got == args, (fn, got)
    assert julia_struct_fields("NewtonResult") == ["Omega_p", "gamma", "omega", "det_val",
                                                   "iterations", "converged", "history"]
    assert julia_struct_fields("DevicePrecomputed")[:11] == [
        "m", "G", "N_alpha", "n_l", "NPh", "W_l_mat", "DJ_vec", "F0l_all", "Omega_res",
        "N_kernel", "d_alpha"]


@pytest.mark.gpu
def test_shim_call_sequence_on_the_device(small_cfg_paths):
    """The shim's ABI calls, in its order, on the published small grid: the
    root of convergence_table.txt:13 within 1e-10, dets equal to the oracle's
    (BLAS form on the same W) to 1e-11 of the det scale."""
    import oracle
    from oracle import blas_form
    from kalnajs_log_spirals_amd import _ffi
    L = _ffi.lib()
    po = oracle.precompute(oracle.load_toml_params(*small_cfg_paths))
    NPh, N, n_l = po["W"].shape
    DJ = np.ascontiguousarray(po["DJ"], dtype=np.float64)
    F0l = np.asfortranarray(po["F0l"], dtype=np.float64)
    Ores = np.asfortranarray(po["Omega_res"], dtype=np.float64)
    Nk = np.ascontiguousarray(po["N_kernel"], dtype=np.float64)
    W = np.asfortranarray(po["W"], dtype=np.complex128)
    D = ct.POINTER(ct.c_double)
    p = lambda a: a.ctypes.data_as(D)   # noqa: E731  (column-major storage, as Julia passes it)
    h = ct.c_void_p()
    assert L.klsp_create(0, ct.byref(h)) == 0
    try:
        prob = _ffi.klsp_problem(NPh, N, n_l, 2, 1.0, p(DJ), p(F0l), p(Ores), p(Nk))
        assert L.klsp_upload_problem(h, ct.byref(prob)) == 0
        assert L.klsp_upload_W(h, p(W)) == 0
        assert L.klsp_prepare(h, _ffi.PREC_FP64) == 0
        # compute_determinant / compute_determinant_batched
        oms = np.array([2 * 0.44 + 0.12j, 0.8 + 0.1j, 0.95 + 0.2j])
        dets = np.empty(3, dtype=np.complex128)
        assert L.klsp_det(h, p(oms[:1].copy()), 1, _ffi.PREC_FP64, p(dets)) == 0
        assert L.klsp_det(h, p(oms), 3, _ffi.PREC_FP64, p(dets)) == 0
        for k, om in enumerate(oms):
            ref = blas_form.det(W, DJ, F0l, Ores, Nk, 1.0, om)
            sc = max(abs(ref), abs(blas_form.det(W, DJ, F0l, Ores, Nk, 1.0, om + 1e-3)))
            assert abs(dets[k] - ref) <= 1e-11 * sc
        # build_M_matrix (device)
        M = np.empty((N, N), dtype=np.complex128, order="F")
        assert L.klsp_build_M(h, oms[1].real, oms[1].imag, _ffi.PREC_FP64, p(M)) == 0
        Mb = blas_form.build_M(W, DJ, F0l, Ores, Nk, 1.0, oms[1])
        assert np.max(np.abs(M - Mb)) <= 1e-12 * np.max(np.abs(Mb))
        # grid_scan
        sp = _ffi.klsp_scan_params(6, 5, 0.40, 0.48, 0.08, 0.16, _ffi.PREC_FP64)
        grid = np.empty((5, 6), dtype=np.complex128, order="F")
        op0, g0 = ct.c_double(), ct.c_double()
        assert L.klsp_grid_scan(h, ct.byref(sp), p(grid), ct.byref(op0), ct.byref(g0)) == 0
        dgo, opo, go = oracle.grid_scan(po, 6, 5, 0.40, 0.48, 0.08, 0.16)
        assert np.max(np.abs(grid - dgo) / np.abs(dgo)) <= 1e-11 and (op0.value, g0.value) == (opo, go)
        # newton_search -> NewtonResult
        npar = _ffi.klsp_newton_params(50, 1e-12, 1e-6, 0, 10, 0.44, 0.125, _ffi.PREC_FP64)
        res = _ffi.klsp_newton_result()
        hist = np.empty((50, 3), order="F")
        assert L.klsp_newton(h, ct.byref(npar), ct.byref(res), p(hist)) == 0
        assert res.converged == 1
        assert abs(res.Omega_p - 0.4397142267) <= 1e-10 and abs(res.gamma - 0.1230663944) <= 1e-10
        assert np.allclose(hist[:res.iterations, :2], oracle.newton(po, 0.44, 0.125)["history"][:, :2],
                           rtol=0, atol=1e-10)
    finally:
        assert L.klsp_destroy(h) == 0
