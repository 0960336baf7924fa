"""The two work decompositions of the real-basis BF16x3 contraction
(klsp_set_bf16_form): folded column-block pairs (1) and folded two-column
strips (2).  Both approximate the FP64 build to the BF16x3 bound, agree with
each other to FP32 accumulation order, and give the mixed search the same
root; the automatic choice takes one of them.  N covers one strip, odd and
even block counts, and strip pairs that need two workgroups."""
import numpy as np
import pytest

from kalnajs_log_spirals_amd import _ffi
from kalnajs_log_spirals_amd.BasisFunctions import from_arrays

pytestmark = pytest.mark.gpu

AUTO, FOLD, STRIP = 0, 1, 2


def relnorm(a, b) -> float:
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def mirror_symmetric_W(rng, NPh, N, n_l, scale):
    """Random W with W[:, N-1-a] = conj W[:, a] and a real centre column."""
    W = (rng.standard_normal((NPh, N, n_l)) + 1j * rng.standard_normal((NPh, N, n_l))) * scale
    h = N // 2
    for a in range(h):
        W[:, N - 1 - a, :] = np.conj(W[:, a, :])
    if N % 2:
        W[:, h, :] = W[:, h, :].real
    return W


@pytest.mark.parametrize("N", [5, 17, 33, 130, 241, 301, 330, 401])
def test_both_forms_match_fp64_and_each_other(N):
    rng = np.random.default_rng(7000 + N)
    NPh, n_l = 173, 2
    W = mirror_symmetric_W(rng, NPh, N, n_l, 1.0 / np.sqrt(NPh * n_l))
    DJ, F0l = rng.uniform(0.2, 1.0, NPh), rng.standard_normal((NPh, n_l))
    Ores, Nk = rng.uniform(-2, 2, (NPh, n_l)), rng.uniform(0.5, 1.5, N)
    pre = from_arrays(2, 0.8, W, DJ, F0l, Ores, Nk)
    ctx = pre.ctx
    try:
        assert ctx.symmetric_basis()
        om = 0.4 + 0.25j
        oms = [om, 0.9 + 0.1j, 0.2 + 0.3j]
        M64 = ctx.build_M(om, _ffi.PREC_FP64)
        T64 = ctx.build_taylor(om, 3, _ffi.PREC_FP64)
        d64 = ctx.det(oms, _ffi.PREC_FP64)
        got = {}
        for form in (FOLD, STRIP, AUTO):
            ctx.set_bf16_form(form)
            Mb = ctx.build_M(om, _ffi.PREC_BF16X3)
            T3 = ctx.build_taylor(om, 3, _ffi.PREC_BF16X3)
            T4 = ctx.build_taylor(om, 4, _ffi.PREC_BF16X3)
            db = ctx.det(oms, _ffi.PREC_BF16X3)
            print(f"N={N} form={form}: M {relnorm(Mb, M64):.2e} det {db}", flush=True)
            assert relnorm(Mb, M64) <= 1e-4
            for j in range(3):
                assert relnorm(T3[j], T64[j]) <= 1e-4, (form, j)
                assert relnorm(T4[j], T64[j]) <= 1e-4, (form, j)
            for k in range(3):
                assert abs(db[k] - d64[k]) <= 1e-3 * max(abs(d64[k]), 1.0), (form, k)
            got[form] = (Mb, T3, T4)
        # the forms differ only in how FP32 partial sums are split over gangs
        for a, b in zip(got[FOLD], got[STRIP]):
            assert relnorm(b, a) <= 2e-6
        for a, b in zip(got[AUTO], got[FOLD] if relnorm(got[AUTO][0], got[FOLD][0]) == 0
                        else got[STRIP]):
            assert np.array_equal(a, b)
    finally:
        ctx.set_bf16_form(AUTO)
        ctx.close()


def test_mixed_search_same_root_in_both_forms(oracle_small):
    """The mixed search's Taylor passes (three full terms and a hi x hi fourth)
    in either form reach the FP64 root within 1e-10."""
    o = oracle_small
    pre = from_arrays(o["m"], o["G"], o["W"], o["DJ"], o["F0l"], o["Omega_res"], o["N_kernel"])
    try:
        assert pre.ctx.symmetric_basis()
        r64, _ = pre.ctx.newton(0.45, 0.11, max_iter=50, tol=1e-12, precision=_ffi.PREC_FP64)
        for form in (FOLD, STRIP):
            pre.ctx.set_bf16_form(form)
            r, _ = pre.ctx.newton(0.45, 0.11, max_iter=50, tol=1e-12, precision=_ffi.PREC_MIXED)
            d = abs(complex(*r.omega[:2]) - complex(*r64.omega[:2]))
            print(f"form={form}: |d omega| {d:.2e}, builds {r.n_fp64_builds} FP64 "
                  f"{r.n_lp_builds} low", flush=True)
            assert r.converged and d <= 1e-10
    finally:
        pre.ctx.close()


@pytest.mark.parametrize("form", [FOLD, STRIP])
def test_graph_scan_in_each_form(oracle_small, form):
    """A BF16x3 scan captured as one HIP graph equals the same scan launched
    kernel by kernel, bit for bit, in each form, on repeated graph launches
    with eager dets between them; changing the form instantiates a new
    graph."""
    o = oracle_small
    pre = from_arrays(o["m"], o["G"], o["W"], o["DJ"], o["F0l"], o["Omega_res"], o["N_kernel"])
    ctx = pre.ctx
    try:
        win = (7, 5, 0.30, 0.60, 0.02, 0.25)
        ctx.set_bf16_form(form)
        ctx.set_scan_graph(False)
        ref = ctx.grid_scan(*win, precision=_ffi.PREC_BF16X3)
        ctx.set_scan_graph(True)
        b0, _ = ctx.scan_graph_stats()
        for k in range(3):
            ctx.det([0.88 + 0.12j, 0.7 + 0.05j], _ffi.PREC_BF16X3)
            dg, op0, g0 = ctx.grid_scan(*win, precision=_ffi.PREC_BF16X3)
            assert np.array_equal(dg, ref[0]), (form, k, np.max(np.abs(dg - ref[0])))
            assert (op0, g0) == (ref[1], ref[2])
        b1, _ = ctx.scan_graph_stats()
        assert b1 - b0 == 1
        ctx.set_bf16_form(FOLD if form == STRIP else STRIP)
        ctx.grid_scan(*win, precision=_ffi.PREC_BF16X3)
        assert ctx.scan_graph_stats()[0] == b1 + 1
    finally:
        ctx.close()


def test_bf16_form_rejects_unknown_values(oracle_small):
    o = oracle_small
    pre = from_arrays(o["m"], o["G"], o["W"], o["DJ"], o["F0l"], o["Omega_res"], o["N_kernel"])
    try:
        for bad in (-1, 3):
            with pytest.raises(_ffi.KlspError):
                pre.ctx.set_bf16_form(bad)
    finally:
        pre.ctx.close()
