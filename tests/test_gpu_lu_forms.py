"""The two device LU forms (klsp_set_lu_form): the step-wise LU (one launch
per 16-column panel, a workgroup per column block) and the one-workgroup LU
give the same bits, and both match LAPACK zgetrf (numpy) on dense matrices
with real pivoting; NaN, singular and model-point (Taylor coefficient)
inputs take the same path in both."""
import numpy as np
import pytest

import oracle
from kalnajs_log_spirals_amd import _ffi
from kalnajs_log_spirals_amd.BasisFunctions import from_arrays

pytestmark = pytest.mark.gpu

AUTO, SINGLE, STEPS = 0, 1, 2


def dense_problem(N, seed):
    """Full-rank M (K = 2N): I - M is a general dense complex matrix."""
    rng = np.random.default_rng(seed)
    NPh, n_l = N, 2
    W = 2.0 * (rng.standard_normal((NPh, N, n_l)) + 1j * rng.standard_normal((NPh, N, n_l)))
    W /= np.sqrt(NPh * n_l)
    DJ, F0l = rng.uniform(0.1, 1, NPh), rng.standard_normal((NPh, n_l))
    Ores, Nk = rng.uniform(-2, 2, (NPh, n_l)), rng.uniform(0.5, 1.5, N)
    return from_arrays(2, 0.7, W, DJ, F0l, Ores, Nk)


OMS = [0.3 + 0.4j, -1.1 + 0.2j, 0.8 + 0.05j]


@pytest.mark.parametrize("N", [1, 2, 15, 16, 17, 31, 33, 100, 301, 401, 511, 512])
def test_step_lu_matches_single_workgroup_lu_and_lapack(N):
    pre = dense_problem(N, 100 + N)
    ctx = pre.ctx
    try:
        ctx.set_lu_form(SINGLE)
        d1 = ctx.det(OMS, _ffi.PREC_FP64)
        ctx.set_lu_form(STEPS)
        d2 = ctx.det(OMS, _ffi.PREC_FP64)
        d2b = np.array([ctx.det([om], _ffi.PREC_FP64)[0] for om in OMS])
        ctx.set_lu_form(AUTO)
        for i, om in enumerate(OMS):
            A = np.eye(N) - ctx.build_M(om, _ffi.PREC_FP64)
            ref = np.linalg.det(A)
            print(f"N={N} om={om}: step {d2[i]} single {d1[i]} lapack {ref}", flush=True)
            assert d2[i] == d1[i], (i, d2[i], d1[i])
            assert d2b[i] == d2[i]
            assert abs(d2[i] - ref) <= 1e-9 * abs(ref)
            assert abs(d2[i] - oracle.det_matrix(A)) <= 1e-9 * abs(ref)
    finally:
        ctx.close()


@pytest.mark.parametrize("N", [8, 37, 301])
def test_step_lu_special_matrices(N):
    """det(I) = 1 exactly, an exactly zero pivot column gives 0, and a NaN in
    W gives NaN, in both forms (izamax: a NaN diagonal stays the pivot)."""
    rng = np.random.default_rng(N)
    om = 1.5 + 0.0j
    Ores = rng.uniform(-2.0, 0.0, (N, 1))
    DJ, Nk, w = rng.uniform(0.2, 1.0, N), rng.uniform(0.5, 1.5, N), rng.uniform(0.5, 1.5, N)
    W = np.zeros((N, N, 1), dtype=np.complex128)
    W[np.arange(N), np.arange(N), 0] = w
    cases = []
    pre = from_arrays(2, 1.0, W, DJ, np.zeros((N, 1)), Ores, Nk)
    cases.append(("identity", pre, lambda d: d == 1.0))
    F0l = rng.uniform(-0.5, 0.5, (N, 1))
    a = N // 3
    F1, D1, O1, N1, W1 = F0l.copy(), DJ.copy(), Ores.copy(), Nk.copy(), W.copy()
    O1[a, 0], D1[a], F1[a, 0], N1[a], W1[a, a, 0] = 0.5, 1.0, 1.0, 1.0, 1.0
    cases.append(("singular", from_arrays(2, 1.0, W1, D1, F1, O1, N1), lambda d: d == 0.0))
    W2 = 0.1 * (rng.standard_normal((N, N, 1)) + 1j * rng.standard_normal((N, N, 1)))
    W2[N // 2, :, 0] = np.nan
    cases.append(("nan", from_arrays(2, 1.0, W2, DJ, F0l, Ores, Nk),
                  lambda d: np.isnan(d.real) and np.isnan(d.imag)))
    for name, pre, ok in cases:
        try:
            out = []
            for form in (SINGLE, STEPS):
                pre.ctx.set_lu_form(form)
                out.append(pre.ctx.det([om], _ffi.PREC_FP64)[0])
            print(name, N, out, flush=True)
            assert ok(out[0]) and ok(out[1]), (name, out)
            if name == "nan":
                A = np.eye(N) - pre.ctx.build_M(om, _ffi.PREC_FP64)
                assert np.isnan(oracle.det_matrix(A))
        finally:
            pre.ctx.close()


def test_step_lu_in_the_mixed_search(oracle_small):
    """Model-point LUs (Taylor coefficients) in both forms: the mixed search
    gives the same root, history and builds."""
    o = oracle_small
    res = []
    for form in (SINGLE, STEPS):
        pre = from_arrays(o["m"], o["G"], o["W"], o["DJ"], o["F0l"], o["Omega_res"],
                          o["N_kernel"])
        try:
            pre.ctx.set_lu_form(form)
            r, H = pre.ctx.newton(0.45, 0.11, max_iter=50, tol=1e-12, precision=_ffi.PREC_MIXED)
            res.append((r.omega[0], r.omega[1], r.n_fp64_builds, r.n_lp_builds, H.copy()))
        finally:
            pre.ctx.close()
    assert res[0][:4] == res[1][:4]
    assert np.array_equal(res[0][4], res[1][4])
