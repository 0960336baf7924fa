"""GPU parity: libklsp (HIP, gfx950) against the FP64 CPU oracle.

Tolerances (FP64 path): W and M normwise relative <= 1e-12, det relative
<= 1e-11, Newton roots |d omega| <= 1e-10.  The FP32 path is checked against
the FP64 device result.
"""
import numpy as np
import pytest

import oracle
import kalnajs_log_spirals_amd as K
from kalnajs_log_spirals_amd import _ffi
from kalnajs_log_spirals_amd.BasisFunctions import from_arrays

pytestmark = pytest.mark.gpu


def cfg_from_params(p) -> K.KalnajsConfig:
    c = K.KalnajsConfig()
    c.physics.m, c.physics.G = p.m, p.G
    c.model.L0, c.model.n_zang, c.model.q1 = p.L0, p.n_zang, p.q1
    for k in ("NR", "Ne", "Rc_min", "Rc_max", "v_min", "N_alpha", "alpha_max", "l_min", "l_max",
              "nw", "nwa"):
        setattr(c.grid, k, getattr(p, k))
    c.gpu.precision_double = c.cpu.precision_double = True
    return c


def relnorm(a, b) -> float:
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def device_pre(pre_oracle):
    return from_arrays(pre_oracle["m"], pre_oracle["G"], pre_oracle["W"], pre_oracle["DJ"],
                       pre_oracle["F0l"], pre_oracle["Omega_res"], pre_oracle["N_kernel"])


OMEGAS = [2 * 0.4397142267 + 0.1230663944j, 0.6 + 0.05j, 1.1 + 0.2j, 0.9 - 0.1j]


def det_scale(pre_o, om) -> float:
    """Size of det near omega: the first OMEGAS entry is the published root,
    where det itself is ~1e-13, so errors are measured against the det one
    1e-3 away (an error e there moves the root by ~e / |det'| ~ 1e-3 e / scale)."""
    return max(abs(oracle.compute_det(pre_o, om)), abs(oracle.compute_det(pre_o, om + 1e-3)))


def test_library_reports_gfx950():
    print("library:", _ffi.version(), "devices:", _ffi.device_count(), flush=True)
    assert _ffi.device_count() >= 1
    assert K.detect_gpu_type() == K.GPUType.AMD


def test_W_quadrature_matches_oracle(oracle_tiny):
    """compute_W_l (BasisFunctions.jl:199) on FP64 MFMA vs the direct sum."""
    p = oracle_tiny["params"]
    cfg = cfg_from_params(p)
    pre = K.precompute_all(cfg, K.create_toomre_zang_model(cfg), K.GPUType.AMD)
    W = pre.W_l_mat
    assert W.shape == oracle_tiny["W"].shape
    assert relnorm(W, oracle_tiny["W"]) <= 1e-12
    np.testing.assert_allclose(pre.DJ_vec, oracle_tiny["DJ"], rtol=1e-13, atol=0)
    np.testing.assert_allclose(pre.F0l_all, oracle_tiny["F0l"], rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(pre.N_kernel, oracle_tiny["N_kernel"], rtol=1e-14)
    # the standalone entry point returns the same array
    W2 = K.compute_W_l(cfg, pre.orbit_data, pre.alpha_arr, pre.l_arr, K.GPUType.AMD)
    assert np.array_equal(W2, W)


def test_W_quadrature_many_harmonics():
    """n_l = 301 (l = -150..150, the uneven batch's largest): more l-blocks than
    one workgroup covers, so the grid's third dimension is used."""
    import oracle
    p = oracle.Params(NR=5, Ne=3, Rc_min=0.1, Rc_max=10.0, v_min=0.05, N_alpha=9,
                      alpha_max=5.0, l_min=-150, l_max=150, nw=401, nwa=41)
    po = oracle.precompute(p)
    cfg = cfg_from_params(p)
    pre = K.precompute_all(cfg, K.create_toomre_zang_model(cfg), K.GPUType.AMD)
    W = pre.W_l_mat
    assert W.shape == po["W"].shape == (15, 9, 301)
    for il in (0, 150, 255, 256, 300):
        assert relnorm(W[:, :, il], po["W"][:, :, il]) <= 1e-12, il
    assert relnorm(W, po["W"]) <= 1e-12


def test_W_quadrature_small_grid(oracle_small, small_cfg_paths):
    cfg = K.load_model(K.load_config(small_cfg_paths[0]), small_cfg_paths[1])
    pre = K.precompute_all(cfg, K.create_toomre_zang_model(cfg), K.GPUType.AMD)
    W = pre.W_l_mat
    for il in range(W.shape[2]):
        assert relnorm(W[:, :, il], oracle_small["W"][:, :, il]) <= 1e-12, il


@pytest.mark.parametrize("case", ["gaps_large_l", "large_alpha", "uneven_alpha",
                                  "even_symmetric", "wide_even_symmetric",
                                  "uneven_symmetric"])
def test_W_quadrature_general_l_and_large_arguments(case):
    """compute_W with l values that are not one consecutive run (every cos
    evaluated directly), with alphas that are not evenly spaced (every
    exp(-i alpha ln r) evaluated directly) and with sine/cosine arguments
    beyond 2^20 (the
    library's sincos instead of the short reduction; both kinds of orbit in
    one launch for large alpha).  With alpha up to 5e5 the phase alpha ln r
    inherits alpha times the one-ulp difference between the device's and the
    host's ln, so that case is held to 4 max|alpha| eps max|ln r| instead of
    1e-12.  The symmetric cases (alpha[j] == -alpha[N-1-j] exactly, N even)
    take the kernel's mirrored form, which computes the columns N/2.. and
    writes W(-alpha) = conj W(alpha) into the other half: 20 and 150 evenly
    spaced alphas (one and two alpha-workgroups) and 36 unevenly spaced ones."""
    p = oracle.Params(NR=9, Ne=7, Rc_min=0.05, Rc_max=20.0, v_min=0.01, N_alpha=21,
                      alpha_max=8.0, l_min=-3, l_max=3, nw=401, nwa=41)
    g = oracle.compute_grids(p)
    o = oracle.compute_orbits(p, g)
    if case == "gaps_large_l":
        l_arr = np.array([-400000, -3, 0, 2, 5, 350001, 17], dtype=np.int32)
        alpha = oracle.linrange(-8.0, 8.0, 21)
    elif case == "uneven_alpha":
        l_arr = np.arange(-3, 4, dtype=np.int32)
        alpha = np.sort(np.random.default_rng(7).uniform(-8.0, 8.0, 37))
    elif case == "even_symmetric":
        l_arr = np.arange(-3, 4, dtype=np.int32)
        alpha = oracle.linrange(-8.0, 8.0, 20)
    elif case == "wide_even_symmetric":
        l_arr = np.arange(-3, 4, dtype=np.int32)
        alpha = oracle.linrange(-8.0, 8.0, 150)
    elif case == "uneven_symmetric":
        l_arr = np.arange(-3, 4, dtype=np.int32)
        h = np.sort(np.random.default_rng(11).uniform(0.01, 8.0, 18))
        alpha = np.concatenate([-h[::-1], h])
    else:
        l_arr = np.arange(-3, 4, dtype=np.int32)
        alpha = oracle.linrange(-5e5, 5e5, 23)
    if case.endswith("_symmetric"):
        assert alpha.size % 2 == 0 and np.array_equal(alpha, -alpha[::-1])
    Wo = oracle.compute_W(p, o, alpha, l_arr)
    NPh, N, n_l = p.NPh, alpha.size, l_arr.size
    ctx = _ffi.Context(0)
    try:
        ctx.upload_problem(NPh, N, n_l, p.m, 1.0, np.zeros(NPh), np.zeros((NPh, n_l)),
                           np.zeros((NPh, n_l)), np.zeros(N))
        # the oracle's [NPh, nwa] row-major samples are the library's [nwa, NPh]
        W = ctx.compute_W(o["w1"], o["ra"], o["pha"], o["Om1"], o["Om2"], alpha, l_arr,
                          download=True)
    finally:
        ctx.close()
    tol = 1e-12
    if case == "large_alpha":
        tol = 4 * np.max(np.abs(alpha)) * np.finfo(float).eps * np.nanmax(np.abs(np.log(o["ra"])))
    for il in range(n_l):
        err = relnorm(W[:, :, il], Wo[:, :, il])
        print(case, "l", l_arr[il], f"relnorm {err:.2e} (tolerance {tol:.1e})", flush=True)
        assert err <= tol, il


@pytest.mark.parametrize("omega", OMEGAS)
def test_M_fp64_matches_oracle(oracle_tiny, omega):
    """build_M_matrix (MatrixBuilder.jl:27) on FP64 MFMA."""
    pre = device_pre(oracle_tiny)
    M = K.build_M_matrix(omega, pre)
    Mo = oracle.build_M(oracle_tiny, omega)
    assert relnorm(M, Mo) <= 1e-12


@pytest.mark.parametrize("omega", OMEGAS)
def test_det_fp64_matches_oracle_small(oracle_small, omega):
    pre = device_pre(oracle_small)
    dev = K.to_device(pre, K.GPUType.AMD, np.float64)
    d = K.compute_determinant(omega, pre, dev, K.GPUType.AMD, np.float64)
    do = oracle.compute_det(oracle_small, omega)
    assert abs(d - do) <= 1e-11 * det_scale(oracle_small, omega)
    M = K.build_M_matrix(omega, pre)
    assert relnorm(M, oracle.build_M(oracle_small, omega)) <= 1e-12


def test_lu_matches_lapack(oracle_small):
    """The device LU vs LAPACK zgetrf on the same device-built M."""
    pre = device_pre(oracle_small)
    for om in OMEGAS:
        M = pre.ctx.build_M(om, _ffi.PREC_FP64)
        d = pre.ctx.det([om], _ffi.PREC_FP64)[0]
        ref = np.linalg.det(np.eye(M.shape[0]) - M)
        sc = det_scale(oracle_small, om)
        assert abs(d - ref) <= 1e-12 * sc
        assert abs(d - oracle.det_matrix(np.eye(M.shape[0]) - M)) <= 1e-12 * sc


RAGGED = [(1, 1, 1), (3, 5, 2), (13, 16, 1), (13, 17, 3), (7, 37, 2), (5, 64, 2), (11, 65, 3),
          (4, 130, 1)]


@pytest.mark.parametrize("NPh,N,n_l", RAGGED)
def test_random_problem_ragged_sizes(NPh, N, n_l):
    """Sizes that are not multiples of the 16x16x4 MFMA tiles or the 64 tile."""
    rng = np.random.default_rng(1000 * NPh + 10 * N + n_l)
    W = rng.standard_normal((NPh, N, n_l)) + 1j * rng.standard_normal((NPh, N, n_l))
    pre_o = dict(m=2, G=0.7, W=W / np.sqrt(NPh * n_l), DJ=rng.uniform(0.1, 1, NPh),
                 F0l=rng.standard_normal((NPh, n_l)), Omega_res=rng.uniform(-2, 2, (NPh, n_l)),
                 N_kernel=rng.uniform(0.05, 0.3, N))
    pre = device_pre(pre_o)
    for om in (0.3 + 0.4j, -1.2 + 0.05j):
        assert relnorm(K.build_M_matrix(om, pre), oracle.build_M(pre_o, om)) <= 1e-12
        d = pre.ctx.det([om], _ffi.PREC_FP64)[0]
        do = oracle.compute_det(pre_o, om)
        assert abs(d - do) <= 1e-11 * max(abs(do), 1e-300)


@pytest.mark.parametrize("N,n_l", [(301, 5), (45, 1)])
def test_balanced_work_tables(N, n_l):
    """The contraction's work tables: N = 301 (19 x 19 blocks: 4 x 4 tiles of
    16, 12 and 9 blocks get 45 / 34 / 26 K pieces) and N = 45 (3 x 3 blocks,
    one edge class); every FP64 and FP32 launch shape (1, 2, 3 weight sets)
    against numpy ZGEMMs per l (the BLAS form), batched FP32 dets against
    single ones."""
    from oracle import blas_form
    rng = np.random.default_rng(N + 7 * n_l)
    NPh = 2501          # K / 8 >= 32 x 45: the full piece counts are reached
    W = (rng.standard_normal((NPh, N, n_l)) + 1j * rng.standard_normal((NPh, N, n_l)))
    W /= np.sqrt(NPh * n_l)
    DJ, F0l = rng.uniform(0.1, 1, NPh), rng.standard_normal((NPh, n_l))
    Ores, Nk = rng.uniform(-2, 2, (NPh, n_l)), rng.uniform(0.05, 0.3, N)
    pre = from_arrays(2, 0.7, W, DJ, F0l, Ores, Nk)
    oms = [0.3 + 0.4j, -1.2 + 0.05j, 0.9 + 0.2j]
    for om in oms:
        ref = blas_form.build_M(W, DJ, F0l, Ores, Nk, 0.7, om)
        assert relnorm(pre.ctx.build_M(om, _ffi.PREC_FP64), ref) <= 1e-12
        assert relnorm(pre.ctx.build_M(om, _ffi.PREC_FP32), ref) <= 2e-6
    # FP32 dets of 2 and 3 omegas per pass (the batched kernels) vs one per pass
    single = [pre.ctx.det([om], _ffi.PREC_FP32)[0] for om in oms]
    for nb in (2, 3):
        db = pre.ctx.det(oms[:nb], _ffi.PREC_FP32)
        for i in range(nb):
            assert abs(db[i] - single[i]) <= 1e-5 * abs(single[i]), (nb, i)
    # Taylor terms T_j = G sum_l W_l^T diag(DJ F0l / d^j) conj(W_l) diag(N), d = w_c - Omega_res:
    # the batched FP64 kernels (2 and 3 weight sets) and the FP32 ones
    wc = 0.9 + 0.2j
    ref = []
    for j in (1, 2, 3):
        T = np.zeros((N, N), complex)
        for il in range(n_l):
            c = DJ * F0l[:, il] / (wc - Ores[:, il]) ** j
            T += W[:, :, il].T @ (c[:, None] * W[:, :, il].conj())
        ref.append(0.7 * T * Nk[None, :])
    for nt in (2, 3):
        T64 = pre.ctx.build_taylor(wc, nt, _ffi.PREC_FP64)
        T32 = pre.ctx.build_taylor(wc, nt, _ffi.PREC_FP32)
        for j in range(nt):
            assert relnorm(T64[j], ref[j]) <= 1e-12, (nt, j)
            assert relnorm(T32[j], ref[j]) <= 1e-5, (nt, j)
    pre.ctx.close()


@pytest.mark.parametrize("N", [100, 300, 520, 1000, 1024])
def test_lu_dense_pivoting(N):
    """Full-rank M (K = 2N) scaled so that I - M is a general dense complex
    matrix: the blocked LU pivots in most columns; det vs LAPACK on the same
    device-built M and vs the BLAS-form build."""
    from oracle import blas_form
    rng = np.random.default_rng(N)
    NPh, n_l = N, 2
    W = 2.0 * (rng.standard_normal((NPh, N, n_l)) + 1j * rng.standard_normal((NPh, N, n_l)))
    W /= np.sqrt(NPh * n_l)
    DJ, F0l = rng.uniform(0.1, 1, NPh), rng.standard_normal((NPh, n_l))
    Ores, Nk = rng.uniform(-2, 2, (NPh, n_l)), rng.uniform(0.5, 1.5, N)
    pre = from_arrays(2, 0.7, W, DJ, F0l, Ores, Nk)
    om = 0.3 + 0.4j
    M = pre.ctx.build_M(om, _ffi.PREC_FP64)
    Mb = blas_form.build_M(W, DJ, F0l, Ores, Nk, 0.7, om)
    assert relnorm(M, Mb) <= 1e-12
    A = np.eye(N) - M
    assert np.linalg.cond(A) < 1e8
    d = pre.ctx.det([om], _ffi.PREC_FP64)[0]
    ref = np.linalg.det(A)
    # det is a product of N pivots: relative agreement ~ N eps cond
    assert abs(d - ref) <= 1e-9 * abs(ref)
    # the pivoting is real: LAPACK interchanges rows in a third of the columns or more
    import scipy.linalg
    piv = scipy.linalg.lu_factor(A)[1]
    assert np.count_nonzero(piv != np.arange(N)) >= N // 3


@pytest.mark.parametrize("N", [8, 37, 301])
def test_lu_identity_diagonal_and_singular(N):
    """Special cases of the device LU on matrices built through the normal
    path: F0l = 0 gives M = 0 and det(I) = 1 exactly; W diagonal
    (W[k, a] = w_a delta_ka, one l) gives a diagonal M and
    det(I - M) = prod_a (1 - G N_a c_a w_a^2) in closed form; with
    c = N = G = w = 1 on one diagonal entry (every operation exact) I - M
    has an exactly zero column, where LAPACK's zgetrf reports a zero pivot
    and LinearAlgebra's det returns exactly 0 (MatrixBuilder.jl:85)."""
    rng = np.random.default_rng(N)
    om = 1.5 + 0.0j
    G = 1.0
    Ores = rng.uniform(-2.0, 0.0, (N, 1))
    DJ = rng.uniform(0.2, 1.0, N)
    Nk = rng.uniform(0.5, 1.5, N)
    w = rng.uniform(0.5, 1.5, N)
    W = np.zeros((N, N, 1), dtype=np.complex128)
    W[np.arange(N), np.arange(N), 0] = w
    # identity: F0l = 0
    pre = from_arrays(2, G, W, DJ, np.zeros((N, 1)), Ores, Nk)
    d = pre.ctx.det([om], _ffi.PREC_FP64)[0]
    pre.ctx.close()
    assert d == 1.0, d
    # diagonal, closed form
    F0l = rng.uniform(-0.5, 0.5, (N, 1))
    pre = from_arrays(2, G, W, DJ, F0l, Ores, Nk)
    M = pre.ctx.build_M(om, _ffi.PREC_FP64)
    c = DJ * F0l[:, 0] / (om - Ores[:, 0])
    diag = G * Nk * c * w * w
    assert np.count_nonzero(M - np.diag(np.diag(M))) == 0
    np.testing.assert_allclose(np.diag(M), diag, rtol=1e-14, atol=0)
    d = pre.ctx.det([om], _ffi.PREC_FP64)[0]
    pre.ctx.close()
    ref = np.prod(1.0 - diag)
    assert abs(d - ref) <= 4 * N * np.finfo(float).eps * abs(ref), (d, ref)
    # exactly singular: entry a has d = om - Ores = 1, DJ = F0l = N = w = 1
    a = N // 3
    Ores[a, 0], DJ[a], F0l[a, 0], Nk[a], W[a, a, 0] = 0.5, 1.0, 1.0, 1.0, 1.0
    pre = from_arrays(2, G, W, DJ, F0l, Ores, Nk)
    M = pre.ctx.build_M(om, _ffi.PREC_FP64)
    assert M[a, a] == 1.0
    d = pre.ctx.det([om], _ffi.PREC_FP64)[0]
    pre.ctx.close()
    assert d == 0.0, d


def test_resonance_clamp():
    """|omega - Omega_res| < 1e-12 -> denominator 1e-12 + 1e-12 i (MatrixBuilder.jl:47)."""
    rng = np.random.default_rng(7)
    NPh, N, n_l = 6, 9, 2
    W = (rng.standard_normal((NPh, N, n_l)) + 1j * rng.standard_normal((NPh, N, n_l))) * 1e-7
    Ores = rng.uniform(0.5, 1.5, (NPh, n_l))
    om = complex(Ores[2, 1], 0.0)
    Ores[4, 0] = om.real + 3e-13
    pre_o = dict(m=2, G=1.0, W=W, DJ=np.ones(NPh), F0l=np.ones((NPh, n_l)), Omega_res=Ores,
                 N_kernel=np.ones(N))
    pre = device_pre(pre_o)
    M = K.build_M_matrix(om, pre)
    Mo = oracle.build_M(pre_o, om)
    assert relnorm(M, Mo) <= 1e-12
    assert np.max(np.abs(M)) > 1e-3  # the clamped terms dominate


def test_invalid_orbit_is_skipped():
    """An orbit with a non-finite sample gives W = 0 (BasisFunctions.jl:224-226)."""
    p = oracle.Params(NR=5, Ne=5, Rc_min=0.1, Rc_max=10.0, N_alpha=17, alpha_max=5.0,
                      l_min=-2, l_max=2, nw=201, nwa=21)
    g = oracle.compute_grids(p)
    o = oracle.compute_orbits(p, g)
    o["ra"][7, 3] = np.nan
    o["pha"][11, 0] = np.inf
    alpha = oracle.linrange(-5, 5, 17)
    l_arr = np.arange(-2, 3, dtype=np.int32)
    Wo = oracle.compute_W(p, o, alpha, l_arr)
    assert np.all(Wo[7] == 0) and np.all(Wo[11] == 0)
    ctx = _ffi.Context(0)
    ctx.upload_problem(p.NPh, 17, 5, p.m, 1.0, np.zeros(p.NPh), np.zeros((p.NPh, 5)),
                       np.zeros((p.NPh, 5)), np.zeros(17))
    W = ctx.compute_W(o["w1"], o["ra"], o["pha"], o["Om1"], o["Om2"], alpha, l_arr, download=True)
    assert np.all(W[7] == 0) and np.all(W[11] == 0)
    assert relnorm(W, Wo) <= 1e-12
    ctx.close()


def test_conjugate_symmetry_small(oracle_small):
    """det(I - M(conj w)) = conj det(I - M(w)) while the clamp is inactive."""
    pre = device_pre(oracle_small)
    for om in OMEGAS[:3]:
        d1, d2 = pre.ctx.det([om, np.conj(om)], _ffi.PREC_FP64)
        assert abs(d2 - np.conj(d1)) <= 1e-11 * det_scale(oracle_small, om)


def test_batched_equals_single(oracle_small):
    pre = device_pre(oracle_small)
    oms = np.array(OMEGAS + [0.7 + 0.3j, 0.8 + 0.11j, 0.95 + 0.02j])
    db = pre.ctx.det(oms, _ffi.PREC_FP64)
    for i, om in enumerate(oms):
        assert abs(pre.ctx.det([om], _ffi.PREC_FP64)[0] - db[i]) <= 1e-13 * abs(db[i])


def test_newton_small_grid_known_answer(small_cfg_paths):
    """Full product pipeline on the published grid: the root of
    convergence_table.txt:13 (Omega_p = 0.4397142267, gamma = 0.1230663944)."""
    cfg = K.load_model(K.load_config(small_cfg_paths[0]), small_cfg_paths[1])
    model = K.create_toomre_zang_model(cfg)
    pre = K.precompute_all(cfg, model, K.GPUType.AMD)
    dev = K.to_device(pre, K.GPUType.AMD, np.float64)
    r = K.newton_search(cfg, pre, dev, K.GPUType.AMD, np.float64, 0.44, 0.125)
    print("root", r.Omega_p, r.gamma, r.iterations, r.det_val, flush=True)
    assert r.converged
    assert abs(r.Omega_p - 0.4397142267) <= 1e-10
    assert abs(r.gamma - 0.1230663944) <= 1e-10
    pre_o = oracle.precompute(oracle.load_toml_params(*small_cfg_paths))
    ro = oracle.newton(pre_o, 0.44, 0.125)
    assert r.iterations == ro["iterations"]
    assert abs(complex(2 * r.Omega_p, r.gamma) - ro["omega"]) <= 1e-10
    np.testing.assert_allclose(r.history[:, :2], ro["history"][:, :2], rtol=0, atol=1e-10)


def test_newton_line_search_and_bounds(oracle_small):
    pre = device_pre(oracle_small)
    res, H = pre.ctx.newton(0.45, 0.11, max_iter=20, tol=1e-12, use_line_search=True,
                            max_line_search=10)
    ro = oracle.newton(oracle_small, 0.45, 0.11, max_iter=20, use_line_search=True)
    assert res.converged and ro["converged"]
    assert abs(complex(res.omega[0], res.omega[1]) - ro["omega"]) <= 1e-10
    assert res.iterations == ro["iterations"]


def test_grid_scan_small(oracle_small):
    pre = device_pre(oracle_small)
    dg, op0, g0 = pre.ctx.grid_scan(6, 5, 0.40, 0.48, 0.08, 0.16)
    dgo, opo, go = oracle.grid_scan(oracle_small, 6, 5, 0.40, 0.48, 0.08, 0.16)
    assert np.max(np.abs(dg - dgo) / np.abs(dgo)) <= 1e-11
    assert (op0, g0) == (opo, go)
    assert K.find_minimum(dg, oracle.linrange(0.40, 0.48, 6), oracle.linrange(0.08, 0.16, 5)) \
        == (op0, g0)


def test_fp32_path_close_to_fp64(oracle_small):
    """complex64 W + FP32 MFMA with FP64 K-tile sums vs the FP64 device path."""
    pre = device_pre(oracle_small)
    for om in OMEGAS:
        M64 = pre.ctx.build_M(om, _ffi.PREC_FP64)
        M32 = pre.ctx.build_M(om, _ffi.PREC_FP32)
        assert relnorm(M32, M64) <= 2e-6
        d64 = pre.ctx.det([om], _ffi.PREC_FP64)[0]
        d32 = pre.ctx.det([om], _ffi.PREC_FP32)[0]
        assert abs(d32 - d64) <= 1e-4 * abs(d64) + 1e-9


# --------------------------------------------------------------------------
# Mixed-precision Newton (KLSP_PREC_MIXED) and its Taylor building block
# --------------------------------------------------------------------------
@pytest.mark.parametrize("prec", [_ffi.PREC_FP64, _ffi.PREC_FP32])
def test_taylor_terms_expand_M(oracle_small, prec):
    """T_1 = M(w_c), and M(w_c + s) - (T_1 - s T_2 + s^2 T_3) = O(|s|^3),
    with M from the oracle at the shifted points."""
    pre = device_pre(oracle_small)
    wc = 2 * 0.44 + 0.12j
    T = pre.ctx.build_taylor(wc, 3, prec)
    M0 = oracle.build_M(oracle_small, wc)
    tol = 1e-12 if prec == _ffi.PREC_FP64 else 2e-6
    assert relnorm(T[0], M0) <= tol
    errs = []
    for s in (4e-3, 2e-3 + 2e-3j, 1e-3j):
        Ms = oracle.build_M(oracle_small, wc + s)
        errs.append(relnorm(T[0] - s * T[1] + s * s * T[2], Ms))
        # third-order remainder: |s|^3 |T_4| ~ |s / Im w_c|^3 |M|
        assert errs[-1] <= 20 * (abs(s) / wc.imag) ** 3 + tol, (s, errs[-1])
    if prec == _ffi.PREC_FP64:
        # the remainder shrinks like |s|^3: halving s divides it by ~8
        e1 = relnorm(T[0] - 2e-3 * T[1] + 4e-6 * T[2], oracle.build_M(oracle_small, wc + 2e-3))
        e2 = relnorm(T[0] - 1e-3 * T[1] + 1e-6 * T[2], oracle.build_M(oracle_small, wc + 1e-3))
        assert 5.0 <= e1 / e2 <= 11.0


@pytest.mark.parametrize("om", OMEGAS)
def test_bf16x3_build_close_to_fp64(oracle_small, om):
    """BF16 hi/lo split products, FP32 accumulation: M within 1e-4 normwise."""
    pre = device_pre(oracle_small)
    M64 = pre.ctx.build_M(om, _ffi.PREC_FP64)
    Mb = pre.ctx.build_M(om, _ffi.PREC_BF16X3)
    assert relnorm(Mb, M64) <= 1e-4
    d64 = pre.ctx.det([om], _ffi.PREC_FP64)[0]
    db = pre.ctx.det([om], _ffi.PREC_BF16X3)[0]
    assert abs(db - d64) <= 1e-3 * det_scale(oracle_small, om)


@pytest.mark.parametrize("NPh,N,n_l", [(37, 5, 2), (61, 17, 3), (100, 130, 2), (40, 330, 1)])
def test_bf16x3_ragged_sizes(NPh, N, n_l):
    """Sizes off the kernel's tiles (N = 330 needs two workgroup rows of beta)."""
    rng = np.random.default_rng(NPh * 1000 + N)
    W = (rng.standard_normal((NPh, N, n_l)) + 1j * rng.standard_normal((NPh, N, n_l))) / np.sqrt(NPh)
    DJ = rng.uniform(0.5, 1.5, NPh)
    F0l = rng.standard_normal((NPh, n_l))
    Ores = rng.uniform(-1, 1, (NPh, n_l))
    Nk = rng.uniform(0.5, 1.0, N)
    pre = from_arrays(2, 1.0, W, DJ, F0l, Ores, Nk)
    om = 0.3 + 0.2j
    M64 = pre.ctx.build_M(om, _ffi.PREC_FP64)
    Mb = pre.ctx.build_M(om, _ffi.PREC_BF16X3)
    assert relnorm(Mb, M64) <= 1e-4
    T64 = pre.ctx.build_taylor(om, 3, _ffi.PREC_FP64)
    Tb = pre.ctx.build_taylor(om, 3, _ffi.PREC_BF16X3)
    for j in range(3):
        assert relnorm(Tb[j], T64[j]) <= 1e-4


def test_mixed_newton_small_grid_known_answer(small_cfg_paths):
    """Mixed search on the published grid lands on the FP64 root of the
    reference iteration (convergence_table.txt:13) within 1e-10."""
    cfg = K.load_model(K.load_config(small_cfg_paths[0]), small_cfg_paths[1])
    model = K.create_toomre_zang_model(cfg)
    pre = K.precompute_all(cfg, model, K.GPUType.AMD)
    dev = K.to_device(pre, K.GPUType.AMD, np.float64)
    r = K.newton_search(cfg, pre, dev, K.GPUType.AMD, np.float64, 0.44, 0.125,
                        precision="mixed")
    print("mixed root", r.Omega_p, r.gamma, r.n_lp_builds, r.n_fp64_builds, r.det_val, flush=True)
    assert r.converged and abs(r.det_val) < cfg.newton.tol
    pre_o = oracle.precompute(oracle.load_toml_params(*small_cfg_paths))
    ro = oracle.newton(pre_o, 0.44, 0.125)
    assert abs(complex(2 * r.Omega_p, r.gamma) - ro["omega"]) <= 1e-10
    assert abs(r.Omega_p - 0.4397142267) <= 1e-10
    assert abs(r.gamma - 0.1230663944) <= 1e-10
    assert r.n_fp64_builds <= 2 and r.n_lp_builds <= 4
    assert r.n_det_evals == r.n_fp64_builds + r.n_lp_builds
    assert len(r.history) == r.iterations


@pytest.mark.parametrize("mode", [_ffi.PREC_MIXED, _ffi.PREC_MIXED_FP32])
@pytest.mark.parametrize("start", [(0.45, 0.11), (0.43, 0.14), (0.4397, 0.1231)])
def test_mixed_newton_matches_fp64_search(oracle_small, start, mode):
    pre = device_pre(oracle_small)
    r64, _ = pre.ctx.newton(*start, max_iter=50, tol=1e-12, precision=_ffi.PREC_FP64)
    rmx, H = pre.ctx.newton(*start, max_iter=50, tol=1e-12, precision=mode)
    assert r64.converged and rmx.converged
    w64 = complex(r64.omega[0], r64.omega[1])
    wmx = complex(rmx.omega[0], rmx.omega[1])
    assert abs(wmx - w64) <= 1e-10
    assert rmx.n_fp64_builds < r64.n_fp64_builds
    assert H.shape == (rmx.iterations, 3) and np.all(np.isfinite(H))


def test_mixed_newton_respects_bounds(oracle_small):
    """A start below the gamma bound is lifted to 1e-4 and every move stays in
    bounds, as in the reference iteration."""
    pre = device_pre(oracle_small)
    res, H = pre.ctx.newton(0.44, 1e-6, max_iter=30, tol=1e-12, precision=_ffi.PREC_MIXED)
    assert np.all(H[:, 1] >= 1e-4) and np.all((H[:, 0] >= 0) & (H[:, 0] <= 1))
    assert res.gamma >= 1e-4


# --------------------------------------------------------------------------
# Batched omega-grid scan (64 x 64, Re w/m in [0.30, 0.60], Im w in
# [0.02, 0.25]) in the low precisions against the FP64 scan
# --------------------------------------------------------------------------
@pytest.mark.parametrize("prec,tol", [(_ffi.PREC_FP32, 1e-4), (_ffi.PREC_BF16X3, 2e-3)])
def test_grid_scan_64x64_low_precision(small_cfg_paths, prec, tol):
    from kalnajs_log_spirals_amd.prng import SplitMix64
    cfg = K.load_model(K.load_config(small_cfg_paths[0]), small_cfg_paths[1])
    model = K.create_toomre_zang_model(cfg)
    pre = K.precompute_all(cfg, model, K.GPUType.AMD)
    ctx = pre.ctx
    win = (64, 64, 0.30, 0.60, 0.02, 0.25)
    g64, op64, ga64 = ctx.grid_scan(*win, precision=_ffi.PREC_FP64)
    glp, oplp, galp = ctx.grid_scan(*win, precision=prec)
    assert g64.shape == (64, 64) and np.all(np.isfinite(g64)) and np.all(np.isfinite(glp))
    # error relative to the local |det| scale (the larger of the point and
    # the map's median, so that near-zeros of det do not dominate)
    scale = np.maximum(np.abs(g64), np.median(np.abs(g64)))
    err = np.abs(glp - g64) / scale
    assert np.max(err) <= tol, np.max(err)
    rng = SplitMix64(0)
    idx = [(rng.below(64), rng.below(64)) for _ in range(64)]
    for i, j in idx:
        assert abs(glp[i, j] - g64[i, j]) <= tol * scale[i, j]
    # the same minimum of |det| on the grid
    assert (oplp, galp) == (op64, ga64)


# --------------------------------------------------------------------------
# Uneven batch of Newton searches (draws of the 48-draw batch, grids scaled
# down to test size): the mixed search lands on the FP64 search's root
# --------------------------------------------------------------------------
def test_uneven_batch_mixed_roots_are_fp64_roots():
    """Every converged mixed root is a root of the FP64 det to 1e-10: three
    FP64 Newton steps with tol = 0 from it move it by at most 1e-10.  Where
    the reference FP64 iteration from the same start reaches the same root,
    the two agree to 1e-10 after the same polish (the reference stops at
    |det| < tol, which on a coarse grid, |det'| ~ 1e-8, leaves ~1e-4 of
    omega; Newton basins differ with the path, so the two may also end at
    different roots)."""
    from kalnajs_log_spirals_amd.UnevenBatch import config_for, draw_batch, run_batch
    draws = draw_batch(48, 0)[:4]
    rm = run_batch(draws, precision="mixed", scale=0.15, log=print)
    n_same = 0
    assert sum(a.converged for a in rm) >= 3
    for d, a in zip(draws, rm):
        cfg = config_for(d, scale=0.15)
        pre = K.precompute_all(cfg, K.create_toomre_zang_model(cfg), K.GPUType.AMD)
        wm = complex(d.m * a.Omega_p, a.gamma)
        if a.converged:
            pm, _ = pre.ctx.newton(a.Omega_p, a.gamma, max_iter=3, tol=0.0,
                                   precision=_ffi.PREC_FP64)
            assert abs(complex(pm.omega[0], pm.omega[1]) - wm) <= 1e-10
        r64, _ = pre.ctx.newton(d.Omega_p0, d.gamma0, max_iter=50, tol=1e-12,
                                precision=_ffi.PREC_FP64)
        if a.converged and r64.converged:
            pol, _ = pre.ctx.newton(r64.Omega_p, r64.gamma, max_iter=3, tol=0.0,
                                    precision=_ffi.PREC_FP64)
            wp = complex(pol.omega[0], pol.omega[1])
            print("draw", d.index, "fp64", complex(r64.omega[0], r64.omega[1]), "polished", wp,
                  "mixed", wm, flush=True)
            if abs(wp - wm) < 1e-6:
                n_same += 1
                assert abs(wp - wm) <= 1e-10
                assert a.n_fp64_builds < r64.n_det_evals
        pre.ctx.close()
    assert n_same >= 1


def test_bf16x3_four_taylor_terms(oracle_small):
    """The 4-term BF16x3 pass: T_1..T_3 as the FP64 3-term build and
    T_4 = -dT_3/dw / 3 (d/dw (c/d^j) = -j c/d^(j+1)), the derivative taken by
    central differences of FP64 builds (truncation ~eps^2, far below 1e-4)."""
    pre = device_pre(oracle_small)
    wc, eps = 2 * 0.44 + 0.12j, 1e-4
    Tb = pre.ctx.build_taylor(wc, 4, _ffi.PREC_BF16X3)
    T64 = pre.ctx.build_taylor(wc, 3, _ffi.PREC_FP64)
    for j in range(3):
        assert relnorm(Tb[j], T64[j]) <= 1e-4
    Tp = pre.ctx.build_taylor(wc + eps, 3, _ffi.PREC_FP64)[2]
    Tm = pre.ctx.build_taylor(wc - eps, 3, _ffi.PREC_FP64)[2]
    T4 = -(Tp - Tm) / (2 * eps) / 3
    assert relnorm(Tb[3], T4) <= 1e-4
    with pytest.raises(_ffi.KlspError):
        pre.ctx.build_taylor(wc, 4, _ffi.PREC_FP64)


def test_context_reuse_across_problems():
    """One context given problems of different sizes in turn (as the uneven
    batch does): its W buffers and low-precision copies are reused when the
    new W fits and replaced when it does not, and every det equals the one a
    fresh context gives (same kernels, same data: bit-identical), in FP64,
    FP32 and BF16x3."""
    rng = np.random.default_rng(5)

    def problem(NPh, N, n_l):
        W = (rng.standard_normal((NPh, N, n_l)) + 1j * rng.standard_normal((NPh, N, n_l)))
        W *= 0.5 / np.sqrt(NPh * n_l)
        return dict(NPh=NPh, N=N, n_l=n_l, W=W, DJ=rng.uniform(0.1, 1, NPh),
                    F0l=rng.standard_normal((NPh, n_l)),
                    Ores=rng.uniform(-2, 2, (NPh, n_l)), Nk=rng.uniform(0.5, 1.5, N))

    def load(ctx, q):
        ctx.upload_problem(q["NPh"], q["N"], q["n_l"], 2, 0.7, q["DJ"], q["F0l"], q["Ores"],
                           q["Nk"])
        ctx.upload_W(q["W"])

    oms = [0.3 + 0.4j, 0.9 + 0.1j, 0.5 + 0.25j]
    precs = (_ffi.PREC_FP64, _ffi.PREC_FP32, _ffi.PREC_BF16X3)
    # large, smaller (reuse), larger than both (new buffers), small again
    probs = [problem(700, 96, 5), problem(300, 40, 3), problem(900, 130, 6),
             problem(65, 17, 2)]
    shared = _ffi.Context(0)
    try:
        for q in probs:
            load(shared, q)
            got = {p: shared.det(oms, p) for p in precs}
            fresh = _ffi.Context(0)
            try:
                load(fresh, q)
                for p in precs:
                    ref = fresh.det(oms, p)
                    print("problem", q["NPh"], q["N"], q["n_l"], "precision", p, flush=True)
                    assert np.array_equal(got[p], ref), (p, got[p], ref)
            finally:
                fresh.close()
    finally:
        shared.close()


def test_closed_context_raises(oracle_small):
    """Calls on a closed context, and to_device on data whose context was
    closed, raise KlspError instead of passing freed memory to the library;
    closing twice is harmless and other contexts keep working."""
    pre = device_pre(oracle_small)
    ctx = pre.ctx
    d0 = ctx.det([0.6 + 0.05j], _ffi.PREC_FP64)
    ctx.close()
    assert ctx.closed
    ctx.close()
    for call in (lambda: ctx.det([0.6 + 0.05j], _ffi.PREC_FP64),
                 lambda: ctx.prepare(_ffi.PREC_FP64),
                 lambda: ctx.newton(0.45, 0.11, precision=_ffi.PREC_MIXED),
                 lambda: K.to_device(pre, K.GPUType.AMD, np.float64)):
        with pytest.raises(_ffi.KlspError, match="closed"):
            call()
    again = device_pre(oracle_small)
    try:
        assert np.array_equal(again.ctx.det([0.6 + 0.05j], _ffi.PREC_FP64), d0)
    finally:
        again.ctx.close()


# --------------------------------------------------------------------------
# Non-finite input and the convergence flag of the mixed search
# --------------------------------------------------------------------------
def test_nan_in_W_gives_nan_det_and_no_convergence():
    """A NaN in W reaches row and column a of M, so I - M has a NaN on its
    diagonal: zgetrf's izamax takes it as the pivot and det is NaN (as
    LinearAlgebra/numpy and the oracle give), not 0; neither the FP64 nor the
    mixed Newton search may report convergence there."""
    rng = np.random.default_rng(11)
    NPh, N, n_l = 40, 17, 2
    W = (rng.standard_normal((NPh, N, n_l)) + 1j * rng.standard_normal((NPh, N, n_l))) / 10
    W[5, 3, 1] = complex(np.nan, 0.0)
    DJ, F0l = rng.uniform(0.5, 1.5, NPh), rng.standard_normal((NPh, n_l))
    Ores, Nk = rng.uniform(-1, 1, (NPh, n_l)), rng.uniform(0.5, 1.0, N)
    pre_o = dict(m=2, G=1.0, W=W, DJ=DJ, F0l=F0l, Omega_res=Ores, N_kernel=Nk)
    om = 0.8 + 0.12j
    Mo = oracle.build_M(pre_o, om)
    do = oracle.det_matrix(np.eye(N) - Mo)
    assert np.isnan(do) and np.isnan(np.linalg.det(np.eye(N) - Mo))
    pre = device_pre(pre_o)
    try:
        d = pre.ctx.det([om], _ffi.PREC_FP64)[0]
        print("det with a NaN in W:", d, "oracle:", do, flush=True)
        assert np.isnan(d)
        for prec in (_ffi.PREC_FP64, _ffi.PREC_MIXED):
            res, _ = pre.ctx.newton(0.4, 0.12, max_iter=5, tol=1e-12, precision=prec)
            assert res.converged == 0, prec
    finally:
        pre.ctx.close()


@pytest.mark.parametrize("max_iter", [1, 2, 3, 50])
@pytest.mark.parametrize("start", [(0.30, 0.20), (0.55, 0.05), (0.44, 0.125)])
def test_mixed_newton_converged_flag_is_honest(oracle_small, start, max_iter):
    """converged from the mixed search means the returned omega is an FP64
    root: three FP64 Newton steps with tol = 0 move it by at most 1e-10 and
    its FP64 |det| is below 10 tol (the det's rounding floor there is
    ~1e-13).  Far starts with small iteration budgets
    end between anchors, where the flag must be 0."""
    pre = device_pre(oracle_small)
    try:
        res, _ = pre.ctx.newton(*start, max_iter=max_iter, tol=1e-12, precision=_ffi.PREC_MIXED)
        w = complex(res.omega[0], res.omega[1])
        print("start", start, "max_iter", max_iter, "converged", res.converged, "omega", w,
              flush=True)
        if res.converged:
            d64 = pre.ctx.det([w], _ffi.PREC_FP64)[0]
            pol, _ = pre.ctx.newton(res.Omega_p, res.gamma, max_iter=3, tol=0.0,
                                    precision=_ffi.PREC_FP64)
            assert abs(complex(pol.omega[0], pol.omega[1]) - w) <= 1e-10
            assert abs(d64) < 1e-12 * 10
    finally:
        pre.ctx.close()


# --------------------------------------------------------------------------
# Real basis of a symmetric alpha grid (one-omega FP64 builds): the same M
# as the complex product and the BLAS-form oracle, on every tile shape
# --------------------------------------------------------------------------
def mirror_symmetric_W(rng, NPh, N, n_l, scale=1.0):
    """Random W with W[:, N-1-a] = conj W[:, a] and a real centre column."""
    W = (rng.standard_normal((NPh, N, n_l)) + 1j * rng.standard_normal((NPh, N, n_l))) * scale
    h = N // 2
    for a in range(h):
        W[:, N - 1 - a, :] = np.conj(W[:, a, :])
    if N % 2:
        W[:, h, :] = W[:, h, :].real
    return W


@pytest.mark.parametrize("N", [1, 2, 5, 16, 17, 31, 32, 33, 64, 65, 79, 130, 301])
def test_symmetric_basis_matches_complex_product(N):
    """Symmetric W: the real-basis build equals the complex product and the
    oracle's BLAS-form build to 1e-12 normwise, and the dets agree (N covers
    odd and even grids, tiles cut by the edge of the 16-column blocks, and
    one to five tile rows)."""
    from oracle import blas_form
    rng = np.random.default_rng(1000 + N)
    NPh, n_l = 97, 3
    W = mirror_symmetric_W(rng, NPh, N, n_l, 1.0 / np.sqrt(NPh * n_l))
    DJ, F0l = rng.uniform(0.2, 1.0, NPh), rng.standard_normal((NPh, n_l))
    Ores, Nk = rng.uniform(-2, 2, (NPh, n_l)), rng.uniform(0.5, 1.5, N)
    pre = from_arrays(2, 0.8, W, DJ, F0l, Ores, Nk)
    try:
        assert pre.ctx.symmetric_basis()
        for om in (0.3 + 0.4j, 1.1 - 0.2j):
            Ms = pre.ctx.build_M(om, _ffi.PREC_FP64)
            ds = pre.ctx.det([om], _ffi.PREC_FP64)[0]
            pre.ctx.set_symmetric_basis(False)
            Mc = pre.ctx.build_M(om, _ffi.PREC_FP64)
            dc = pre.ctx.det([om], _ffi.PREC_FP64)[0]
            pre.ctx.set_symmetric_basis(True)
            Mb = blas_form.build_M(W, DJ, F0l, Ores, Nk, 0.8, om)
            assert relnorm(Ms, Mb) <= 1e-12, relnorm(Ms, Mb)
            assert relnorm(Ms, Mc) <= 1e-12
            assert abs(ds - dc) <= 1e-12 * max(abs(dc), 1.0)
    finally:
        pre.ctx.close()


def test_symmetric_basis_detection():
    """An uploaded W is used in the real basis only if it is exactly mirror
    symmetric: one entry off by one ulp, or an imaginary part in the centre
    column, keeps the complex product (which then still matches the oracle)."""
    from oracle import blas_form
    rng = np.random.default_rng(21)
    NPh, N, n_l = 50, 21, 2
    W = mirror_symmetric_W(rng, NPh, N, n_l, 0.1)
    DJ, F0l = rng.uniform(0.2, 1.0, NPh), rng.standard_normal((NPh, n_l))
    Ores, Nk = rng.uniform(-2, 2, (NPh, n_l)), rng.uniform(0.5, 1.5, N)
    om = 0.5 + 0.3j
    for case in ("symmetric", "ulp", "centre"):
        V = W.copy()
        if case == "ulp":
            V[7, 3, 1] = complex(np.nextafter(V[7, 3, 1].real, 1.0), V[7, 3, 1].imag)
        elif case == "centre":
            V[11, N // 2, 0] += 1e-300j
        pre = from_arrays(2, 1.0, V, DJ, F0l, Ores, Nk)
        try:
            assert pre.ctx.symmetric_basis() == (case == "symmetric"), case
            M = pre.ctx.build_M(om, _ffi.PREC_FP64)
            assert relnorm(M, blas_form.build_M(V, DJ, F0l, Ores, Nk, 1.0, om)) <= 1e-12
        finally:
            pre.ctx.close()


def test_quadrature_W_is_symmetric_on_the_reference_grid(oracle_tiny):
    """W from the device quadrature on range(-alpha_max, alpha_max, N) is
    mirror symmetric with a real alpha = 0 column (as the reference's
    e^(-i 0 ln r) = 1 makes it), so FP64 builds take the real basis."""
    p = oracle_tiny["params"]
    cfg = cfg_from_params(p)
    pre = K.precompute_all(cfg, K.create_toomre_zang_model(cfg), K.GPUType.AMD)
    try:
        W = pre.W_l_mat
        N = W.shape[1]
        assert np.array_equal(W[:, ::-1, :], np.conj(W))
        if N % 2:
            assert np.all(W[:, N // 2, :].imag == 0)
        assert pre.ctx.symmetric_basis()
        om = 0.8 + 0.1j
        Mo = oracle.build_M(oracle_tiny, om)
        assert relnorm(pre.ctx.build_M(om, _ffi.PREC_FP64), Mo) <= 1e-12
    finally:
        pre.ctx.close()


@pytest.mark.parametrize("N", [33, 130, 301])
def test_symmetric_basis_every_variant(N):
    """Every compiled variant of the real-basis kernel (prefetch depth,
    diagonal-tile form) gives the oracle's M to 1e-12; the autotuner times
    them all and keeps one of them."""
    from oracle import blas_form
    rng = np.random.default_rng(2000 + N)
    NPh, n_l = 211, 2
    W = mirror_symmetric_W(rng, NPh, N, n_l, 1.0 / np.sqrt(NPh * n_l))
    DJ, F0l = rng.uniform(0.2, 1.0, NPh), rng.standard_normal((NPh, n_l))
    Ores, Nk = rng.uniform(-2, 2, (NPh, n_l)), rng.uniform(0.5, 1.5, N)
    pre = from_arrays(2, 0.8, W, DJ, F0l, Ores, Nk)
    try:
        om = 0.45 + 0.15j
        Mb = blas_form.build_M(W, DJ, F0l, Ores, Nk, 0.8, om)
        times, chosen = pre.ctx.autotune(_ffi.PREC_FP64, reps=1)
        print("variants", times, "chosen", chosen, flush=True)
        assert len(times) >= 2 and chosen in times
        assert relnorm(pre.ctx.build_M(om, _ffi.PREC_FP64), Mb) <= 1e-12
        for v in range(len(times)):
            pre.ctx.set_variant(v)
            assert relnorm(pre.ctx.build_M(om, _ffi.PREC_FP64), Mb) <= 1e-12, v
        with pytest.raises(_ffi.KlspError):
            pre.ctx.set_variant(len(times))
    finally:
        pre.ctx.close()


@pytest.mark.parametrize("N", [5, 17, 33, 130, 301, 330, 401])
def test_symmetric_basis_bf16x3(N):
    """BF16x3 builds in the real basis (folded column pairs; one, two or
    several workgroups per pair as N grows): M within 1e-4 normwise of the
    FP64 build, the three- and four-term Taylor passes within 1e-4 of the FP64
    terms (T_4 by the single hi x hi product: 1e-2), and the same det as the
    complex BF16x3 kernel to the precision of either."""
    rng = np.random.default_rng(3000 + N)
    NPh, n_l = 157, 2
    W = mirror_symmetric_W(rng, NPh, N, n_l, 1.0 / np.sqrt(NPh * n_l))
    DJ, F0l = rng.uniform(0.2, 1.0, NPh), rng.standard_normal((NPh, n_l))
    Ores, Nk = rng.uniform(-2, 2, (NPh, n_l)), rng.uniform(0.5, 1.5, N)
    pre = from_arrays(2, 0.8, W, DJ, F0l, Ores, Nk)
    try:
        assert pre.ctx.symmetric_basis()
        om = 0.4 + 0.25j
        M64 = pre.ctx.build_M(om, _ffi.PREC_FP64)
        Mb = pre.ctx.build_M(om, _ffi.PREC_BF16X3)
        assert relnorm(Mb, M64) <= 1e-4, relnorm(Mb, M64)
        T64 = pre.ctx.build_taylor(om, 3, _ffi.PREC_FP64)
        for nt in (2, 3):
            Tb = pre.ctx.build_taylor(om, nt, _ffi.PREC_BF16X3)
            for j in range(nt):
                assert relnorm(Tb[j], T64[j]) <= 1e-4, (nt, j)
        oms = [om, 0.9 + 0.1j, 0.2 + 0.3j]
        db = pre.ctx.det(oms, _ffi.PREC_BF16X3)
        d64 = pre.ctx.det(oms, _ffi.PREC_FP64)
        pre.ctx.set_symmetric_basis(False)
        dc = pre.ctx.det(oms, _ffi.PREC_BF16X3)
        pre.ctx.set_symmetric_basis(True)
        for k in range(3):
            sc = max(abs(d64[k]), 1.0)
            assert abs(db[k] - d64[k]) <= 1e-3 * sc
            assert abs(dc[k] - d64[k]) <= 1e-3 * sc
    finally:
        pre.ctx.close()


@pytest.mark.parametrize("N", [1, 2, 5, 16, 17, 31, 33, 64, 65, 130, 301])
def test_symmetric_basis_fp32(N):
    """FP32 builds in the real basis (4 x 2 tiles over the upper triangle,
    K-tile sums in FP64): M within 2e-6 normwise of the FP64 build, as the
    complex FP32 build is; the three Taylor terms of an FP32 pass alike; dets
    within 1e-5 of FP64; one launch per weight set in three-omega batches."""
    rng = np.random.default_rng(5000 + N)
    NPh, n_l = 131, 3
    W = mirror_symmetric_W(rng, NPh, N, n_l, 1.0 / np.sqrt(NPh * n_l))
    DJ, F0l = rng.uniform(0.2, 1.0, NPh), rng.standard_normal((NPh, n_l))
    Ores, Nk = rng.uniform(-2, 2, (NPh, n_l)), rng.uniform(0.5, 1.5, N)
    pre = from_arrays(2, 0.8, W, DJ, F0l, Ores, Nk)
    try:
        om = 0.4 + 0.25j
        M64 = pre.ctx.build_M(om, _ffi.PREC_FP64)
        Mf = pre.ctx.build_M(om, _ffi.PREC_FP32)
        assert relnorm(Mf, M64) <= 2e-6, relnorm(Mf, M64)
        T64 = pre.ctx.build_taylor(om, 3, _ffi.PREC_FP64)
        Tf = pre.ctx.build_taylor(om, 3, _ffi.PREC_FP32)
        for j in range(3):
            assert relnorm(Tf[j], T64[j]) <= 2e-6, (j, relnorm(Tf[j], T64[j]))
        oms = [om, 0.9 + 0.1j, 0.2 + 0.3j, 0.6 + 0.2j]
        df = pre.ctx.det(oms, _ffi.PREC_FP32)
        d64 = pre.ctx.det(oms, _ffi.PREC_FP64)
        for a, b in zip(df, d64):
            assert abs(a - b) <= 1e-5 * max(abs(b), 1e-300), (a, b)
        pre.ctx.set_symmetric_basis(False)
        Mc = pre.ctx.build_M(om, _ffi.PREC_FP32)
        pre.ctx.set_symmetric_basis(True)
        assert relnorm(Mc, M64) <= 2e-6
    finally:
        pre.ctx.close()
