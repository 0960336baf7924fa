"""The small plain-FP64 case end to end on the device against the oracle:
Toomre-Zang n_zang = 4, q1 = 7, m = 2, G = 1; NR = 101, Ne = 21,
N_alpha = 101 on alpha in [-30, 30], l = -25..25 (nw = 100001, nwa = 2001);
W precompute, one det at omega = 2 * 0.4397 + 0.1231i, then Newton
(delta 1e-6, tol 1e-12, at most 50 iterations).

The oracle's direct W sum over all 2121 orbits would take minutes, so W is
checked on 24 seeded orbits (normwise relative <= 1e-12 per orbit); the
det and the Newton iteration are checked against the oracle's BLAS-form
det (numpy ZGEMM per l + LAPACK) and the oracle's Newton driver built on the
device's W (det relative <= 1e-11 of the det scale, root |d omega| <= 1e-10,
the same iteration count).
"""
import os

import numpy as np
import pytest

import oracle
from oracle import blas_form
import kalnajs_log_spirals_amd as K
from kalnajs_log_spirals_amd import _ffi
from kalnajs_log_spirals_amd.prng import SplitMix64

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "kalnajs_log_spirals_amd")


def small_fp64_config():
    cfg = K.load_model(K.load_config(os.path.join(PKG, "configs", "default.toml")),
                       os.path.join(PKG, "models", "kalnajs_zang4.toml"))
    g = cfg.grid
    g.NR, g.Ne, g.N_alpha, g.alpha_max = 101, 21, 101, 30.0
    g.l_min, g.l_max = -25, 25
    return cfg


def oracle_params(cfg) -> oracle.Params:
    g = cfg.grid
    return oracle.Params(m=cfg.physics.m, G=cfg.physics.G, L0=cfg.model.L0,
                         n_zang=cfg.model.n_zang, q1=cfg.model.q1, NR=g.NR, Ne=g.Ne,
                         Rc_min=g.Rc_min, Rc_max=g.Rc_max, v_min=g.v_min, N_alpha=g.N_alpha,
                         alpha_max=g.alpha_max, l_min=g.l_min, l_max=g.l_max, nw=g.nw, nwa=g.nwa)


def test_small_fp64_case_end_to_end():
    cfg = small_fp64_config()
    p = oracle_params(cfg)
    print("grid", p, flush=True)
    po = oracle.precompute(p, with_W=False)
    pre = K.precompute_all(cfg, K.create_toomre_zang_model(cfg), K.GPUType.AMD)
    try:
        assert (pre.NPh, pre.N_alpha, pre.n_l) == (2121, 101, 51)
        np.testing.assert_allclose(pre.DJ_vec, po["DJ"], rtol=1e-13, atol=0)
        np.testing.assert_allclose(pre.F0l_all, po["F0l"], rtol=1e-13, atol=1e-300)
        np.testing.assert_allclose(pre.N_kernel, po["N_kernel"], rtol=1e-14)

        W = pre.W_l_mat
        # W on 24 seeded orbits against the oracle's direct sum
        rng = SplitMix64(1)
        idx = sorted({rng.below(p.NPh) for _ in range(24)})
        o = po["orbits"]
        sub = {k: np.ascontiguousarray(o[k][idx]) for k in ("w1", "ra", "pha", "Om1", "Om2")}
        ps = oracle.Params(**{**p.__dict__, "NR": len(idx), "Ne": 1})
        Wo = oracle.compute_W(ps, sub, po["alpha"], po["l_arr"])
        for j, k in enumerate(idx):
            ref = Wo[j]
            scale = np.max(np.abs(ref))
            if scale == 0.0:
                assert np.max(np.abs(W[k])) == 0.0
                continue
            err = float(np.max(np.abs(W[k] - ref)) / scale)
            assert err <= 1e-12, (k, err)

        dev = K.to_device(pre, K.GPUType.AMD, np.float64)
        args = (W, po["DJ"], po["F0l"], po["Omega_res"], po["N_kernel"], p.G)
        om0 = complex(2 * 0.4397, 0.1231)
        d = K.compute_determinant(om0, pre, dev, K.GPUType.AMD, np.float64)
        ref = blas_form.det(*args, om0)
        scale = max(abs(ref), abs(blas_form.det(*args, om0 + 1e-3)))
        print(f"det {d} oracle {ref} |d|/scale {abs(d - ref) / scale:.2e}", flush=True)
        assert abs(d - ref) <= 1e-11 * scale

        r = K.newton_search(cfg, pre, dev, K.GPUType.AMD, np.float64, 0.4397, 0.1231)
        ro = oracle.newton_fn(lambda w: blas_form.det(*args, w), p.m, 0.4397, 0.1231)
        w_dev = complex(p.m * r.Omega_p, r.gamma)
        w_ora = complex(p.m * ro["Omega_p"], ro["gamma"])
        print("root", r.Omega_p, r.gamma, r.iterations, "oracle", ro["Omega_p"], ro["gamma"],
              ro["iterations"], flush=True)
        assert r.converged and ro["converged"]
        assert r.iterations == ro["iterations"]
        assert abs(w_dev - w_ora) <= 1e-10
    finally:
        pre.ctx.close()


# --------------------------------------------------------------------------
# The default grid (NR = 501, Ne = 51, N_alpha = 301, alpha in [-30, 30],
# l = -100..100, K = 5 135 751): the drop-in FP64 path, the mixed-precision
# search, the 64 x 64 omega-grid scan and the 48-draw uneven batch at their
# stated sizes.  The oracle's direct W sum is checked on seeded orbits, the
# det against the oracle's BLAS-form det on the device's W (numpy ZGEMMs per
# l + LAPACK, ~10 s of host BLAS), and everything else through properties
# that hold at any size: H(conj w) = H(w)^H, det(conj w) = conj det(w), and
# roots that FP64 Newton steps with tol = 0 no longer move.
# --------------------------------------------------------------------------
OMEGA0 = (0.4397, 0.1231)      # the runner's start, models/kalnajs_zang4.toml:20-21


def default_config():
    return K.load_model(K.load_config(os.path.join(PKG, "configs", "default.toml")),
                        os.path.join(PKG, "models", "kalnajs_zang4.toml"))


@pytest.fixture(scope="module")
def default_grid():
    cfg = default_config()
    g = cfg.grid
    assert (g.NR, g.Ne, g.N_alpha, g.alpha_max, g.l_min, g.l_max) == (501, 51, 301, 30.0, -100, 100)
    pre = K.precompute_all(cfg, K.create_toomre_zang_model(cfg), K.GPUType.AMD)
    dev = K.to_device(pre, K.GPUType.AMD, np.float64)
    try:
        yield cfg, pre, dev
    finally:
        pre.ctx.close()


@pytest.fixture(scope="module")
def default_fp64_root(default_grid):
    cfg, pre, dev = default_grid
    r = K.newton_search(cfg, pre, dev, K.GPUType.AMD, np.float64, *OMEGA0)
    print("FP64 root", r.Omega_p, r.gamma, r.iterations, r.n_det_evals, abs(r.det_val), flush=True)
    return r


def test_config2_default_grid_W_and_det_vs_oracle(default_grid):
    """W on 24 seeded orbits vs the oracle's direct sum (<= 1e-12 per orbit),
    the omega-independent arrays vs the oracle, and the det at omega_0 vs the
    oracle's BLAS-form det on the device's W (<= 1e-11 of the det scale)."""
    cfg, pre, dev = default_grid
    p = oracle_params(cfg)
    po = oracle.precompute(p, with_W=False)
    np.testing.assert_allclose(pre.DJ_vec, po["DJ"], rtol=1e-13, atol=0)
    # F0l = (l Omega_1 + m Omega_2) F_E + m F_L cancels near one l per orbit
    # (a few entries in 5e6 are 1e2 below their row): rounding is relative to
    # the row's largest term
    row = np.max(np.abs(po["F0l"]), axis=1, keepdims=True)
    assert np.all(np.abs(pre.F0l_all - po["F0l"]) <= 1e-13 * np.abs(po["F0l"]) + 1e-15 * row)
    np.testing.assert_allclose(pre.N_kernel, po["N_kernel"], rtol=1e-14)
    W = pre.ctx.download_W()
    try:
        rng = SplitMix64(2)
        idx = sorted({rng.below(p.NPh) for _ in range(24)})
        o = po["orbits"]
        sub = {k: np.ascontiguousarray(o[k][idx]) for k in ("w1", "ra", "pha", "Om1", "Om2")}
        ps = oracle.Params(**{**p.__dict__, "NR": len(idx), "Ne": 1})
        Wo = oracle.compute_W(ps, sub, po["alpha"], po["l_arr"])
        for j, k in enumerate(idx):
            ref = Wo[j]
            scale = np.max(np.abs(ref))
            if scale == 0.0:
                assert np.max(np.abs(W[k])) == 0.0
                continue
            err = float(np.max(np.abs(W[k] - ref)) / scale)
            assert err <= 1e-12, (k, err)
        args = (W, po["DJ"], po["F0l"], po["Omega_res"], po["N_kernel"], p.G)
        om0 = complex(2 * OMEGA0[0], OMEGA0[1])
        d = dev.ctx.det([om0, om0 + 1e-3], _ffi.PREC_FP64)
        ref = blas_form.det(*args, om0)
        scale = max(abs(ref), abs(d[1]))
        print(f"default grid det {d[0]} BLAS-form {ref} |d|/scale {abs(d[0] - ref) / scale:.2e}",
              flush=True)
        assert abs(d[0] - ref) <= 1e-11 * scale
    finally:
        del W


def test_config2_default_grid_conjugate_symmetry(default_grid):
    """H = M diag(N)^-1 / G satisfies H(conj w) = H(w)^H (W c conj W with
    c(conj w) = conj c(w) while the resonance clamp is inactive), so
    det(I - M(conj w)) = conj det(I - M(w))."""
    cfg, pre, dev = default_grid
    Nk, G = pre.N_kernel, cfg.physics.G
    for om in (complex(2 * OMEGA0[0], OMEGA0[1]), 0.7 + 0.05j):
        M1 = dev.ctx.build_M(om, _ffi.PREC_FP64)
        M2 = dev.ctx.build_M(np.conj(om), _ffi.PREC_FP64)
        H1, H2 = M1 / (G * Nk[None, :]), M2 / (G * Nk[None, :])
        assert np.max(np.abs(H2 - H1.conj().T)) <= 1e-12 * np.max(np.abs(H1))
        d1, d2 = dev.ctx.det([om, np.conj(om)], _ffi.PREC_FP64)
        scale = max(abs(d1), abs(dev.ctx.det([om + 1e-3], _ffi.PREC_FP64)[0]))
        assert abs(d2 - np.conj(d1)) <= 1e-11 * scale


def test_config2_default_grid_fp64_newton(default_grid, default_fp64_root):
    """The reference iteration (delta 1e-6, tol 1e-12, <= 50 iterations) from
    omega_0 converges, and two further FP64 steps with tol = 0 move the root
    by at most 1e-10."""
    cfg, pre, dev = default_grid
    r = default_fp64_root
    assert r.converged and r.iterations < cfg.newton.max_iter
    pol, _ = dev.ctx.newton(r.Omega_p, r.gamma, max_iter=2, tol=0.0, precision=_ffi.PREC_FP64)
    move = abs(complex(pol.omega[0], pol.omega[1]) - r.omega)
    print(f"FP64 root polish move {move:.2e}", flush=True)
    assert move <= 1e-10


def test_config3_default_grid_mixed_root(default_grid, default_fp64_root):
    """The mixed-precision search (BF16x3 Taylor passes, FP64 anchored finish)
    lands within 1e-10 of the FP64 root with one BF16x3 pass and one FP64
    build (the pass's zero lies 5e-7 off, inside the first anchor's bound)."""
    cfg, pre, dev = default_grid
    rm = K.newton_search(cfg, pre, dev, K.GPUType.AMD, np.float64, *OMEGA0, precision="mixed")
    dw = abs(rm.omega - default_fp64_root.omega)
    print(f"mixed root {rm.Omega_p} {rm.gamma}: {rm.n_lp_builds} lp + {rm.n_fp64_builds} FP64 "
          f"builds, |d omega| {dw:.2e}", flush=True)
    assert rm.converged
    assert dw <= 1e-10
    assert rm.n_fp64_builds == 1 and rm.n_lp_builds == 1


# starts (Omega_p, gamma) from which the mixed search reaches roots with
# gamma ~ 0.015-0.017; the first is a seeded random start whose root a bound
# on Frobenius norms of the Taylor terms once accepted 1.2e-10 off
SMALL_GAMMA_STARTS = [(0.38502504623867634, 0.12373476766636918), (0.3978, 0.0999),
                      (0.3229, 0.1477), (0.3621, 0.2108)]


@pytest.mark.parametrize("start", SMALL_GAMMA_STARTS)
def test_config3_default_grid_mixed_small_gamma_roots(default_grid, start):
    """Near roots with small gamma, |v^H T_3 u| / |v^H T_2 u| (u, v the null
    vectors of I - M) is ~10 times what the term norms suggest; the mixed
    search's roots there are still FP64 roots: two FP64 Newton steps with
    tol = 0 move them by at most 1e-10."""
    cfg, pre, dev = default_grid
    rm, _ = dev.ctx.newton(*start, max_iter=50, tol=1e-12, precision=_ffi.PREC_MIXED)
    assert rm.converged and rm.gamma < 0.02
    pol, _ = dev.ctx.newton(rm.Omega_p, rm.gamma, max_iter=2, tol=0.0, precision=_ffi.PREC_FP64)
    move = abs(complex(pol.omega[0], pol.omega[1]) - complex(rm.omega[0], rm.omega[1]))
    print(f"start {start}: root ({rm.Omega_p:.10f}, {rm.gamma:.10f}), {rm.n_lp_builds} low + "
          f"{rm.n_fp64_builds} FP64 builds, polish move {move:.2e}", flush=True)
    assert move <= 1e-10


def test_config4_default_grid_64x64_scan(default_grid):
    """The 64 x 64 window (Re w/m in linspace(0.30, 0.60, 64), Im w in
    linspace(0.02, 0.25, 64)) in BF16x3 against FP64 dets at 64 SplitMix64
    points (<= 1e-3 of the local |det| scale), and the same argmin: the BF16x3
    argmin cell has the smallest FP64 |det| of its 5 x 5 neighbourhood and of
    the 64 sampled points."""
    cfg, pre, dev = default_grid
    ctx = dev.ctx
    grid, Op0, g0 = ctx.grid_scan(64, 64, 0.30, 0.60, 0.02, 0.25, precision=_ffi.PREC_BF16X3)
    assert grid.shape == (64, 64) and np.all(np.isfinite(grid))
    Ops, gams = np.linspace(0.30, 0.60, 64), np.linspace(0.02, 0.25, 64)
    rng = SplitMix64(0)
    pts = [(rng.below(64), rng.below(64)) for _ in range(64)]
    ref = ctx.det([complex(2 * Ops[j], gams[i]) for i, j in pts], _ffi.PREC_FP64)
    med = float(np.median(np.abs(ref)))
    errs = [abs(grid[i, j] - ref[k]) / max(abs(ref[k]), med) for k, (i, j) in enumerate(pts)]
    print(f"scan: max error at 64 points {max(errs):.2e}, argmin Omega_p={Op0} gamma={g0}",
          flush=True)
    assert max(errs) <= 1e-3
    ia, ja = int(np.argmin(np.abs(gams - g0))), int(np.argmin(np.abs(Ops - Op0)))
    assert abs(Ops[ja] - Op0) <= 1e-14 and abs(gams[ia] - g0) <= 1e-14
    nb = [(i, j) for i in range(max(0, ia - 2), min(64, ia + 3))
          for j in range(max(0, ja - 2), min(64, ja + 3))]
    dn = np.abs(ctx.det([complex(2 * Ops[j], gams[i]) for i, j in nb], _ffi.PREC_FP64))
    d_arg = dn[nb.index((ia, ja))]
    assert d_arg == dn.min()
    assert d_arg <= np.abs(ref).min()


def test_config5_uneven_batch_full_size():
    """All 48 draws of the uneven batch (seed 0) at full size in mixed
    precision: every draw converges, and two FP64 Newton steps with tol = 0
    move each root by at most 1e-10."""
    from kalnajs_log_spirals_amd.UnevenBatch import draw_batch, run_batch
    draws = draw_batch(48, 0)
    rm = run_batch(draws, precision="mixed", polish=2, orbits="device", log=print)
    bad = [(r.draw.index, r.converged, r.polish_move) for r in rm
           if not (r.converged and r.polish_move <= 1e-10)]
    print(f"uneven batch: {sum(r.converged for r in rm)}/48 converged, largest polish move "
          f"{max(r.polish_move for r in rm if r.converged):.2e}", flush=True)
    assert not bad, bad
