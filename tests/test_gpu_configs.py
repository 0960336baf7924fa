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
