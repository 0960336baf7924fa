"""The linear generalized eigenproblem omega D C = (E + B) C
(Matlab/Matrix_LogExp.m:373-530) with its blocks D_l, E_l, F_l built on the
device (klsp_build_pencil), against the oracle's restatement."""
import numpy as np
import pytest

import oracle
from oracle import linear_form as lf
import kalnajs_log_spirals_amd as K
from kalnajs_log_spirals_amd.BasisFunctions import from_arrays

pytestmark = pytest.mark.gpu


def relnorm(a, b) -> float:
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


@pytest.mark.parametrize("NPh,N,n_l", [(37, 5, 3), (2121, 17, 4), (211, 37, 2), (8, 8, 5)])
def test_pencil_blocks_match_oracle(NPh, N, n_l):
    """Slabs of any length (NPh not a multiple of the kernel's 8-row steps;
    the rows past a slab meet zero weights) and any N."""
    rng = np.random.default_rng(NPh + 100 * N)
    W = (rng.standard_normal((NPh, N, n_l)) + 1j * rng.standard_normal((NPh, N, n_l))) / np.sqrt(NPh)
    DJ, F0l = rng.uniform(0.2, 1.0, NPh), rng.standard_normal((NPh, n_l))
    Ores, Nk = rng.uniform(-2, 2, (NPh, n_l)), rng.uniform(0.5, 1.5, N)
    pre = from_arrays(2, 0.7, W, DJ, F0l, Ores, Nk)
    try:
        b = K.pencil_blocks(pre)
        Do, Eo, Fo = lf.blocks(W, DJ, F0l, Ores)
        for got, ref in ((b.D, Do), (b.E, Eo), (b.F, Fo)):
            for il in range(n_l):
                assert relnorm(got[:, :, il], ref[:, :, il]) <= 1e-12, il
    finally:
        pre.ctx.close()


def test_linear_form_on_square_problem_matches_oracle_and_det_zeros():
    """Square invertible W_l: the product's assembly and solve on the device
    blocks give the oracle's eigenvalues, and each is a zero of the device's
    det(I - M)."""
    rng = np.random.default_rng(9)
    N, n_l = 6, 3
    W = (rng.standard_normal((N, N, n_l)) + 1j * rng.standard_normal((N, N, n_l))) / np.sqrt(N)
    DJ, F0l = rng.uniform(0.3, 1.2, N), rng.standard_normal((N, n_l))
    Ores, Nk = rng.uniform(-2, 2, (N, n_l)), rng.uniform(0.5, 1.5, N)
    pre = from_arrays(2, 0.8, W, DJ, F0l, Ores, Nk)
    try:
        top, w, n_inf = K.linear_form_modes(pre)
        wo, n_info = lf.solve(*lf.assemble(*lf.blocks(W, DJ, F0l, Ores), Nk, 0.8))
        assert n_inf == n_info == 0 and w.size == wo.size == N * n_l
        dist = np.abs(w[:, None] - wo[None, :])
        assert dist.min(axis=1).max() <= 1e-9 and dist.min(axis=0).max() <= 1e-9
        d0 = np.abs(pre.ctx.det(w))
        d1 = np.abs(pre.ctx.det(w + 1e-3))
        assert np.all(d0 <= 1e-8 * d1), (d0, d1)
        assert len(top) == 10 and top[0][1] == pytest.approx(w[0].imag)
    finally:
        pre.ctx.close()


def test_linear_form_published_small_grid(small_cfg_paths, oracle_small):
    """configs/small.toml + models/kalnajs_zang4.toml (convergence_table.txt:13):
    the device blocks equal the oracle's; the pencil is solved (QZ on the host,
    order 2121) and its behaviour printed: D is singular (W_l are numerically
    rank-deficient), so there are infinite eigenvalues and finite ones that
    are not modes; the distance of the nearest eigenvalue to the det root
    0.4397142267 / 0.1230663944 is reported."""
    cfg = K.load_model(K.load_config(small_cfg_paths[0]), small_cfg_paths[1])
    pre = K.precompute_all(cfg, K.create_toomre_zang_model(cfg), K.GPUType.AMD)
    try:
        b = K.pencil_blocks(pre)
        Do, Eo, Fo = lf.blocks(oracle_small["W"], oracle_small["DJ"], oracle_small["F0l"],
                               oracle_small["Omega_res"])
        for got, ref in ((b.D, Do), (b.E, Eo), (b.F, Fo)):
            assert relnorm(got, ref) <= 1e-12
        A, Dg = K.assemble_pencil(b, pre.N_kernel, pre.G)
        l0 = b.D.shape[2] // 2                      # l = 0 (l = -l_max .. l_max)
        ev = np.linalg.eigvalsh(0.5 * (b.D[:, :, l0] + b.D[:, :, l0].conj().T))
        w, n_inf = K.solve_pencil(A, Dg)
        root = complex(2 * 0.4397142267, 0.1230663944)
        near = w[np.argmin(np.abs(w - root))]
        print(f"pencil order {A.shape[0]}: {w.size} finite eigenvalues, {n_inf} infinite; "
              f"D_l (l = 0) eigenvalues {ev.min():.3e} .. {ev.max():.3e}; largest |omega| "
              f"{np.max(np.abs(w)):.3e}; top growth rates (Omega_p, gamma): "
              f"{[(round(x.real / 2, 6), round(x.imag, 6)) for x in w[:5]]}; nearest to the det "
              f"root: Omega_p={near.real / 2:.6f} gamma={near.imag:.6f} "
              f"(|d omega| {abs(near - root):.3e})", flush=True)
        assert w.size + n_inf == A.shape[0]
        assert np.all(np.isfinite(w))
    finally:
        pre.ctx.close()
