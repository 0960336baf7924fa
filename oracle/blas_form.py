"""BLAS form of det(I - M(omega)) on numpy/OpenBLAS -- TEST INFRASTRUCTURE.

Restates NL_julia_direct.jl:299-310 (the clearest statement of the
reference's BLAS-form det): per radial harmonic l one ZGEMM
M += W_l^T (c_l .* conj(W_l)), then M *= G, M[:, a] *= N[a] and
det(I - M) by LAPACK (zgetrf).  Used (a) as an independent cross-check of
the C oracle's triple-loop build and (b) as the timed CPU baseline.
"""
from __future__ import annotations

import numpy as np


def weights(omega: complex, DJ: np.ndarray, F0l_l: np.ndarray, Ores_l: np.ndarray) -> np.ndarray:
    """c = DJ*F0l / (omega - Omega_res) with the 1e-12 clamp (MatrixBuilder.jl:43-51)."""
    d = omega - Ores_l.astype(np.complex128)
    d[np.abs(d) < 1e-12] = 1e-12 + 1e-12j
    return (DJ * F0l_l) / d


def build_M(W: np.ndarray, DJ, F0l, Ores, Nk, G: float, omega: complex) -> np.ndarray:
    """W [NPh, N, n_l] complex128; returns M [N, N]."""
    NPh, N, n_l = W.shape
    M = np.zeros((N, N), dtype=np.complex128)
    for il in range(n_l):
        Wl = W[:, :, il]
        c = weights(omega, DJ, F0l[:, il], Ores[:, il])
        M += Wl.T @ (c[:, None] * Wl.conj())
    M *= G
    M *= np.asarray(Nk)[None, :]
    return M


def det(W, DJ, F0l, Ores, Nk, G, omega) -> complex:
    M = build_M(W, DJ, F0l, Ores, Nk, G, omega)
    return complex(np.linalg.det(np.eye(M.shape[0]) - M))


def time_sample(NPh: int, N: int, n_l: int, n_slabs: int, n_distinct: int = 8, seed: int = 0):
    """Time `n_slabs` per-l ZGEMM accumulations on synthetic [NPh, N] l-slabs
    (cycling over `n_distinct` buffers, each larger than the host caches)
    plus the median of 5 LUs of N x N; return (seconds per full det scaled
    to n_l slabs, seconds per slab, seconds per LU)."""
    import time

    rng = np.random.default_rng(seed)
    nd = max(1, min(n_distinct, n_slabs))
    slabs = [(rng.standard_normal((NPh, N)) + 1j * rng.standard_normal((NPh, N)))
             for _ in range(nd)]
    DJ = rng.uniform(0.1, 1.0, NPh)
    F0l = rng.standard_normal(NPh)
    Ores = rng.uniform(-3, 3, NPh)
    M = np.zeros((N, N), dtype=np.complex128)
    t0 = time.perf_counter()
    for i in range(n_slabs):
        Wl = slabs[i % nd]
        c = weights(0.88 + 0.12j, DJ, F0l, Ores)
        M += Wl.T @ (c[:, None] * Wl.conj())
    per_slab = (time.perf_counter() - t0) / n_slabs
    A = np.eye(N) - 1e-3 * M / max(np.abs(M).max(), 1.0)
    lus = []
    for _ in range(5):
        t0 = time.perf_counter()
        np.linalg.det(A)
        lus.append(time.perf_counter() - t0)
    lu = float(np.median(lus))
    return per_slab * n_l + lu, per_slab, lu
