"""Linear generalized eigenproblem omega D C = (E + B) C -- TEST INFRASTRUCTURE.

Restates Matlab/Matrix_LogExp.m:373-530 on numpy/LAPACK in this build's
discretisation (W, DJ, F0l, Omega_res and N_kernel as the src/ path defines
them; N_kernel already carries the alpha weights, as the script's
N_kernel .* S_alpha does; WNORM = 1 as set at :30):
  D_l = W_l^T diag(DJ) conj W_l, E_l = W_l^T diag(DJ Omega_res_l) conj W_l,
  F_l = W_l^T diag(DJ F0l_l) conj W_l                                (:385-403)
  B[l, l'] = G F_l diag(N_kernel) D_l'                               (:414-430)
  D, E block diagonal in l                                           (:434-449)
  eig(E + B, D), eigenvalues sorted by imaginary part, descending    (:493-512)
"""
from __future__ import annotations

import numpy as np


def blocks(W, DJ, F0l, Ores):
    """D_l, E_l, F_l, each [N, N, n_l] complex (Matrix_LogExp.m:385-403)."""
    NPh, N, n_l = W.shape
    D = np.empty((N, N, n_l), dtype=np.complex128)
    E = np.empty_like(D)
    F = np.empty_like(D)
    for il in range(n_l):
        Wl = W[:, :, il]
        Wc = Wl.conj()
        D[:, :, il] = Wl.T @ (DJ[:, None] * Wc)
        E[:, :, il] = Wl.T @ ((DJ * Ores[:, il])[:, None] * Wc)
        F[:, :, il] = Wl.T @ ((DJ * F0l[:, il])[:, None] * Wc)
    return D, E, F


def assemble(D, E, F, Nk, G):
    """(E_global + B, D_global), each [N n_l, N n_l] (Matrix_LogExp.m:414-449)."""
    N, _, n_l = D.shape
    nt = N * n_l
    A = np.zeros((nt, nt), dtype=np.complex128)
    Dg = np.zeros((nt, nt), dtype=np.complex128)
    for i in range(n_l):
        FN = G * F[:, :, i] * np.asarray(Nk)[None, :]
        for j in range(n_l):
            A[i * N:(i + 1) * N, j * N:(j + 1) * N] = FN @ D[:, :, j]
        A[i * N:(i + 1) * N, i * N:(i + 1) * N] += E[:, :, i]
        Dg[i * N:(i + 1) * N, i * N:(i + 1) * N] = D[:, :, i]
    return A, Dg


def solve(A, Dg):
    """Finite eigenvalues of A C = omega Dg C, sorted by imaginary part
    (descending), and the number of infinite ones (Matrix_LogExp.m:499-506)."""
    import scipy.linalg
    w = scipy.linalg.eigvals(A, Dg)
    fin = w[np.isfinite(w)]
    return fin[np.argsort(-fin.imag, kind="stable")], int(w.size - fin.size)
