"""LinearForm -- the linear generalized eigenproblem omega D C = (E + B) C
(Matlab/Matrix_LogExp.m:373-530; the paper's Eqs. 21-31).

The per-l blocks D_l, E_l, F_l = W_l^T diag(.) conj W_l are the same
contraction as M(omega) with omega-independent weights (DJ, DJ Omega_res,
DJ F0l) and are built on the device (klsp_build_pencil).  B[l, l'] =
G F_l diag(N_kernel) D_l' and the block-diagonal D and E are assembled here,
and the pencil is solved by LAPACK's QZ (zggev) on the host, as the reference
script does with MATLAB's eig (Matrix_LogExp.m:503): the pencil has order
N_alpha n_l (2121 on the published small grid, 60501 at the default grid,
where the dense QZ is out of reach for either), so this form is for small
grids.  The W_l of these problems are numerically rank-deficient (the script
prints their ranks, :394-409), so D is singular: the pencil has infinite
eigenvalues and finite ones that are not modes (solve reports both).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .BasisFunctions import DevicePrecomputed, PrecomputedData

__all__ = ["PencilBlocks", "pencil_blocks", "assemble_pencil", "solve_pencil",
           "linear_form_modes"]


@dataclass
class PencilBlocks:
    D: np.ndarray        # [N, N, n_l]  W_l^T diag(DJ) conj W_l
    E: np.ndarray        # [N, N, n_l]  W_l^T diag(DJ Omega_res_l) conj W_l
    F: np.ndarray        # [N, N, n_l]  W_l^T diag(DJ F0l_l) conj W_l


def pencil_blocks(dev) -> PencilBlocks:
    """Matrix_LogExp.m:385-403 on the device.  dev: DevicePrecomputed or
    PrecomputedData with a device context."""
    ctx = dev.ctx
    if ctx is None:
        raise RuntimeError("pencil_blocks: no device context (this build has no CPU path)")
    return PencilBlocks(*ctx.build_pencil())


def assemble_pencil(b: PencilBlocks, N_kernel, G: float):
    """(E + B, D) of order N n_l (Matrix_LogExp.m:414-449)."""
    N, _, n_l = b.D.shape
    nt = N * n_l
    A = np.zeros((nt, nt), dtype=np.complex128)
    Dg = np.zeros((nt, nt), dtype=np.complex128)
    Fn = [G * b.F[:, :, i] * np.asarray(N_kernel)[None, :] for i in range(n_l)]
    Dcat = np.concatenate([b.D[:, :, j] for j in range(n_l)], axis=1)   # [N, N n_l]
    for i in range(n_l):
        A[i * N:(i + 1) * N, :] = Fn[i] @ Dcat
        A[i * N:(i + 1) * N, i * N:(i + 1) * N] += b.E[:, :, i]
        Dg[i * N:(i + 1) * N, i * N:(i + 1) * N] = b.D[:, :, i]
    return A, Dg


def solve_pencil(A, Dg):
    """Finite eigenvalues of (E + B) C = omega D C sorted by growth rate
    (imaginary part, descending; Matrix_LogExp.m:503-506) and the count of
    infinite ones."""
    import scipy.linalg
    w = scipy.linalg.eigvals(A, Dg)
    fin = w[np.isfinite(w)]
    return fin[np.argsort(-fin.imag, kind="stable")], int(w.size - fin.size)


def linear_form_modes(precomputed: PrecomputedData, dev: DevicePrecomputed | None = None,
                      n_display: int = 10, verbose: bool = False):
    """The script's output (Matrix_LogExp.m:508-516): the n_display
    eigenvalues with the largest growth rates as (Omega_p, gamma) pairs, with
    all finite eigenvalues and the infinite count."""
    b = pencil_blocks(dev if dev is not None else precomputed)
    A, Dg = assemble_pencil(b, precomputed.N_kernel, precomputed.G)
    w, n_inf = solve_pencil(A, Dg)
    m = precomputed.m
    top = [(float(x.real / m), float(x.imag)) for x in w[:n_display]]
    if verbose:
        print("Omega_p (pattern speed) | gamma (growth rate)")
        for op, g in top:
            print(f"{op:12.6f}           | {g:12.6f}")
    return top, w, n_inf
