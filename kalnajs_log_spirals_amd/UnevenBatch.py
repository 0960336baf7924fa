"""UnevenBatch -- a batch of Newton searches over Toomre-Zang disk variants of
uneven sizes (no counterpart in the reference, whose runner does one model
per process: run_kalnajs_gpu.jl:276-307).

Each draw is one disk model (n_zang, q1, m) on its own grid (NR, Ne,
N_alpha, l = -l_max..l_max) with a starting omega; the draws come from
SplitMix64 so that the batch is the same everywhere.  Parameters a draw does
not set take the values of configs/default.toml.

The batch runs one draw at a time on the device, all of them in one shared
context: its W buffers grow to the largest draw's and are reused by every
later draw that fits (no per-draw allocation and release of tens of GB), so
the largest draw's W bounds device memory.  With orbits = "host" a worker
thread computes the next draw's host half (grids and orbits) while the device
works on the current one; with orbits = "device" the orbits are integrated on
the device too and only the O(NR * Ne) grid and distribution-function work
stays on the host.
"""
from __future__ import annotations

import os
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import asdict, dataclass
from typing import List, Optional

import numpy as np

from . import GPUBackend, _ffi
from .BasisFunctions import precompute_device, precompute_device_orbits, precompute_host, to_device
from .Configuration import KalnajsConfig, load_config
from .GPUBackend import GPUType
from .Models import create_toomre_zang_model
from .NewtonSolver import newton_search
from .prng import SplitMix64

__all__ = ["Draw", "DrawResult", "draw_batch", "config_for", "run_batch"]

_N_ZANG = (2, 3, 4, 5, 6)
_Q1 = (4, 6, 7, 9, 12)
_M = (1, 2, 3)


@dataclass
class Draw:
    index: int
    n_zang: int
    q1: int
    m: int
    NR: int
    Ne: int
    N_alpha: int
    l_max: int
    Omega_p0: float
    gamma0: float


@dataclass
class DrawResult:
    draw: Draw
    precision: str
    converged: bool
    Omega_p: float
    gamma: float
    abs_det: float
    iterations: int
    n_fp64_builds: int
    n_lp_builds: int
    host_s: float        # grids + orbits (worker thread, overlapped); 0 with device orbits
    W_s: float           # device W quadrature (wall, incl. orbit upload); with device
                         # orbits: the whole precompute (grids, orbits, W, DF arrays)
    newton_s: float      # to_device + the search (incl. a low-precision W copy's first build)
    polished: complex = complex("nan")   # root after `polish` FP64 Newton steps, tol = 0
    polish_move: float = float("nan")    # |polished - omega|
    device_ms: float = float("nan")      # the search's contraction + LU device time
    setup_s: float = float("nan")        # part of newton_s: to_device + the low-precision W copy


# the W copy a mixed search builds its low-precision passes from
_LP_COPY = {"mixed": _ffi.PREC_BF16X3, "mixed_fp32": _ffi.PREC_FP32}


def draw_batch(n: int = 48, seed: int = 0) -> List[Draw]:
    """n draws: n_zang in {2..6}, q1 in {4, 6, 7, 9, 12}, m in {1, 2, 3},
    NR in [201, 801], Ne in [21, 101], N_alpha in [101, 401], l_max in
    [50, 150], Re omega / m in [0.3, 0.6), Im omega in [0.05, 0.2)."""
    g = SplitMix64(seed)
    out = []
    for i in range(n):
        out.append(Draw(index=i, n_zang=_N_ZANG[g.below(5)], q1=_Q1[g.below(5)],
                        m=_M[g.below(3)], NR=201 + g.below(601), Ne=21 + g.below(81),
                        N_alpha=101 + g.below(301), l_max=50 + g.below(101),
                        Omega_p0=0.3 + 0.3 * g.uniform(), gamma0=0.05 + 0.15 * g.uniform()))
    return out


def _default_config() -> KalnajsConfig:
    return load_config(os.path.join(os.path.dirname(__file__), "configs", "default.toml"))


def config_for(d: Draw, scale: float = 1.0, base: Optional[KalnajsConfig] = None) -> KalnajsConfig:
    """The configuration of a draw.  scale < 1 shrinks the grid (for tests):
    NR, Ne, N_alpha and l_max are multiplied by it (floors 11, 5, 11, 3) and
    nw / nwa become 10001 / 201."""
    import copy
    c = copy.deepcopy(base if base is not None else _default_config())
    c.model.n_zang, c.model.q1, c.physics.m = d.n_zang, d.q1, d.m
    if scale >= 1.0:
        c.grid.NR, c.grid.Ne, c.grid.N_alpha = d.NR, d.Ne, d.N_alpha
        c.grid.l_min, c.grid.l_max = -d.l_max, d.l_max
    else:
        c.grid.NR = max(11, int(round(d.NR * scale)))
        c.grid.Ne = max(5, int(round(d.Ne * scale)))
        c.grid.N_alpha = max(11, int(round(d.N_alpha * scale)))
        lm = max(3, int(round(d.l_max * scale)))
        c.grid.l_min, c.grid.l_max = -lm, lm
        c.grid.nw, c.grid.nwa = 10001, 201
    return c


def run_batch(draws: List[Draw], precision: str = "mixed", scale: float = 1.0,
              base: Optional[KalnajsConfig] = None, log=None, polish: int = 0,
              orbits: str = "host", symmetric_basis: bool = True) -> List[DrawResult]:
    """Newton search (NewtonSolver.jl:36 semantics, `precision` as in
    newton_search) from each draw's starting omega.  With polish > 0 every
    converged root is then refined by that many FP64 Newton steps with
    tol = 0 (an a-posteriori check: the move is the root's FP64 error).
    orbits = "host" | "device" chooses where the orbits are integrated;
    symmetric_basis = False makes every build use the complex product
    (klsp_set_symmetric_basis)."""
    if orbits not in ("host", "device"):
        raise ValueError(f"run_batch: orbits must be 'host' or 'device', not {orbits!r}")
    cfgs = [config_for(d, scale, base) for d in draws]

    def host(i):
        if orbits == "device":
            return None, 0.0
        t0 = time.perf_counter()
        h = precompute_host(cfgs[i], create_toomre_zang_model(cfgs[i]))
        return h, time.perf_counter() - t0

    results = []
    # one context for the whole batch: its W buffers grow to the largest
    # draw's and are reused, instead of tens of GB allocated and freed per draw
    ctx = _ffi.Context(GPUBackend.current_device())
    ctx.set_symmetric_basis(symmetric_basis)
    try:
        with ThreadPoolExecutor(max_workers=1) as pool:
            nxt = pool.submit(host, 0) if draws else None
            for i, d in enumerate(draws):
                h, t_host = nxt.result()
                nxt = pool.submit(host, i + 1) if i + 1 < len(draws) else None
                t0 = time.perf_counter()
                if orbits == "device":
                    pre = precompute_device_orbits(cfgs[i], create_toomre_zang_model(cfgs[i]),
                                                   download_orbits=False, ctx=ctx)
                else:
                    pre = precompute_device(cfgs[i], h, ctx=ctx)
                t_W = time.perf_counter() - t0
                del h
                t0 = time.perf_counter()
                dev = to_device(pre, GPUType.AMD, np.float64)
                if precision in _LP_COPY:
                    pre.ctx.prepare(_LP_COPY[precision])
                t_setup = time.perf_counter() - t0
                # device-time accounting (one event wait per build, microseconds)
                pre.ctx.set_profiling(True)
                r = newton_search(cfgs[i], pre, dev, GPUType.AMD, np.float64, d.Omega_p0, d.gamma0,
                                  precision=precision)
                t_n = time.perf_counter() - t0
                tm = pre.ctx.last_timing()
                res = DrawResult(d, precision, bool(r.converged), float(r.Omega_p), float(r.gamma),
                                 float(abs(r.det_val)), int(r.iterations), int(r.n_fp64_builds),
                                 int(r.n_lp_builds), t_host, t_W, t_n,
                                 device_ms=float(tm["contract_ms"] + tm["lu_ms"]),
                                 setup_s=t_setup)
                if polish > 0 and r.converged:
                    pr, _ = pre.ctx.newton(r.Omega_p, r.gamma, max_iter=polish, tol=0.0,
                                           precision=_ffi.PREC_FP64)
                    res.polished = complex(pr.omega[0], pr.omega[1])
                    res.polish_move = abs(res.polished - complex(d.m * r.Omega_p, r.gamma))
                results.append(res)
                if log:
                    log(f"[uneven] {i:2d} m={d.m} n={d.n_zang} q1={d.q1} NR={cfgs[i].grid.NR} "
                        f"Ne={cfgs[i].grid.Ne} N={cfgs[i].grid.N_alpha} l={cfgs[i].grid.l_max}: "
                        f"{'converged' if res.converged else 'not converged'} "
                        f"Op={res.Omega_p:.13f} g={res.gamma:.13f} host {t_host:.1f}s "
                        f"W {t_W:.2f}s newton {t_n:.2f}s ({res.n_lp_builds} lp + "
                        f"{res.n_fp64_builds} fp64 builds, {res.device_ms / 1e3:.2f}s on device, "
                        f"setup {t_setup:.2f}s)"
                        + (f" polish move {res.polish_move:.2e}" if polish > 0 else ""))
    finally:
        ctx.close()
    return results


def to_records(results: List[DrawResult]) -> list:
    return [dict(asdict(r), draw=asdict(r.draw)) for r in results]
