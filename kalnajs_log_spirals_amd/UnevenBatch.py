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

Packed (run_batch(..., workers=k)): k host threads each drive a context and
stream of their own, so draws overlap on the device (one draw's orbit and W
quadrature kernels, or its latency-bound LU launches, beside another's
contractions).  Packer admits draws largest first while the device memory
the contexts hold (klsp_device_bytes) plus the estimate of the next draw
(footprint_bytes) fits a budget, and has a worker release its context when
nothing fits beside the others.
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

__all__ = ["Draw", "DrawResult", "draw_batch", "config_for", "run_batch", "footprint_bytes",
           "Packer"]

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


def _search(i: int, d: Draw, cfg: KalnajsConfig, ctx, h, t_host: float, precision: str,
            polish: int, orbits: str, log) -> DrawResult:
    """One draw on context ctx: device precompute (orbits on the device, or W
    from the host orbits h), the Newton search and the optional polish."""
    t0 = time.perf_counter()
    if orbits == "device":
        pre = precompute_device_orbits(cfg, create_toomre_zang_model(cfg),
                                       download_orbits=False, ctx=ctx)
    else:
        pre = precompute_device(cfg, h, ctx=ctx)
    t_W = time.perf_counter() - t0
    t0 = time.perf_counter()
    dev = to_device(pre, GPUType.AMD, np.float64)
    if precision in _LP_COPY:
        pre.ctx.prepare(_LP_COPY[precision])
    t_setup = time.perf_counter() - t0
    # device-time accounting (one event wait per build, microseconds)
    pre.ctx.set_profiling(True)
    r = newton_search(cfg, pre, dev, GPUType.AMD, np.float64, d.Omega_p0, d.gamma0,
                      precision=precision)
    t_n = time.perf_counter() - t0
    tm = pre.ctx.last_timing()
    res = DrawResult(d, precision, bool(r.converged), float(r.Omega_p), float(r.gamma),
                     float(abs(r.det_val)), int(r.iterations), int(r.n_fp64_builds),
                     int(r.n_lp_builds), t_host, t_W, t_n,
                     device_ms=float(tm["contract_ms"] + tm["lu_ms"]), setup_s=t_setup)
    if polish > 0 and r.converged:
        pr, _ = pre.ctx.newton(r.Omega_p, r.gamma, max_iter=polish, tol=0.0,
                               precision=_ffi.PREC_FP64)
        res.polished = complex(pr.omega[0], pr.omega[1])
        res.polish_move = abs(res.polished - complex(d.m * r.Omega_p, r.gamma))
    if log:
        log(f"[uneven] {i:2d} m={d.m} n={d.n_zang} q1={d.q1} NR={cfg.grid.NR} "
            f"Ne={cfg.grid.Ne} N={cfg.grid.N_alpha} l={cfg.grid.l_max}: "
            f"{'converged' if res.converged else 'not converged'} "
            f"Op={res.Omega_p:.13f} g={res.gamma:.13f} host {t_host:.1f}s "
            f"W {t_W:.2f}s newton {t_n:.2f}s ({res.n_lp_builds} lp + "
            f"{res.n_fp64_builds} fp64 builds, {res.device_ms / 1e3:.2f}s on device, "
            f"setup {t_setup:.2f}s)"
            + (f" polish move {res.polish_move:.2e}" if polish > 0 else ""))
    return res


def footprint_bytes(cfg: KalnajsConfig, precision: str = "mixed") -> tuple:
    """(resident, transient) device bytes of one draw on a context of its
    own: resident = what the context holds between calls (W, the real basis
    and its low-precision copy, the weights, the K-piece partials, matrix
    slots, problem arrays), transient = the orbit samples and quadrature
    temporaries that exist only during the precompute.  An upper estimate
    from the grid sizes (klsp_device_bytes gives the held bytes exactly once
    the context has run the draw)."""
    g = cfg.grid
    NPh = g.NR * g.Ne
    n_l = g.l_max - g.l_min + 1
    K = NPh * n_l
    N = g.N_alpha
    Np = (N + 15) // 16 * 16
    Kp = (K + 7) // 8 * 8
    rows = Kp + 64
    res = rows * Np * 16 + rows * Np * 8                      # W, real basis R
    if precision in ("mixed", "bf16x3"):
        res += ((Kp + 63) // 64 * 64) * Np * 4                # bf16 hi/lo planes of R
    if precision in ("mixed_fp32", "fp32"):
        res += rows * Np * 8                                  # complex64 W
    res += 4 * rows * 24                                      # weights (complex128 + complex64)
    res += 4 * 1100 * Np * Np * 16                            # K-piece partials
    res += 2 * (256 << 20) + 8 * N * N * 16                   # matrix slots, LU copies, S
    res += (NPh + 2 * K + N) * 8
    transient = 7 * NPh * g.nwa * 8                           # orbit samples + quadrature temporaries
    return int(res * 1.05) + (256 << 20), int(transient * 1.05)


class Packer:
    """Which draw each of `workers` contexts runs next, so that the device
    memory the contexts hold stays within `budget` bytes.

    Draws are taken largest first.  A worker whose context holds `held`
    bytes may start draw i when the others' held and in-flight bytes plus
    max(held, resident_i) + transient_i fit the budget; a context grows to
    its largest draw and keeps that memory (buffers are reused: reallocating
    tens of GB costs seconds) until the worker releases it.  When nothing
    fits beside the others, a worker waits for another to finish while any
    other worker is running, and is told to release its context (RELEASE)
    only when all others are idle.  With every worker idle and
    empty, the largest remaining draw starts even if its estimate exceeds the
    budget (the estimate is an upper bound).  Pure logic: runs on the CPU."""

    RELEASE = -1

    def __init__(self, sizes, budget: int, workers: int):
        import threading
        self.sizes = [tuple(int(x) for x in s) for s in sizes]
        self.budget = int(budget)
        self.workers = int(workers)
        self.remaining = sorted(range(len(self.sizes)), key=lambda i: -sum(self.sizes[i]))
        self.held = [0] * self.workers       # bytes a worker's context holds
        self.extra = [0] * self.workers      # in-flight bytes of its current draw beyond held
        self.running = [False] * self.workers
        self.peak = 0                        # the largest sum of held + in-flight seen
        self.cond = threading.Condition()

    def _others(self, w):
        return sum(self.held[j] + self.extra[j] for j in range(self.workers) if j != w)

    def next(self, w: int):
        """The next draw for worker w, RELEASE, or None when no draws remain."""
        with self.cond:
            while True:
                if not self.remaining:
                    return None
                others = self._others(w)
                for i in self.remaining:
                    r, t = self.sizes[i]
                    need = max(self.held[w], r) + t
                    if others + need <= self.budget:
                        return self._start(w, i, need)
                others_running = any(self.running[j] for j in range(self.workers) if j != w)
                if self.held[w] > 0 and not others_running:
                    return Packer.RELEASE
                if not any(self.running) and all(h == 0 for h in self.held):
                    i = self.remaining[0]
                    r, t = self.sizes[i]
                    return self._start(w, i, r + t)
                self.cond.wait()

    def _start(self, w, i, need):
        self.remaining.remove(i)
        self.extra[w] = need - self.held[w]
        self.running[w] = True
        self.peak = max(self.peak, sum(self.held) + sum(self.extra))
        return i

    def done(self, w: int, held_bytes: int):
        """Worker w finished its draw; its context now holds held_bytes."""
        with self.cond:
            self.held[w] = int(held_bytes)
            self.extra[w] = 0
            self.running[w] = False
            self.cond.notify_all()

    def released(self, w: int):
        with self.cond:
            self.held[w] = 0
            self.cond.notify_all()


def run_batch(draws: List[Draw], precision: str = "mixed", scale: float = 1.0,
              base: Optional[KalnajsConfig] = None, log=None, polish: int = 0,
              orbits: str = "host", symmetric_basis: bool = True, workers: int = 1,
              budget: Optional[int] = None) -> List[DrawResult]:
    """Newton search (NewtonSolver.jl:36 semantics, `precision` as in
    newton_search) from each draw's starting omega.  With polish > 0 every
    converged root is then refined by that many FP64 Newton steps with
    tol = 0 (an a-posteriori check: the move is the root's FP64 error).
    orbits = "host" | "device" chooses where the orbits are integrated;
    symmetric_basis = False makes every build use the complex product
    (klsp_set_symmetric_basis).

    workers > 1 packs draws onto the device (orbits on the device only):
    that many host threads each drive a context of its own on its own
    stream, so one draw's precompute or LU-bound search overlaps another's
    contractions; Packer admits draws largest first while the estimated
    (footprint_bytes) and held (klsp_device_bytes) device memory of the
    contexts fits `budget` (default: 85% of the device's memory).  Results
    come back in draw order and equal the one-at-a-time batch's."""
    if orbits not in ("host", "device"):
        raise ValueError(f"run_batch: orbits must be 'host' or 'device', not {orbits!r}")
    cfgs = [config_for(d, scale, base) for d in draws]
    if workers > 1:
        if orbits != "device":
            raise ValueError("run_batch: packed batches integrate the orbits on the device")
        return _run_packed(draws, cfgs, precision, log, polish, symmetric_basis, workers, budget)

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
                results.append(_search(i, d, cfgs[i], ctx, h, t_host, precision, polish,
                                       orbits, log))
                del h
    finally:
        ctx.close()
    return results


def _run_packed(draws, cfgs, precision, log, polish, symmetric_basis, workers, budget):
    import threading
    dev_id = GPUBackend.current_device()
    if budget is None:
        budget = int(0.85 * GPUBackend.gpu_total_memory(dev_id))
    packer = Packer([footprint_bytes(c, precision) for c in cfgs], budget, workers)
    results: List[Optional[DrawResult]] = [None] * len(draws)
    errors = []

    def work(w):
        ctx = None
        try:
            while True:
                i = packer.next(w)
                if i is None:
                    break
                if i == Packer.RELEASE:
                    ctx.close()
                    ctx = None
                    packer.released(w)
                    continue
                held = 0
                try:
                    if ctx is None:
                        ctx = _ffi.Context(dev_id)
                        ctx.set_symmetric_basis(symmetric_basis)
                    results[i] = _search(i, draws[i], cfgs[i], ctx, None, 0.0, precision,
                                         polish, "device", log)
                    held = ctx.device_bytes()
                finally:
                    packer.done(w, held)
        except BaseException as e:      # reported after every worker has ended
            errors.append(e)
            with packer.cond:
                packer.remaining.clear()
                packer.cond.notify_all()
        finally:
            if ctx is not None:
                ctx.close()

    threads = [threading.Thread(target=work, args=(w,), name=f"uneven-{w}")
               for w in range(workers)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    if log:
        log(f"[uneven] packed on {workers} contexts, budget {budget / 1e9:.0f} GB, "
            f"peak estimate {packer.peak / 1e9:.0f} GB")
    return results


def to_records(results: List[DrawResult]) -> list:
    return [dict(asdict(r), draw=asdict(r.draw)) for r in results]
