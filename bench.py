#!/usr/bin/env python3
"""Benchmark of the det(I - M(omega)) hot path on one MI355X.

Workload (the drop-in FP64 GPU configuration): Toomre-Zang n=4, q1=7, m=2 on
the default grid NR=501, Ne=51, N_alpha=301, alpha_max=30, l in [-100, 100]
(K = 201*501*51 = 5 135 751), nw/nwa = 100001/2001.  A step is one
det(I - M(omega)) evaluation with W resident in HBM: resonance weights, the
K-long contraction (FP64 MFMA, in the real basis of the symmetric alpha
grid: two real symmetric products), the G*N epilogue and the FP64 complex
LU.  The contraction variant is chosen by the library's autotuner first.
Also measured, with the other configurations of the workload set: Newton
roots/s (FP64 reference iteration from omega0 = 0.8794 + 0.1231i), the
mixed-precision search (BF16x3 Taylor-model passes, FP64 finish; also with
FP32 passes) and its distance to the FP64 root, the FP32 and BF16x3 det
rates, the whole 64 x 64 omega-grid scan in BF16x3 checked against FP64 dets
at 64 SplitMix64 points, the 48-draw uneven batch of Newton searches in
mixed precision and with the FP64 reference iteration (roots counted on
both sides only where two FP64 Newton steps with tol = 0 move them by at
most 1e-10), the W quadrature and orbit kernels, and the CPU baseline
(numpy/OpenBLAS BLAS-form det: one det's worth of l-slab ZGEMMs over 8
synthetic slabs plus one LU).

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("det(I−M(ω)) evals/s + Newton roots/s, mixed vs FP64 GPU at "
          "|δω|≤1e-10")
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the real-basis FP64 contraction at the default grid,
# from two rocprofv3 --pmc passes (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0,
# calibrated on a copy kernel of known size); profiles/det_default_grid.md
# holds the run.
# kernel name -> FETCH + WRITE bytes per launch (the autotuner may keep any
# variant; each was profiled in the same run)
PROFILED_TRAFFIC_BYTES = {
    "klsp::k_contract_sym64<4, false>": 37.51e9 + 0.056e9,
    "klsp::k_contract_sym64<4, true>": 39.69e9 + 0.062e9,
    "klsp::k_contract_sym64<6, false>": 35.95e9 + 0.056e9,
    "klsp::k_contract_sym64<6, true>": 35.82e9 + 0.062e9,
    "klsp::k_contract_sym64<8, false>": 35.97e9 + 0.056e9,
    "klsp::k_contract_sym64<8, true>": 36.04e9 + 0.062e9,
}


def sym_kernel_name(variant_name: str) -> str:
    """Kernel symbol of an FP64 real-basis variant ("sym64 4x4 d8 [fulldiag]")."""
    depth = variant_name.split()[2][1:]
    return f"klsp::k_contract_sym64<{depth}, {'true' if 'fulldiag' in variant_name else 'false'}>"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="default", help="configs/<name>.toml")
    ap.add_argument("--cpu-slabs", type=int, default=-1,
                    help="l-slab ZGEMMs timed for the CPU baseline (-1 = all n_l, 0 = skip)")
    ap.add_argument("--no-newton", action="store_true")
    ap.add_argument("--uneven", type=int, default=48,
                    help="draws of the 48-draw uneven batch of Newton searches run in mixed "
                         "precision and with the FP64 reference iteration (0 = skip)")
    ap.add_argument("--no-uneven-fp64", action="store_true",
                    help="skip the FP64 reference iteration over the uneven draws")
    ap.add_argument("--uneven-orbits", choices=("host", "device"), default="device",
                    help="where the uneven draws' orbits are integrated")
    ap.add_argument("--uneven-workers", type=int, default=1,
                    help="contexts the uneven draws are packed onto (1: one at a time)")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-scan", action="store_true",
                    help="skip the whole 64 x 64 omega-grid scan in BF16x3 (checked against FP64 "
                         "dets at 64 SplitMix64-drawn grid points)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default variant of the FP64 contraction")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the omega-sharded and K-sharded multi-GPU forms are not built",
              file=sys.stderr)
        return 2

    import kalnajs_log_spirals_amd as K
    from kalnajs_log_spirals_amd import _ffi

    cfg_path = os.path.join(ROOT, "kalnajs_log_spirals_amd", "configs", f"{args.config}.toml")
    model_path = os.path.join(ROOT, "kalnajs_log_spirals_amd", "models", "kalnajs_zang4.toml")
    cfg = K.load_model(K.load_config(cfg_path), model_path)
    g = cfg.grid
    n_l = g.l_max - g.l_min + 1
    NPh = g.NR * g.Ne
    Kdim = n_l * NPh
    N = g.N_alpha
    log(f"[bench] library {_ffi.version()}, config {args.config}: NR={g.NR} Ne={g.Ne} "
        f"N={N} l=[{g.l_min},{g.l_max}] K={Kdim}")

    model = K.create_toomre_zang_model(cfg)
    t0 = time.perf_counter()
    grids = K.compute_grids(cfg, model)
    od = K.compute_orbits(cfg, model, grids)
    t_orb = time.perf_counter() - t0
    log(f"[bench] host orbits: {t_orb:.2f} s ({_ffi.host_threads()} threads)")
    t0 = time.perf_counter()
    pre = K.precompute_all(cfg, model, K.GPUType.AMD)
    t_pre = time.perf_counter() - t0
    W_ms = pre.ctx.last_W_ms()
    log(f"[bench] precompute_all: {t_pre:.2f} s (W quadrature on device {W_ms:.1f} ms)")
    # the same precompute with the orbits integrated on the device
    t0 = time.perf_counter()
    pre_d = K.precompute_device_orbits(cfg, model)
    t_pre_d = time.perf_counter() - t0
    orbit_ms, W_ms_d = pre_d.ctx.last_orbit_ms(), pre_d.ctx.last_W_ms()
    od_h = pre.orbit_data
    d_om = max(float(np.max(np.abs(pre_d.orbit_data.Omega_1 - od_h.Omega_1)
                            / np.abs(od_h.Omega_1))),
               float(np.max(np.abs(pre_d.orbit_data.Omega_2 - od_h.Omega_2)
                            / np.maximum(np.abs(od_h.Omega_2), 1e-300))))
    d_samples = max(float(np.max(np.abs(getattr(pre_d.orbit_data, k) - getattr(od_h, k))))
                    for k in ("w1", "ra", "pha"))
    log(f"[bench] device orbits: precompute {t_pre_d:.2f} s (orbit kernel {orbit_ms:.1f} ms, "
        f"W quadrature {W_ms_d:.1f} ms); max rel |d Omega| vs host {d_om:.2e}, max |d sample| "
        f"{d_samples:.2e}")
    dev = K.to_device(pre, K.GPUType.AMD, np.float64)
    ctx = dev.ctx
    det_h, det_d = ctx.det([complex(2 * 0.4397, 0.1231)])[0], pre_d.ctx.det(
        [complex(2 * 0.4397, 0.1231)])[0]
    d_det = abs(det_d - det_h) / abs(det_h)
    log(f"[bench] det(I - M) with device vs host orbits: relative difference {d_det:.2e}")
    pre_d.ctx.close()
    del pre_d
    ctx.set_profiling(True)
    sym = ctx.symmetric_basis()
    tune = {}
    chosen = _ffi.lib().klsp_variant_name(_ffi.PREC_FP64, 2).decode() if sym else "complex"
    if sym and not args.no_autotune:
        t0 = time.perf_counter()
        tune, chosen = ctx.autotune(_ffi.PREC_FP64, reps=2)
        log(f"[bench] autotune ({time.perf_counter() - t0:.2f} s): " +
            ", ".join(f"{k} {v:.2f} ms" for k, v in tune.items()) + f" -> {chosen}")

    om0 = complex(2 * 0.4397, 0.1231)
    oms = [om0 + 0.001 * i * (1 + 1j) for i in range(args.warmup + max(args.steps, 3))]
    for i in range(args.warmup):
        ctx.det([oms[i]], _ffi.PREC_FP64)
    log(f"[bench] warmup done ({args.warmup} det)")
    contract_ms = lu_ms = weights_ms = reduce_ms = 0.0
    launches = 0
    dets = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        dets.append(ctx.det([oms[args.warmup + i]], _ffi.PREC_FP64)[0])
        t = ctx.last_timing()
        contract_ms += t["contract_ms"]
        lu_ms += t["lu_ms"]
        weights_ms += t["weights_ms"]
        reduce_ms += t["reduce_ms"]
        launches += t["contract_launches"]
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    det_per_s = args.steps / elapsed
    avg_contract = contract_ms / max(launches, 1)
    # complex product of the reference form (8 K N^2) and the real-basis work
    # on a symmetric grid: two real symmetric rank-K products, upper
    # triangles, 2 K N (N + 1) flop (the MFMAs also cover the 16-block padding
    # of the triangle: nbk (nbk + 1) / 2 blocks of 16 x 16 for P and Q)
    flop_complex = 8.0 * Kdim * N * N
    nbk = -(-N // 16)
    flop_sym = 2.0 * Kdim * N * (N + 1)
    flop_sym_exec = 2.0 * 2.0 * Kdim * 256 * nbk * (nbk + 1) / 2
    flop_build = flop_sym if sym else flop_complex
    achieved_tflops = flop_build / (avg_contract * 1e-3) / 1e12
    log(f"[bench] FP64 det: {ms_per_step:.2f} ms/det ({det_per_s:.3f} det/s); contraction "
        f"{avg_contract:.2f} ms = {achieved_tflops:.1f} TFLOP/s algorithmic "
        f"({'real basis' if sym else 'complex'}; {flop_complex / (avg_contract * 1e-3) / 1e12:.0f} "
        f"TFLOP/s of the complex form); LU {lu_ms / args.steps:.2f} ms; "
        f"weights {weights_ms / args.steps:.3f} ms; reduce {reduce_ms / args.steps:.3f} ms")

    extra = {"det_per_s_fp64": det_per_s, "contract_ms_fp64": avg_contract,
             "fp64_symmetric_basis": sym, "fp64_contract_variant": chosen,
             "fp64_autotune_ms": tune,
             "fp64_contract_tflops_complex_form": flop_complex / (avg_contract * 1e-3) / 1e12,
             "fp64_contract_tflops_executed_mfma": (flop_sym_exec if sym else flop_complex)
                                                   / (avg_contract * 1e-3) / 1e12,
             "lu_ms": lu_ms / args.steps, "weights_ms": weights_ms / args.steps,
             "reduce_ms": reduce_ms / args.steps, "W_quadrature_ms": W_ms,
             "host_orbits_s": t_orb, "precompute_s": t_pre,
             "device_orbits_ms": orbit_ms, "precompute_device_orbits_s": t_pre_d,
             "device_orbits_max_rel_domega": d_om, "device_orbits_max_dsample": d_samples,
             "device_orbits_det_rel_diff": d_det,
             # algorithmic 4 n_l nwa N flop per orbit; on the symmetric alpha grid
             # only the upper half of the columns is computed (W(-a) = conj W(a)),
             # in workgroups of 64 columns
             "W_quadrature_tflops_algorithmic": 4.0 * n_l * g.nwa * N * NPh / (W_ms * 1e-3) / 1e12,
             "W_quadrature_tflops_executed": 4.0 * n_l * g.nwa * (-(-(N - N // 2) // 64) * 64)
                                             * NPh / (W_ms * 1e-3) / 1e12}

    if not args.no_fp32:
        # FP32 builds: the real basis (what FP32 dets use on this symmetric
        # grid), then the complex form (asymmetric grids)
        for form, real_basis in (("", True), ("_complex_form", False)):
            ctx.set_symmetric_basis(real_basis)
            ctx.prepare(_ffi.PREC_FP32)
            ctx.det([om0], _ffi.PREC_FP32)
            n32 = min(max(3, args.steps // 2), len(oms) - args.warmup)
            c32 = 0.0
            t0 = time.perf_counter()
            d32 = []
            for i in range(n32):
                d32.append(ctx.det([oms[args.warmup + i]], _ffi.PREC_FP32)[0])
                c32 += ctx.last_timing()["contract_ms"]
            e32 = time.perf_counter() - t0
            rel = max(abs(d32[i] - dets[i]) / abs(dets[i]) for i in range(min(n32, len(dets))))
            flop32 = flop_sym if real_basis and sym else flop_complex
            extra.update({f"det_per_s_fp32{form}": n32 / e32, f"contract_ms_fp32{form}": c32 / n32,
                          f"fp32_contract_tflops{form}": flop32 / (c32 / n32 * 1e-3) / 1e12,
                          f"fp32_det_rel_err_vs_fp64{form}": rel})
            log(f"[bench] FP32 det{form.replace('_', ' ')}: {1e3 * e32 / n32:.2f} ms/det, "
                f"contraction {c32 / n32:.2f} ms, det rel err vs FP64 {rel:.2e}")
        ctx.set_symmetric_basis(True)

    if not args.no_fp32:
        ctx.prepare(_ffi.PREC_BF16X3)
        ctx.det([om0], _ffi.PREC_BF16X3)
        nb = min(max(3, args.steps // 2), len(oms) - args.warmup)
        cb = 0.0
        t0 = time.perf_counter()
        db = []
        for i in range(nb):
            db.append(ctx.det([oms[args.warmup + i]], _ffi.PREC_BF16X3)[0])
            cb += ctx.last_timing()["contract_ms"]
        eb = time.perf_counter() - t0
        relb = max(abs(db[i] - dets[i]) / abs(dets[i]) for i in range(min(nb, len(dets))))
        ctx.build_taylor(om0, 3, _ffi.PREC_BF16X3)
        tb3 = ctx.last_timing()["contract_ms"]
        ctx.build_taylor(om0, 4, _ffi.PREC_BF16X3)
        tb4 = ctx.last_timing()["contract_ms"]
        # bf16 MFMA work of the three-term split products (3 per real product)
        flop_bf = 3.0 * (flop_sym_exec if sym else flop_complex)
        extra.update({"det_per_s_bf16x3": nb / eb, "contract_ms_bf16x3": cb / nb,
                      "bf16x3_contract_tflops_mfma": flop_bf / (cb / nb * 1e-3) / 1e12,
                      "bf16x3_det_rel_err_vs_fp64": relb,
                      "taylor3_contract_ms_bf16x3": tb3, "taylor4_contract_ms_bf16x3": tb4})
        log(f"[bench] BF16x3 det: {1e3 * eb / nb:.2f} ms/det, contraction {cb / nb:.2f} ms "
            f"({flop_bf / (cb / nb * 1e-3) / 1e15:.2f} PFLOP/s of bf16 MFMA), det rel err vs FP64 "
            f"{relb:.2e}; 3-term Taylor contraction {tb3:.2f} ms, 4-term {tb4:.2f} ms")

    if not args.no_scan:
        # the whole window: Re w/m in linspace(0.30, 0.60, 64), Im w in linspace(0.02, 0.25, 64)
        from kalnajs_log_spirals_amd.prng import SplitMix64
        t0 = time.perf_counter()
        grid, Op0, g0 = ctx.grid_scan(64, 64, 0.30, 0.60, 0.02, 0.25, precision=_ffi.PREC_BF16X3)
        tsf = time.perf_counter() - t0
        Ops, gams = np.linspace(0.30, 0.60, 64), np.linspace(0.02, 0.25, 64)
        rng = SplitMix64(0)
        pts = [(rng.below(64), rng.below(64)) for _ in range(64)]
        ref = ctx.det([complex(2 * Ops[j], gams[i]) for i, j in pts], _ffi.PREC_FP64)
        med = float(np.median(np.abs(ref)))
        errf = max(abs(grid[i, j] - ref[k]) / max(abs(ref[k]), med) for k, (i, j) in enumerate(pts))
        extra.update({"scan_64x64_s_bf16x3": tsf, "scan_64x64_dets_per_s_bf16x3": 4096 / tsf,
                      "scan_64x64_rel_err_vs_fp64_64pts": errf,
                      "scan_64x64_argmin": [Op0, g0]})
        log(f"[bench] BF16x3 64 x 64 scan: {tsf:.1f} s ({4096 / tsf:.1f} det/s), rel err vs FP64 at "
            f"64 drawn points {errf:.2e}, argmin |det| at Omega_p={Op0:.4f} gamma={g0:.4f}")

    if not args.no_newton:
        t0 = time.perf_counter()
        r = K.newton_search(cfg, pre, dev, K.GPUType.AMD, np.float64, 0.4397, 0.1231)
        tn = time.perf_counter() - t0
        extra.update({"newton_roots_per_s_fp64": 1.0 / tn, "newton_s_fp64": tn,
                      "newton_iterations": r.iterations, "newton_det_evals": r.n_det_evals,
                      "newton_converged": r.converged, "root_Omega_p": r.Omega_p,
                      "root_gamma": r.gamma, "root_abs_det": abs(r.det_val)})
        log(f"[bench] FP64 Newton: {tn:.2f} s, {r.iterations} iterations, {r.n_det_evals} dets, "
            f"root Omega_p={r.Omega_p:.12f} gamma={r.gamma:.12f} |det|={abs(r.det_val):.2e}")
        w64 = r.omega
        # mixed precision with FP32 passes, for comparison
        ctx.prepare(_ffi.PREC_FP32)
        t0 = time.perf_counter()
        rf = K.newton_search(cfg, pre, dev, K.GPUType.AMD, np.float64, 0.4397, 0.1231,
                             precision="mixed_fp32")
        tf = time.perf_counter() - t0
        extra.update({"newton_s_mixed_fp32": tf,
                      "newton_mixed_fp32_abs_domega_vs_fp64": abs(rf.omega - w64)})
        log(f"[bench] mixed (FP32 passes) Newton: {tf:.3f} s, {rf.n_lp_builds} passes + "
            f"{rf.n_fp64_builds} FP64, |d omega| {abs(rf.omega - w64):.2e}")
        # mixed precision (BF16x3 passes; the planes of W were made by the BF16 leg or here)
        ctx.prepare(_ffi.PREC_BF16X3)
        t0 = time.perf_counter()
        rm = K.newton_search(cfg, pre, dev, K.GPUType.AMD, np.float64, 0.4397, 0.1231,
                             precision="mixed")
        tm = time.perf_counter() - t0
        tmx = ctx.last_timing()
        dw = abs(rm.omega - w64)
        extra.update({"newton_roots_per_s_mixed": (1.0 / tm) if dw <= 1e-10 else 0.0,
                      "newton_s_mixed": tm, "newton_mixed_speedup": tn / tm,
                      "newton_mixed_lp_builds": rm.n_lp_builds,
                      "newton_mixed_fp64_builds": rm.n_fp64_builds,
                      "newton_mixed_converged": rm.converged,
                      "newton_mixed_abs_domega_vs_fp64": dw,
                      "newton_mixed_dOmega_p": rm.Omega_p - r.Omega_p,
                      "newton_mixed_dgamma": rm.gamma - r.gamma,
                      "newton_mixed_abs_det": abs(rm.det_val),
                      "newton_mixed_contract_ms": tmx["contract_ms"],
                      "newton_mixed_lu_ms": tmx["lu_ms"]})
        log(f"[bench] mixed Newton: {tm:.3f} s ({tn / tm:.2f}x FP64), {rm.n_lp_builds} BF16x3 "
            f"Taylor passes + {rm.n_fp64_builds} FP64 dets, |d omega| vs FP64 root {dw:.2e}, "
            f"|det|={abs(rm.det_val):.2e}; contractions {tmx['contract_ms']:.1f} ms, "
            f"LUs {tmx['lu_ms']:.1f} ms")

    if args.uneven > 0:
        # the uneven batch: mixed-precision searches, then the FP64 reference
        # iteration from the same starts; on both sides a root counts only if
        # two FP64 Newton steps with tol = 0 move it by at most 1e-10 (it is
        # an FP64 root to that tolerance)
        from kalnajs_log_spirals_amd.UnevenBatch import draw_batch, run_batch
        ctx.close()
        draws = draw_batch(48, 0)[:args.uneven]

        def good(rs):
            return [r for r in rs if r.converged and r.polish_move <= 1e-10]

        t0 = time.perf_counter()
        rm = run_batch(draws, precision="mixed", log=log, polish=2, orbits=args.uneven_orbits,
                       workers=args.uneven_workers)
        wall = time.perf_counter() - t0
        gm = good(rm)
        tn = sum(r.newton_s for r in rm)
        extra.update({"uneven_draws": len(draws),
                      "uneven_converged_mixed": sum(r.converged for r in rm),
                      "uneven_roots_mixed": len(gm),
                      "uneven_wall_s_mixed": wall, "uneven_newton_s_mixed": tn,
                      "uneven_roots_per_s_mixed": len(gm) / tn if tn > 0 else 0.0,
                      "uneven_fp64_builds_mixed": sum(r.n_fp64_builds for r in rm),
                      "uneven_lp_builds_mixed": sum(r.n_lp_builds for r in rm),
                      "uneven_max_fp64_polish_move_mixed":
                          max((r.polish_move for r in rm if r.converged), default=None),
                      "uneven_precompute_s": sum(r.W_s for r in rm),
                      "uneven_orbits": args.uneven_orbits,
                      "uneven_workers": args.uneven_workers})
        log(f"[bench] uneven batch (mixed): {len(gm)}/{len(draws)} FP64-grade roots, Newton "
            f"{tn:.1f} s ({len(gm) / tn:.2f} roots/s), wall {wall:.1f} s")
        if not args.no_uneven_fp64:
            t0 = time.perf_counter()
            r64 = run_batch(draws, precision="fp64", log=log, polish=2,
                            orbits=args.uneven_orbits, workers=args.uneven_workers)
            wall64 = time.perf_counter() - t0
            g64 = good(r64)
            tn64 = sum(r.newton_s for r in r64)
            both = [(a, b) for a, b in zip(rm, r64) if a.converged and b.converged]
            # the same root: the polished roots within 1e-6 of each other
            same = [(a, b) for a, b in both if abs(a.polished - b.polished) < 1e-6]
            dws = [abs(complex(a.draw.m * a.Omega_p, a.gamma) - b.polished) for a, b in same]
            rps64 = len(g64) / tn64 if tn64 > 0 else 0.0
            rpsm = extra["uneven_roots_per_s_mixed"]
            extra.update({"uneven_converged_fp64": sum(r.converged for r in r64),
                          "uneven_roots_fp64": len(g64),
                          "uneven_newton_s_fp64": tn64, "uneven_wall_s_fp64": wall64,
                          "uneven_roots_per_s_fp64": rps64,
                          "uneven_roots_per_s_ratio": rpsm / rps64 if rps64 > 0 else None,
                          "uneven_fp64_builds_fp64": sum(r.n_fp64_builds for r in r64),
                          "uneven_converged_both": len(both), "uneven_same_root": len(same),
                          "uneven_max_abs_domega_same_root": max(dws) if dws else None,
                          "uneven_max_fp64_polish_move_fp64":
                              max((r.polish_move for r in r64 if r.converged), default=None)})
            log(f"[bench] uneven batch (FP64 reference iteration): {sum(r.converged for r in r64)} "
                f"converged, {len(g64)} FP64-grade roots, Newton {tn64:.1f} s ({rps64:.3f} roots/s); "
                f"mixed/FP64 roots/s {rpsm / rps64 if rps64 > 0 else float('nan'):.2f}x; "
                f"{len(same)} of {len(both)} on the same root, max |d omega| there "
                f"{max(dws) if dws else float('nan'):.2e}")

    cpu = None
    if args.cpu_slabs < 0:
        args.cpu_slabs = n_l
    if args.cpu_slabs > 0:
        from oracle import blas_form
        from threadpoolctl import threadpool_info, threadpool_limits
        blas = [i for i in threadpool_info() if i.get("user_api") == "blas"]
        cores = max((i.get("num_threads", 1) for i in blas), default=os.cpu_count())
        t0 = time.perf_counter()
        s_det, s_slab, s_lu = blas_form.time_sample(NPh, N, n_l, args.cpu_slabs)
        log(f"[bench] CPU BLAS-form: {s_slab * 1e3:.1f} ms per l-slab, LU {s_lu * 1e3:.2f} ms -> "
            f"{s_det:.2f} s per det ({time.perf_counter() - t0:.1f} s of CPU work)")
        cpu = {"value": 1.0 / s_det, "unit": "det/s", "cores": int(cores), "kind": "port",
               "sample": f"numpy/OpenBLAS BLAS-form det: {args.cpu_slabs} per-l ZGEMMs "
                         f"W_l^T (c*conj W_l) over 8 synthetic [{NPh}x{N}] complex128 l-slabs + "
                         f"one {N}x{N} LU, scaled to {n_l} slabs per det",
               "cpu_model": _cpu_model(),
               "blas": [{k: i.get(k) for k in ("internal_api", "version", "architecture",
                                               "num_threads")} for i in blas]}
        # the reference's runner caps BLAS at 4 threads (configs/default.toml:28)
        n4 = min(args.cpu_slabs, 48)
        t0 = time.perf_counter()
        with threadpool_limits(limits=4, user_api="blas"):
            s4_det, s4_slab, _ = blas_form.time_sample(NPh, N, n_l, n4)
        log(f"[bench] CPU BLAS-form at 4 threads: {s4_slab * 1e3:.1f} ms per l-slab -> "
            f"{s4_det:.2f} s per det ({time.perf_counter() - t0:.1f} s of CPU work, {n4} slabs)")
        cpu["value_4_threads"] = 1.0 / s4_det
        cpu["sample_4_threads"] = f"{n4} l-slabs at 4 BLAS threads, scaled to {n_l} slabs per det"
        if "newton_det_evals" in extra:
            # the CPU path runs the same reference iteration: its roots/s is the
            # det rate over the builds of the FP64 search above (not run on the CPU)
            evals = extra["newton_det_evals"]
            cpu["newton_roots_per_s_derived"] = cpu["value"] / max(evals, 1)
            cpu["newton_roots_note"] = (f"derived: det/s over the {evals} builds the reference "
                                        "iteration takes from the same start")

    kernel = sym_kernel_name(chosen) if sym else "klsp::k_contract<true, 4, 4, 1>"
    out = {
        "metric": METRIC, "value": det_per_s, "unit": "det/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "deterministic Toomre-Zang model (no external data)",
        "config": {"workload": f"FP64 det(I-M) at the {args.config} grid, W resident in HBM",
                   "NR": g.NR, "Ne": g.Ne, "N_alpha": N, "l_min": g.l_min, "l_max": g.l_max,
                   "K": Kdim, "nwa": g.nwa, "m": cfg.physics.m, "n_zang": cfg.model.n_zang,
                   "q1": cfg.model.q1},
        "roofline": {"bound": "mfma", "achieved": achieved_tflops, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_FP64_TFLOPS,
                     "traffic": PROFILED_TRAFFIC_BYTES.get(kernel) if args.config == "default"
                                else None,
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE+WRITE_SIZE)",
                     "kernel": kernel, "flop_per_launch": flop_build,
                     "flop_note": ("2 K N (N + 1): the two real symmetric rank-K products of the "
                                   "real basis" if sym else "8 K N^2: the complex product"),
                     "algorithmic_bytes_per_launch": (8.0 if sym else 16.0) * Kdim * N},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
