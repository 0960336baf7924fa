// klsp_contract.hpp -- the M(omega) contraction kernels (shared templates).
//
//   M_w[beta, alpha] = sum_K W[K][beta] c_w[K] conj(W[K][alpha])
//   (MatrixBuilder.jl:119-141: sum over l of W_l^T diag(c_l) conj(W_l))
//
// Device layout: W[K][Np] ("rows of K", K = l_index*NPh + k) with alpha
// padded to Np = roundup(N, 16) and K to Kp = roundup(K, 8), plus kRowPad
// zero rows; c_w[K] in rows of Kp + kRowPad.  A 16x16x4 MFMA operand (16
// columns x 4 rows of K) is four 256-byte row segments, so every operand
// load is coalesced and the same array feeds both sides of the product.
//
// Work decomposition ("gangs").  A wave owns an MB x NB grid of 16x16
// output blocks for NW omegas.  A gang is the set of single-wave workgroups
// whose wave tiles cover the whole Np x Np output (wrows x wcols waves).  All
// waves of a gang walk the same contiguous K range on one XCD (workgroups
// are dealt round-robin over the 8 XCDs: linear block L runs on XCD L % 8),
// so rows read by one wave are often still in that XCD's L2 when the gang's
// other waves reach them.  (Forcing lockstep with a gang barrier was
// measured 2.8x slower: it pins every wave to the pace of the slowest.)
// Each gang writes one partial M for its K range; the partials are summed in
// FP64 by k_reduce.
//
// Balanced launches (FP64 and FP32).  Where Np/16 is not a multiple of the
// tile size, edge tiles hold fewer blocks than full ones, and with one wave
// per SIMD and equal K ranges their SIMDs idle for the difference (12% at
// 19 x 19 blocks in 4 x 4 tiles).  A work table instead gives each tile a
// number of K pieces proportional to its block count, so that every wave
// does the same work; the host packs the (tile, piece) units onto the 8 XCDs
// within their resident wave slots, units of the same K range together, and
// workgroup L reads its unit from the table (XCD = L mod 8).  Partial slice
// `slot` of a tile holds its piece `slot`; k_reduce sums each output block
// over its tile's piece count.
//
// Inner loop.  Two K-steps per trip on two register operand sets
// (ping-pong): the loads of step k+1 are issued before the MFMAs of step k
// (scheduling barriers pin the order), so memory latency hides behind the
// MFMAs.  Complex products use four real MFMAs per block:
//   Cr += Ar Br + (-Ai) Bi,   Ci += Ar Bi + Ai Br,   B = c * conj(W_alpha).
// A wave whose tile hangs over Np runs a compile-time specialisation with
// fewer blocks, so the loop body has no per-block branches.  FP64
// accumulator lanes hold rows kk + 4v of a block, FP32 ones rows 4kk + v.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>

#include "klsp_internal.hpp"

namespace klsp {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kF32Tile = 256;        // K rows per FP32 accumulation tile

struct MfmaF64 {
    __device__ __forceinline__ v4d operator()(double a, double b, v4d c) const
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
struct MfmaF32 {
    __device__ __forceinline__ v4f operator()(float a, float b, v4f c) const
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// One K-step (4 rows of K): B = c conj(X) in registers, then 4*NW*MB*NB
// MFMAs ordered so that consecutive MFMAs hit different accumulators.
template <int MB, int NB, int NW, typename V, typename S2, typename ACC, typename MF>
__device__ __forceinline__ void cstep(const S2 (&A)[MB], const S2 (&X)[NB], const double2 (&C)[NW],
                                      ACC (&cr)[NW][MB][NB], ACC (&ci)[NW][MB][NB], MF mfma)
{
    S2 Bv[NW][NB];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const V c0 = (V)C[w].x, c1 = (V)C[w].y;
#pragma unroll
        for (int n = 0; n < NB; ++n) {
            // c conj(x) with one multiply and one fused multiply-add per part
            // (VALU work shares the FP64 MFMA's pipeline)
            Bv[w][n].x = __builtin_fma(c0, X[n].x, c1 * X[n].y);
            Bv[w][n].y = __builtin_fma(c1, X[n].x, -(c0 * X[n].y));
        }
    }
    V nAy[MB];
#pragma unroll
    for (int m = 0; m < MB; ++m) nAy[m] = -A[m].y;
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int n = 0; n < NB; ++n) cr[w][m][n] = mfma(A[m].x, Bv[w][n].x, cr[w][m][n]);
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int n = 0; n < NB; ++n) ci[w][m][n] = mfma(A[m].x, Bv[w][n].y, ci[w][m][n]);
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int n = 0; n < NB; ++n) cr[w][m][n] = mfma(nAy[m], Bv[w][n].y, cr[w][m][n]);
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int n = 0; n < NB; ++n) ci[w][m][n] = mfma(A[m].y, Bv[w][n].x, ci[w][m][n]);
}

// Operand fetch of one K-step: MB row segments of the beta side, NB of the
// alpha side, NW weights.
template <int MB, int NB, int NW, typename S2>
__device__ __forceinline__ void cload(const S2* pa, const S2* px, const double2* pc, int64_t Kpc,
                                      S2 (&A)[MB], S2 (&X)[NB], double2 (&C)[NW])
{
#pragma unroll
    for (int m = 0; m < MB; ++m) A[m] = pa[16 * m];
#pragma unroll
    for (int n = 0; n < NB; ++n) X[n] = px[16 * n];
#pragma unroll
    for (int w = 0; w < NW; ++w) C[w] = pc[w * Kpc];
}

// One workgroup of a balanced launch: wave tile widx over rows [Kb, Ke) of
// K into partial slice `slot`; widx < 0 is an idle slot.  Kb and Ke are
// multiples of the kernel's rows per loop trip (kRowsPerTrip), except Ke = Kp.
template <bool F64, int NW> constexpr int kRowsPerTrip = (F64 && NW == 1) ? 16 : 8;
struct WorkUnit {
    int64_t Kb, Ke;
    int widx, slot;
};

// Arguments shared by every contraction launch.
struct ContractArgs {
    const void* W;         // double2 or float2 rows of K, or paired bf16 rows
    const double2* cw;     // weights [NW][Kp + kRowPad]
    const float2* cwf;     // the same in FP32 (BF16 kernel)
    double2* P;            // partials [NW][n_gangs][Np][Np] (alpha-major)
    int64_t Kp;
    int Np;
    int64_t Kg;            // K rows per gang (multiple of kF32Tile)
    int n_gangs, wrows, wcols;
    const WorkUnit* units; // balanced launch: one unit per workgroup (else null)
    int n_units;           // workgroups of a balanced launch
};

template <int MB, int NB, int NW>
__device__ __forceinline__ void wave_f64(const ContractArgs& a, int64_t Kb, int64_t Ke, int b0,
                                         int a0, int gang)
{
    const double2* __restrict__ W = (const double2*)a.W;
    const int lane = threadIdx.x & 63, i = lane & 15, kk = lane >> 4;
    const int Np = a.Np;
    const int64_t Kpc = a.Kp + kRowPad;
    v4d cr[NW][MB][NB], ci[NW][MB][NB];
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int n = 0; n < NB; ++n) {
                cr[w][m][n] = (v4d){0, 0, 0, 0};
                ci[w][m][n] = (v4d){0, 0, 0, 0};
            }
    const int64_t rstep = (int64_t)kKStep * Np;
    const double2* pa = W + (Kb + kk) * Np + b0 + i;
    const double2* px = W + (Kb + kk) * Np + a0 + i;
    const double2* pc = a.cw + Kb + kk;
    const MfmaF64 mf;
    if constexpr (NW == 1) {
        // One omega (accumulators in AGPRs, VGPRs to spare): four operand
        // sets, loads three K-steps ahead of the MFMAs (with one wave per
        // SIMD nothing else hides their latency), four steps (16 rows) a
        // trip.  [Kb, Ke) is a multiple of 16 rows except where Ke = Kp (a
        // multiple of 8); there the last trip runs past Ke by at most 8 rows
        // and its loads by at most 19, inside the kRowPad zero rows.
        double2 A0[MB], X0[NB], C0[1], A1[MB], X1[NB], C1[1];
        double2 A2[MB], X2[NB], C2[1], A3[MB], X3[NB], C3[1];
        // a 32-bit trip count keeps the loop test in scalar registers.  The
        // prologue sits behind the same test as the loop: otherwise the
        // compiler sinks its loads past that test, reorders them, and the
        // wait at the loop head (which must hold for the first trip too)
        // then covers part of the second operand set.
        const int ntrip = (int)((Ke - Kb + 4 * kKStep - 1) / (4 * kKStep));
        if (ntrip > 0) {
        cload<MB, NB, 1>(pa, px, pc, Kpc, A0, X0, C0);
        __builtin_amdgcn_sched_barrier(0);
        cload<MB, NB, 1>(pa + rstep, px + rstep, pc + kKStep, Kpc, A1, X1, C1);
        __builtin_amdgcn_sched_barrier(0);
        cload<MB, NB, 1>(pa + 2 * rstep, px + 2 * rstep, pc + 2 * kKStep, Kpc, A2, X2, C2);
        __builtin_amdgcn_sched_barrier(0);
        for (int trip = 0; trip < ntrip; ++trip) {
            cload<MB, NB, 1>(pa + 3 * rstep, px + 3 * rstep, pc + 3 * kKStep, Kpc, A3, X3, C3);
            __builtin_amdgcn_sched_barrier(0);
            cstep<MB, NB, 1, double>(A0, X0, C0, cr, ci, mf);
            __builtin_amdgcn_sched_barrier(0);
            cload<MB, NB, 1>(pa + 4 * rstep, px + 4 * rstep, pc + 4 * kKStep, Kpc, A0, X0, C0);
            __builtin_amdgcn_sched_barrier(0);
            cstep<MB, NB, 1, double>(A1, X1, C1, cr, ci, mf);
            __builtin_amdgcn_sched_barrier(0);
            cload<MB, NB, 1>(pa + 5 * rstep, px + 5 * rstep, pc + 5 * kKStep, Kpc, A1, X1, C1);
            __builtin_amdgcn_sched_barrier(0);
            cstep<MB, NB, 1, double>(A2, X2, C2, cr, ci, mf);
            __builtin_amdgcn_sched_barrier(0);
            cload<MB, NB, 1>(pa + 6 * rstep, px + 6 * rstep, pc + 6 * kKStep, Kpc, A2, X2, C2);
            __builtin_amdgcn_sched_barrier(0);
            cstep<MB, NB, 1, double>(A3, X3, C3, cr, ci, mf);
            __builtin_amdgcn_sched_barrier(0);
            pa += 4 * rstep;
            px += 4 * rstep;
            pc += 4 * kKStep;
        }
        }
    } else {
    double2 A0[MB], X0[NB], C0[NW], A1[MB], X1[NB], C1[NW];
    cload<MB, NB, NW>(pa, px, pc, Kpc, A0, X0, C0);
    for (int64_t K0 = Kb; K0 < Ke; K0 += 2 * kKStep) {
        cload<MB, NB, NW>(pa + rstep, px + rstep, pc + kKStep, Kpc, A1, X1, C1);
        __builtin_amdgcn_sched_barrier(0);
        cstep<MB, NB, NW, double>(A0, X0, C0, cr, ci, mf);
        __builtin_amdgcn_sched_barrier(0);
        pa += 2 * rstep;
        px += 2 * rstep;
        pc += 2 * kKStep;
        cload<MB, NB, NW>(pa, px, pc, Kpc, A0, X0, C0);
        __builtin_amdgcn_sched_barrier(0);
        cstep<MB, NB, NW, double>(A1, X1, C1, cr, ci, mf);
        __builtin_amdgcn_sched_barrier(0);
    }
    }
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        double2* Pw = a.P + ((int64_t)w * a.n_gangs + gang) * (int64_t)Np * Np;
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int n = 0; n < NB; ++n) {
                double2* dst = Pw + (int64_t)(a0 + 16 * n + i) * Np + b0 + 16 * m + kk;
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    dst[4 * v] = make_double2(cr[w][m][n][v], ci[w][m][n][v]);
            }
    }
}

template <int MB, int NB, int NW>
__device__ __forceinline__ void wave_f32(const ContractArgs& a, int64_t Kb, int64_t Ke, int b0,
                                         int a0, int gang)
{
    const float2* __restrict__ W = (const float2*)a.W;
    const int lane = threadIdx.x & 63, i = lane & 15, kk = lane >> 4;
    const int Np = a.Np;
    const int64_t Kpc = a.Kp + kRowPad;
    v4d dr[NW][MB][NB], di[NW][MB][NB];
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int n = 0; n < NB; ++n) {
                dr[w][m][n] = (v4d){0, 0, 0, 0};
                di[w][m][n] = (v4d){0, 0, 0, 0};
            }
    const int64_t rstep = (int64_t)kKStep * Np;
    const float2* pa = W + (Kb + kk) * Np + b0 + i;
    const float2* px = W + (Kb + kk) * Np + a0 + i;
    const double2* pc = a.cw + Kb + kk;
    float2 A0[MB], X0[NB], A1[MB], X1[NB];
    double2 C0[NW], C1[NW];
    cload<MB, NB, NW>(pa, px, pc, Kpc, A0, X0, C0);
    const MfmaF32 mf;
    for (int64_t T0 = Kb; T0 < Ke; T0 += kF32Tile) {
        const int npairs = (int)((min(Ke, T0 + kF32Tile) - T0) / (2 * kKStep));
        v4f cr[NW][MB][NB], ci[NW][MB][NB];
#pragma unroll
        for (int w = 0; w < NW; ++w)
#pragma unroll
            for (int m = 0; m < MB; ++m)
#pragma unroll
                for (int n = 0; n < NB; ++n) {
                    cr[w][m][n] = (v4f){0, 0, 0, 0};
                    ci[w][m][n] = (v4f){0, 0, 0, 0};
                }
        for (int st = 0; st < npairs; ++st) {
            cload<MB, NB, NW>(pa + rstep, px + rstep, pc + kKStep, Kpc, A1, X1, C1);
            __builtin_amdgcn_sched_barrier(0);
            cstep<MB, NB, NW, float>(A0, X0, C0, cr, ci, mf);
            __builtin_amdgcn_sched_barrier(0);
            pa += 2 * rstep;
            px += 2 * rstep;
            pc += 2 * kKStep;
            cload<MB, NB, NW>(pa, px, pc, Kpc, A0, X0, C0);
            __builtin_amdgcn_sched_barrier(0);
            cstep<MB, NB, NW, float>(A1, X1, C1, cr, ci, mf);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int w = 0; w < NW; ++w)
#pragma unroll
            for (int m = 0; m < MB; ++m)
#pragma unroll
                for (int n = 0; n < NB; ++n)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        dr[w][m][n][v] += (double)cr[w][m][n][v];
                        di[w][m][n][v] += (double)ci[w][m][n][v];
                    }
    }
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        double2* Pw = a.P + ((int64_t)w * a.n_gangs + gang) * (int64_t)Np * Np;
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int n = 0; n < NB; ++n) {
                double2* dst = Pw + (int64_t)(a0 + 16 * n + i) * Np + b0 + 16 * m + 4 * kk;
#pragma unroll
                for (int v = 0; v < 4; ++v) dst[v] = make_double2(dr[w][m][n][v], di[w][m][n][v]);
            }
    }
}

// Selects the specialisation with exactly mv x nv valid blocks.
template <bool F64, int MB, int NB, int NW>
__device__ __forceinline__ void dispatch(int mv, int nv, const ContractArgs& a, int64_t Kb,
                                         int64_t Ke, int b0, int a0, int gang)
{
    if constexpr (MB > 1) {
        if (mv < MB) {
            dispatch<F64, MB - 1, NB, NW>(mv, nv, a, Kb, Ke, b0, a0, gang);
            return;
        }
    }
    if constexpr (NB > 1) {
        if (nv < NB) {
            dispatch<F64, MB, NB - 1, NW>(mv, nv, a, Kb, Ke, b0, a0, gang);
            return;
        }
    }
    if constexpr (F64) wave_f64<MB, NB, NW>(a, Kb, Ke, b0, a0, gang);
    else wave_f32<MB, NB, NW>(a, Kb, Ke, b0, a0, gang);
}

template <bool F64, int MB, int NB, int NW>
__global__ __launch_bounds__(64) void k_contract(ContractArgs a)
{
    const int L = blockIdx.x;
    int gang, widx;
    int64_t Kb, Ke;
    if (a.units) {
        const WorkUnit u = a.units[L];
        if (u.widx < 0) return;
        widx = u.widx;
        gang = u.slot;
        Kb = u.Kb;
        Ke = u.Ke;
    } else {
        const int xcd = L & 7, j = L >> 3;
        const int gw = a.wrows * a.wcols;
        gang = xcd + 8 * (j / gw);
        widx = j % gw;
        if (gang >= a.n_gangs) return;
        Kb = min((int64_t)gang * a.Kg, a.Kp);
        Ke = min(a.Kp, Kb + a.Kg);
    }
    const int b0 = (widx / a.wcols) * MB * 16, a0 = (widx % a.wcols) * NB * 16;
    const int mv = min(MB, (a.Np - b0) / 16), nv = min(NB, (a.Np - a0) / 16);
    dispatch<F64, MB, NB, NW>(mv, nv, a, Kb, Ke, b0, a0, gang);
}

// Host-side entry points of the two contraction translation units.
struct ContractShape {
    int MB, NB, waves_per_cu;
};
ContractShape contract_f64_shape(int nw);
ContractShape contract_f32_shape(int nw);
hipError_t contract_f64_launch(int nw, const ContractArgs& a, hipStream_t stream);
hipError_t contract_f32_launch(int nw, const ContractArgs& a, hipStream_t stream);

// FP64 contraction in the real basis of a symmetric alpha grid
// (klsp_contract_sym64.hip): the upper-triangle tiles of P = R^T diag(Re c) R
// and Q = R^T diag(Im c) R for one omega, partial slice P + iQ per K piece;
// square tiles of shape().MB blocks, work units from a balanced table whose
// K pieces are multiples of quantum() rows.  `variant` (0 .. variants() - 1)
// selects a compiled form: prefetch depth and whether diagonal tiles compute
// all their blocks (full_diag) or the upper triangle only.
int contract_sym64_variants();
ContractShape contract_sym64_shape(int variant);
int contract_sym64_quantum(int variant);
bool contract_sym64_full_diag(int variant);
const char* contract_sym64_name(int variant);
hipError_t contract_sym64_launch(int variant, const ContractArgs& a, hipStream_t stream);

// Real-basis FP32 contraction (klsp_contract_sym32.hip): 4 x 2 tiles on or
// above the diagonal, one weight set per launch (a.cwf, a.P point at it).
ContractShape contract_sym32_shape();
int contract_sym32_quantum();
hipError_t contract_sym32_launch(const ContractArgs& a, hipStream_t stream);
// FP32 real basis of `rows` rows of W into R32 ([rows][Np] floats).
hipError_t split_R_f32(const double2* W64, int64_t rows, int N, int Np, float* R32,
                       hipStream_t stream);

// BF16 three-term contraction (klsp_contract_bf16.hip): workgroups of
// kBf16Waves waves stacked along beta, kBf16MB blocks per wave, kBf16NB alpha
// blocks per workgroup; K ranges per gang are multiples of 16 rows.
constexpr int kBf16Waves = 4, kBf16MB = 5, kBf16NB = 1;
struct Bf16Geometry {
    int wrows, wcols, wg_per_cu;   // workgroup tiles per gang, resident workgroups per CU
    int mb = 0;                    // block slots per wave (real-basis kernel)
};
// nf: the first nf of the nw weight sets take the three-term products, the
// others the single hi x hi product (nf = nw except in the mixed search).
Bf16Geometry contract_bf16_geometry(int Np, int nw, int nf);
hipError_t contract_bf16_launch(int nw, int nf, const ContractArgs& a, hipStream_t stream);
// W complex128 rows of K (`rows` present) -> bf16 hi and lo planes of nq
// rows of four K (2 x nq x Np x 16 bytes at Wb).
hipError_t split_W_bf16(const double2* W64, int64_t rows, int Np, int64_t nq, void* Wb,
                        hipStream_t stream);

// BF16 three-term contraction in the real basis of a symmetric alpha grid
// (klsp_contract_symbf16.hip), in one of two work decompositions of the upper
// triangle: kBf16Fold (folded column-block pairs, up to four weight sets a
// launch) or kBf16Strip (folded pairs of two-column strips, up to two); both
// give the same bits for the same gang plan.  wcols column groups x wrows
// workgroups each, partial slice P + iQ per gang and weight set; K ranges
// per gang are multiples of 64 rows.  The real basis is held as bf16 hi and
// lo planes of symbf16_records(Kp) records of eight K rows
// (2 x records x Np x 16 bytes), made from W by split_R_bf16.
constexpr int kBf16Fold = 1, kBf16Strip = 2;
Bf16Geometry contract_symbf16_geometry(int form, int Np, int nw, int nf);
int contract_symbf16_max_sets(int form);
// Share of a launch's block slots (nw weight sets) holding a block of the
// upper triangle (the rest idle through the launch).
double contract_symbf16_slot_use(int form, int Np, int nw);
hipError_t contract_symbf16_launch(int form, int nw, int nf, int mb, const ContractArgs& a,
                                   hipStream_t stream);
int64_t symbf16_records(int64_t Kp);
hipError_t split_R_bf16(const double2* W64, int64_t rows, int N, int Np, int64_t nq, void* Rb,
                        hipStream_t stream);

}  // namespace klsp
