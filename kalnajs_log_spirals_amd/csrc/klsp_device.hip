// klsp_device.hip -- CDNA4 (gfx950) kernels of the det(I - M(omega)) path and
// the C-ABI drivers above them.
//
// Stage map (reference -> kernel):
//   compute_W_l        BasisFunctions.jl:199  -> k_orbit_prep + k_W_f64
//   weights c(omega)   MatrixBuilder.jl:124   -> k_weights
//   sum_l W^T c conj W MatrixBuilder.jl:119   -> k_contract_f64 / k_contract_f32
//   *G, *N_kernel      MatrixBuilder.jl:144   -> k_reduce
//   det(I - M)         MatrixBuilder.jl:182   -> k_lu_det
//
// Device layout of W ("rows of K"): W[K][Np] with K = l_index*NPh + k
// (the reference's column order of W_l_mat, flattened over l) and the alpha
// axis padded to Np = roundup(N, 16) with zeros; K is padded to Kp =
// roundup(K, 4) with zero rows.  A 16x16x4 MFMA operand (16 alphas x 4 K)
// is then four 256-byte row segments, so every operand load is coalesced,
// and the same array feeds both operands of M = W^T diag(c) conj(W).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "klsp.h"
#include "klsp_internal.hpp"

using namespace klsp;

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

#define KLSP_VERSION_STRING "klsp 0.1.0 gfx950"

// =====================================================================
// Device helpers
// =====================================================================

// 1/(c + i d) by Smith's algorithm with the same range scaling as Julia's
// inv(::ComplexF64); the weights use it so c(omega) rounds like the reference.
__device__ __forceinline__ double2 d_inv(double c, double d)
{
    double s = 1.0;
    const double cd = fmax(fabs(c), fabs(d));
    const double eps = 2.220446049250313e-16;
    if (cd >= 0.5 * 1.7976931348623157e308) { c *= 0.5; d *= 0.5; s *= 0.5; }
    if (cd <= 2.2250738585072014e-308 * 2.0 / eps) {
        const double bs = 2.0 / (eps * eps);
        c *= bs; d *= bs; s *= bs;
    }
    double p, q;
    if (fabs(d) <= fabs(c)) {
        const double r = d / c, t = 1.0 / (c + d * r);
        p = t; q = -r * t;
    } else {
        const double r = c / d, t = 1.0 / (d + c * r);
        p = r * t; q = -t;
    }
    return make_double2(p * s, q * s);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// =====================================================================
// Stage 1: W_l(alpha) orbit quadrature (BasisFunctions.jl:199-266)
// =====================================================================

// Per-orbit preprocessing, one workgroup per orbit k: validity (any
// non-finite w1/ra/pha skips the orbit, :224-226), trapezoid weights of the
// non-uniform w1 samples (:229), phi = (Om2/Om1) w1 - pha (:232-235).  Emits
// per sample: u = ln r, wt = Sw1 / (pi sqrt r) (0 where the integrand is
// non-finite, :248), w1 and m*phi, padded to nwa_p samples with wt = 0.
__global__ __launch_bounds__(256) void k_orbit_prep(
    const double* __restrict__ w1, const double* __restrict__ ra,
    const double* __restrict__ pha, const double* __restrict__ Om1,
    const double* __restrict__ Om2, int nwa, int nwa_p, int m, double* __restrict__ u,
    double* __restrict__ wt, double* __restrict__ w1o, double* __restrict__ mph,
    int* __restrict__ valid)
{
    const int64_t k = blockIdx.x;
    const double* W1 = w1 + k * nwa;
    const double* RA = ra + k * nwa;
    const double* PH = pha + k * nwa;
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < nwa; t += blockDim.x)
        if (!isfinite(W1[t]) || !isfinite(RA[t]) || !isfinite(PH[t])) atomicOr(&bad, 1);
    __syncthreads();
    const bool ok = (bad == 0);
    const double r21 = Om2[k] / Om1[k];
    const double inv_pi = 1.0 / 3.14159265358979323846;
    for (int t = threadIdx.x; t < nwa_p; t += blockDim.x) {
        double lu = 0.0, w = 0.0, x1 = 0.0, mp = 0.0;
        if (ok && t < nwa) {
            double sw;
            if (t == 0) sw = (W1[1] - W1[0]) / 2;
            else if (t == nwa - 1) sw = (W1[nwa - 1] - W1[nwa - 2]) / 2;
            else sw = (W1[t + 1] - W1[t - 1]) / 2;
            const double r = RA[t];
            const double phi = r21 * W1[t] - PH[t];
            x1 = W1[t];
            mp = (double)m * phi;
            lu = log(r);
            w = sw / sqrt(r) * inv_pi;
            if (!(r > 0.0) || !isfinite(lu) || !isfinite(w)) { lu = 0.0; w = 0.0; }
        }
        const int64_t o = k * nwa_p + t;
        u[o] = lu; wt[o] = w; w1o[o] = x1; mph[o] = mp;
    }
    if (threadIdx.x == 0) valid[k] = ok ? 1 : 0;
}

// W_k = A_k B_k per orbit on FP64 MFMA: A[l, t] = wt_t cos(l w1_t + m phi_t)
// (real, n_l x nwa), B[t, a] = exp(-i alpha_a ln r_t) (nwa x N).
// Workgroup = (orbit k, 32 alphas); wave w owns l-blocks w, w+4, w+8, w+12.
// Operands are generated in registers in the MFMA lane layout:
//   A (16 l x 4 t): lane = i + 16*kk holds A[l0+i][t0+kk]
//   B (4 t x 16 a): lane = i + 16*kk holds B[t0+kk][a0+i]
//   D (16 l x 16 a): lane holds D[kk + 4*v][i], v = 0..3 (the FP64 MFMA's
//   accumulator layout; the FP32 16x16x4 MFMA uses D[4*kk + v][i])
constexpr int kWMaxLB = 4;

__global__ __launch_bounds__(256) void k_W_f64(
    const double* __restrict__ u, const double* __restrict__ wt,
    const double* __restrict__ w1o, const double* __restrict__ mph,
    const int* __restrict__ valid, const double* __restrict__ alpha,
    const int* __restrict__ l_arr, int64_t NPh, int N, int Np, int n_l, int nwa_p,
    double2* __restrict__ W)
{
    const int64_t k = blockIdx.x;
    if (!valid[k]) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = lane & 15, kk = lane >> 4;
    const int a0 = blockIdx.y * 32;
    const int nLB = (n_l + 15) / 16;
    const bool bj0 = a0 < N, bj1 = a0 + 16 < N;
    double al[2];
    for (int j = 0; j < 2; ++j) {
        const int a = a0 + 16 * j + i;
        al[j] = a < N ? alpha[a] : 0.0;
    }
    double lv[kWMaxLB];
    for (int q = 0; q < kWMaxLB; ++q) {
        const int li = (wave + 4 * q) * 16 + i;
        lv[q] = li < n_l ? (double)l_arr[li] : 0.0;
    }
    v4d accr[kWMaxLB][2], acci[kWMaxLB][2];
    for (int q = 0; q < kWMaxLB; ++q)
        for (int j = 0; j < 2; ++j) { accr[q][j] = (v4d){0, 0, 0, 0}; acci[q][j] = accr[q][j]; }

    const double* U = u + k * nwa_p;
    const double* WT = wt + k * nwa_p;
    const double* X1 = w1o + k * nwa_p;
    const double* MP = mph + k * nwa_p;
    for (int t0 = 0; t0 < nwa_p; t0 += 4) {
        const int t = t0 + kk;
        const double ut = U[t], wtt = WT[t], w1t = X1[t], mp = MP[t];
        double br[2], bi[2];
        for (int j = 0; j < 2; ++j) {
            const double y = -al[j] * ut;
            sincos(y, &bi[j], &br[j]);
        }
#pragma unroll
        for (int q = 0; q < kWMaxLB; ++q) {
            if (wave + 4 * q >= nLB) break;
            const double a = wtt * cos(lv[q] * w1t + mp);
            if (bj0) {
                accr[q][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, br[0], accr[q][0], 0, 0, 0);
                acci[q][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bi[0], acci[q][0], 0, 0, 0);
            }
            if (bj1) {
                accr[q][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, br[1], accr[q][1], 0, 0, 0);
                acci[q][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bi[1], acci[q][1], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kWMaxLB; ++q) {
        const int lb = wave + 4 * q;
        if (lb >= nLB) break;
        for (int j = 0; j < 2; ++j) {
            const int a = a0 + 16 * j + i;
            if (a >= N) continue;
            for (int v = 0; v < 4; ++v) {
                const int li = lb * 16 + kk + 4 * v;   // f64 D layout: row = lane/16 + 4v
                if (li < n_l)
                    W[((int64_t)li * NPh + k) * Np + a] = make_double2(accr[q][j][v], acci[q][j][v]);
            }
        }
    }
}

// Device rows-of-K layout <-> reference [NPh, N, n_l] for one l-slab.
__global__ void k_W_slab_to_ref(const double2* __restrict__ W, int64_t NPh, int N, int Np,
                                int li, double2* __restrict__ out)
{
    const int64_t n = NPh * N;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = x % NPh, a = x / NPh;
        out[x] = W[((int64_t)li * NPh + k) * Np + a];
    }
}

__global__ void k_W_ref_to_slab(const double2* __restrict__ in, int64_t NPh, int N, int Np,
                                int li, double2* __restrict__ W)
{
    const int64_t n = NPh * N;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = x % NPh, a = x / NPh;
        W[((int64_t)li * NPh + k) * Np + a] = in[x];
    }
}

__global__ void k_W64_to_32(const double2* __restrict__ W, int64_t n, float2* __restrict__ Wf)
{
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x) {
        const double2 w = W[x];
        Wf[x] = make_float2((float)w.x, (float)w.y);
    }
}

// =====================================================================
// Stage 2: resonance weights c[b][K] (MatrixBuilder.jl:43-51)
// d = omega - Omega_res (|d| < 1e-12 -> 1e-12 + 1e-12 i); c = DJ*F0l / d.
// =====================================================================
__global__ void k_weights(const double* __restrict__ DJ, const double* __restrict__ F0l,
                          const double* __restrict__ Ores, int64_t NPh, int64_t K, int64_t Kp,
                          const double* __restrict__ om, int nb, double2* __restrict__ c,
                          int* __restrict__ counters)
{
    if (blockIdx.x == 0 && threadIdx.x < 8) counters[threadIdx.x] = 0;
    const int64_t n = (int64_t)nb * Kp;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int b = (int)(x / Kp);
        const int64_t q = x - (int64_t)b * Kp;
        double2 r = make_double2(0.0, 0.0);
        if (q < K) {
            double dr = om[2 * b] - Ores[q], di = om[2 * b + 1] - 0.0;
            if (hypot(dr, di) < 1e-12) { dr = 1e-12; di = 1e-12; }
            const double xw = DJ[q % NPh] * F0l[q];
            const double2 iv = d_inv(dr, di);
            r = make_double2(xw * iv.x, xw * iv.y);
        }
        c[x] = r;
    }
}

// =====================================================================
// Stage 3: the contraction M_w[beta, alpha] = sum_K W[K][beta] c_w[K]
// conj(W[K][alpha]) (MatrixBuilder.jl:119-141), split over K.
//
// Work decomposition.  A workgroup is 4 waves arranged 2x2; a wave owns an
// MB x NB grid of 16x16 output blocks for NW omegas over one K-split, so the
// workgroup tile is (32*MB) x (32*NB).  Waves read their MFMA operands
// straight from global memory -- four 256-byte row segments per 16x4
// operand, shared with the partner wave through L1 -- and prefetch the next
// K-step into registers before issuing the current step's MFMAs, so in
// steady state the matrix pipe never waits on a load.  A wave whose tile
// hangs over Np runs a compile-time specialisation with fewer blocks (no
// per-block branches in the inner loop).  Complex products use four real
// MFMAs per block:  Cr += Ar Br + (-Ai) Bi,  Ci += Ar Bi + Ai Br,  with
// B = c * conj(W_alpha) formed in registers.
//
// Grid: T*S workgroups, T output tiles x S K-splits (S % 8 == 0).  The
// linear block index is remapped so that the T tiles of one K-split land on
// the same XCD (blocks are dealt round-robin over the 8 XCDs), which keeps
// the W rows they share in that XCD's L2.  Partials go to
// P[w][s][alpha][beta] (column-major M per split).  FP64 accumulator lanes
// hold rows kk + 4v, FP32 ones rows 4kk + v.  Kp is a multiple of 8 and W
// and c carry kRowPad zero rows past Kp, so the loads issued for the step
// after the last one stay in bounds.
// =====================================================================

// One K-step (4 rows of K) of the wave tile: B = c conj(X) in registers, then
// 4*NW*MB*NB MFMAs, ordered so consecutive MFMAs hit different accumulators.
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
            Bv[w][n].x = c0 * X[n].x + c1 * X[n].y;
            Bv[w][n].y = c1 * X[n].x - c0 * X[n].y;
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

// Operand fetch of one K-step for a wave: MB row segments of the beta side,
// NB of the alpha side, NW weights.
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

// The K loop runs two K-steps per trip on two operand sets (ping-pong): the
// loads of step k+1 are issued before the MFMAs of step k (scheduling
// barriers keep the order), so a step's memory latency hides behind the
// previous step's MFMAs.  K ranges are multiples of 8 rows.
template <int MB, int NB, int NW>
__device__ __forceinline__ void wave_f64(const double2* __restrict__ W,
                                         const double2* __restrict__ cw,
                                         double2* __restrict__ P, int64_t Kpc, int Np,
                                         int64_t Kb, int64_t Ke, int b0, int a0, int s, int S,
                                         int lane)
{
    const int i = lane & 15, kk = lane >> 4;
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
    const double2* pc = cw + Kb + kk;
    double2 A0[MB], X0[NB], C0[NW], A1[MB], X1[NB], C1[NW];
    cload<MB, NB, NW>(pa, px, pc, Kpc, A0, X0, C0);
    const MfmaF64 mf;
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
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        double2* Pw = P + ((int64_t)w * S + s) * (int64_t)Np * Np;
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

// Picks the specialisation with exactly mv x nv valid blocks (1 <= mv <= MB).
template <int MB, int NB, int NW>
__device__ __forceinline__ void dispatch_f64(int mv, int nv, const double2* W, const double2* cw,
                                             double2* P, int64_t Kpc, int Np, int64_t Kb,
                                             int64_t Ke, int b0, int a0, int s, int S, int lane)
{
    if constexpr (MB > 1) {
        if (mv < MB) {
            dispatch_f64<MB - 1, NB, NW>(mv, nv, W, cw, P, Kpc, Np, Kb, Ke, b0, a0, s, S, lane);
            return;
        }
    }
    if constexpr (NB > 1) {
        if (nv < NB) {
            dispatch_f64<MB, NB - 1, NW>(mv, nv, W, cw, P, Kpc, Np, Kb, Ke, b0, a0, s, S, lane);
            return;
        }
    }
    wave_f64<MB, NB, NW>(W, cw, P, Kpc, Np, Kb, Ke, b0, a0, s, S, lane);
}

// Persistent form: the grid is exactly the resident workgroup count; each
// workgroup pulls (tile, split) units from the counter of its XCD (block
// index mod 8) until the XCD's units run out.  Unit u of XCD x is split
// x + 8*(u / T), tile u % T, so the tiles of a split run together on one XCD.
// counters[0..7] are zeroed by k_weights, launched just before on the stream.
template <int MB, int NB, int NW>
__global__ __launch_bounds__(256) void k_contract_f64(
    const double2* __restrict__ W, const double2* __restrict__ cw, double2* __restrict__ P,
    int64_t Kp, int Np, int64_t Ks, int S, int tiles_a, int T, int* __restrict__ counters)
{
    const int xcd = blockIdx.x & 7;
    const int units = (S / 8) * T;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __shared__ int s_unit;
    for (;;) {
        if (threadIdx.x == 0) s_unit = atomicAdd(&counters[xcd], 1);
        __syncthreads();
        const int u = s_unit;
        __syncthreads();
        if (u >= units) break;
        const int tile = u % T, s = xcd + 8 * (u / T);
        const int tb = tile / tiles_a, ta = tile % tiles_a;
        const int b0 = (tb * 2 + (wave >> 1)) * MB * 16;
        const int a0 = (ta * 2 + (wave & 1)) * NB * 16;
        const int mv = min(MB, (Np - b0) / 16), nv = min(NB, (Np - a0) / 16);
        if (mv <= 0 || nv <= 0) continue;
        const int64_t Kb = min((int64_t)s * Ks, Kp);
        const int64_t Ke = min(Kp, Kb + Ks);
        dispatch_f64<MB, NB, NW>(mv, nv, W, cw, P, Kp + kRowPad, Np, Kb, Ke, b0, a0, s, S, lane);
    }
}

// FP32 variant: W in complex64, weights rounded to FP32 after being formed
// in FP64, FP32 MFMA over K-tiles of kF32Tile, K-tile partials added into
// FP64 accumulators.
constexpr int kF32Tile = 256;

template <int MB, int NB, int NW>
__device__ __forceinline__ void wave_f32(const float2* __restrict__ W,
                                         const double2* __restrict__ cw,
                                         double2* __restrict__ P, int64_t Kpc, int Np,
                                         int64_t Kb, int64_t Ke, int b0, int a0, int s, int S,
                                         int lane)
{
    const int i = lane & 15, kk = lane >> 4;
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
    const double2* pc = cw + Kb + kk;
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
        double2* Pw = P + ((int64_t)w * S + s) * (int64_t)Np * Np;
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

template <int MB, int NB, int NW>
__device__ __forceinline__ void dispatch_f32(int mv, int nv, const float2* W, const double2* cw,
                                             double2* P, int64_t Kpc, int Np, int64_t Kb,
                                             int64_t Ke, int b0, int a0, int s, int S, int lane)
{
    if constexpr (MB > 1) {
        if (mv < MB) {
            dispatch_f32<MB - 1, NB, NW>(mv, nv, W, cw, P, Kpc, Np, Kb, Ke, b0, a0, s, S, lane);
            return;
        }
    }
    if constexpr (NB > 1) {
        if (nv < NB) {
            dispatch_f32<MB, NB - 1, NW>(mv, nv, W, cw, P, Kpc, Np, Kb, Ke, b0, a0, s, S, lane);
            return;
        }
    }
    wave_f32<MB, NB, NW>(W, cw, P, Kpc, Np, Kb, Ke, b0, a0, s, S, lane);
}

template <int MB, int NB, int NW>
__global__ __launch_bounds__(256) void k_contract_f32(
    const float2* __restrict__ W, const double2* __restrict__ cw, double2* __restrict__ P,
    int64_t Kp, int Np, int64_t Ks, int S, int tiles_a, int T, int* __restrict__ counters)
{
    const int xcd = blockIdx.x & 7;
    const int units = (S / 8) * T;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __shared__ int s_unit;
    for (;;) {
        if (threadIdx.x == 0) s_unit = atomicAdd(&counters[xcd], 1);
        __syncthreads();
        const int u = s_unit;
        __syncthreads();
        if (u >= units) break;
        const int tile = u % T, s = xcd + 8 * (u / T);
        const int tb = tile / tiles_a, ta = tile % tiles_a;
        const int b0 = (tb * 2 + (wave >> 1)) * MB * 16;
        const int a0 = (ta * 2 + (wave & 1)) * NB * 16;
        const int mv = min(MB, (Np - b0) / 16), nv = min(NB, (Np - a0) / 16);
        if (mv <= 0 || nv <= 0) continue;
        const int64_t Kb = min((int64_t)s * Ks, Kp);
        const int64_t Ke = min(Kp, Kb + Ks);
        dispatch_f32<MB, NB, NW>(mv, nv, W, cw, P, Kp + kRowPad, Np, Kb, Ke, b0, a0, s, S, lane);
    }
}

// Wave-tile shapes (MB x NB blocks) per precision and omega-batch size.
template <int NW> struct F64Shape;
#ifndef KLSP_F64_MB1
#define KLSP_F64_MB1 4
#define KLSP_F64_NB1 2
#endif
template <> struct F64Shape<1> { static constexpr int MB = KLSP_F64_MB1, NB = KLSP_F64_NB1; };
template <> struct F64Shape<2> { static constexpr int MB = 2, NB = 4; };
template <> struct F64Shape<3> { static constexpr int MB = 2, NB = 2; };
template <int NW> struct F32Shape;
template <> struct F32Shape<1> { static constexpr int MB = 4, NB = 2; };
template <> struct F32Shape<2> { static constexpr int MB = 2, NB = 2; };

// Sum of the S K-split partials (fixed order, FP64), then *G and *N[alpha]
// (MatrixBuilder.jl:144-149).  M is [nb][N, N] column-major.
__global__ void k_reduce(const double2* __restrict__ P, int S, int Np, int N, int nb, double G,
                         const double* __restrict__ Nk, double2* __restrict__ M)
{
    const int64_t NN = (int64_t)N * N, n = nb * NN;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int b = (int)(x / NN);
        const int64_t r = x - b * NN;
        const int beta = (int)(r % N), alpha = (int)(r / N);
        const double2* src = P + (int64_t)b * S * Np * Np + (int64_t)alpha * Np + beta;
        double sr = 0.0, si = 0.0;
        for (int s = 0; s < S; ++s) {
            const double2 v = src[(int64_t)s * Np * Np];
            sr += v.x;
            si += v.y;
        }
        const double g = Nk[alpha];
        M[x] = make_double2(sr * G * g, si * G * g);
    }
}

// =====================================================================
// Stage 4: det(I - M) by complex LU with partial pivoting, one workgroup
// per matrix (MatrixBuilder.jl:182 -> LinearAlgebra det -> LAPACK zgetrf).
// Pivot = first row maximising |re|+|im| (izamax), multipliers scaled by the
// reciprocal pivot, rank-1 trailing update; det = prod U_jj * (-1)^swaps
// accumulated in LinearAlgebra's order.  The working copy A = I - M lives in
// global memory (L2-resident); the pivot column and row are cached in LDS.
// =====================================================================
__device__ __forceinline__ void argmax_merge(double& v, int& idx, double ov, int oi)
{
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
}

__global__ __launch_bounds__(1024) void k_lu_det(const double2* __restrict__ M,
                                                  double2* __restrict__ Awork, int N,
                                                  double2* __restrict__ det_out)
{
    const int b = blockIdx.x;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int64_t NN = (int64_t)N * N;
    const double2* Mb = M + b * NN;
    double2* A = Awork + b * NN;
    __shared__ double2 colj[kMaxN];
    __shared__ double2 rowj[kMaxN];
    __shared__ double redv[32];
    __shared__ int redi[32];
    __shared__ int s_piv;

    for (int64_t x = tid; x < NN; x += nth) {
        const double2 m = Mb[x];
        const int r = (int)(x % N), c = (int)(x / N);
        A[x] = make_double2((r == c ? 1.0 : 0.0) - m.x, 0.0 - m.y);
    }
    __syncthreads();

    double2 det = make_double2(1.0, 0.0);
    int swaps = 0;
    bool singular = false;
    for (int j = 0; j < N; ++j) {
        // pivot search over rows j..N-1 of column j
        double bv = -1.0;
        int bi = N;
        for (int r = j + tid; r < N; r += nth) {
            const double2 a = A[r + (int64_t)N * j];
            const double v = fabs(a.x) + fabs(a.y);
            if (v > bv) { bv = v; bi = r; }
        }
        for (int off = 32; off >= 1; off >>= 1) {
            const double ov = __shfl_xor(bv, off);
            const int oi = __shfl_xor(bi, off);
            argmax_merge(bv, bi, ov, oi);
        }
        const int wid = tid >> 6;
        if ((tid & 63) == 0) { redv[wid] = bv; redi[wid] = bi; }
        __syncthreads();
        if (tid == 0) {
            double v = redv[0];
            int idx = redi[0];
            for (int w = 1; w < (nth + 63) / 64; ++w) argmax_merge(v, idx, redv[w], redi[w]);
            s_piv = (v > 0.0) ? idx : -1;
        }
        __syncthreads();
        const int p = s_piv;
        if (p < 0) { singular = true; __syncthreads(); continue; }
        if (p != j) {
            ++swaps;
            for (int c = j + tid; c < N; c += nth) {
                const double2 t = A[j + (int64_t)N * c];
                A[j + (int64_t)N * c] = A[p + (int64_t)N * c];
                A[p + (int64_t)N * c] = t;
            }
        }
        __syncthreads();
        const double2 piv = A[j + (int64_t)N * j];
        det = cmul(det, piv);
        const double2 rp = d_inv(piv.x, piv.y);
        for (int r = j + 1 + tid; r < N; r += nth) {
            const double2 l = cmul(A[r + (int64_t)N * j], rp);
            A[r + (int64_t)N * j] = l;
            colj[r] = l;
        }
        for (int c = j + 1 + tid; c < N; c += nth) rowj[c] = A[j + (int64_t)N * c];
        __syncthreads();
        const int n = N - j - 1;
        const int64_t nn = (int64_t)n * n;
        for (int64_t x = tid; x < nn; x += nth) {
            const int r = j + 1 + (int)(x % n), c = j + 1 + (int)(x / n);
            const double2 lu = cmul(colj[r], rowj[c]);
            double2* a = A + r + (int64_t)N * c;
            const double2 av = *a;
            *a = make_double2(av.x - lu.x, av.y - lu.y);
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (singular) det = make_double2(0.0, 0.0);
        else if (swaps & 1) det = make_double2(-det.x, -det.y);
        det_out[b] = det;
    }
}

// =====================================================================
// Host side: context, buffers, drivers
// =====================================================================
struct klsp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int n_cu = 256;
    // problem
    bool have_problem = false;
    int64_t NPh = 0, N = 0, n_l = 0, K = 0, Kp = 0;
    int Np = 0, m = 0;
    double G = 0.0;
    double *d_DJ = nullptr, *d_F0l = nullptr, *d_Ores = nullptr, *d_Nk = nullptr;
    // W
    double2* d_W64 = nullptr;
    float2* d_W32 = nullptr;
    bool have_W = false;
    // workspaces
    double* d_om = nullptr;
    double2* d_c = nullptr;   size_t c_cap = 0;
    double2* d_P = nullptr;   size_t P_cap = 0;
    double2* d_M = nullptr;   size_t M_cap = 0;
    double2* d_A = nullptr;   size_t A_cap = 0;
    double2* d_det = nullptr;
    int last_S = 0;           // K-splits of the last contraction launch
    int* d_counters = nullptr;  // per-XCD work counters of the persistent contraction
    std::vector<std::pair<const void*, int>> occ_cache;
    // profiling
    bool prof = false;
    hipEvent_t ev[6] = {};
    klsp_timing last{};
    double last_W_ms = 0.0;
};

static const char* g_noctx_err = "klsp: null context";

#define KLSP_HIP(ctx, call)                                                                 \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (ctx)->err = std::string("HIP error in ") + #call + ": " + hipGetErrorString(e_); \
            return e_ == hipErrorOutOfMemory ? KLSP_ERR_OOM : KLSP_ERR_HIP;                 \
        }                                                                                   \
    } while (0)

static int fail(klsp_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

// Grow-only device workspace.
template <typename T>
static int ensure(klsp_ctx* ctx, T*& ptr, size_t& cap, size_t n)
{
    if (n <= cap) return KLSP_OK;
    if (ptr) KLSP_HIP(ctx, hipFree(ptr));
    ptr = nullptr;
    cap = 0;
    KLSP_HIP(ctx, hipMalloc((void**)&ptr, n * sizeof(T)));
    cap = n;
    return KLSP_OK;
}

static int grid_for(int64_t n, int block = 256)
{
    const int64_t g = (n + block - 1) / block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 65536));
}

// K-split count and K per split for the contraction.
struct SplitPlan {
    int tiles_a, T, S;
    int64_t Ks;
};

// Output tiles of (32*MB) x (32*NB) and S K-splits (S % 8 == 0, K per split
// a multiple of kF32Tile), sized so the persistent grid sees ~6 units per
// resident workgroup (the tail of the last units stays short).
static SplitPlan plan_splits(const klsp_ctx* ctx, int MB, int NB)
{
    SplitPlan sp;
    const int tiles_b = (int)((ctx->Np + 32 * MB - 1) / (32 * MB));
    sp.tiles_a = (int)((ctx->Np + 32 * NB - 1) / (32 * NB));
    sp.T = tiles_b * sp.tiles_a;
    const int64_t target_units = (int64_t)ctx->n_cu * 2 * 6;
    int64_t S = (target_units + sp.T - 1) / sp.T;
    S = std::min<int64_t>(S, std::max<int64_t>(1, ctx->Kp / kF32Tile));
    S = round_up(std::max<int64_t>(S, 1), 8);
    sp.Ks = round_up((ctx->Kp + S - 1) / S, kF32Tile);
    const int64_t used = (ctx->Kp + sp.Ks - 1) / sp.Ks;
    sp.S = (int)round_up(std::max<int64_t>(used, 1), 8);
    return sp;
}

// Largest split count any shape can ask for (sizes the partials buffer).
static int max_splits(const klsp_ctx* ctx)
{
    int S = 0;
    const int shapes[][2] = {{4, 4}, {2, 4}, {2, 2}, {4, 2}};
    for (auto& sh : shapes) S = std::max(S, plan_splits(ctx, sh[0], sh[1]).S);
    return S;
}

static int free_problem(klsp_ctx* ctx)
{
    hipFree(ctx->d_DJ); hipFree(ctx->d_F0l); hipFree(ctx->d_Ores); hipFree(ctx->d_Nk);
    hipFree(ctx->d_W64); hipFree(ctx->d_W32);
    ctx->d_DJ = ctx->d_F0l = ctx->d_Ores = ctx->d_Nk = nullptr;
    ctx->d_W64 = nullptr;
    ctx->d_W32 = nullptr;
    ctx->have_problem = ctx->have_W = false;
    return KLSP_OK;
}

extern "C" {

const char* klsp_version(void) { return KLSP_VERSION_STRING; }

int klsp_device_count(int* n_out)
{
    if (!n_out) return KLSP_ERR_INVALID_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *n_out = n;
    return KLSP_OK;
}

int klsp_create(int device, klsp_ctx** ctx_out)
{
    if (!ctx_out) return KLSP_ERR_INVALID_ARG;
    *ctx_out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return KLSP_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return KLSP_ERR_INVALID_ARG;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return KLSP_ERR_HIP;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return KLSP_ERR_NO_DEVICE;
    klsp_ctx* ctx = new klsp_ctx();
    ctx->device = device;
    ctx->n_cu = prop.multiProcessorCount;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return KLSP_ERR_HIP;
    }
    for (auto& e : ctx->ev) hipEventCreate(&e);
    if (hipMalloc((void**)&ctx->d_om, sizeof(double) * 2 * kMaxBatch) != hipSuccess ||
        hipMalloc((void**)&ctx->d_det, sizeof(double2) * kMaxBatch) != hipSuccess ||
        hipMalloc((void**)&ctx->d_counters, sizeof(int) * 8) != hipSuccess) {
        delete ctx;
        return KLSP_ERR_OOM;
    }
    *ctx_out = ctx;
    return KLSP_OK;
}

int klsp_destroy(klsp_ctx* ctx)
{
    if (!ctx) return KLSP_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    free_problem(ctx);
    hipFree(ctx->d_om); hipFree(ctx->d_c); hipFree(ctx->d_P); hipFree(ctx->d_M);
    hipFree(ctx->d_A); hipFree(ctx->d_det); hipFree(ctx->d_counters);
    for (auto& e : ctx->ev) hipEventDestroy(e);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return KLSP_OK;
}

const char* klsp_last_error(const klsp_ctx* ctx) { return ctx ? ctx->err.c_str() : g_noctx_err; }

int klsp_upload_problem(klsp_ctx* ctx, const klsp_problem* pr)
{
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (!pr || pr->NPh < 1 || pr->N_alpha < 1 || pr->n_l < 1 || !pr->DJ || !pr->F0l ||
        !pr->Omega_res || !pr->N_kernel)
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_upload_problem: bad problem description");
    if (pr->N_alpha > kMaxN)
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_upload_problem: N_alpha exceeds 1024");
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    free_problem(ctx);
    ctx->NPh = pr->NPh;
    ctx->N = pr->N_alpha;
    ctx->n_l = pr->n_l;
    ctx->K = pr->NPh * pr->n_l;
    ctx->Kp = round_up(ctx->K, 2 * kKStep);
    ctx->Np = (int)round_up(ctx->N, kColPad);
    ctx->m = pr->m;
    ctx->G = pr->G;
    const size_t nK = (size_t)ctx->K;
    KLSP_HIP(ctx, hipMalloc((void**)&ctx->d_DJ, sizeof(double) * ctx->NPh));
    KLSP_HIP(ctx, hipMalloc((void**)&ctx->d_F0l, sizeof(double) * nK));
    KLSP_HIP(ctx, hipMalloc((void**)&ctx->d_Ores, sizeof(double) * nK));
    KLSP_HIP(ctx, hipMalloc((void**)&ctx->d_Nk, sizeof(double) * ctx->N));
    KLSP_HIP(ctx, hipMemcpy(ctx->d_DJ, pr->DJ, sizeof(double) * ctx->NPh, hipMemcpyHostToDevice));
    KLSP_HIP(ctx, hipMemcpy(ctx->d_F0l, pr->F0l, sizeof(double) * nK, hipMemcpyHostToDevice));
    KLSP_HIP(ctx, hipMemcpy(ctx->d_Ores, pr->Omega_res, sizeof(double) * nK, hipMemcpyHostToDevice));
    KLSP_HIP(ctx, hipMemcpy(ctx->d_Nk, pr->N_kernel, sizeof(double) * ctx->N, hipMemcpyHostToDevice));
    ctx->have_problem = true;
    return KLSP_OK;
}

static int alloc_W64(klsp_ctx* ctx)
{
    if (ctx->d_W32) { KLSP_HIP(ctx, hipFree(ctx->d_W32)); ctx->d_W32 = nullptr; }
    if (!ctx->d_W64)
        KLSP_HIP(ctx, hipMalloc((void**)&ctx->d_W64,
                                sizeof(double2) * (ctx->Kp + kRowPad) * ctx->Np));
    KLSP_HIP(ctx, hipMemsetAsync(ctx->d_W64, 0, sizeof(double2) * (ctx->Kp + kRowPad) * ctx->Np,
                                 ctx->stream));
    return KLSP_OK;
}

int klsp_compute_W(klsp_ctx* ctx, int64_t nwa, const double* w1, const double* ra,
                   const double* pha, const double* Om1, const double* Om2,
                   const double* alpha, const int32_t* l_arr, double* W_out)
{
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (!ctx->have_problem)
        return fail(ctx, KLSP_ERR_NOT_READY, "klsp_compute_W: upload the problem first");
    if (nwa < 2 || !w1 || !ra || !pha || !Om1 || !Om2 || !alpha || !l_arr)
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_compute_W: bad arguments");
    if (ctx->n_l > 16 * 4 * kWMaxLB)
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_compute_W: n_l exceeds 256");
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t NPh = ctx->NPh, nwa_p = round_up(nwa, 4);
    const size_t no = (size_t)NPh * nwa, np = (size_t)NPh * nwa_p;
    double *d_w1 = nullptr, *d_ra = nullptr, *d_pha = nullptr, *d_o1 = nullptr, *d_o2 = nullptr;
    double *d_u = nullptr, *d_wt = nullptr, *d_x1 = nullptr, *d_mp = nullptr, *d_al = nullptr;
    int *d_valid = nullptr, *d_l = nullptr;
    auto cleanup = [&]() {
        hipFree(d_w1); hipFree(d_ra); hipFree(d_pha); hipFree(d_o1); hipFree(d_o2);
        hipFree(d_u); hipFree(d_wt); hipFree(d_x1); hipFree(d_mp); hipFree(d_al);
        hipFree(d_valid); hipFree(d_l);
    };
    hipError_t e = hipSuccess;
#define KLSP_TRY(call) if (e == hipSuccess) e = (call)
    KLSP_TRY(hipMalloc((void**)&d_w1, no * 8));
    KLSP_TRY(hipMalloc((void**)&d_ra, no * 8));
    KLSP_TRY(hipMalloc((void**)&d_pha, no * 8));
    KLSP_TRY(hipMalloc((void**)&d_o1, NPh * 8));
    KLSP_TRY(hipMalloc((void**)&d_o2, NPh * 8));
    KLSP_TRY(hipMalloc((void**)&d_u, np * 8));
    KLSP_TRY(hipMalloc((void**)&d_wt, np * 8));
    KLSP_TRY(hipMalloc((void**)&d_x1, np * 8));
    KLSP_TRY(hipMalloc((void**)&d_mp, np * 8));
    KLSP_TRY(hipMalloc((void**)&d_al, ctx->N * 8));
    KLSP_TRY(hipMalloc((void**)&d_valid, NPh * 4));
    KLSP_TRY(hipMalloc((void**)&d_l, ctx->n_l * 4));
    KLSP_TRY(hipMemcpy(d_w1, w1, no * 8, hipMemcpyHostToDevice));
    KLSP_TRY(hipMemcpy(d_ra, ra, no * 8, hipMemcpyHostToDevice));
    KLSP_TRY(hipMemcpy(d_pha, pha, no * 8, hipMemcpyHostToDevice));
    KLSP_TRY(hipMemcpy(d_o1, Om1, NPh * 8, hipMemcpyHostToDevice));
    KLSP_TRY(hipMemcpy(d_o2, Om2, NPh * 8, hipMemcpyHostToDevice));
    KLSP_TRY(hipMemcpy(d_al, alpha, ctx->N * 8, hipMemcpyHostToDevice));
    KLSP_TRY(hipMemcpy(d_l, l_arr, ctx->n_l * 4, hipMemcpyHostToDevice));
#undef KLSP_TRY
    if (e != hipSuccess) {
        cleanup();
        return fail(ctx, e == hipErrorOutOfMemory ? KLSP_ERR_OOM : KLSP_ERR_HIP,
                    std::string("klsp_compute_W: ") + hipGetErrorString(e));
    }
    int rc = alloc_W64(ctx);
    if (rc != KLSP_OK) { cleanup(); return rc; }
    hipEventRecord(ctx->ev[0], ctx->stream);
    hipLaunchKernelGGL(k_orbit_prep, dim3((unsigned)NPh), dim3(256), 0, ctx->stream, d_w1, d_ra,
                       d_pha, d_o1, d_o2, (int)nwa, (int)nwa_p, ctx->m, d_u, d_wt, d_x1, d_mp,
                       d_valid);
    hipLaunchKernelGGL(k_W_f64, dim3((unsigned)NPh, (unsigned)((ctx->N + 31) / 32)), dim3(256), 0,
                       ctx->stream, d_u, d_wt, d_x1, d_mp, d_valid, d_al, d_l, NPh, (int)ctx->N,
                       ctx->Np, (int)ctx->n_l, (int)nwa_p, ctx->d_W64);
    hipEventRecord(ctx->ev[1], ctx->stream);
    e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipGetLastError();
    float ms = 0.f;
    hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    ctx->last_W_ms = ms;
    cleanup();
    if (e != hipSuccess)
        return fail(ctx, KLSP_ERR_HIP, std::string("klsp_compute_W: ") + hipGetErrorString(e));
    ctx->have_W = true;
    if (W_out) return klsp_download_W(ctx, W_out);
    return KLSP_OK;
}

int klsp_upload_W(klsp_ctx* ctx, const double* W)
{
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (!ctx->have_problem)
        return fail(ctx, KLSP_ERR_NOT_READY, "klsp_upload_W: upload the problem first");
    if (!W) return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_upload_W: null W");
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    int rc = alloc_W64(ctx);
    if (rc) return rc;
    const int64_t slab = ctx->NPh * ctx->N;
    double2* d_slab = nullptr;
    KLSP_HIP(ctx, hipMalloc((void**)&d_slab, sizeof(double2) * slab));
    for (int64_t li = 0; li < ctx->n_l; ++li) {
        hipError_t e = hipMemcpyAsync(d_slab, W + 2 * slab * li, sizeof(double2) * slab,
                                      hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) { hipFree(d_slab); KLSP_HIP(ctx, e); }
        hipLaunchKernelGGL(k_W_ref_to_slab, dim3(grid_for(slab)), dim3(256), 0, ctx->stream,
                           d_slab, ctx->NPh, (int)ctx->N, ctx->Np, (int)li, ctx->d_W64);
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    hipFree(d_slab);
    KLSP_HIP(ctx, e);
    ctx->have_W = true;
    return KLSP_OK;
}

int klsp_download_W(klsp_ctx* ctx, double* W_out)
{
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (!ctx->have_W) return fail(ctx, KLSP_ERR_NOT_READY, "klsp_download_W: no W on device");
    if (!W_out) return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_download_W: null output");
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t slab = ctx->NPh * ctx->N;
    double2* d_slab = nullptr;
    KLSP_HIP(ctx, hipMalloc((void**)&d_slab, sizeof(double2) * slab));
    for (int64_t li = 0; li < ctx->n_l; ++li) {
        hipLaunchKernelGGL(k_W_slab_to_ref, dim3(grid_for(slab)), dim3(256), 0, ctx->stream,
                           ctx->d_W64, ctx->NPh, (int)ctx->N, ctx->Np, (int)li, d_slab);
        hipError_t e = hipMemcpyAsync(W_out + 2 * slab * li, d_slab, sizeof(double2) * slab,
                                      hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) { hipFree(d_slab); KLSP_HIP(ctx, e); }
    }
    hipFree(d_slab);
    return KLSP_OK;
}

int klsp_prepare(klsp_ctx* ctx, int precision)
{
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (precision != KLSP_PREC_FP64 && precision != KLSP_PREC_FP32)
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_prepare: unknown precision");
    if (!ctx->have_W) return fail(ctx, KLSP_ERR_NOT_READY, "klsp_prepare: no W on device");
    if (precision == KLSP_PREC_FP32 && !ctx->d_W32) {
        KLSP_HIP(ctx, hipSetDevice(ctx->device));
        const int64_t n = (ctx->Kp + kRowPad) * ctx->Np;
        KLSP_HIP(ctx, hipMalloc((void**)&ctx->d_W32, sizeof(float2) * n));
        hipLaunchKernelGGL(k_W64_to_32, dim3(grid_for(n)), dim3(256), 0, ctx->stream,
                           ctx->d_W64, n, ctx->d_W32);
        KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return KLSP_OK;
}

}  // extern "C"

// Resident workgroups of a 256-thread kernel on the whole device (a multiple
// of 8, so every XCD counter has workgroups pulling from it).
static int resident_grid(klsp_ctx* ctx, const void* kernel)
{
    for (auto& e : ctx->occ_cache)
        if (e.first == kernel) return e.second;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    const int g = (int)round_up((int64_t)ctx->n_cu * per_cu, 8);
    ctx->occ_cache.emplace_back(kernel, g);
    return g;
}

template <int NW>
static void launch_f64(klsp_ctx* ctx)
{
    constexpr int MB = F64Shape<NW>::MB, NB = F64Shape<NW>::NB;
    const SplitPlan sp = plan_splits(ctx, MB, NB);
    const void* kfn = (const void*)k_contract_f64<MB, NB, NW>;
    hipLaunchKernelGGL((k_contract_f64<MB, NB, NW>), dim3((unsigned)resident_grid(ctx, kfn)),
                       dim3(256), 0, ctx->stream, ctx->d_W64, ctx->d_c, ctx->d_P, ctx->Kp, ctx->Np,
                       sp.Ks, sp.S, sp.tiles_a, sp.T, ctx->d_counters);
    ctx->last_S = sp.S;
}

template <int NW>
static void launch_f32(klsp_ctx* ctx)
{
    constexpr int MB = F32Shape<NW>::MB, NB = F32Shape<NW>::NB;
    const SplitPlan sp = plan_splits(ctx, MB, NB);
    const void* kfn = (const void*)k_contract_f32<MB, NB, NW>;
    hipLaunchKernelGGL((k_contract_f32<MB, NB, NW>), dim3((unsigned)resident_grid(ctx, kfn)),
                       dim3(256), 0, ctx->stream, ctx->d_W32, ctx->d_c, ctx->d_P, ctx->Kp, ctx->Np,
                       sp.Ks, sp.S, sp.tiles_a, sp.T, ctx->d_counters);
    ctx->last_S = sp.S;
}

// Builds M for nb <= kMaxBatch omegas into ctx->d_M ([nb][N,N]).
static int build_batch(klsp_ctx* ctx, const double* om, int nb, int precision)
{
    const int64_t Kpc = ctx->Kp + kRowPad;
    int rc;
    if ((rc = ensure(ctx, ctx->d_c, ctx->c_cap, (size_t)kMaxBatch * Kpc))) return rc;
    if ((rc = ensure(ctx, ctx->d_P, ctx->P_cap,
                     (size_t)kMaxBatch * max_splits(ctx) * ctx->Np * ctx->Np)))
        return rc;
    if ((rc = ensure(ctx, ctx->d_M, ctx->M_cap, (size_t)kMaxBatch * ctx->N * ctx->N))) return rc;
    KLSP_HIP(ctx, hipMemcpyAsync(ctx->d_om, om, sizeof(double) * 2 * nb, hipMemcpyHostToDevice,
                                 ctx->stream));
    if (ctx->prof) hipEventRecord(ctx->ev[0], ctx->stream);
    hipLaunchKernelGGL(k_weights, dim3(grid_for((int64_t)nb * Kpc)), dim3(256), 0, ctx->stream,
                       ctx->d_DJ, ctx->d_F0l, ctx->d_Ores, ctx->NPh, ctx->K, Kpc, ctx->d_om, nb,
                       ctx->d_c, ctx->d_counters);
    if (ctx->prof) hipEventRecord(ctx->ev[1], ctx->stream);
    if (precision == KLSP_PREC_FP64) {
        if (nb == 1) launch_f64<1>(ctx);
        else if (nb == 2) launch_f64<2>(ctx);
        else launch_f64<3>(ctx);
    } else {
        if (nb == 1) launch_f32<1>(ctx);
        else launch_f32<2>(ctx);
    }
    if (ctx->prof) hipEventRecord(ctx->ev[2], ctx->stream);
    hipLaunchKernelGGL(k_reduce, dim3(grid_for((int64_t)nb * ctx->N * ctx->N)), dim3(256), 0,
                       ctx->stream, ctx->d_P, ctx->last_S, ctx->Np, (int)ctx->N, nb, ctx->G,
                       ctx->d_Nk, ctx->d_M);
    if (ctx->prof) hipEventRecord(ctx->ev[3], ctx->stream);
    KLSP_HIP(ctx, hipGetLastError());
    return KLSP_OK;
}

static void add_timing(klsp_ctx* ctx, bool with_lu, int nb)
{
    if (!ctx->prof) return;
    hipEventSynchronize(with_lu ? ctx->ev[4] : ctx->ev[3]);
    float a = 0, b = 0, c = 0, d = 0;
    hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]);
    hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]);
    hipEventElapsedTime(&c, ctx->ev[2], ctx->ev[3]);
    if (with_lu) hipEventElapsedTime(&d, ctx->ev[3], ctx->ev[4]);
    ctx->last.weights_ms += a;
    ctx->last.contract_ms += b;
    ctx->last.reduce_ms += c;
    ctx->last.lu_ms += d;
    ctx->last.n_omega += nb;
    ctx->last.contract_launches += 1;
}

static int check_ready(klsp_ctx* ctx, int precision, const char* who)
{
    if (!ctx->have_problem || !ctx->have_W)
        return fail(ctx, KLSP_ERR_NOT_READY, std::string(who) + ": problem and W must be on device");
    if (precision != KLSP_PREC_FP64 && precision != KLSP_PREC_FP32)
        return fail(ctx, KLSP_ERR_INVALID_ARG, std::string(who) + ": unknown precision");
    if (precision == KLSP_PREC_FP32 && !ctx->d_W32) return klsp_prepare(ctx, precision);
    return KLSP_OK;
}

// det(I - M(omega_i)) for n omegas; the workhorse of klsp_det/newton/scan.
static int det_many(klsp_ctx* ctx, const double* om, int64_t n, int precision, double* out)
{
    int rc = check_ready(ctx, precision, "klsp_det");
    if (rc) return rc;
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = ensure(ctx, ctx->d_A, ctx->A_cap, (size_t)kMaxBatch * ctx->N * ctx->N))) return rc;
    const int maxb = precision == KLSP_PREC_FP64 ? 3 : 2;
    for (int64_t i0 = 0; i0 < n; i0 += maxb) {
        const int nb = (int)std::min<int64_t>(maxb, n - i0);
        if ((rc = build_batch(ctx, om + 2 * i0, nb, precision))) return rc;
        hipLaunchKernelGGL(k_lu_det, dim3(nb), dim3(1024), 0, ctx->stream, ctx->d_M, ctx->d_A,
                           (int)ctx->N, ctx->d_det);
        if (ctx->prof) hipEventRecord(ctx->ev[4], ctx->stream);
        KLSP_HIP(ctx, hipGetLastError());
        KLSP_HIP(ctx, hipMemcpyAsync(out + 2 * i0, ctx->d_det, sizeof(double2) * nb,
                                     hipMemcpyDeviceToHost, ctx->stream));
        KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        add_timing(ctx, true, nb);
    }
    return KLSP_OK;
}

static void reset_timing(klsp_ctx* ctx) { ctx->last = klsp_timing{}; }

extern "C" {

int klsp_build_M(klsp_ctx* ctx, double om_re, double om_im, int precision, double* M_out)
{
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (!M_out) return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_build_M: null output");
    int rc = check_ready(ctx, precision, "klsp_build_M");
    if (rc) return rc;
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    reset_timing(ctx);
    const double om[2] = {om_re, om_im};
    if ((rc = build_batch(ctx, om, 1, precision))) return rc;
    KLSP_HIP(ctx, hipMemcpyAsync(M_out, ctx->d_M, sizeof(double2) * ctx->N * ctx->N,
                                 hipMemcpyDeviceToHost, ctx->stream));
    KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    add_timing(ctx, false, 1);
    return KLSP_OK;
}

int klsp_det(klsp_ctx* ctx, const double* omegas, int64_t n, int precision, double* dets_out)
{
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (n < 0 || (n > 0 && (!omegas || !dets_out)))
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_det: bad arguments");
    reset_timing(ctx);
    if (n == 0) return KLSP_OK;
    return det_many(ctx, omegas, n, precision, dets_out);
}

// newton_search, NewtonSolver.jl:36-172, with the forward-difference pair
// (omega + m*delta, omega + i*delta) evaluated as one two-omega batch.
int klsp_newton(klsp_ctx* ctx, const klsp_newton_params* np, klsp_newton_result* res,
                double* history)
{
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (!np || !res || np->max_iter < 0 || np->max_line_search < 0)
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_newton: bad arguments");
    int rc = check_ready(ctx, np->precision, "klsp_newton");
    if (rc) return rc;
    reset_timing(ctx);
    const double m = (double)ctx->m;
    const int prec = np->precision;
    double Op = np->Omega_p_init, gv = np->gamma_init;
    const double tol = np->tol, delta = np->delta;
    int conv = 0, iter_done = 0, n_eval = 0;
    klsp_timing acc{};
    auto det1 = [&](double wr, double wi, double* d) -> int {
        const double om[2] = {wr, wi};
        ++n_eval;
        int r = det_many(ctx, om, 1, prec, d);
        acc.weights_ms += ctx->last.weights_ms; acc.contract_ms += ctx->last.contract_ms;
        acc.reduce_ms += ctx->last.reduce_ms; acc.lu_ms += ctx->last.lu_ms;
        acc.n_omega += ctx->last.n_omega; acc.contract_launches += ctx->last.contract_launches;
        ctx->last = klsp_timing{};
        return r;
    };
    auto det2 = [&](const double* om4, double* d4) -> int {
        n_eval += 2;
        int r = det_many(ctx, om4, 2, prec, d4);
        acc.weights_ms += ctx->last.weights_ms; acc.contract_ms += ctx->last.contract_ms;
        acc.reduce_ms += ctx->last.reduce_ms; acc.lu_ms += ctx->last.lu_ms;
        acc.n_omega += ctx->last.n_omega; acc.contract_launches += ctx->last.contract_launches;
        ctx->last = klsp_timing{};
        return r;
    };
    for (int iter = 1; iter <= np->max_iter; ++iter) {
        iter_done = iter;
        double d[2];
        if ((rc = det1(m * Op, gv, d))) return rc;
        const double res_norm = std::hypot(d[0], d[1]);
        if (history) {
            history[iter - 1] = Op;
            history[iter - 1 + np->max_iter] = gv;
            history[iter - 1 + 2 * np->max_iter] = std::hypot(d[0], d[1]);
        }
        if (res_norm < tol) { conv = 1; break; }
        const double om4[4] = {m * (Op + delta), gv, m * Op, gv + delta};
        double d4[4];
        if ((rc = det2(om4, d4))) return rc;
        const double J11 = (d4[0] - d[0]) / delta, J12 = (d4[2] - d[0]) / delta;
        const double J21 = (d4[1] - d[1]) / delta, J22 = (d4[3] - d[1]) / delta;
        // delta_x = -J \ F by 2x2 LU with partial pivoting
        double a11 = J11, a12 = J12, a21 = J21, a22 = J22, b1 = d[0], b2 = d[1];
        if (std::fabs(a21) > std::fabs(a11)) {
            std::swap(a11, a21); std::swap(a12, a22); std::swap(b1, b2);
        }
        const double l21 = a21 * (1.0 / a11);
        const double x2 = (b2 - l21 * b1) / (a22 - l21 * a12);
        const double x1 = (b1 - a12 * x2) / a11;
        const double dx1 = -x1, dx2 = -x2;
        if (np->use_line_search) {
            double al = 1.0;
            for (int ls = 1; ls <= np->max_line_search; ++ls) {
                const double Opn = Op + al * dx1, gn = gv + al * dx2;
                if (gn < 1e-4 || Opn < 0.0 || Opn > 1.0) { al *= 0.5; continue; }
                double dn[2];
                if ((rc = det1(m * Opn, gn, dn))) return rc;
                if (std::hypot(dn[0], dn[1]) < res_norm) { Op = Opn; gv = gn; break; }
                al *= 0.5;
                if (ls == np->max_line_search) { Op += dx1; gv += dx2; }
            }
        } else {
            Op += dx1;
            gv += dx2;
        }
        if (!std::isnan(gv) && gv < 1e-4) gv = 1e-4;
        if (Op > 1.0) Op = 1.0; else if (Op < 0.0) Op = 0.0;
    }
    double df[2];
    if ((rc = det1(m * Op, gv, df))) return rc;
    res->Omega_p = Op;
    res->gamma = gv;
    res->omega[0] = m * Op;
    res->omega[1] = gv;
    res->det[0] = df[0];
    res->det[1] = df[1];
    res->iterations = iter_done;
    res->converged = conv;
    res->n_det_evals = n_eval;
    ctx->last = acc;
    return KLSP_OK;
}

int klsp_grid_scan(klsp_ctx* ctx, const klsp_scan_params* sp, double* det_grid,
                   double* Op_init, double* g_init)
{
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (!sp || sp->n_Op < 1 || sp->n_gamma < 1 || !det_grid)
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_grid_scan: bad arguments");
    int rc = check_ready(ctx, sp->precision, "klsp_grid_scan");
    if (rc) return rc;
    reset_timing(ctx);
    std::vector<double> Opg(sp->n_Op), gg(sp->n_gamma);
    klsp_linrange(sp->Omega_p_min, sp->Omega_p_max, sp->n_Op, Opg.data());
    klsp_linrange(sp->gamma_min, sp->gamma_max, sp->n_gamma, gg.data());
    const int64_t n = (int64_t)sp->n_Op * sp->n_gamma;
    std::vector<double> om(2 * n);
    for (int j = 0; j < sp->n_Op; ++j)
        for (int i = 0; i < sp->n_gamma; ++i) {
            const int64_t x = i + (int64_t)sp->n_gamma * j;
            om[2 * x] = ctx->m * Opg[j];
            om[2 * x + 1] = gg[i];
        }
    if ((rc = det_many(ctx, om.data(), n, sp->precision, det_grid))) return rc;
    int64_t best = 0;
    double bv = INFINITY;
    for (int64_t x = 0; x < n; ++x) {
        const double v = std::hypot(det_grid[2 * x], det_grid[2 * x + 1]);
        if (v < bv) { bv = v; best = x; }
    }
    if (Op_init) *Op_init = Opg[best / sp->n_gamma];
    if (g_init) *g_init = gg[best % sp->n_gamma];
    return KLSP_OK;
}

int klsp_set_profiling(klsp_ctx* ctx, int on)
{
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    ctx->prof = on != 0;
    return KLSP_OK;
}

int klsp_last_timing(klsp_ctx* ctx, klsp_timing* t)
{
    if (!ctx || !t) return KLSP_ERR_INVALID_ARG;
    *t = ctx->last;
    return KLSP_OK;
}

int klsp_last_W_ms(klsp_ctx* ctx, double* ms)
{
    if (!ctx || !ms) return KLSP_ERR_INVALID_ARG;
    *ms = ctx->last_W_ms;
    return KLSP_OK;
}

}  // extern "C"
