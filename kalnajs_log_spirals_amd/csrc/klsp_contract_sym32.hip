// klsp_contract_sym32.hip -- the M(omega) contraction on a symmetric alpha
// grid in the real basis (see klsp_contract_sym64.hip), on FP32 MFMA
// 16x16x4, for FP32 builds.
//
// R is stored in FP32 (K x Np floats, 4 bytes per entry); B = Re c * x and
// Im c * x are formed in FP32 from the complex64 weights.  Products
// accumulate in FP32 over K-tiles of kF32Tile rows and the tile sums in FP64,
// as in the complex FP32 kernel (wave_f32 in klsp_contract.hpp), so an FP32
// build keeps that kernel's accuracy with half its MFMAs per output block
// and half its bytes per entry, on the upper triangle only.
//
// Work decomposition: single-wave workgroups own a 4 x 2 tile of 16 x 16
// blocks of P and Q (16 FP32 tile accumulators and 16 FP64 sums: the
// register budget of wave_f32) over a K range read from the balanced work
// table.  Tile (r, c) covers row blocks 4r..4r+3 and column blocks 2c, 2c+1
// and exists when any of its blocks lies on or above the diagonal
// (4r <= 2c + 1); its blocks below the diagonal are computed and never read
// (the reduce reads blocks bi <= bj).  Operands are loaded kSym32Depth - 1
// K-steps ahead.
#include <algorithm>

// Build-time knobs (defaults are the measured best at the default grid):
// operand sets in flight, tile shape in 16 x 16 blocks, and an optional
// waves-per-SIMD request (2 fits the 4 x 2 tile in 247 VGPRs without AGPRs;
// measured no faster, 13.8 vs 13.7 ms).
#ifndef KLSP_SYM32_DEPTH
#define KLSP_SYM32_DEPTH 4
#endif
#ifndef KLSP_SYM32_MB
#define KLSP_SYM32_MB 4
#endif
#ifndef KLSP_SYM32_NB
#define KLSP_SYM32_NB 2
#endif

#include "klsp_contract.hpp"

namespace klsp {
namespace {

constexpr int kS32MB = KLSP_SYM32_MB, kS32NB = KLSP_SYM32_NB;
constexpr int kSym32Depth = KLSP_SYM32_DEPTH;

template <int MB, int NB>
__device__ __forceinline__ void s32_load(const float* pa, const float* px, const float2* pc,
                                         float (&A)[MB], float (&X)[NB], float2& C)
{
#pragma unroll
    for (int m = 0; m < MB; ++m) A[m] = pa[16 * m];
#pragma unroll
    for (int n = 0; n < NB; ++n) X[n] = px[16 * n];
    C = *pc;
}

// One K-step (4 rows): the P MFMAs, then the Q MFMAs (consecutive MFMAs hit
// different accumulators).
template <int MB, int NB>
__device__ __forceinline__ void s32_step(const float (&A)[MB], const float (&X)[NB], float2 C,
                                         v4f (&P)[MB][NB], v4f (&Q)[MB][NB])
{
    float Bp[NB], Bq[NB];
#pragma unroll
    for (int n = 0; n < NB; ++n) {
        Bp[n] = C.x * X[n];
        Bq[n] = C.y * X[n];
    }
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n)
            P[m][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[m], Bp[n], P[m][n], 0, 0, 0);
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n)
            Q[m][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[m], Bq[n], Q[m][n], 0, 0, 0);
}

// Rows [Kb, Ke) of R into partial slice `slot` for the tile whose row blocks
// start at b0 and column blocks at a0.  D operand sets are in flight (loads
// D - 1 K-steps ahead of the MFMAs); Kb and Ke are multiples of D * kKStep
// rows except Ke = Kp (a multiple of 8): the last trip then computes at most
// 4 D - 8 rows past Ke and loads at most 8 D - 13 rows past it, inside the
// kRowPad zero rows of R and c.  The FP32 tile accumulators are added into
// the FP64 sums every kF32Tile rows (a whole number of trips) and at the end.
template <int D, int MB, int NB>
__device__ __forceinline__ void wave_sym32(const ContractArgs& a, int64_t Kb, int64_t Ke, int b0,
                                           int a0, int slot)
{
    static_assert(8 * D - 13 < kRowPad, "prefetch past Kp exceeds the zero padding rows");
    static_assert(kF32Tile % (D * kKStep) == 0, "an FP32 tile is a whole number of trips");
    const float* __restrict__ R = (const float*)a.W;
    const int lane = threadIdx.x & 63, i = lane & 15, kk = lane >> 4;
    const int Np = a.Np;
    v4d dp[MB][NB], dq[MB][NB];
    v4f cp[MB][NB], cq[MB][NB];
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n) {
            dp[m][n] = (v4d){0, 0, 0, 0};
            dq[m][n] = (v4d){0, 0, 0, 0};
            cp[m][n] = (v4f){0, 0, 0, 0};
            cq[m][n] = (v4f){0, 0, 0, 0};
        }
    const int64_t rstep = (int64_t)kKStep * Np;
    const float* pa = R + (Kb + kk) * Np + b0 + i;
    const float* px = R + (Kb + kk) * Np + a0 + i;
    const float2* pc = a.cwf + Kb + kk;
    float A[D][MB], X[D][NB];
    float2 C[D];
    constexpr int tile_trips = kF32Tile / (D * kKStep);
    const int ntrip = (int)((Ke - Kb + D * kKStep - 1) / (D * kKStep));
    if (ntrip > 0) {
#pragma unroll
        for (int s = 0; s < D - 1; ++s) {
            s32_load<MB, NB>(pa + s * rstep, px + s * rstep, pc + s * kKStep, A[s], X[s], C[s]);
            __builtin_amdgcn_sched_barrier(0);
        }
        int tt = 0;
        for (int trip = 0; trip < ntrip; ++trip) {
#pragma unroll
            for (int s = 0; s < D; ++s) {
                constexpr int ahead = D - 1;
                const int ls = (s + ahead) % D;
                s32_load<MB, NB>(pa + (s + ahead) * rstep, px + (s + ahead) * rstep,
                                 pc + (s + ahead) * kKStep, A[ls], X[ls], C[ls]);
                __builtin_amdgcn_sched_barrier(0);
                s32_step<MB, NB>(A[s], X[s], C[s], cp, cq);
                __builtin_amdgcn_sched_barrier(0);
            }
            pa += D * rstep;
            px += D * rstep;
            pc += D * kKStep;
            if (++tt == tile_trips || trip + 1 == ntrip) {
                tt = 0;
#pragma unroll
                for (int m = 0; m < MB; ++m)
#pragma unroll
                    for (int n = 0; n < NB; ++n) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            dp[m][n][v] += (double)cp[m][n][v];
                            dq[m][n][v] += (double)cq[m][n][v];
                        }
                        cp[m][n] = (v4f){0, 0, 0, 0};
                        cq[m][n] = (v4f){0, 0, 0, 0};
                    }
            }
        }
    }
    // FP32 accumulator lanes hold rows 4 kk + v of a block; slices are
    // column-major (column index = the B-side column of R)
    double2* Ps = a.P + (int64_t)slot * Np * Np;
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n) {
            double2* dst = Ps + (int64_t)(a0 + 16 * n + i) * Np + b0 + 16 * m + 4 * kk;
#pragma unroll
            for (int v = 0; v < 4; ++v) dst[v] = make_double2(dp[m][n][v], dq[m][n][v]);
        }
}

// compile-time specialisations with exactly mv x nv valid blocks
template <int D, int MB, int NB>
__device__ __forceinline__ void dispatch32(int mv, int nv, const ContractArgs& a, int64_t Kb,
                                           int64_t Ke, int b0, int a0, int slot)
{
    if constexpr (MB > 1) {
        if (mv < MB) {
            dispatch32<D, MB - 1, NB>(mv, nv, a, Kb, Ke, b0, a0, slot);
            return;
        }
    }
    if constexpr (NB > 1) {
        if (nv < NB) {
            dispatch32<D, MB, NB - 1>(mv, nv, a, Kb, Ke, b0, a0, slot);
            return;
        }
    }
    wave_sym32<D, MB, NB>(a, Kb, Ke, b0, a0, slot);
}

#ifdef KLSP_SYM32_WPE
#define KLSP_SYM32_ATTR __attribute__((amdgpu_waves_per_eu(KLSP_SYM32_WPE, KLSP_SYM32_WPE)))
#else
#define KLSP_SYM32_ATTR
#endif
template <int D>
__global__ __launch_bounds__(64) KLSP_SYM32_ATTR void k_contract_sym32(ContractArgs a)
{
    const WorkUnit u = a.units[blockIdx.x];
    if (u.widx < 0) return;
    const int tr = u.widx / a.wcols, tc = u.widx % a.wcols;
    const int nb = a.Np / 16;
    const int b0 = tr * kS32MB * 16, a0 = tc * kS32NB * 16;
    const int mv = min(kS32MB, nb - tr * kS32MB), nv = min(kS32NB, nb - tc * kS32NB);
    dispatch32<D, kS32MB, kS32NB>(mv, nv, a, u.Kb, u.Ke, b0, a0, u.slot);
}

// The real basis in FP32: R[k][j] = Re W[k][h + j] for j < nU = N - h,
// Im W[k][j] for nU <= j < N, 0 in the padding columns (k_W64_to_R's layout).
__global__ void k_W64_to_R32(const double2* __restrict__ W, int64_t rows, int N, int Np,
                             float* __restrict__ R)
{
    const int h = N >> 1, nU = N - h;
    const int64_t n = rows * Np;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = x / Np;
        const int j = (int)(x - k * Np);
        const double2* row = W + k * Np;
        R[x] = (float)(j < nU ? row[h + j].x : (j < N ? row[j].y : 0.0));
    }
}

}  // namespace

ContractShape contract_sym32_shape()
{
    static std::atomic<int> occ{0};
    if (occ == 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, (const void*)k_contract_sym32<kSym32Depth>, 64, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 8;
        occ = per_cu;
    }
    return {kS32MB, kS32NB, occ};
}

int contract_sym32_quantum() { return kSym32Depth * kKStep; }

hipError_t contract_sym32_launch(const ContractArgs& a, hipStream_t stream)
{
    ContractArgs args = a;
    void* params[] = {(void*)&args};
    return hipLaunchKernel((const void*)k_contract_sym32<kSym32Depth>, dim3((unsigned)a.n_units),
                           dim3(64), params, 0, stream);
}

hipError_t split_R_f32(const double2* W64, int64_t rows, int N, int Np, float* R32,
                       hipStream_t stream)
{
    const int64_t n = rows * Np;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_W64_to_R32, dim3(grid), dim3(256), 0, stream, W64, rows, N, Np, R32);
    return hipGetLastError();
}

}  // namespace klsp
