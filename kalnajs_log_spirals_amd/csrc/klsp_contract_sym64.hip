// klsp_contract_sym64.hip -- the M(omega) contraction on a symmetric alpha
// grid, in a real basis, on FP64 MFMA 16x16x4 (one omega per pass).
//
// The reference's alpha grid range(-alpha_max, alpha_max, N) is symmetric,
// and there W_l(-alpha) = conj W_l(alpha) (BasisFunctions.jl:240-245: the
// l-side factor is real).  With h = N / 2 and nU = N - h, the real K x N
// basis
//   R[:, j] = Re W[:, h + j]   (j < nU: alpha >= 0)
//   R[:, j] = Im W[:, j]       (nU <= j < N: alpha > 0)
// gives every column of W as W[:, a] = R[:, u(a)] + i s(a) R[:, v(a)]
// (a' = max(a, N - 1 - a), u = a' - h, v = a' if a' >= nU (no such term for
// the real alpha = 0 column), s = +1 for a = a', -1 for the mirrored half).
// Then the contraction of MatrixBuilder.jl:119-141,
//   H = sum_K W[K, :]^T c[K] conj W[K, :],
// is H[b, a] = S[ub, ua] + i sb S[vb, ua] - i sa S[ub, va] + sb sa S[vb, va]
// with S = P + i Q and two real symmetric products over K,
//   P = R^T diag(Re c) R,   Q = R^T diag(Im c) R.
// Only the upper-triangle blocks of P and Q are computed: per pair of 16 x 16
// blocks and K-step 2 MFMAs (P and Q share the A operand) instead of the 4 of
// the complex product, on about half the block pairs, and R is 8 bytes per
// entry instead of 16.  The expansion of S into M is the reduce's epilogue.
//
// Work decomposition as in klsp_contract.hpp: single-wave workgroups own a
// 4 x 4 tile of blocks of P and Q (32 accumulator blocks: the 256 AGPRs) over
// a K range read from the balanced work table; only tiles on or above the
// diagonal exist, and a diagonal tile computes its blocks m <= n and loads
// one operand set (its row and column ranges of R are the same).  Operands
// are loaded D - 1 K-steps ahead of the MFMAs.  The compiled variants (D and
// the diagonal tiles' form) are chosen per problem by the autotuner
// (klsp_autotune) or by the default.
#include "klsp_contract.hpp"

namespace klsp {
namespace {

constexpr int kSymMB = 4;      // tile edge in blocks

// DM: 0 = off-diagonal tile; 1 = diagonal tile, blocks m <= n; 2 = diagonal
// tile, all blocks.  Diagonal tiles load one operand set (X is A).
template <int MB, int NB, int DM>
__device__ __forceinline__ void sym_load(const double* pa, const double* px, const double2* pc,
                                         double (&A)[MB], double (&X)[NB], double2& C)
{
#pragma unroll
    for (int m = 0; m < MB; ++m) A[m] = pa[16 * m];
    if constexpr (DM == 0) {
#pragma unroll
        for (int n = 0; n < NB; ++n) X[n] = px[16 * n];
    }
    C = *pc;
}

// One K-step (4 rows): Bp = Re c * x, Bq = Im c * x, then the P MFMAs and
// the Q MFMAs (consecutive MFMAs hit different accumulators).
template <int MB, int NB, int DM>
__device__ __forceinline__ void sym_step(const double (&A)[MB], const double (&X)[NB], double2 C,
                                         v4d (&P)[MB][NB], v4d (&Q)[MB][NB])
{
    double Bp[NB], Bq[NB];
#pragma unroll
    for (int n = 0; n < NB; ++n) {
        double x;
        if constexpr (DM != 0) x = A[n];
        else x = X[n];
        Bp[n] = C.x * x;
        Bq[n] = C.y * x;
    }
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n)
            if (DM != 1 || m <= n)
                P[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[m], Bp[n], P[m][n], 0, 0, 0);
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n)
            if (DM != 1 || m <= n)
                Q[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[m], Bq[n], Q[m][n], 0, 0, 0);
}

// Rows [Kb, Ke) of R into partial slice `slot` for the tile whose row blocks
// start at b0 and column blocks at a0 (diagonal modes: a0 == b0, MB == NB).  Kb and Ke
// are multiples of D * kKStep rows except Ke = Kp (a multiple of 8):
// the last trip then computes at most 4 D - 8 rows past Ke and loads at most
// 8 D - 13 rows past it (8 and 19 at D = 4), inside the kRowPad zero rows of
// R and c.
template <int D, int MB, int NB, int DM>
__device__ __forceinline__ void wave_sym64(const ContractArgs& a, int64_t Kb, int64_t Ke, int b0,
                                           int a0, int slot)
{
    static_assert(DM == 0 || MB == NB, "a diagonal tile is square");
    static_assert(8 * D - 13 < kRowPad, "prefetch past Kp exceeds the zero padding rows");
    const double* __restrict__ R = (const double*)a.W;
    const int lane = threadIdx.x & 63, i = lane & 15, kk = lane >> 4;
    const int Np = a.Np;
    v4d P[MB][NB], Q[MB][NB];
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n) {
            P[m][n] = (v4d){0, 0, 0, 0};
            Q[m][n] = (v4d){0, 0, 0, 0};
        }
    const int64_t rstep = (int64_t)kKStep * Np;
    const double* pa = R + (Kb + kk) * Np + b0 + i;
    const double* px = R + (Kb + kk) * Np + a0 + i;
    const double2* pc = a.cw + Kb + kk;
    double A[D][MB], X[D][NB];
    double2 C[D];
    // the prologue behind the trip test, as in the complex kernel, so the
    // wait at the loop head covers exactly one operand set
    const int ntrip = (int)((Ke - Kb + D * kKStep - 1) / (D * kKStep));
    if (ntrip > 0) {
#pragma unroll
        for (int s = 0; s < D - 1; ++s) {
            sym_load<MB, NB, DM>(pa + s * rstep, px + s * rstep, pc + s * kKStep, A[s], X[s], C[s]);
            __builtin_amdgcn_sched_barrier(0);
        }
        for (int trip = 0; trip < ntrip; ++trip) {
#pragma unroll
            for (int s = 0; s < D; ++s) {
                constexpr int ahead = D - 1;
                const int ls = (s + ahead) % D;
                sym_load<MB, NB, DM>(pa + (s + ahead) * rstep, px + (s + ahead) * rstep,
                                       pc + (s + ahead) * kKStep, A[ls], X[ls], C[ls]);
                __builtin_amdgcn_sched_barrier(0);
                sym_step<MB, NB, DM>(A[s], X[s], C[s], P, Q);
                __builtin_amdgcn_sched_barrier(0);
            }
            pa += D * rstep;
            px += D * rstep;
            pc += D * kKStep;
        }
    }
    // FP64 accumulator lanes hold rows kk + 4v of a block; slices are
    // column-major (column index = the B-side column of R)
    double2* Ps = a.P + (int64_t)slot * Np * Np;
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n) {
            if (DM == 1 && m > n) continue;
            double2* dst = Ps + (int64_t)(a0 + 16 * n + i) * Np + b0 + 16 * m + kk;
#pragma unroll
            for (int v = 0; v < 4; ++v) dst[4 * v] = make_double2(P[m][n][v], Q[m][n][v]);
        }
}

// compile-time specialisations with exactly mv x nv valid blocks
template <int D, int MB, int NB>
__device__ __forceinline__ void dispatch_off(int mv, int nv, const ContractArgs& a, int64_t Kb,
                                             int64_t Ke, int b0, int a0, int slot)
{
    if constexpr (MB > 1) {
        if (mv < MB) {
            dispatch_off<D, MB - 1, NB>(mv, nv, a, Kb, Ke, b0, a0, slot);
            return;
        }
    }
    if constexpr (NB > 1) {
        if (nv < NB) {
            dispatch_off<D, MB, NB - 1>(mv, nv, a, Kb, Ke, b0, a0, slot);
            return;
        }
    }
    wave_sym64<D, MB, NB, 0>(a, Kb, Ke, b0, a0, slot);
}

template <int D, int MB, int DM>
__device__ __forceinline__ void dispatch_diag(int mv, const ContractArgs& a, int64_t Kb,
                                              int64_t Ke, int b0, int slot)
{
    if constexpr (MB > 1) {
        if (mv < MB) {
            dispatch_diag<D, MB - 1, DM>(mv, a, Kb, Ke, b0, slot);
            return;
        }
    }
    wave_sym64<D, MB, MB, DM>(a, Kb, Ke, b0, b0, slot);
}

// D: operand sets in flight (loads D - 1 K-steps ahead of the MFMAs).
// FULLDIAG: diagonal tiles compute all their blocks (those below the
// diagonal are written but never read): more MFMA work, but every unit then
// covers a tile's worth of blocks, so all tiles share one K partition and R
// is fetched from HBM fewer times (measured 25 vs 36 GB per launch at the
// default grid, where that lets the clock run ~7% higher).
template <int D, bool FULLDIAG>
__global__ __launch_bounds__(64) void k_contract_sym64(ContractArgs a)
{
    const WorkUnit u = a.units[blockIdx.x];
    if (u.widx < 0) return;
    const int tr = u.widx / a.wcols, tc = u.widx % a.wcols;
    const int nb = a.Np / 16;
    const int b0 = tr * kSymMB * 16, a0 = tc * kSymMB * 16;
    const int mv = min(kSymMB, nb - tr * kSymMB), nv = min(kSymMB, nb - tc * kSymMB);
    if (tr == tc) dispatch_diag<D, kSymMB, FULLDIAG ? 2 : 1>(mv, a, u.Kb, u.Ke, b0, u.slot);
    else dispatch_off<D, kSymMB, kSymMB>(mv, nv, a, u.Kb, u.Ke, b0, a0, u.slot);
}

struct SymVariant {
    const void* fn;
    int depth;
    bool full_diag;
};

const SymVariant kVariants[] = {
    {(const void*)k_contract_sym64<4, false>, 4, false},
    {(const void*)k_contract_sym64<6, false>, 6, false},
    {(const void*)k_contract_sym64<8, false>, 8, false},
    {(const void*)k_contract_sym64<4, true>, 4, true},
    {(const void*)k_contract_sym64<6, true>, 6, true},
    {(const void*)k_contract_sym64<8, true>, 8, true},
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

}  // namespace

int contract_sym64_variants() { return kNumVariants; }

static int clamp_variant(int v) { return v < 0 || v >= kNumVariants ? 0 : v; }

ContractShape contract_sym64_shape(int variant)
{
    static std::atomic<int> occ[kNumVariants] = {};
    const int v = clamp_variant(variant);
    if (occ[v] == 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kVariants[v].fn, 64, 0) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 4;
        occ[v] = per_cu;
    }
    return {kSymMB, kSymMB, occ[v]};
}

int contract_sym64_quantum(int variant) { return kVariants[clamp_variant(variant)].depth * kKStep; }

bool contract_sym64_full_diag(int variant) { return kVariants[clamp_variant(variant)].full_diag; }

const char* contract_sym64_name(int variant)
{
    static const char* names[kNumVariants] = {"sym64 4x4 d4", "sym64 4x4 d6", "sym64 4x4 d8",
                                              "sym64 4x4 d4 fulldiag", "sym64 4x4 d6 fulldiag",
                                              "sym64 4x4 d8 fulldiag"};
    return names[clamp_variant(variant)];
}

hipError_t contract_sym64_launch(int variant, const ContractArgs& a, hipStream_t stream)
{
    ContractArgs args = a;
    void* params[] = {(void*)&args};
    return hipLaunchKernel(kVariants[clamp_variant(variant)].fn, dim3((unsigned)a.n_units),
                           dim3(64), params, 0, stream);
}

}  // namespace klsp
