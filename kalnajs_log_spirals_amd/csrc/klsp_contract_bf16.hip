// klsp_contract_bf16.hip -- the M(omega) contraction on BF16 MFMA with
// three-term split products, for low-precision builds (Taylor passes of the
// mixed-precision Newton search, BF16x3 dets).
//
// Storage.  W is split once into bf16 hi/lo parts, x = hi + lo with
// hi = bf16(x) and lo = bf16(x - hi) (16 significant bits), each dword
// holding re | im << 16, in two planes of rows of four K:
//   Wh[K/4][Np] = {hi(4q), hi(4q+1), hi(4q+2), hi(4q+3)}   (16 bytes)
//   Wl[K/4][Np] = the same for lo,
// so one 16-byte load is exactly one MFMA operand of 16x16x32 bf16 (lane
// group g of a stage of 16 K rows holds rows 4g..4g+3, two bf16 each).
//
// Products.  B = c conj(W_alpha) is formed in FP32 from hi + lo and split
// the same way; a real product x y is taken as xh yh + xl yh + xh yl (the
// dropped xl yl is ~2^-16 relative).  With the B dwords
//   ReH = (Br_h, -Bi_h), ImH = (Bi_h, Br_h), ReL, ImL alike for the lo parts,
// each pair of slots of a dword gives Ar Br - Ai Bi resp. Ar Bi + Ai Br, and
// 16 K rows take three MFMAs per output part:
//   Re += H . ReH + L . ReH + H . ReL,   Im += H . ImH + L . ImH + H . ImL.
// Accumulation is FP32 over a gang's K range; the gang partials are summed
// in FP64 by k_reduce.  Weight sets from NF on take the single product
// Re += H . ReH, Im += H . ImH (relative error ~2^-8 instead of ~2^-16): the
// mixed Newton search uses this for the Taylor term T_4, which enters M
// multiplied by s^3 and the anchored FP64 finish multiplied by
// 3 s_a^2 (s - s_a) (|s_a| / Im w ~ 0.03).
//
// Work decomposition.  A workgroup is 4 waves stacked along beta (5 blocks
// of 16 rows per wave, 320 rows) over one alpha block (16 columns).  Per
// stage of 16 K rows its 256 threads form B for the 16 columns x 16 rows x NW
// weight sets once (one thread = one K row x one column), split it and store
// it in LDS in operand order; each wave then reads its B operands from LDS
// and its A operands straight from global memory (every wave owns distinct
// beta rows), with the next stage's loads in flight meanwhile.
// Gangs (the workgroups covering the whole output) walk contiguous K ranges
// on one XCD, as in the FP64 and FP32 kernels.
#include <algorithm>

#include "klsp_contract.hpp"

namespace klsp {
namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef unsigned v2u __attribute__((ext_vector_type(2)));

constexpr int kMB = kBf16MB;            // beta blocks per wave
constexpr int kNB = kBf16NB;            // alpha blocks per workgroup
constexpr int kTA = kNB * 16;           // alpha columns per workgroup
constexpr int kMaxNW = kMaxW;
constexpr int kThreads = kBf16Waves * 64;
static_assert(kThreads == 16 * kTA, "one B-producing thread per K row and column");

__device__ __forceinline__ unsigned pk_bf16(float a, float b)
{
    const v2f f = {a, b};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(f, v2bf));
}
__device__ __forceinline__ float lo_f(unsigned u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float hi_f(unsigned u) { return __uint_as_float(u & 0xffff0000u); }
__device__ __forceinline__ unsigned swap16(unsigned u) { return __builtin_amdgcn_alignbit(u, u, 16); }

__device__ __forceinline__ v4f mfma(v4u a, v4u b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                   __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}

// B operands of one stage, in operand order: rec[w][part][g][col] holds the
// dwords of K rows 4g..4g+3 for part ReH, ImH, ReL, ImL.
struct StageB {
    v4u rec[kMaxNW][4][4][kTA];
};

// One K row (k = 0..15 within the stage) x one column of B for NW weight
// sets into LDS; x = (hi, lo) of W at that row and column.
template <int NW, int NF>
__device__ __forceinline__ void produce_B(StageB& sb, v2u x, const v2f (&cw)[NW], int k, int c)
{
    const float xr = lo_f(x.x) + lo_f(x.y), xi = hi_f(x.x) + hi_f(x.y);
    const int g = k >> 2, t = k & 3;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const float cr = cw[w].x, ci = cw[w].y;
        const float br = cr * xr + ci * xi;      // c conj(x)
        const float bi = ci * xr - cr * xi;
        const unsigned hp = pk_bf16(br, bi);
        ((unsigned*)&sb.rec[w][0][g][c])[t] = hp ^ 0x80000000u;
        ((unsigned*)&sb.rec[w][1][g][c])[t] = swap16(hp);
        if (w < NF) {
            const unsigned lp = pk_bf16(br - lo_f(hp), bi - hi_f(hp));
            ((unsigned*)&sb.rec[w][2][g][c])[t] = lp ^ 0x80000000u;
            ((unsigned*)&sb.rec[w][3][g][c])[t] = swap16(lp);
        }
    }
}

// Stores to LDS done and visible, then the workgroup barrier.  Global loads
// stay in flight (a __syncthreads fence would wait for them).
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <int MB, int NB, int NW>
struct Acc {
    static constexpr int M = MB > 0 ? MB : 1;
    v4f r[NW][M][NB], i[NW][M][NB];
};

template <int MB>
__device__ __forceinline__ void load_A(const v4u* __restrict__ Wh, const v4u* __restrict__ Wl,
                                       int64_t q0, int Np, int b0, int g, int i,
                                       v4u (&A)[MB > 0 ? MB : 1][2])
{
#pragma unroll
    for (int m = 0; m < MB; ++m) {
        const int64_t x = (q0 + g) * Np + b0 + 16 * m + i;
        A[m][0] = Wh[x];
        A[m][1] = Wl[x];
    }
}

template <int MB, int NB, int NW, int NF>
__device__ __forceinline__ void mfma_stage(const StageB& sb, const v4u (&A)[MB > 0 ? MB : 1][2],
                                           Acc<MB, NB, NW>& acc, int g, int i)
{
#pragma unroll
    for (int n = 0; n < NB; ++n)
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const int col = 16 * n + i;
            const v4u rh = sb.rec[w][0][g][col], ih = sb.rec[w][1][g][col];
#pragma unroll
            for (int m = 0; m < MB; ++m) acc.r[w][m][n] = mfma(A[m][0], rh, acc.r[w][m][n]);
#pragma unroll
            for (int m = 0; m < MB; ++m) acc.i[w][m][n] = mfma(A[m][0], ih, acc.i[w][m][n]);
            if (w >= NF) {
                __builtin_amdgcn_sched_barrier(0);
                continue;
            }
            const v4u rl = sb.rec[w][2][g][col], il = sb.rec[w][3][g][col];
#pragma unroll
            for (int m = 0; m < MB; ++m) acc.r[w][m][n] = mfma(A[m][1], rh, acc.r[w][m][n]);
#pragma unroll
            for (int m = 0; m < MB; ++m) acc.i[w][m][n] = mfma(A[m][1], ih, acc.i[w][m][n]);
#pragma unroll
            for (int m = 0; m < MB; ++m) acc.r[w][m][n] = mfma(A[m][0], rl, acc.r[w][m][n]);
#pragma unroll
            for (int m = 0; m < MB; ++m) acc.i[w][m][n] = mfma(A[m][0], il, acc.i[w][m][n]);
            // keep the B operands of one (n, w) live at a time
            __builtin_amdgcn_sched_barrier(0);
        }
}

// One workgroup's walk over [Kb, Ke) (multiples of 16) for a wave owning
// MB beta blocks from b0 and the workgroup's NB alpha blocks from a0.
template <int MB, int NB, int NW, int NF>
__device__ __forceinline__ void wg_walk(const ContractArgs& a, StageB* sb, int64_t Kb, int64_t Ke,
                                        int b0, int a0, int gang)
{
    constexpr int MA = MB > 0 ? MB : 1;
    const int Np = a.Np;
    const int64_t nq = (a.Kp + 15) / 16 * 4;          // rows of four K per plane
    const v4u* __restrict__ Wh = (const v4u*)a.W;
    const v4u* __restrict__ Wl = Wh + nq * Np;
    const int tid = threadIdx.x, lane = tid & 63, i = lane & 15, g = lane >> 4;
    const int64_t Kpc = a.Kp + kRowPad;
    const int kr = tid / kTA, c = tid % kTA;        // B-producing role
    const bool cvalid = a0 + c < Np;
    Acc<MB, NB, NW> acc;
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int m = 0; m < MA; ++m)
#pragma unroll
            for (int n = 0; n < NB; ++n) {
                acc.r[w][m][n] = (v4f){0, 0, 0, 0};
                acc.i[w][m][n] = (v4f){0, 0, 0, 0};
            }
    // A operands are loaded once per stage (the other workgroup on the CU
    // covers their latency); the B inputs of the next stage are prefetched.
    v4u A[MA][2];
    v2u X, Xn;
    v2f C[NW], Cn[NW];
    auto load_B_in = [&](int64_t K0, v2u& Xo, v2f (&Co)[NW]) {
        const int64_t q0 = K0 >> 2;
        if (cvalid) {
            const int64_t x = ((q0 + (kr >> 2)) * Np + a0 + c) * 4 + (kr & 3);
            Xo = (v2u){((const unsigned*)Wh)[x], ((const unsigned*)Wl)[x]};
        } else {
            Xo = (v2u){0, 0};
        }
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const float2 t = a.cwf[w * Kpc + K0 + kr];
            Co[w] = (v2f){t.x, t.y};
        }
    };
    if (Kb < Ke) load_B_in(Kb, X, C);
    int buf = 0;
    for (int64_t K0 = Kb; K0 < Ke; K0 += 16) {
        load_A<MB>(Wh, Wl, K0 >> 2, Np, b0, g, i, A);
        produce_B<NW, NF>(sb[buf], X, C, kr, c);
        load_B_in(K0 + 16 < Ke ? K0 + 16 : K0, Xn, Cn);
        lds_barrier();
        mfma_stage<MB, NB, NW, NF>(sb[buf], A, acc, g, i);
        X = Xn;
#pragma unroll
        for (int w = 0; w < NW; ++w) C[w] = Cn[w];
        buf ^= 1;
    }
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        double2* Pw = a.P + ((int64_t)w * a.n_gangs + gang) * (int64_t)Np * Np;
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int n = 0; n < NB; ++n) {
                double2* dst = Pw + (int64_t)(a0 + 16 * n + i) * Np + b0 + 16 * m + 4 * g;
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    dst[v] = make_double2((double)acc.r[w][m][n][v], (double)acc.i[w][m][n][v]);
            }
    }
}


template <int NW, int NF>
__global__ __launch_bounds__(kThreads, 2) void k_contract_bf16(ContractArgs a)
{
    __shared__ StageB sb[2];
    const int L = blockIdx.x;
    const int xcd = L & 7, j = L >> 3;
    const int gw = a.wrows * a.wcols;
    const int gang = xcd + 8 * (j / gw), widx = j % gw;
    if (gang >= a.n_gangs) return;          // whole workgroup
    const int wave = threadIdx.x >> 6;
    const int b0 = ((widx / a.wcols) * kBf16Waves + wave) * kMB * 16;
    const int a0 = (widx % a.wcols) * kTA;
    const int mv = max(0, min(kMB, (a.Np - b0) / 16));
    const int64_t Kp16 = (a.Kp + 15) / 16 * 16;
    const int64_t Kb = min((int64_t)gang * a.Kg, Kp16);
    const int64_t Ke = min(Kp16, Kb + a.Kg);
    // every wave runs the same stages (the barriers need it); waves past the
    // last beta block only produce B
    switch (mv) {
    case 5: wg_walk<5, kNB, NW, NF>(a, sb, Kb, Ke, b0, a0, gang); break;
    case 4: wg_walk<4, kNB, NW, NF>(a, sb, Kb, Ke, b0, a0, gang); break;
    case 3: wg_walk<3, kNB, NW, NF>(a, sb, Kb, Ke, b0, a0, gang); break;
    case 2: wg_walk<2, kNB, NW, NF>(a, sb, Kb, Ke, b0, a0, gang); break;
    case 1: wg_walk<1, kNB, NW, NF>(a, sb, Kb, Ke, b0, a0, gang); break;
    default: wg_walk<0, kNB, NW, NF>(a, sb, Kb, Ke, b0, a0, gang); break;
    }
}
static_assert(kMB == 5, "the switch above enumerates the wave tile heights");

// W (complex128 rows of K, `rows` rows present) -> the Wh, Wl planes.
__global__ void k_W64_to_bf16(const double2* __restrict__ W, int64_t rows, int Np, int64_t nq,
                              v4u* __restrict__ Wb)
{
    const int64_t n = nq * Np;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = x / Np;
        const int col = (int)(x - q * Np);
        unsigned h[4], l[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int64_t k = 4 * q + t;
            const double2 v = k < rows ? W[k * Np + col] : make_double2(0.0, 0.0);
            h[t] = pk_bf16((float)v.x, (float)v.y);
            l[t] = pk_bf16((float)(v.x - (double)lo_f(h[t])), (float)(v.y - (double)hi_f(h[t])));
        }
        Wb[x] = (v4u){h[0], h[1], h[2], h[3]};
        Wb[n + x] = (v4u){l[0], l[1], l[2], l[3]};
    }
}

template <int NW, int NF>
const void* kernel_bf16()
{
    return (const void*)k_contract_bf16<NW, NF>;
}

// nf < nw only for the four Taylor terms of the mixed search (three full)
const void* kernel_for(int nw, int nf)
{
    if (nw >= 4 && nf < 4) return kernel_bf16<4, 3>();
    return nw <= 1 ? kernel_bf16<1, 1>() : nw == 2 ? kernel_bf16<2, 2>()
         : nw == 3 ? kernel_bf16<3, 3>() : kernel_bf16<4, 4>();
}

}  // namespace

Bf16Geometry contract_bf16_geometry(int Np, int nw, int nf)
{
    static std::atomic<int> occ[kMaxNW + 1][2] = {};
    const int i = nw < 1 ? 1 : (nw > kMaxNW ? kMaxNW : nw);
    const int h = (i >= 4 && nf < 4) ? 1 : 0;
    if (occ[i][h] == 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel_for(i, nf), kThreads, 0) !=
                hipSuccess || per_cu < 1)
            per_cu = 1;
        occ[i][h] = per_cu;
    }
    const int nb = Np / 16;
    Bf16Geometry g;
    g.wrows = (nb + kBf16Waves * kMB - 1) / (kBf16Waves * kMB);
    g.wcols = (nb + kNB - 1) / kNB;
    g.wg_per_cu = occ[i][h];
    return g;
}

hipError_t contract_bf16_launch(int nw, int nf, const ContractArgs& a, hipStream_t stream)
{
    const unsigned grid = (unsigned)(a.n_gangs * a.wrows * a.wcols);
    ContractArgs args = a;
    void* params[] = {(void*)&args};
    return hipLaunchKernel(kernel_for(nw, nf), dim3(grid), dim3(kThreads), params, 0, stream);
}

hipError_t split_W_bf16(const double2* W64, int64_t rows, int Np, int64_t nq, void* Wb,
                        hipStream_t stream)
{
    const int64_t n = nq * Np;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_W64_to_bf16, dim3(grid), dim3(256), 0, stream, W64, rows, Np, nq,
                       (v4u*)Wb);
    return hipGetLastError();
}

}  // namespace klsp
