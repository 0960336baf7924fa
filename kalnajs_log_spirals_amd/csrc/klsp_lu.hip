// klsp_lu.hip -- det(I - M) by complex LU with partial pivoting, one launch
// per 16-column panel (MatrixBuilder.jl:182 -> LinearAlgebra det -> LAPACK
// zgetrf), for batches of a few matrices.
//
// The single-workgroup LU (k_lu_det in klsp_device.hip) keeps one CU per
// matrix, so a lone det leaves the rest of the device idle and its time is a
// chain of dependent panel steps on that CU.  Here launch k gives each column
// block b >= k of every matrix its own workgroup:
//   * it applies panel k-1's row interchanges (a list of at most 32 moved
//     rows) to its 16 columns, solves U12 = L11^-1 A12 and subtracts the
//     rank-16 update L21 U12 from rows below the panel;
//   * the workgroup of block k then factors it: the panel's rows live in
//     registers (one row per thread, N <= 512), each column's pivot is found
//     by a DPP wave maximum and one workgroup barrier (the LDS slots of the
//     reduction alternate between columns, so no second barrier is needed),
//     and the factored block goes back to global memory for the next launch.
// The launch boundary is the only synchronisation between workgroups.
//
// Arithmetic is that of k_lu_det step for step: the same pivot (first
// largest |re| + |im|, a NaN diagonal stays the pivot, an exactly zero
// column gives det 0), multipliers scaled by d_inv of the pivot, cfms in
// ascending panel-column order for the in-panel elimination, the U12 solve
// and the trailing update, and det = prod U_jj in column order; the two
// kernels give the same bits.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "klsp_cplx.hpp"
#include "klsp_internal.hpp"
#include "klsp_lu.hpp"

namespace klsp {

namespace {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
constexpr int kB = 16;                 // panel and column-block width
constexpr int kMaxMoved = 2 * kB;      // rows one panel's interchanges can move
static_assert(kMaxMoved * kB == kThreads, "one (moved row, column) pair per thread");

// Per-matrix bookkeeping carried from one launch to the next.
struct LuAux {
    double2 det;                       // product of U's diagonal so far
    int nmoved[2];                     // moved rows of the last panels of even / odd index
    int swaps;                         // row interchanges so far
    int singular;                      // some pivot column was exactly zero
    int moved[2][kMaxMoved][2];        // (destination row, source row)
};

// v from the lane DPP control kCtrl names (lanes it leaves out keep v)
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ double dpp_or_self(double v)
{
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)x, (int)x, kCtrl, kRowMask, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(x >> 32), (int)(x >> 32), kCtrl, kRowMask,
                                               0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The wave's largest v (no NaN among them) in every lane: within rows of 16
// lanes by quad permutes and row rotations, then across rows by the row
// broadcasts of lanes 15 and 31, which leave the total in lane 63.
__device__ __forceinline__ double wave_max(double v)
{
    v = fmax(v, dpp_or_self<0xb1>(v));          // quad_perm [1,0,3,2]
    v = fmax(v, dpp_or_self<0x4e>(v));          // quad_perm [2,3,0,1]
    v = fmax(v, dpp_or_self<0x124>(v));         // row_ror:4
    v = fmax(v, dpp_or_self<0x128>(v));         // row_ror:8
    v = fmax(v, dpp_or_self<0x142, 0xa>(v));    // row_bcast:15 into rows 1 and 3
    v = fmax(v, dpp_or_self<0x143, 0xc>(v));    // row_bcast:31 into rows 2 and 3
    return bcast(v, 63);
}

// Launch k of the LU of matrix blockIdx.y; workgroup blockIdx.x owns column
// block b = k + blockIdx.x.  N <= 512: one panel row per thread.
__global__ __launch_bounds__(kThreads) void k_lu_step(const double2* __restrict__ M,
                                                      const double2* __restrict__ coef, int nt,
                                                      double2* __restrict__ Awork, int N, int k,
                                                      LuAux* __restrict__ auxw,
                                                      double2* __restrict__ det_out)
{
    const int b = k + blockIdx.x;
    const int mat = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int nb = (N + kB - 1) / kB;
    const int64_t NN = (int64_t)N * N;
    double2* A = Awork + mat * NN;
    LuAux* ax = auxw + mat;
    const int c0 = b * kB, nc = min(kB, N - c0);

    __shared__ double2 Ush[kB][kB];          // U12 of this block [panel row][column]
    __shared__ double2 Lsh[kB][kB];          // L11 of the last panel [row][column]
    __shared__ int mvd[kMaxMoved], mvs[kMaxMoved];
    __shared__ int nmv;
    __shared__ double red_v[2][kWaves];      // per wave: largest candidate,
    __shared__ int red_i[2][kWaves];         // its position;
    __shared__ double2 rows[2][kWaves + 1][kB];   // the candidate rows (rotated as the
    __shared__ int rows_o[2][kWaves + 1];         // registers), slot kWaves: the diagonal row

    // this thread's row of the block (position tid of the panel when b == k)
    double2 a[kB];
#pragma unroll
    for (int c = 0; c < kB; ++c) a[c] = make_double2(0.0, 0.0);

    if (k == 0) {
        // this block's columns of I - M (or of I - the model point)
        for (int64_t x = tid; x < (int64_t)N * nc; x += kThreads) {
            const int r = (int)(x % N), c = c0 + (int)(x / N);
            const int64_t e = r + (int64_t)N * c;
            double2 m;
            if (coef == nullptr) {
                m = M[mat * NN + e];
            } else {
                m = make_double2(0.0, 0.0);
                for (int j = 0; j < nt; ++j) {
                    const double2 t = cmul(coef[mat * nt + j], M[j * NN + e]);
                    m = make_double2(m.x + t.x, m.y + t.y);
                }
            }
            A[e] = make_double2((r == c ? 1.0 : 0.0) - m.x, 0.0 - m.y);
        }
        __syncthreads();
    } else {
        const int p0 = (k - 1) * kB;         // the last panel: rows / columns p0 .. p0 + 15
        const int par = (k - 1) & 1;
        // 1. its row interchanges: every moved row is read before any is written
        if (tid == 0) nmv = ax->nmoved[par];
        if (tid < kMaxMoved) {
            mvd[tid] = ax->moved[par][tid][0];
            mvs[tid] = ax->moved[par][tid][1];
        }
        __syncthreads();
        {
            const int e = tid / kB, c = tid % kB;
            const bool act = e < nmv && c < nc;
            double2 v = make_double2(0.0, 0.0);
            if (act) v = A[mvs[e] + (int64_t)N * (c0 + c)];
            __syncthreads();
            if (act) A[mvd[e] + (int64_t)N * (c0 + c)] = v;
        }
        if (tid < kB * kB) {
            const int t = tid / kB, s = tid % kB;
            Lsh[t][s] = A[(p0 + t) + (int64_t)N * (p0 + s)];
        }
        __syncthreads();
        if (tid < kB * kB) {
            const int t = tid / kB, c = tid % kB;
            Ush[t][c] = c < nc ? A[(p0 + t) + (int64_t)N * (c0 + c)] : make_double2(0.0, 0.0);
        }
        __syncthreads();
        // 2. U12 = L11^-1 A12 (unit lower L11): thread c solves column c, one
        // column of L11 at a time (its reads issued together)
        if (tid < nc) {
            double2 u[kB];
#pragma unroll
            for (int t = 0; t < kB; ++t) u[t] = Ush[t][tid];
#pragma unroll
            for (int s = 0; s < kB - 1; ++s) {
                // an offset the compiler cannot see through keeps each
                // column's reads here (hoisted, the 120 values would spill)
                int o = 0;
                asm volatile("" : "+v"(o));
                const double2* Lc = &Lsh[0][s] + o;
                double2 lc[kB];
#pragma unroll
                for (int t = s + 1; t < kB; ++t) lc[t] = Lc[t * kB];
#pragma unroll
                for (int t = s + 1; t < kB; ++t) u[t] = cfms(u[t], lc[t], u[s]);
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int t = 0; t < kB; ++t) Ush[t][tid] = u[t];
        }
        __syncthreads();
        // 3. rank-16 update of rows p0 + 16 .. N - 1: a thread per row and
        // the block's columns in registers; per panel column t its L21 entry
        // (loaded one step ahead) and U12 row t from LDS (all lanes read the
        // same addresses).  The t loop stays rolled: unrolled, the compiler
        // hoists every U12 read and spills.
        const int r = p0 + kB + tid;         // one row a thread (N <= 512)
        if (r < N) {
#pragma unroll
            for (int c = 0; c < kB; ++c)
                a[c] = c < nc ? A[r + (int64_t)N * (c0 + c)] : make_double2(0.0, 0.0);
            double2 lt = A[r + (int64_t)N * p0];
#pragma unroll 1
            for (int t = 0; t < kB; ++t) {
                const double2 ln = t + 1 < kB ? A[r + (int64_t)N * (p0 + t + 1)] : lt;
                double2 ut[kB];
#pragma unroll
                for (int c = 0; c < kB; ++c) ut[c] = Ush[t][c];
#pragma unroll
                for (int c = 0; c < kB; ++c) a[c] = cfms(a[c], lt, ut[c]);
                lt = ln;
            }
            // the next panel stays in registers for its factorisation
            if (b != k) {
#pragma unroll
                for (int c = 0; c < kB; ++c)
                    if (c < nc) A[r + (int64_t)N * (c0 + c)] = a[c];
            }
        }
        if (b != k) return;
    }
    if (b != k) return;

    // 4. factor block k: rows k0 .. N - 1 (positions 0 .. m - 1), thread tid
    // holding position tid.  Columns go in groups of four: within a group
    // the column index into the row registers is a compile-time constant,
    // and after it the registers rotate by four, so slot u of group g is
    // column 4g + u, the slots after it the columns still to come and the
    // last slots the multipliers of earlier groups.  (Unrolled over all 16
    // columns the code outgrows the instruction cache it runs through once
    // per launch; rotating after every column costs 64 moves a column.)
    const int k0 = c0, m = N - k0, r = tid;
    int orow = r;                            // position the held row came from
    if (k == 0) {
#pragma unroll
        for (int c = 0; c < kB; ++c)
            a[c] = (r < m && c < nc) ? A[(k0 + r) + (int64_t)N * (k0 + c)]
                                     : make_double2(0.0, 0.0);
    }
    // waves without panel rows leave (a barrier waits only for waves that
    // have not ended), so fewer waves share the SIMDs in the column loop
    const int nwa = (m + 63) / 64;
    if (wid >= nwa) return;
    int swaps = 0;
    bool singular = false;
    __shared__ double2 dgl[kB];              // U_jj of this panel
    auto rotate4 = [&]() {
        double2 t[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) t[c] = a[c];
#pragma unroll
        for (int c = 0; c < kB - 4; ++c) a[c] = a[c + 4];
#pragma unroll
        for (int c = 0; c < 4; ++c) a[kB - 4 + c] = t[c];
    };
    int g = 0;
    for (; g < nc; g += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int jj = g + u;
            if (jj < nc) {
                const int par = jj & 1;
                const int rem = nc - 1 - jj;     // columns right of jj
                // this thread's candidate (a NaN never wins, as in izamax)
                const double v = fabs(a[u].x) + fabs(a[u].y);
                const double bv = (r >= jj && r < m && v > -1.0) ? v : -1.0;
                if (r == jj) {
#pragma unroll
                    for (int c = 0; c < kB; ++c) rows[par][kWaves][c] = a[c];
                    rows_o[par][kWaves] = orow;
                }
                // the wave's candidate: the lowest lane holding the wave maximum
                const double wv = wave_max(bv);
                const unsigned long long hit = __ballot(bv == wv);
                const int src = hit ? __ffsll((long long)hit) - 1 : 0;
                if (lane == src) {
                    red_v[par][wid] = wv;
                    red_i[par][wid] = r;
#pragma unroll
                    for (int c = 0; c < kB; ++c) rows[par][wid][c] = a[c];
                    rows_o[par][wid] = orow;
                }
                __syncthreads();
                // the pivot: the first wave with the largest candidate (waves
                // hold ascending positions, so this is the first largest row);
                // a NaN diagonal stays the pivot
                double gv = -1.0;
                int gw = 0, gi = 0;
#pragma unroll
                for (int w = 0; w < kWaves; ++w) {
                    const double vw = red_v[par][w];
                    const int iw = red_i[par][w];
                    if (w < nwa && vw > gv) { gv = vw; gw = w; gi = iw; }
                }
                const double2 dd = rows[par][kWaves][u];
                const bool nan_diag = isnan(fabs(dd.x) + fabs(dd.y));
                if (!nan_diag && gv == 0.0) {
                    // an exactly zero pivot column: zgetf2 reports info > 0 and
                    // LinearAlgebra's det(::LU) returns 0 for that factorisation
                    singular = true;
                    if (tid == 0) dgl[jj] = make_double2(0.0, 0.0);
                } else {
                    const int sel = nan_diag ? kWaves : gw;    // the pivot row's slot
                    const int p = nan_diag ? jj : gi;
                    double2 pr[kB];
#pragma unroll
                    for (int c = 0; c < kB; ++c) pr[c] = rows[par][sel][c];
                    if (p != jj) {
                        ++swaps;
                        if (r == p) {
#pragma unroll
                            for (int c = 0; c < kB; ++c) a[c] = rows[par][kWaves][c];
                            orow = rows_o[par][kWaves];
                        } else if (r == jj) {
#pragma unroll
                            for (int c = 0; c < kB; ++c) a[c] = pr[c];
                            orow = rows_o[par][sel];
                        }
                    }
                    const double2 piv = pr[u];
                    if (tid == 0) dgl[jj] = piv;
                    const double2 rp = d_inv(piv.x, piv.y);
                    if (r > jj && r < m) {
                        const double2 l = cmul(a[u], rp);
                        a[u] = l;
#pragma unroll
                        for (int c = u + 1; c < kB; ++c)
                            if (c - u <= rem) a[c] = cfms(a[c], l, pr[c]);
                    }
                }
            }
        }
        rotate4();
    }
    // back to a[c] = column c
    for (; g < kB; g += 4) rotate4();
    // the factored block (L below the diagonal) for the next launch
    if (r < m) {
#pragma unroll
        for (int c = 0; c < kB; ++c)
            if (c < nc) A[(k0 + r) + (int64_t)N * (k0 + c)] = a[c];
    }
    // the rows this panel moved, for the next launch's interchanges
    if (tid == 0) nmv = 0;
    __syncthreads();
    if (r < m && orow != r) {
        const int q = atomicAdd(&nmv, 1);
        ax->moved[k & 1][q][0] = k0 + r;
        ax->moved[k & 1][q][1] = k0 + orow;
    }
    __syncthreads();
    if (tid == 0) {
        ax->nmoved[k & 1] = nmv;
        const int sw = (k == 0 ? 0 : ax->swaps) + swaps;
        const int sg = (k == 0 ? 0 : ax->singular) | (singular ? 1 : 0);
        ax->swaps = sw;
        ax->singular = sg;
        // prod U_jj in column order, carried from panel to panel
        double2 det = k == 0 ? make_double2(1.0, 0.0) : ax->det;
        for (int j = 0; j < nc; ++j) det = cmul(det, dgl[j]);
        ax->det = det;
        if (k == nb - 1) {
            if (sg) det = make_double2(0.0, 0.0);
            else if (sw & 1) det = make_double2(-det.x, -det.y);
            det_out[mat] = det;
        }
    }
}

}  // namespace

size_t lu_step_aux_bytes() { return sizeof(LuAux); }

hipError_t launch_lu_steps(hipStream_t st, const double2* M, const double2* coef, int nt,
                           double2* A, int N, int nd, void* aux, double2* det_out)
{
    if (N < 1 || N > kLuStepMaxN || nd < 1) return hipErrorInvalidValue;
    static_assert(kLuStepMaxN <= kThreads, "one panel row per thread");
    const int nb = (N + kB - 1) / kB;
    LuAux* ax = static_cast<LuAux*>(aux);
    for (int k = 0; k < nb; ++k)
        hipLaunchKernelGGL(k_lu_step, dim3(nb - k, nd), dim3(kThreads), 0, st, M, coef, nt, A,
                           N, k, ax, det_out);
    return hipGetLastError();
}

}  // namespace klsp
