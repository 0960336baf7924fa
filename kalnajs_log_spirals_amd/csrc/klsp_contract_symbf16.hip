// klsp_contract_symbf16.hip -- the M(omega) contraction in the real basis of
// a symmetric alpha grid on BF16 MFMA with three-term split products, for
// low-precision builds (Taylor passes of the mixed-precision Newton search,
// BF16x3 dets and scans).
//
// Real basis (see klsp_contract_sym64.hip): S = P + iQ with
// P = R^T diag(Re c) R and Q = R^T diag(Im c) R over K, upper-triangle blocks
// only; the reduce expands S into M.
//
// Storage.  R is split once into bf16 hi/lo parts, x = hi + lo, hi = bf16(x),
// lo = bf16(x - hi), in two planes of records of eight K rows:
//   Rh[K/8][Np] = {hi(8q), ..., hi(8q+7)}   (16 bytes), Rl alike,
// so one 16-byte load is one 16x16x32 bf16 MFMA operand (lane group g of a
// stage of 32 K rows holds rows 8g..8g+7).
//
// Products.  B = Re c * x and Im c * x are formed in FP32 from hi + lo of the
// B-side columns and split the same way; a real product x y is taken as
// xh yh + xl yh + xh yl (the dropped xl yl is ~2^-16 relative), so a block
// and K-step of 32 rows take three MFMAs for P and three for Q.  Weight sets
// from NF on take the single product xh yh (the mixed search's Taylor term
// T_4, as in the complex BF16 kernel).  Accumulation is FP32 over a gang's K
// range; gang partials are summed in FP64 by the reduce.
//
// Two work decompositions of the upper triangle of nb x nb blocks; both
// stage the B operands of a workgroup's column blocks in LDS (formed once a
// stage by all its threads, in operand order), read A (the rows of a wave's
// slots) from global memory one stage ahead (two for one weight set), and
// walk a gang's contiguous K range on one XCD.  For one block and stage both
// issue the same MFMAs in the same order, so with the same gang plan they
// give the same bits.
//
// Fold form.  Column blocks c and nb - 1 - c form a pair holding
// (c + 1) + (nb - c) = nb + 1 blocks (the middle column of an odd nb alone
// holds (nb + 1) / 2), so equal slot counts cover the triangle with little
// idle work.  A workgroup is 4 waves of MB block slots (one block each) over
// one pair; a pass of up to four weight sets is one launch.
//
// Strip form.  Strip j is column blocks 2j and 2j + 1 (one block for the last
// strip of an odd nb) with its rows 0 .. min(2j + 1, nb - 1); a row slot is
// one row of a strip, both its blocks, so one A load feeds the MFMAs of two
// column blocks and a wave reads its strip's B operands from LDS once a
// stage.  Strips j and ns - 1 - j fold into one workgroup of 8 waves of kMR
// row slots, each wave in one strip.  Half the A loads and less than half the
// LDS reads of the fold form, but its slots fill the triangle well only for
// some nb (79% at nb = 19); a launch takes at most two weight sets.
#include <algorithm>

#include "klsp_contract.hpp"

namespace klsp {
namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef unsigned v2u __attribute__((ext_vector_type(2)));

constexpr int kKq = 64;            // K rows per pair of stages (a gang's range is a multiple)

__device__ __forceinline__ unsigned pk_bf16(float a, float b)
{
    const v2f f = {a, b};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(f, v2bf));
}
__device__ __forceinline__ float lo_f(unsigned u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float hi_f(unsigned u) { return __uint_as_float(u & 0xffff0000u); }

__device__ __forceinline__ v4f mfma(v4u a, v4u b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                   __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}

// B operands of one stage in operand order: rec[w][part][g][col], part =
// P hi, P lo, Q hi, Q lo; K rows 8g..8g+7 of staged column col (C columns).
template <int NW, int C>
struct StageB {
    v4u rec[NW][4][4][C];
};

// B inputs of one stage for this thread: rows 4 rg .. 4 rg + 3 of its staged
// column, hi and lo (4 bf16 each), and their weights.
template <int NW>
struct BIn {
    v2u h, l;
    v2f c[NW][4];
};

// Stage loads address this gang's planes through a uniform (scalar) base
// that moves by a fixed stride per stage, plus per-lane byte offsets fixed
// for the whole K range: no address arithmetic per stage in vector registers.
__device__ __forceinline__ v4u ld16(const char* base, unsigned off)
{
    return *(const v4u*)(base + off);
}

// B inputs of a stage from the plane record row at `hb` (hi; lo at hb +
// plane) and the weight rows at `wb` (set w at wb + w * wstride bytes).
template <int NW>
__device__ __forceinline__ void load_B_in(const char* hb, int64_t plane, const char* wb,
                                          int64_t wstride, unsigned offB, unsigned offW,
                                          BIn<NW>& o)
{
    o.h = *(const v2u*)(hb + offB);
    o.l = *(const v2u*)(hb + plane + offB);
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const v4f c01 = *(const v4f*)(wb + w * wstride + offW);
        const v4f c23 = *(const v4f*)(wb + w * wstride + offW + 16);
        o.c[w][0] = (v2f){c01.x, c01.y};
        o.c[w][1] = (v2f){c01.z, c01.w};
        o.c[w][2] = (v2f){c23.x, c23.y};
        o.c[w][3] = (v2f){c23.z, c23.w};
    }
}

template <int NW, int NF, int C>
__device__ __forceinline__ void produce_B(StageB<NW, C>& sb, const BIn<NW>& in, int rg, int col)
{
    float x[4];
    x[0] = lo_f(in.h.x) + lo_f(in.l.x);
    x[1] = hi_f(in.h.x) + hi_f(in.l.x);
    x[2] = lo_f(in.h.y) + lo_f(in.l.y);
    x[3] = hi_f(in.h.y) + hi_f(in.l.y);
    const int g = rg >> 1, d = 2 * (rg & 1);
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        float bp[4], bq[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            bp[t] = in.c[w][t].x * x[t];
            bq[t] = in.c[w][t].y * x[t];
        }
        const unsigned ph0 = pk_bf16(bp[0], bp[1]), ph1 = pk_bf16(bp[2], bp[3]);
        const unsigned qh0 = pk_bf16(bq[0], bq[1]), qh1 = pk_bf16(bq[2], bq[3]);
        unsigned* r0 = (unsigned*)&sb.rec[w][0][g][col];
        unsigned* r2 = (unsigned*)&sb.rec[w][2][g][col];
        r0[d] = ph0;
        r0[d + 1] = ph1;
        r2[d] = qh0;
        r2[d + 1] = qh1;
        if (w < NF) {
            unsigned* r1 = (unsigned*)&sb.rec[w][1][g][col];
            unsigned* r3 = (unsigned*)&sb.rec[w][3][g][col];
            r1[d] = pk_bf16(bp[0] - lo_f(ph0), bp[1] - hi_f(ph0));
            r1[d + 1] = pk_bf16(bp[2] - lo_f(ph1), bp[3] - hi_f(ph1));
            r3[d] = pk_bf16(bq[0] - lo_f(qh0), bq[1] - hi_f(qh0));
            r3[d + 1] = pk_bf16(bq[2] - lo_f(qh1), bq[3] - hi_f(qh1));
        }
    }
}

// Stores to LDS done and visible, then the workgroup barrier.  Global loads
// stay in flight (a __syncthreads fence would wait for them).
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// A operands of a stage: slot s's record at hb + offA[s] (hi) and the same
// offset in the lo plane.  Idle slots read a valid block (their MFMAs are
// skipped).
template <int MB>
__device__ __forceinline__ void load_A(const char* hb, int64_t plane, const unsigned (&offA)[MB],
                                       v4u (&Ah)[MB], v4u (&Al)[MB])
{
#pragma unroll
    for (int s = 0; s < MB; ++s) {
        Ah[s] = ld16(hb, offA[s]);
        Al[s] = ld16(hb + plane, offA[s]);
    }
}

// Per-gang K range and plane addressing shared by both forms.
struct GangRange {
    int64_t Kb, Ke, plane, stage, wstride;
    const char* hb;
    const char* wb;
};

__device__ __forceinline__ GangRange gang_range(const ContractArgs& a, int gang)
{
    GangRange r;
    const int Np = a.Np;
    // a gang's K range is a whole number of stage pairs (kKq rows)
    const int64_t Kq = (a.Kp + kKq - 1) / kKq * kKq;
    r.Kb = min((int64_t)gang * a.Kg, Kq);
    r.Ke = min(Kq, r.Kb + a.Kg);
    r.plane = Kq / 8 * Np * 16;                 // hi plane -> lo plane, bytes
    r.stage = (int64_t)Np * 64;                 // 4 records of Np columns
    r.wstride = (a.Kp + kRowPad) * (int64_t)sizeof(float2);
    r.hb = (const char*)a.W + (r.Kb >> 3) * (int64_t)Np * 16;
    r.wb = (const char*)(a.cwf + r.Kb);
    return r;
}

// The K loop of both forms: two stages a trip on two register sets; a
// stage's B inputs are loaded one stage ahead, its A operands one stage ahead
// or (deep: one weight set, whose stage is too short to cover a load issued
// one stage ahead) two stages ahead; the trip has no exit between its stages
// (the range holds an even number of stages).  On the last trip the loads of
// the next trip still issue, from this trip's rows (never used), so that
// every trip issues the same loads in the same order and the compiler's
// waits count only the loads issued after the one waited for.  `stage_fn`
// runs a stage's MFMAs: (B buffer, A hi, A lo, address of the stage two
// ahead for the deep reload).
template <int NW, int NF, int C, int MB, bool deep, typename StageFn>
__device__ __forceinline__ void k_loop(const GangRange& r, StageB<NW, C> (&sb)[2],
                                       const unsigned (&offA)[MB], unsigned offB, unsigned offW,
                                       int rg, int col, StageFn stage_fn)
{
    if (r.Kb >= r.Ke) return;
    v4u A0h[MB], A0l[MB], A1h[MB], A1l[MB];
    BIn<NW> b0, b1;
    load_A<MB>(r.hb, r.plane, offA, A0h, A0l);
    if (deep) load_A<MB>(r.hb + r.stage, r.plane, offA, A1h, A1l);
    load_B_in<NW>(r.hb, r.plane, r.wb, r.wstride, offB, offW, b0);
    for (int64_t K0 = r.Kb; K0 < r.Ke; K0 += kKq) {
        const char* h0 = r.hb + ((K0 - r.Kb) >> 5) * r.stage;
        const char* w0 = r.wb + (K0 - r.Kb) * (int64_t)sizeof(float2);
        const bool more = K0 + kKq < r.Ke;
        const int64_t ahead = more ? 2 * r.stage : 0;
        const char* w2 = w0 + (more ? 64 * sizeof(float2) : 0);
        produce_B<NW, NF, C>(sb[0], b0, rg, col);
        load_B_in<NW>(h0 + r.stage, r.plane, w0 + 32 * sizeof(float2), r.wstride, offB, offW, b1);
        if (!deep) load_A<MB>(h0 + r.stage, r.plane, offA, A1h, A1l);
        lds_barrier();
        stage_fn(sb[0], A0h, A0l, h0 + ahead);
        produce_B<NW, NF, C>(sb[1], b1, rg, col);
        load_B_in<NW>(h0 + ahead, r.plane, w2, r.wstride, offB, offW, b0);
        if (!deep) load_A<MB>(h0 + ahead, r.plane, offA, A0h, A0l);
        lds_barrier();
        stage_fn(sb[1], A1h, A1l, h0 + ahead + r.stage);
    }
}

// ---------------------------------------------------------------- fold form

namespace fold {

constexpr int kWaves = 4;
constexpr int kThreads = kWaves * 64;
constexpr int kCols = 32;          // B-side columns staged per workgroup (two blocks)
constexpr int kMaxMB = 5;          // block slots per wave
constexpr int kMaxNW = kMaxW;

// The pair of column blocks of workgroup slice `pair` and its slot list:
// slot s < na is block (s, ca), then (s - na, cb) for the second column.
struct Pair {
    int ca, cb, na, nslots;
};

__device__ __forceinline__ Pair pair_of(int p, int nb)
{
    Pair q;
    q.ca = p;
    q.cb = nb - 1 - p;
    q.na = p + 1;
    q.nslots = q.ca == q.cb ? q.na : nb + 1;
    return q;
}

// The MFMAs of a stage.  With Next, each slot's A operands are reloaded
// from the stage two ahead (`hn`) as soon as the slot's last MFMA has
// issued, so those loads have about two stages to arrive.
template <int MB, int NW, int NF, bool Next>
__device__ __forceinline__ void mfma_stage(const StageB<NW, kCols>& sb, v4u (&Ah)[MB],
                                           v4u (&Al)[MB], const int (&csel)[MB],
                                           v4f (&P)[NW][MB], v4f (&Q)[NW][MB], int g, int i,
                                           const char* hn, int64_t plane,
                                           const unsigned (&offA)[MB])
{
#pragma unroll
    for (int s = 0; s < MB; ++s) {
        if (csel[s] >= 0) {
            const int col = 16 * csel[s] + i;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const v4u ph = sb.rec[w][0][g][col], qh = sb.rec[w][2][g][col];
                P[w][s] = mfma(Ah[s], ph, P[w][s]);
                Q[w][s] = mfma(Ah[s], qh, Q[w][s]);
                if (w < NF) {
                    const v4u pl = sb.rec[w][1][g][col], ql = sb.rec[w][3][g][col];
                    P[w][s] = mfma(Al[s], ph, P[w][s]);
                    Q[w][s] = mfma(Al[s], qh, Q[w][s]);
                    P[w][s] = mfma(Ah[s], pl, P[w][s]);
                    Q[w][s] = mfma(Ah[s], ql, Q[w][s]);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if constexpr (Next) {
            Ah[s] = ld16(hn, offA[s]);
            Al[s] = ld16(hn + plane, offA[s]);
        }
    }
}

template <int MB, int NW, int NF>
__global__ __launch_bounds__(kThreads) void k_contract_symbf16_fold(ContractArgs a)
{
    __shared__ StageB<NW, kCols> sb[2];
    const int L = blockIdx.x;
    const int xcd = L & 7, j = L >> 3;
    const int gw = a.wrows * a.wcols;         // workgroups per gang: wcols pairs x wrows
    const int gang = xcd + 8 * (j / gw), widx = j % gw;
    if (gang >= a.n_gangs) return;            // whole workgroup
    const int Np = a.Np, nb = Np / 16;
    const Pair pr = pair_of(widx / a.wrows, nb);
    const int sub = widx % a.wrows;           // workgroup of the pair
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, i = lane & 15, g = lane >> 4;
    // this wave's slots: row-block offsets and staged column block (-1: idle)
    int row0[MB], csel[MB];
#pragma unroll
    for (int s = 0; s < MB; ++s) {
        const int sl = (sub * kWaves + wave) * MB + s;
        if (sl < pr.na) { row0[s] = 16 * sl; csel[s] = 0; }
        else if (sl < pr.nslots) { row0[s] = 16 * (sl - pr.na); csel[s] = 1; }
        else { row0[s] = -1; csel[s] = -1; }
    }
    const GangRange r = gang_range(a, gang);
    // B-producing role: staged column col (two blocks: ca, cb), rows 4 rg..
    // (the middle column of an odd nb stages its block twice; no slot reads
    // the second copy)
    const int col = tid % kCols, rg = tid / kCols;
    const int col_g = (col < 16 ? 16 * pr.ca : 16 * pr.cb) + (col & 15);
    unsigned offA[MB];
#pragma unroll
    for (int s = 0; s < MB; ++s) offA[s] = (unsigned)((g * Np + max(row0[s], 0) + i) * 16);
    const unsigned offB = (unsigned)(((rg >> 1) * Np + col_g) * 16 + 8 * (rg & 1));
    const unsigned offW = (unsigned)(4 * rg * (int)sizeof(float2));
    v4f P[NW][MB], Q[NW][MB];
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int s = 0; s < MB; ++s) {
            P[w][s] = (v4f){0, 0, 0, 0};
            Q[w][s] = (v4f){0, 0, 0, 0};
        }
    // One weight set: a stage is 6 MB MFMAs, too few to cover a load issued
    // one stage ahead, so A is loaded two stages ahead (measured 5.06 -> 4.66
    // ms at the default grid).  Three or four weight sets give 18-24 MFMAs a
    // slot; there the deeper prefetch measured 2-5% slower.
    constexpr bool deep = NW == 1;
    k_loop<NW, NF, kCols, MB, deep>(
        r, sb, offA, offB, offW, rg, col,
        [&](const StageB<NW, kCols>& s, v4u(&Ah)[MB], v4u(&Al)[MB], const char* hn) {
            mfma_stage<MB, NW, NF, deep>(s, Ah, Al, csel, P, Q, g, i, hn, r.plane, offA);
        });
    // FP32 accumulator lanes hold rows 4g + v of a block (column i); partial
    // slices are column-major: column = B-side real column
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        double2* Pw = a.P + ((int64_t)w * a.n_gangs + gang) * (int64_t)Np * Np;
#pragma unroll
        for (int s = 0; s < MB; ++s) {
            if (csel[s] < 0) continue;
            const int c0 = 16 * (csel[s] ? pr.cb : pr.ca);
            double2* dst = Pw + (int64_t)(c0 + i) * Np + row0[s] + 4 * g;
#pragma unroll
            for (int v = 0; v < 4; ++v)
                dst[v] = make_double2((double)P[w][s][v], (double)Q[w][s][v]);
        }
    }
}

template <int NW, int NF>
const void* kernel_nw(int mb)
{
    switch (mb) {
    case 1: return (const void*)k_contract_symbf16_fold<1, NW, NF>;
    case 2: return (const void*)k_contract_symbf16_fold<2, NW, NF>;
    case 3: return (const void*)k_contract_symbf16_fold<3, NW, NF>;
    case 4: return (const void*)k_contract_symbf16_fold<4, NW, NF>;
    default: return (const void*)k_contract_symbf16_fold<5, NW, NF>;
    }
}

// nf < nw only for the mixed search's Taylor terms: four with three full,
// or (split into two launches) two with one full
const void* kernel_for(int nw, int nf, int mb)
{
    if (nw >= 4 && nf < 4) return kernel_nw<4, 3>(mb);
    if (nw == 2 && nf < 2) return kernel_nw<2, 1>(mb);
    return nw <= 1 ? kernel_nw<1, 1>(mb) : nw == 2 ? kernel_nw<2, 2>(mb)
         : nw == 3 ? kernel_nw<3, 3>(mb) : kernel_nw<4, 4>(mb);
}

// Slots per wave and workgroups per column pair: the fewest workgroups per
// pair with at most max_mb slots per wave, then the fewest slots that hold
// the pair's nb + 1 blocks.  max_mb keeps the accumulators and the prefetched
// A operands within 256 VGPRs (two workgroups per CU): 5 slots for one or two
// weight sets, 4 for three, 3 for four.
void layout(int nb, int nw, int* wrows, int* wcols, int* mb)
{
    const int need = nb + 1;                  // blocks of a full pair
    const int max_mb = nw <= 2 ? kMaxMB : nw == 3 ? 4 : 3;
    const int per_pair = (need + kWaves * max_mb - 1) / (kWaves * max_mb);
    *mb = (need + kWaves * per_pair - 1) / (kWaves * per_pair);
    *wrows = per_pair;                        // workgroups per pair
    *wcols = (nb + 1) / 2;                    // pairs (the middle column of an odd nb alone)
}

}  // namespace fold

// --------------------------------------------------------------- strip form

namespace strip {

constexpr int kWaves = 8;
constexpr int kThreads = kWaves * 64;
constexpr int kCols = 64;          // B-side columns staged per workgroup (two strips)
constexpr int kMR = 3;             // row slots per wave
constexpr int kMaxNW = 2;          // weight sets per launch

// Strip j: column blocks c0 = 2j and, when it exists, c0 + 1; rows
// 0 .. nrows - 1.  Row r holds block (r, c0) when r <= c0 and (r, c0 + 1)
// when `two`.
struct Strip {
    int c0, nrows;
    bool two;
};

__host__ __device__ inline Strip strip_of(int j, int nb)
{
    Strip s;
    s.c0 = 2 * j;
    s.two = 2 * j + 1 < nb;
    s.nrows = 2 * j + 2 < nb ? 2 * j + 2 : nb;
    return s;
}

__host__ __device__ inline int strip_waves(int j, int nb)
{
    return (strip_of(j, nb).nrows + kMR - 1) / kMR;
}

// The MFMAs of a stage: per weight set the strip's B operands (both column
// blocks) from LDS once, then every slot's blocks (idle blocks skip their
// MFMAs).  With Next, the A operands are reloaded from the stage two ahead
// (`hn`) after the stage's MFMAs have issued.
template <int NW, int NF, bool Next>
__device__ __forceinline__ void mfma_stage(const StageB<NW, kCols>& sb, v4u (&Ah)[kMR],
                                           v4u (&Al)[kMR], int colb, const bool (&on0)[kMR],
                                           const bool (&on1)[kMR], v4f (&P)[NW][kMR][2],
                                           v4f (&Q)[NW][kMR][2], int g, const char* hn,
                                           int64_t plane, const unsigned (&offA)[kMR])
{
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        v4u ph[2], qh[2], pl[2], ql[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            ph[c] = sb.rec[w][0][g][colb + 16 * c];
            qh[c] = sb.rec[w][2][g][colb + 16 * c];
            if (w < NF) {
                pl[c] = sb.rec[w][1][g][colb + 16 * c];
                ql[c] = sb.rec[w][3][g][colb + 16 * c];
            }
        }
#pragma unroll
        for (int s = 0; s < kMR; ++s)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                if (c == 0 ? on0[s] : on1[s]) {
                    P[w][s][c] = mfma(Ah[s], ph[c], P[w][s][c]);
                    Q[w][s][c] = mfma(Ah[s], qh[c], Q[w][s][c]);
                    if (w < NF) {
                        P[w][s][c] = mfma(Al[s], ph[c], P[w][s][c]);
                        Q[w][s][c] = mfma(Al[s], qh[c], Q[w][s][c]);
                        P[w][s][c] = mfma(Ah[s], pl[c], P[w][s][c]);
                        Q[w][s][c] = mfma(Ah[s], ql[c], Q[w][s][c]);
                    }
                }
            }
    }
    if constexpr (Next) load_A<kMR>(hn, plane, offA, Ah, Al);
}

template <int NW, int NF>
__global__ __launch_bounds__(kThreads) void k_contract_symbf16_strip(ContractArgs a)
{
    __shared__ StageB<NW, kCols> sb[2];
    const int L = blockIdx.x;
    const int xcd = L & 7, j = L >> 3;
    const int gw = a.wrows * a.wcols;         // workgroups per gang: wcols strip pairs x wrows
    const int gang = xcd + 8 * (j / gw), widx = j % gw;
    if (gang >= a.n_gangs) return;            // whole workgroup
    const int Np = a.Np, nb = Np / 16, ns = (nb + 1) / 2;
    const int pr = widx / a.wrows, sub = widx % a.wrows;
    const int ja = pr, jb = ns - 1 - pr;      // the folded strips (one when ja == jb)
    const Strip SA = strip_of(ja, nb), SB = strip_of(jb, nb);
    const int wva = strip_waves(ja, nb), wvb = jb != ja ? strip_waves(jb, nb) : 0;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, i = lane & 15, g = lane >> 4;
    // this wave's strip (sel: 0 = first, 1 = second, -1 idle) and row slots
    const int gv = sub * kWaves + wave;
    const int sel = gv < wva ? 0 : gv < wva + wvb ? 1 : -1;
    const Strip S = sel == 1 ? SB : SA;
    const int rbase = (sel == 1 ? gv - wva : gv) * kMR;
    int row[kMR];
    bool on0[kMR], on1[kMR];
#pragma unroll
    for (int s = 0; s < kMR; ++s) {
        row[s] = rbase + s;
        const bool v = sel >= 0 && row[s] < S.nrows;
        on0[s] = v && row[s] <= S.c0;
        on1[s] = v && S.two;
    }
    const int colb = 32 * max(sel, 0) + i;    // staged column of the strip's first block
    const GangRange r = gang_range(a, gang);
    // B-producing role: staged column col (first strip's blocks below 32,
    // the second's from 32), rows 4 rg .. 4 rg + 3.  A strip of one block
    // stages it twice and the middle strip of an odd strip count fills both
    // halves; no MFMA reads those copies.
    const int col = tid % kCols, rg = tid / kCols;
    const Strip SC = col < 32 ? SA : SB;
    const int col_g = 16 * SC.c0 + (SC.two ? (col & 31) : (col & 15));
    unsigned offA[kMR];
#pragma unroll
    for (int s = 0; s < kMR; ++s)
        offA[s] = (unsigned)((g * Np + 16 * min(row[s], nb - 1) + i) * 16);
    const unsigned offB = (unsigned)(((rg >> 1) * Np + col_g) * 16 + 8 * (rg & 1));
    const unsigned offW = (unsigned)(4 * rg * (int)sizeof(float2));
    v4f P[NW][kMR][2], Q[NW][kMR][2];
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int s = 0; s < kMR; ++s)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                P[w][s][c] = (v4f){0, 0, 0, 0};
                Q[w][s][c] = (v4f){0, 0, 0, 0};
            }
    constexpr bool deep = NW == 1;
    k_loop<NW, NF, kCols, kMR, deep>(
        r, sb, offA, offB, offW, rg, col,
        [&](const StageB<NW, kCols>& s, v4u(&Ah)[kMR], v4u(&Al)[kMR], const char* hn) {
            mfma_stage<NW, NF, deep>(s, Ah, Al, colb, on0, on1, P, Q, g, hn, r.plane, offA);
        });
    // FP32 accumulator lanes hold rows 4g + v of a block (column i); partial
    // slices are column-major: column = B-side real column
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        double2* Pw = a.P + ((int64_t)w * a.n_gangs + gang) * (int64_t)Np * Np;
#pragma unroll
        for (int s = 0; s < kMR; ++s)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                if (!(c == 0 ? on0[s] : on1[s])) continue;
                double2* dst = Pw + (int64_t)(16 * (S.c0 + c) + i) * Np + 16 * row[s] + 4 * g;
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    dst[v] = make_double2((double)P[w][s][c][v], (double)Q[w][s][c][v]);
            }
    }
}

const void* kernel_for(int nw, int nf)
{
    if (nw <= 1)
        return nf < 1 ? (const void*)k_contract_symbf16_strip<1, 0>
                      : (const void*)k_contract_symbf16_strip<1, 1>;
    return nf < 1   ? (const void*)k_contract_symbf16_strip<2, 0>
           : nf < 2 ? (const void*)k_contract_symbf16_strip<2, 1>
                    : (const void*)k_contract_symbf16_strip<2, 2>;
}

// Strip pairs and workgroups per pair: the most waves any pair's two strips
// need, in workgroups of kWaves.
void layout(int nb, int* wrows, int* wcols)
{
    const int ns = (nb + 1) / 2;
    int most = 1;
    for (int p = 0; p < (ns + 1) / 2; ++p) {
        const int jb = ns - 1 - p;
        most = std::max(most, strip_waves(p, nb) + (jb != p ? strip_waves(jb, nb) : 0));
    }
    *wrows = (most + kWaves - 1) / kWaves;
    *wcols = (ns + 1) / 2;
}

}  // namespace strip

// W (complex128 rows of K, `rows` rows present) -> the real basis's bf16 hi
// and lo planes of nq records of eight K rows (columns j < nU from Re W[h+j],
// nU <= j < N from Im W[j], zero beyond N).
__global__ void k_W64_to_Rbf16(const double2* __restrict__ W, int64_t rows, int N, int Np,
                               int64_t nq, v4u* __restrict__ Rb)
{
    const int h = N >> 1, nU = N - h;
    const int64_t n = nq * Np;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = x / Np;
        const int col = (int)(x - q * Np);
        double d[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int64_t k = 8 * q + t;
            d[t] = 0.0;
            if (k < rows && col < N) d[t] = col < nU ? W[k * Np + h + col].x : W[k * Np + col].y;
        }
        unsigned hw[4], lw[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            hw[t] = pk_bf16((float)d[2 * t], (float)d[2 * t + 1]);
            lw[t] = pk_bf16((float)(d[2 * t] - (double)lo_f(hw[t])),
                            (float)(d[2 * t + 1] - (double)hi_f(hw[t])));
        }
        Rb[x] = (v4u){hw[0], hw[1], hw[2], hw[3]};
        Rb[n + x] = (v4u){lw[0], lw[1], lw[2], lw[3]};
    }
}

int occupancy(const void* k, int threads)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, threads, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return per_cu;
}

}  // namespace

int contract_symbf16_max_sets(int form)
{
    return form == kBf16Strip ? strip::kMaxNW : fold::kMaxNW;
}

Bf16Geometry contract_symbf16_geometry(int form, int Np, int nw, int nf)
{
    const int nb = Np / 16;
    Bf16Geometry g;
    if (form == kBf16Strip) {
        strip::layout(nb, &g.wrows, &g.wcols);
        g.mb = strip::kMR;
        const int w = nw < 1 ? 1 : (nw > strip::kMaxNW ? strip::kMaxNW : nw);
        const int f = nf < 0 ? 0 : (nf > w ? w : nf);
        static std::atomic<int> occ[strip::kMaxNW + 1][strip::kMaxNW + 1] = {};
        if (occ[w][f] == 0) occ[w][f] = occupancy(strip::kernel_for(w, f), strip::kThreads);
        g.wg_per_cu = occ[w][f];
        return g;
    }
    fold::layout(nb, nw, &g.wrows, &g.wcols, &g.mb);
    static std::atomic<int> occ[fold::kMaxNW + 1][2][fold::kMaxMB + 1] = {};
    const int i = nw < 1 ? 1 : (nw > fold::kMaxNW ? fold::kMaxNW : nw);
    const int hh = nf < i ? 1 : 0;
    if (occ[i][hh][g.mb] == 0)
        occ[i][hh][g.mb] = occupancy(fold::kernel_for(i, nf, g.mb), fold::kThreads);
    g.wg_per_cu = occ[i][hh][g.mb];
    return g;
}

// Share of a launch's block slots that hold an upper-triangle block.
double contract_symbf16_slot_use(int form, int Np, int nw)
{
    const Bf16Geometry g = contract_symbf16_geometry(form, Np, nw, nw);
    const int nb = Np / 16;
    const double slots = form == kBf16Strip
                             ? (double)g.wcols * g.wrows * strip::kWaves * strip::kMR * 2
                             : (double)g.wcols * g.wrows * fold::kWaves * g.mb;
    return (double)nb * (nb + 1) / 2 / slots;
}

hipError_t contract_symbf16_launch(int form, int nw, int nf, int mb, const ContractArgs& a,
                                   hipStream_t stream)
{
    if (nw < 1 || nw > contract_symbf16_max_sets(form)) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)(a.n_gangs * a.wrows * a.wcols);
    ContractArgs args = a;
    void* params[] = {(void*)&args};
    if (form == kBf16Strip)
        return hipLaunchKernel(strip::kernel_for(nw, nf), dim3(grid), dim3(strip::kThreads),
                               params, 0, stream);
    return hipLaunchKernel(fold::kernel_for(nw, nf, mb), dim3(grid), dim3(fold::kThreads), params,
                           0, stream);
}

int64_t symbf16_records(int64_t Kp) { return (Kp + kKq - 1) / kKq * kKq / 8; }

hipError_t split_R_bf16(const double2* W64, int64_t rows, int N, int Np, int64_t nq, void* Rb,
                        hipStream_t stream)
{
    const int64_t n = nq * Np;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_W64_to_Rbf16, dim3(grid), dim3(256), 0, stream, W64, rows, N, Np, nq,
                       (v4u*)Rb);
    return hipGetLastError();
}

}  // namespace klsp
