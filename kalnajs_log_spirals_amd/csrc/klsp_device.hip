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
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "klsp.h"
#include "klsp_contract.hpp"
#include "klsp_cplx.hpp"
#include "klsp_internal.hpp"
#include "klsp_lu.hpp"

using namespace klsp;


#define KLSP_VERSION_STRING "klsp 0.1.0 gfx950"

// =====================================================================
// Device helpers
// =====================================================================


// sin x and cos x for |x| < 2^20, without branches: Cody-Waite reduction by
// pi/2 with pi/2 split into three doubles (x - q P1 - q P2 - q P3, each step
// an FMA), the Taylor series of sin (through x^17) and cos (through x^18) on
// [-pi/4, pi/4] -- the first omitted terms are below 1e-19 there -- and the
// quadrant q mod 4.  Absolute error <= 1.2e-16 (checked against glibc over
// |x| < 2^20 and at multiples of pi/2).
__device__ __forceinline__ void sincos_small(double x, double* sn, double* cs)
{
    const double q = rint(x * 0x1.45f306dc9c883p-1);          // x * 2/pi
    double r = fma(-q, 0x1.921fb54442d18p+0, x);
    r = fma(-q, 0x1.1a62633145c07p-54, r);
    r = fma(-q, -0x1.f1976b7ed8fbcp-110, r);
    const double z = r * r;
    double ps = fma(z, 1.0 / 355687428096000.0, -1.0 / 1307674368000.0);
    ps = fma(z, ps, 1.0 / 6227020800.0);
    ps = fma(z, ps, -1.0 / 39916800.0);
    ps = fma(z, ps, 1.0 / 362880.0);
    ps = fma(z, ps, -1.0 / 5040.0);
    ps = fma(z, ps, 1.0 / 120.0);
    ps = fma(z, ps, -1.0 / 6.0);
    const double S = fma(r * z, ps, r);
    double pc = fma(z, -1.0 / 6402373705728000.0, 1.0 / 20922789888000.0);
    pc = fma(z, pc, -1.0 / 87178291200.0);
    pc = fma(z, pc, 1.0 / 479001600.0);
    pc = fma(z, pc, -1.0 / 3628800.0);
    pc = fma(z, pc, 1.0 / 40320.0);
    pc = fma(z, pc, -1.0 / 720.0);
    pc = fma(z, pc, 1.0 / 24.0);
    pc = fma(z, pc, -0.5);
    const double C = fma(z, pc, 1.0);
    // quadrant n = q mod 4: (sin, cos) = (S, C), (C, -S), (-S, -C), (-C, S);
    // the sign flips are xors of the high words' sign bits
    const int n = (int)q & 3;
    const double ss = (n & 1) ? C : S, cc = (n & 1) ? S : C;
    *sn = __hiloint2double(__double2hiint(ss) ^ ((n & 2) << 30), __double2loint(ss));
    *cs = __hiloint2double(__double2hiint(cc) ^ (((n + 1) & 2) << 30), __double2loint(cc));
}

// Largest |argument| given to sincos_small by the W quadrature.
constexpr double kSmallTrigArg = 1048576.0;   // 2^20


// =====================================================================
// Stage 1: W_l(alpha) orbit quadrature (BasisFunctions.jl:199-266)
// =====================================================================

// Per-orbit preprocessing, one workgroup per orbit k: validity (any
// non-finite w1/ra/pha skips the orbit, :224-226), trapezoid weights of the
// non-uniform w1 samples (:229), phi = (Om2/Om1) w1 - pha (:232-235).  Emits
// per sample: u = ln r, wt = Sw1 / (pi sqrt r) (0 where the integrand is
// non-finite, :248), w1 and m*phi, padded to nwa_p samples with wt = 0.
// valid[k] = 0 skips the orbit; bit 1 marks that every argument of the
// quadrature's sines and cosines (l w1 + m phi with |l| <= lmax, 16 w1,
// alpha ln r with |alpha| <= amax) stays below kSmallTrigArg.
__global__ __launch_bounds__(256) void k_orbit_prep(
    const double* __restrict__ w1, const double* __restrict__ ra,
    const double* __restrict__ pha, const double* __restrict__ Om1,
    const double* __restrict__ Om2, int nwa, int nwa_p, int m, double lmax, double amax,
    double* __restrict__ u, double* __restrict__ wt, double* __restrict__ w1o,
    double* __restrict__ mph, int* __restrict__ valid)
{
    const int64_t k = blockIdx.x;
    const double* W1 = w1 + k * nwa;
    const double* RA = ra + k * nwa;
    const double* PH = pha + k * nwa;
    __shared__ int bad, big;
    if (threadIdx.x == 0) { bad = 0; big = 0; }
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
            if (!(lmax * fabs(x1) + fabs(mp) < kSmallTrigArg) ||
                !(16.0 * fabs(x1) < kSmallTrigArg) || !(amax * fabs(lu) < kSmallTrigArg))
                atomicOr(&big, 1);
        }
        const int64_t o = k * nwa_p + t;
        u[o] = lu; wt[o] = w; w1o[o] = x1; mph[o] = mp;
    }
    __syncthreads();
    if (threadIdx.x == 0) valid[k] = ok ? (big ? 1 : 3) : 0;
}

// compute_orbits (OrbitIntegration.jl:209-302), one workgroup per orbit.
// Circular (e < 1e-10) and epicyclic (e < 0.01) orbits are closed forms;
// general orbits need the cumulative trapezoid sums of drs dw sv over the nw
// phase samples (t = cumsum, :277-281; phase, :287-295).  They run as block
// scans: each thread sums its contiguous share of the terms, storing its
// running sums at the stride-th samples, the shares are prefix-summed, and
// each thread then adds its offset to the samples it stored.  The endpoint extrapolations of sv
// (2 sv[1] - sv[2], ...) are taken from the recomputed neighbours.
__device__ __forceinline__ double orbit_rv(double x, double E, double l2, bool nonradial)
{
    const double q = nonradial ? 2.0 * (E - log(x)) * (x * x) - l2 : 2.0 * (E - log(x));
    return sqrt(fmax(0.0, q));
}

__global__ __launch_bounds__(256) void k_orbits(
    const double* __restrict__ Rc, const double* __restrict__ ecc_a, const double* __restrict__ R1a,
    const double* __restrict__ R2a, const double* __restrict__ Ea, const double* __restrict__ La,
    const double* __restrict__ L2a, const double* __restrict__ SGa, const double* __restrict__ wg,
    const double* __restrict__ cw, const double* __restrict__ sw, int nw, int nwa,
    double* __restrict__ w1, double* __restrict__ ra, double* __restrict__ pha,
    double* __restrict__ Om1, double* __restrict__ Om2)
{
    const int64_t k = blockIdx.x;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int stride = (nw - 1) / (nwa - 1);
    const double pi = 3.14159265358979323846;
    const double Ri = Rc[k], ecc = ecc_a[k], r1 = R1a[k], r2 = R2a[k], Ek = Ea[k], Lk = La[k];
    const double l2 = L2a[k], sg = SGa[k];
    double* W1 = w1 + k * nwa;
    double* RA = ra + k * nwa;
    double* PH = pha + k * nwa;
    if (fabs(ecc) < 1e-10 || ecc < 0.01) {
        const bool circ = fabs(ecc) < 1e-10;
        const double kap = sqrt(2.0) / Ri, Om = 1.0 / Ri, aj = (r2 - r1) / 2;
        const double o1 = kap, o2 = Om * sg, c2 = 2 * Om / kap * aj / Ri;
        for (int a = tid; a < nwa; a += nth) {
            const int i = a * stride;
            W1[a] = wg[i];
            RA[a] = circ ? Ri : Ri - aj * cw[i];
            PH[a] = wg[i] * o2 / o1 + (circ ? 0.0 : c2 * sw[i] * sg);
        }
        if (tid == 0) { Om1[k] = o1; Om2[k] = o2; }
        return;
    }
    const double rs = (r1 + r2) / 2, drs = (r2 - r1) / 2, dw = wg[1] - wg[0];
    const bool nonradial = l2 > 1e-12;
    const bool phase = fabs(1.0 - ecc) > 1e-10 && nonradial;
    auto xs = [&](int i) { return rs - drs * cw[i]; };
    auto sv1_raw = [&](int i) {
        const double x = xs(i), rv = orbit_rv(x, Ek, l2, nonradial);
        return nonradial ? sw[i] * x / rv : sw[i] / rv;
    };
    auto sv2_raw = [&](int i) { return sw[i] / orbit_rv(xs(i), Ek, l2, nonradial); };
    // The first sample is extrapolated before the last one, so at nw = 3 the
    // last extrapolation reads the extrapolated first sample (:262-263).
    auto sv1_first = [&]() { return nonradial ? 2.0 * sv1_raw(1) - sv1_raw(2) : sv1_raw(0); };
    auto sv2_first = [&]() { return 2.0 * sv2_raw(1) - sv2_raw(2); };
    auto d1 = [&](int i) {
        double v;
        if (i == 0) v = sv1_first();
        else if (i == nw - 1) v = 2.0 * sv1_raw(nw - 2) - (nw == 3 ? sv1_first() : sv1_raw(nw - 3));
        else v = sv1_raw(i);
        return drs * dw * v;
    };
    auto d2 = [&](int i) {
        double v;
        if (i == 0) v = sv2_first();
        else if (i == nw - 1) v = 2.0 * sv2_raw(nw - 2) - (nw == 3 ? sv2_first() : sv2_raw(nw - 3));
        else v = sv2_raw(i);
        return drs * dw * v / xs(i);
    };
    // terms j = 1..nw-1; thread t owns [j0, j1).  One pass: the running
    // sums of the thread's own range are stored at its samples a * stride,
    // and after the block scan the thread adds its offset to them.
    const int chunk = (nw - 1 + nth - 1) / nth;
    const int j0 = 1 + tid * chunk, j1 = min(nw, j0 + chunk);
    const int a_lo = (j0 + stride - 1) / stride;          // first stored sample in range
    double s1 = 0.0, s2 = 0.0;
    if (j0 < j1) {
        double p1 = d1(j0 - 1), p2 = phase ? d2(j0 - 1) : 0.0;
        int next = a_lo * stride;
        for (int j = j0; j < j1; ++j) {
            const double c1 = d1(j);
            s1 += (p1 + c1) / 2;
            p1 = c1;
            if (phase) {
                const double c2v = d2(j);
                s2 += (p2 + c2v) / 2;
                p2 = c2v;
            }
            if (j == next) {
                const int a = j / stride;
                W1[a] = s1;
                PH[a] = s2;
                RA[a] = xs(j);
                next += stride;
            }
        }
    }
    __shared__ double sh1[256], sh2[256], tot[2];
    sh1[tid] = s1;
    sh2[tid] = s2;
    __syncthreads();
    if (tid == 0) {
        double a1 = 0.0, a2 = 0.0;
        for (int t = 0; t < nth; ++t) {
            const double v1 = sh1[t], v2 = sh2[t];
            sh1[t] = a1;
            sh2[t] = a2;
            a1 += v1;
            a2 += v2;
        }
        tot[0] = a1;
        tot[1] = a2;
    }
    __syncthreads();
    const double o1 = pi / tot[0];
    if (tid == 0) {
        W1[0] = 0.0 * o1;
        RA[0] = xs(0);
        PH[0] = phase ? Lk * 0.0 : 0.0;
        Om1[k] = o1;
        Om2[k] = o1 * (phase ? Lk * tot[1] : pi / 2 * 1.0 * sg) / pi;
    }
    if (j0 < j1) {
        const double off1 = sh1[tid], off2 = sh2[tid];
        for (int j = a_lo * stride; j < j1; j += stride) {
            const int a = j / stride;
            W1[a] = (off1 + W1[a]) * o1;
            PH[a] = phase ? Lk * (off2 + PH[a]) : pi / 2 * 1.0 * sg;
        }
    }
}

// W_k = A_k B_k per orbit on FP64 MFMA: A[l, t] = wt_t cos(l w1_t + m phi_t)
// (real, n_l x nwa), B[t, a] = exp(-i alpha_a ln r_t) (nwa x N).
// Workgroup = (orbit k, 64 alphas, 128 l) of twelve waves: waves 8-11
// generate the operands of stage s (32 samples t) into one LDS buffer while
// waves 0-7 run the MFMAs of stage s-1 from the other; one barrier per
// stage.  MFMA wave w owns alpha-block w % 4 (re and im) and l-blocks
// 4 (w / 4) .. +3.  On gfx950 FP64 VALU instructions and FP64 MFMAs share
// one pipeline per SIMD (measured: a wave of each on one SIMD takes the sum
// of their times), so the split does not hide operand generation; the time
// is the MFMAs plus the generation's FP64 instructions, which is why the
// generation uses rotations and a branch-free short sincos, and each
// generating thread covers 16 l and 8 alphas per sincos pair.  Lane layouts:
//   A (16 l x 4 t): lane = i + 16*kk holds A[l0+i][t0+kk]
//   B (4 t x 16 a): lane = i + 16*kk holds B[t0+kk][a0+i]
//   D (16 l x 16 a): lane holds D[kk + 4*v][i], v = 0..3 (the FP64 MFMA's
//   accumulator layout; the FP32 16x16x4 MFMA uses D[4*kk + v][i])
// When l_arr is a run of consecutive integers (the reference's l_min:l_max)
// a thread's l = l_b + 8 j for j = 0..15 come from one sincos and rotations
// by exp(8 i w1); evenly spaced alphas a_b + 8 j (j = 0..7) likewise by
// rotations by exp(-8 i dalpha ln r); otherwise every value is evaluated
// directly.
constexpr int kWTk = 32;    // samples t per LDS stage
constexpr int kWLc = 128;   // l per workgroup (8 l-blocks)
constexpr int kWAc = 64;    // alphas per workgroup (4 alpha-blocks)
constexpr int kWMmaWaves = 8;   // MFMA waves; 4 more generate operands

struct WStage {
    double A[kWTk][kWLc];
    double Br[kWTk][kWAc], Bi[kWTk][kWAc];
};

template <bool kSmall>
__device__ __forceinline__ void w_sincos(double x, double* sn, double* cs)
{
    if (kSmall) sincos_small(x, sn, cs);
    else sincos(x, sn, cs);
}

// Operand generation for one stage by the 256 generating threads: sample
// ft = p / 8 of the stage, column group fg = p % 8 (columns fg + 8 j).
template <bool kSmall>
__device__ __forceinline__ void w_fill(
    WStage& st, int p, int nlb, int n_l, const int* __restrict__ l_arr, int lc0, int l_consec,
    double lb, const double* al, double dalpha, double ut, double wtt, double w1t, double mp)
{
    const int ft = p >> 3, fg = p & 7;
    if (l_consec) {
        double s0, c0, sr, cr;
        w_sincos<kSmall>(lb * w1t + mp, &s0, &c0);
        w_sincos<kSmall>(8.0 * w1t, &sr, &cr);
        s0 *= wtt;                           // the rotations carry the weight along
        c0 *= wtt;
#pragma unroll
        for (int j = 0; j < kWLc / 8; ++j) {
            // column fg + 8 j is in l-block j / 2; rows of inactive blocks are
            // never read, rows past n_l only feed output rows that are not
            // stored, so every row is written (no branch in the unrolled loop)
            st.A[ft][fg + 8 * j] = c0;
            const double c1 = c0 * cr - s0 * sr;
            s0 = s0 * cr + c0 * sr;
            c0 = c1;
        }
    } else {
        for (int j = 0; j < kWLc / 8; ++j) {
            if ((j >> 1) >= nlb) break;
            const int l = lc0 + fg + 8 * j;
            double sl = 0.0, cl = 0.0;
            if (l < n_l) w_sincos<kSmall>((double)l_arr[l] * w1t + mp, &sl, &cl);
            st.A[ft][fg + 8 * j] = wtt * cl;
        }
    }
    if (dalpha != 0.0) {
        double sb, cb, sr, cr;
        w_sincos<kSmall>(-al[0] * ut, &sb, &cb);
        w_sincos<kSmall>(-8.0 * dalpha * ut, &sr, &cr);
#pragma unroll
        for (int j = 0; j < kWAc / 8; ++j) {
            st.Br[ft][fg + 8 * j] = cb;
            st.Bi[ft][fg + 8 * j] = sb;
            const double c1 = cb * cr - sb * sr;
            sb = sb * cr + cb * sr;
            cb = c1;
        }
    } else {
        for (int j = 0; j < kWAc / 8; ++j) {
            double sb, cb;
            w_sincos<kSmall>(-al[j] * ut, &sb, &cb);
            st.Br[ft][fg + 8 * j] = cb;
            st.Bi[ft][fg + 8 * j] = sb;
        }
    }
}

// Operand generation by waves 4..: stage s into buffer s & 1, one barrier
// per stage (nst + 1 in all, matching the MFMA waves).  kSmall selects
// sincos_small (every argument of the orbit below kSmallTrigArg) or the
// library's sincos.
template <bool kSmall>
__device__ __forceinline__ void w_produce(
    WStage* st, int64_t k, const double* __restrict__ u, const double* __restrict__ wt,
    const double* __restrict__ w1o, const double* __restrict__ mph,
    const double* __restrict__ alpha, const int* __restrict__ l_arr, int l_consec,
    double dalpha, int N, int n_l, int nwa_p, int lc0, int ac0, int nlb, int nst)
{
    const int p = threadIdx.x - 64 * kWMmaWaves, ft = p >> 3, fg = p & 7;
    const double lb = lc0 + fg < n_l ? (double)l_arr[lc0 + fg] : 0.0;
    double al[kWAc / 8];
#pragma unroll
    for (int j = 0; j < kWAc / 8; ++j) {
        const int a = ac0 + fg + 8 * j;
        al[j] = a < N ? alpha[a] : 0.0;
    }
    const double* U = u + k * nwa_p;
    const double* WT = wt + k * nwa_p;
    const double* X1 = w1o + k * nwa_p;
    const double* MP = mph + k * nwa_p;
    double ut = 0.0, wtt = 0.0, w1t = 0.0, mp = 0.0;
    if (ft < nwa_p) { ut = U[ft]; wtt = WT[ft]; w1t = X1[ft]; mp = MP[ft]; }
    for (int s = 0; s <= nst; ++s) {
        if (s < nst) {
            const int tn = (s + 1) * kWTk + ft;   // next stage's sample, loaded early
            double utn = 0.0, wtn = 0.0, w1n = 0.0, mpn = 0.0;
            if (tn < nwa_p) { utn = U[tn]; wtn = WT[tn]; w1n = X1[tn]; mpn = MP[tn]; }
            w_fill<kSmall>(st[s & 1], p, nlb, n_l, l_arr, lc0, l_consec, lb, al, dalpha, ut, wtt,
                           w1t, mp);
            ut = utn; wtt = wtn; w1t = w1n; mp = mpn;
        }
        __syncthreads();
    }
}

// MFMA wave: alpha-block ab, l-blocks lq0 .. lq0 + NQ - 1 (NQ fixed so the
// MFMA sequence has no branches); the MFMAs of stage s-1 read buffer
// (s-1) & 1 while stage s is generated.
template <int NQ>
__device__ __forceinline__ void w_consume(
    const WStage* st, int64_t k, int nst, int ab, int lq0, int lc0, int ac0, int64_t NPh, int N,
    int Np, int n_l, bool on, bool mirror, double2* __restrict__ W)
{
    const int lane = threadIdx.x & 63, i = lane & 15, kk = lane >> 4;
    v4d accr[NQ], acci[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) { accr[q] = (v4d){0, 0, 0, 0}; acci[q] = accr[q]; }
    for (int s = 0; s <= nst; ++s) {
        if (s >= 1 && on) {
            const WStage& b = st[(s - 1) & 1];
            // operands of k-step ks+1 are read before the MFMAs of ks
            double an[NQ], brn, bin;
#pragma unroll
            for (int q = 0; q < NQ; ++q) an[q] = b.A[kk][16 * (lq0 + q) + i];
            brn = b.Br[kk][16 * ab + i];
            bin = b.Bi[kk][16 * ab + i];
#pragma unroll
            for (int ks = 0; ks < kWTk / 4; ++ks) {
                double a[NQ];
#pragma unroll
                for (int q = 0; q < NQ; ++q) a[q] = an[q];
                const double br = brn, bi = bin;
                if (ks + 1 < kWTk / 4) {
#pragma unroll
                    for (int q = 0; q < NQ; ++q) an[q] = b.A[4 * (ks + 1) + kk][16 * (lq0 + q) + i];
                    brn = b.Br[4 * (ks + 1) + kk][16 * ab + i];
                    bin = b.Bi[4 * (ks + 1) + kk][16 * ab + i];
                }
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    accr[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], br, accr[q], 0, 0, 0);
                    acci[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], bi, acci[q], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    if (!on) return;
    const int a = ac0 + 16 * ab + i;
    if (a >= N) return;
    // symmetric alpha grid: W(-alpha) = conj W(alpha) (the l-side factor is
    // real); its centre column is alpha = 0 exactly, where e^(-i alpha ln r)
    // = 1 and W is real (the rotated sines only approximate that 0)
    const int am = N - 1 - a;
    const bool wm = mirror && am < a;
    const bool centre = mirror && am == a;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        for (int v = 0; v < 4; ++v) {
            const int li = lc0 + 16 * (lq0 + q) + kk + 4 * v;   // f64 D layout: row = lane/16 + 4v
            if (li < n_l) {
                double2* row = W + ((int64_t)li * NPh + k) * Np;
                row[a] = make_double2(accr[q][v], centre ? 0.0 : acci[q][v]);
                if (wm) row[am] = make_double2(accr[q][v], -acci[q][v]);
            }
        }
}

__global__ __launch_bounds__(64 * kWMmaWaves + 256, 1) void k_W_f64(
    const double* __restrict__ u, const double* __restrict__ wt,
    const double* __restrict__ w1o, const double* __restrict__ mph,
    const int* __restrict__ valid, const double* __restrict__ alpha,
    const int* __restrict__ l_arr, int l_consec, double dalpha, int64_t NPh, int N, int Np,
    int n_l, int nwa_p, int a_first, int mirror, double2* __restrict__ W)
{
    __shared__ WStage st[2];
    const int64_t k = blockIdx.x;
    const int f = valid[k];
    if (!f) return;
    const int wave = threadIdx.x >> 6;
    const int ac0 = a_first + blockIdx.y * kWAc, lc0 = blockIdx.z * kWLc;
    const int nlb = min(kWLc / 16, (n_l - lc0 + 15) / 16);   // active l-blocks
    const int nst = (nwa_p + kWTk - 1) / kWTk;
    if (wave >= kWMmaWaves) {
        if (f & 2)
            w_produce<true>(st, k, u, wt, w1o, mph, alpha, l_arr, l_consec, dalpha, N, n_l, nwa_p,
                            lc0, ac0, nlb, nst);
        else
            w_produce<false>(st, k, u, wt, w1o, mph, alpha, l_arr, l_consec, dalpha, N, n_l,
                             nwa_p, lc0, ac0, nlb, nst);
        return;
    }
    // MFMA wave w: alpha-block w % 4, l-blocks 4 (w / 4) .. +3 of the chunk
    const int ab = wave & 3, lq0 = 4 * (wave >> 2);
    const int nq = max(0, min(4, nlb - lq0));
    const bool on = ac0 + 16 * ab < N && nq > 0;
    switch (nq) {
    case 4: w_consume<4>(st, k, nst, ab, lq0, lc0, ac0, NPh, N, Np, n_l, on, mirror != 0, W); break;
    case 3: w_consume<3>(st, k, nst, ab, lq0, lc0, ac0, NPh, N, Np, n_l, on, mirror != 0, W); break;
    case 2: w_consume<2>(st, k, nst, ab, lq0, lc0, ac0, NPh, N, Np, n_l, on, mirror != 0, W); break;
    default: w_consume<1>(st, k, nst, ab, lq0, lc0, ac0, NPh, N, Np, n_l, on, mirror != 0, W); break;
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

// Mirror symmetry of W over its alpha columns, W[:, N-1-a] = conj W[:, a]
// with a real centre column (odd N), checked value by value over `rows`
// rows: any violation (NaN included) stores 1 in *flag.
__global__ void k_W_mirror_check(const double2* __restrict__ W, int64_t rows, int N, int Np,
                                 int* __restrict__ flag)
{
    const int h = N >> 1;
    const int64_t n = rows * (N - h);
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = x / (N - h);
        const int a = h + (int)(x - k * (N - h));   // upper half, centre included
        const double2 w = W[k * Np + a], wm = W[k * Np + (N - 1 - a)];
        if (!(wm.x == w.x && wm.y == -w.y)) *flag = 1;   // the centre: w.y == -w.y iff 0
    }
}

// The real basis of a symmetric alpha grid: R[k][j] = Re W[k][h + j] for
// j < nU = N - h, Im W[k][j] for nU <= j < N, 0 in the padding columns.
__global__ void k_W64_to_R(const double2* __restrict__ W, int64_t rows, int N, int Np,
                           double* __restrict__ R)
{
    const int h = N >> 1, nU = N - h;
    const int64_t n = rows * Np;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = x / Np;
        const int j = (int)(x - k * Np);
        const double2* row = W + k * Np;
        R[x] = j < nU ? row[h + j].x : (j < N ? row[j].y : 0.0);
    }
}

// =====================================================================
// Stage 2: resonance weights c[b][K] (MatrixBuilder.jl:43-51)
// d = omega - Omega_res (|d| < 1e-12 -> 1e-12 + 1e-12 i); c = DJ*F0l / d.
// =====================================================================
__global__ void k_weights(const double* __restrict__ DJ, const double* __restrict__ F0l,
                          const double* __restrict__ Ores, int64_t NPh, int64_t K, int64_t Kp,
                          const double* __restrict__ om, int nb, double2* __restrict__ c,
                          float2* __restrict__ cf)
{
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
        if (cf) cf[x] = make_float2((float)r.x, (float)r.y);
    }
}

// Taylor weights of M around a centre w_c: c[j][K] = DJ*F0l / d^(j+1),
// d = w_c - Omega_res (same clamp), j = 0..nt-1, so that
// M(w_c + s) = sum_j (-s)^j M_j for |s| < min |d| (>= Im w_c).
__global__ void k_weights_taylor(const double* __restrict__ DJ, const double* __restrict__ F0l,
                                 const double* __restrict__ Ores, int64_t NPh, int64_t K,
                                 int64_t Kp, double om_re, double om_im, int nt,
                                 double2* __restrict__ c, float2* __restrict__ cf)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < Kp;
         q += (int64_t)gridDim.x * blockDim.x) {
        double2 r = make_double2(0.0, 0.0), iv = make_double2(0.0, 0.0);
        if (q < K) {
            double dr = om_re - Ores[q], di = om_im;
            if (hypot(dr, di) < 1e-12) { dr = 1e-12; di = 1e-12; }
            const double xw = DJ[q % NPh] * F0l[q];
            iv = d_inv(dr, di);
            r = make_double2(xw * iv.x, xw * iv.y);
        }
        for (int j = 0; j < nt; ++j) {
            c[(int64_t)j * Kp + q] = r;
            if (cf) cf[(int64_t)j * Kp + q] = make_float2((float)r.x, (float)r.y);
            r = cmul(r, iv);
        }
    }
}

// omega-independent weights of the linear form's blocks for radial harmonic
// slab li (Matrix_LogExp.m:385-403): c[0][k] = DJ, c[1][k] = DJ Omega_res,
// c[2][k] = DJ F0l for the slab's orbits k < NPh, zero on the padding rows
// (rows of kstride each).
__global__ void k_weights_pencil(const double* __restrict__ DJ, const double* __restrict__ F0l,
                                 const double* __restrict__ Ores, int64_t NPh, int li,
                                 int64_t kstride, double2* __restrict__ c)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < kstride;
         k += (int64_t)gridDim.x * blockDim.x) {
        double d = 0.0, e = 0.0, f = 0.0;
        if (k < NPh) {
            const int64_t q = (int64_t)li * NPh + k;
            d = DJ[k];
            e = DJ[k] * Ores[q];
            f = DJ[k] * F0l[q];
        }
        c[k] = make_double2(d, 0.0);
        c[kstride + k] = make_double2(e, 0.0);
        c[2 * kstride + k] = make_double2(f, 0.0);
    }
}

__global__ void k_fill(double* __restrict__ x, int64_t n, double v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = v;
}

// =====================================================================
// Stage 3: the contraction M_w = sum_K W^T diag(c_w) conj(W): see
// klsp_contract.hpp (kernels in klsp_contract_f64.hip / klsp_contract_f32.hip).
// =====================================================================

// Sum of the K-split partials (fixed order, FP64), then *G and *N[alpha]
// (MatrixBuilder.jl:144-149).  M is [nb][N, N] column-major.  P holds S
// slices per omega; with cnt (balanced launches) the block (beta/16,
// alpha/16) sums its first cnt[beta/16 + (alpha/16) Np/16] slices.  A
// workgroup takes kRedElems consecutive outputs; its kRedParts waves sum the
// slices s = part (mod kRedParts) of each, and the parts are added in order.
constexpr int kRedElems = 64, kRedParts = 4;

__global__ __launch_bounds__(kRedElems * kRedParts) void k_reduce(
    const double2* __restrict__ P, int S, const int* __restrict__ cnt, int Np, int N, int nb,
    double G, const double* __restrict__ Nk, double2* __restrict__ M)
{
    __shared__ double2 part[kRedParts][kRedElems];
    const int64_t NN = (int64_t)N * N, n = nb * NN;
    const int e = threadIdx.x % kRedElems, j = threadIdx.x / kRedElems;
    for (int64_t x0 = (int64_t)blockIdx.x * kRedElems; x0 < n; x0 += (int64_t)gridDim.x * kRedElems) {
        const int64_t x = x0 + e;
        double sr = 0.0, si = 0.0;
        int beta = 0, alpha = 0;
        if (x < n) {
            const int b = (int)(x / NN);
            const int64_t r = x - b * NN;
            beta = (int)(r % N);
            alpha = (int)(r / N);
            const double2* src = P + (int64_t)b * S * Np * Np + (int64_t)alpha * Np + beta;
            const int ns = cnt ? cnt[(beta >> 4) + (alpha >> 4) * (Np >> 4)] : S;
            for (int s = j; s < ns; s += kRedParts) {
                const double2 v = src[(int64_t)s * Np * Np];
                sr += v.x;
                si += v.y;
            }
        }
        part[j][e] = make_double2(sr, si);
        __syncthreads();
        if (j == 0 && x < n) {
            double tr = 0.0, ti = 0.0;
#pragma unroll
            for (int q = 0; q < kRedParts; ++q) {
                tr += part[q][e].x;
                ti += part[q][e].y;
            }
            const double g = Nk[alpha];
            M[x] = make_double2(tr * G * g, ti * G * g);
        }
        __syncthreads();
    }
}

// Symmetric-basis launches: S_b = P + iQ for weight sets b < nb from the
// partial slices ([nb][nsl][Np][Np]) of the upper-triangle tiles (element
// (i, j), i <= j, sums the first cnt[i/16 + (j/16) Np/16] slices of its
// block, or all nsl without cnt; within a diagonal block the element above
// the diagonal stands for both), stored symmetric as S[i + N j] = S[j + N i].
// Work split as in k_reduce.
__global__ __launch_bounds__(kRedElems * kRedParts) void k_reduce_sym(
    const double2* __restrict__ P, int nsl, const int* __restrict__ cnt, int Np, int N, int nb,
    double2* __restrict__ S)
{
    __shared__ double2 part[kRedParts][kRedElems];
    const int64_t NN = (int64_t)N * N, n = nb * NN;
    const int e = threadIdx.x % kRedElems, q = threadIdx.x / kRedElems;
    for (int64_t x0 = (int64_t)blockIdx.x * kRedElems; x0 < n; x0 += (int64_t)gridDim.x * kRedElems) {
        const int64_t x = x0 + e;
        const int b = (int)(x / NN);
        const int64_t r = x - b * NN;
        const int i = (int)(r % N), j = (int)(r / N);
        const bool on = x < n && i <= j;
        double sr = 0.0, si = 0.0;
        if (on) {
            const double2* src = P + (int64_t)b * nsl * Np * Np + (int64_t)j * Np + i;
            const int ns = cnt ? cnt[(i >> 4) + (j >> 4) * (Np >> 4)] : nsl;
            for (int s = q; s < ns; s += kRedParts) {
                const double2 v = src[(int64_t)s * Np * Np];
                sr += v.x;
                si += v.y;
            }
        }
        part[q][e] = make_double2(sr, si);
        __syncthreads();
        if (q == 0 && on) {
            double tr = 0.0, ti = 0.0;
#pragma unroll
            for (int p = 0; p < kRedParts; ++p) {
                tr += part[p][e].x;
                ti += part[p][e].y;
            }
            S[x] = make_double2(tr, ti);
            S[b * NN + (int64_t)i * N + j] = make_double2(tr, ti);
        }
        __syncthreads();
    }
}

// M_m from S_m for nb matrices (see klsp_contract_sym64.hip): with
// a' = max(a, N-1-a),
// u = a' - N/2, v = a' if a' >= N - N/2, s = +1 (a = a') or -1,
//   H[b, a] = S[ub, ua] + i sb S[vb, ua] - i sa S[ub, va] + sb sa S[vb, va]
// (terms without a v dropped), then M = H * G * N[a] as k_reduce does.
__global__ void k_expand_sym(const double2* __restrict__ Sall, int N, int nb, double G,
                             const double* __restrict__ Nk, double2* __restrict__ Mall)
{
    const int h = N >> 1, nU = N - h;
    const int64_t NN = (int64_t)N * N, n = nb * NN;
    for (int64_t xx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; xx < n;
         xx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t m = xx / NN, x = xx - m * NN;
        const double2* S = Sall + m * NN;
        double2* M = Mall + m * NN;
        const int b = (int)(x % N), a = (int)(x / N);
        const int bp = b >= h ? b : N - 1 - b, ap = a >= h ? a : N - 1 - a;
        const double sb = bp == b ? 1.0 : -1.0, sa = ap == a ? 1.0 : -1.0;
        const int ub = bp - h, ua = ap - h;
        const bool hb = bp >= nU, ha = ap >= nU;
        const double2 t0 = S[(int64_t)ua * N + ub];
        double hr = t0.x, hi = t0.y;
        if (hb) {                       // + i sb S[vb, ua]
            const double2 t = S[(int64_t)ua * N + bp];
            hr -= sb * t.y;
            hi += sb * t.x;
        }
        if (ha) {                       // - i sa S[ub, va]
            const double2 t = S[(int64_t)ap * N + ub];
            hr += sa * t.y;
            hi -= sa * t.x;
        }
        if (hb && ha) {                 // + sb sa S[vb, va]
            const double2 t = S[(int64_t)ap * N + bp];
            hr += sb * sa * t.x;
            hi += sb * sa * t.y;
        }
        const double g = Nk[a];
        M[x] = make_double2(hr * G * g, hi * G * g);
    }
}

// Sum of squares of |A_b - B_b| (B null: of |A_b|) over the NN entries of
// matrix b = blockIdx.x, in FP64 -> out[b] (the Taylor-term norms of the
// mixed search's error bound).
__global__ __launch_bounds__(256) void k_sumsq(const double2* __restrict__ A,
                                               const double2* __restrict__ B, int64_t NN,
                                               double* __restrict__ out)
{
    __shared__ double red[256];
    const double2* a = A + (int64_t)blockIdx.x * NN;
    const double2* b = B ? B + (int64_t)blockIdx.x * NN : nullptr;
    double acc = 0.0;
    for (int64_t x = threadIdx.x; x < NN; x += blockDim.x) {
        double2 v = a[x];
        if (b) {
            const double2 u = b[x];
            v.x -= u.x;
            v.y -= u.y;
        }
        acc = fma(v.x, v.x, fma(v.y, v.y, acc));
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// =====================================================================
// Stage 4: det(I - M) by complex LU with partial pivoting, one workgroup
// per matrix (MatrixBuilder.jl:182 -> LinearAlgebra det -> LAPACK zgetrf).
// Pivot = first row maximising |re|+|im| (izamax), multipliers scaled by the
// reciprocal pivot; det = prod U_jj * (-1)^swaps in LinearAlgebra's order.
// Blocked right-looking form, as zgetrf: a panel of up to kLuMaxNb columns is
// loaded into LDS and factored there column by column by one wave (no
// workgroup barrier per column); then every wave takes
// kLuCols trailing columns at a time, applies the panel's row permutation (one
// gather), solves U12 = L11^-1 A12 with lane broadcasts and subtracts
// L21 U12 (L21 read from LDS), so the trailing matrix in global memory
// (L2-resident) is read and written once per panel.  Only what the
// determinant needs is kept: the L and U12 parts are not written back.
// =====================================================================
constexpr int kLuThreads = 512;       // (1024 threads measured slower: more waves per barrier)
constexpr int kLuMaxNb = 8;           // panel width
constexpr int kLuCols = 4;            // trailing columns per wave pass (U12: 128 VGPRs)
constexpr int kLuPanel = 8192;        // panel elements in LDS (N * width)

__device__ __forceinline__ void argmax_merge(double& v, int& idx, double ov, int oi)
{
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
}



// With coef == nullptr matrix b is I - M[b]; otherwise it is
// I - sum_j coef[b*nt + j] M[j] (a point of the Taylor model of M).
__global__ __launch_bounds__(kLuThreads) void k_lu_det(const double2* __restrict__ M,
                                                        const double2* __restrict__ coef, int nt,
                                                        double2* __restrict__ Awork, int N,
                                                        double2* __restrict__ det_out)
{
    const int b = blockIdx.x;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int lane = tid & 63, wid = tid >> 6, nwv = nth >> 6;
    const int64_t NN = (int64_t)N * N;
    double2* A = Awork + b * NN;
    __shared__ double2 P[kLuPanel];   // panel, column-major, leading dimension m
    __shared__ int perm[kMaxN];
    __shared__ int pv[kLuMaxNb];

    for (int64_t x = tid; x < NN; x += nth) {
        double2 m;
        if (coef == nullptr) {
            m = M[b * NN + x];
        } else {
            m = make_double2(0.0, 0.0);
            for (int j = 0; j < nt; ++j) {
                const double2 t = cmul(coef[b * nt + j], M[j * NN + x]);
                m = make_double2(m.x + t.x, m.y + t.y);
            }
        }
        const int r = (int)(x % N), c = (int)(x / N);
        A[x] = make_double2((r == c ? 1.0 : 0.0) - m.x, 0.0 - m.y);
    }
    __syncthreads();

    const int nbk = max(1, min(kLuMaxNb, kLuPanel / N));
    double2 det = make_double2(1.0, 0.0);
    int swaps = 0;
    bool singular = false;
    for (int k0 = 0; k0 < N; k0 += nbk) {
        const int kb = min(nbk, N - k0), m = N - k0;
        for (int x = tid; x < m * kb; x += nth) {
            const int t = x / m, r = x - t * m;
            P[x] = A[(k0 + r) + (int64_t)N * (k0 + t)];
        }
        __syncthreads();
        // unblocked factorisation of the panel by wave 0 alone: no workgroup
        // barrier per column (a wave's LDS accesses execute in program order)
        if (wid == 0) {
            for (int jj = 0; jj < kb; ++jj) {
                double2* Pj = P + jj * m;
                // izamax: the first entry (the diagonal) seeds the search and
                // a later entry replaces it only if strictly larger, so NaN
                // entries below the diagonal are never chosen, while a NaN
                // diagonal is the pivot and propagates into det
                double bv = -1.0;
                int bi = m;
                for (int r = jj + lane; r < m; r += 64) {
                    const double2 a = Pj[r];
                    const double v = fabs(a.x) + fabs(a.y);
                    if (v > bv) { bv = v; bi = r; }
                }
                for (int off = 32; off >= 1; off >>= 1) {
                    const double ov = __shfl_xor(bv, off);
                    const int oi = __shfl_xor(bi, off);
                    argmax_merge(bv, bi, ov, oi);
                }
                {
                    const double2 d = Pj[jj];
                    if (isnan(fabs(d.x) + fabs(d.y))) { bv = NAN; bi = jj; }
                }
                if (bv == 0.0) {
                    // an exactly zero pivot: zgetf2 reports info > 0 and
                    // LinearAlgebra's det(::LU) returns 0 for that factorisation
                    singular = true;
                    if (lane == 0) pv[jj] = jj;
                    continue;
                }
                const int p = bi;
                if (lane == 0) pv[jj] = p;
                if (p != jj) {
                    ++swaps;
                    if (lane < kb) {
                        const double2 tmp = P[lane * m + jj];
                        P[lane * m + jj] = P[lane * m + p];
                        P[lane * m + p] = tmp;
                    }
                }
                // lane t swapped column t; every lane reads the swapped rows
                // below.  The wave's LDS accesses complete in order, but
                // without this fence the compiler may hand a lane that took
                // no part in the swap a value it loaded before it (Pj[jj]).
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                const double2 piv = Pj[jj];
                det = cmul(det, piv);
                const double2 rp = d_inv(piv.x, piv.y);
                for (int r = jj + 1 + lane; r < m; r += 64) {
                    const double2 l = cmul(Pj[r], rp);
                    Pj[r] = l;
                    for (int t = jj + 1; t < kb; ++t) P[t * m + r] = cfms(P[t * m + r], l, P[t * m + jj]);
                }
            }
        }
        __syncthreads();
        // the panel's row interchanges as one permutation of rows k0..N-1
        for (int r = tid; r < m; r += nth) perm[r] = r;
        __syncthreads();
        if (tid == 0)
            for (int jj = 0; jj < kb; ++jj) {
                const int t = perm[jj];
                perm[jj] = perm[pv[jj]];
                perm[pv[jj]] = t;
            }
        __syncthreads();
        // trailing columns, kLuCols per wave at a time
        const int c0 = k0 + kb, ncol = N - c0;
        for (int cg = wid * kLuCols; cg < ncol; cg += nwv * kLuCols) {
            const int nq = min(kLuCols, ncol - cg);
            double2 u[kLuCols][kLuMaxNb];
            // chunks of 64 rows; the next chunk's gathers are issued before
            // the current one is computed, so their L2 latency overlaps it
            auto gather = [&](int r, double2 (&v)[kLuCols]) {
#pragma unroll
                for (int q = 0; q < kLuCols; ++q)
                    v[q] = (q < nq && r < m) ? A[(k0 + perm[r]) + (int64_t)N * (c0 + cg + q)]
                                             : make_double2(0.0, 0.0);
            };
            double2 x[kLuCols], xn[kLuCols];
            gather(lane, x);
            for (int r0 = 0; r0 < m; r0 += 64) {
                const int r = r0 + lane;
                if (r0 + 64 < m) gather(r + 64, xn);
                if (r0 == 0) {
                    // U12 = L11^-1 A12 (unit lower L11): lane t holds row t.
                    // The L11 entries are read first and the rows broadcast
                    // with readlane, so the dependent chain has no LDS trip.
                    double2 lrow[kLuMaxNb];
#pragma unroll
                    for (int t = 0; t < kLuMaxNb; ++t)
                        lrow[t] = (t < kb && lane > t && lane < kb) ? P[t * m + lane]
                                                                    : make_double2(0.0, 0.0);
#pragma unroll
                    for (int t = 0; t < kLuMaxNb; ++t) {
                        if (t < kb) {
#pragma unroll
                            for (int q = 0; q < kLuCols; ++q)
                                u[q][t] = make_double2(bcast(x[q].x, t), bcast(x[q].y, t));
                            if (lane > t && lane < kb) {
#pragma unroll
                                for (int q = 0; q < kLuCols; ++q)
                                    x[q] = cfms(x[q], lrow[t], u[q][t]);
                            }
                        }
                    }
                }
                if (r >= kb && r < m) {
#pragma unroll
                    for (int t = 0; t < kLuMaxNb; ++t) {
                        if (t >= kb) break;
                        const double2 l = P[t * m + r];
#pragma unroll
                        for (int q = 0; q < kLuCols; ++q) x[q] = cfms(x[q], l, u[q][t]);
                    }
#pragma unroll
                    for (int q = 0; q < kLuCols; ++q)
                        if (q < nq) A[(k0 + r) + (int64_t)N * (c0 + cg + q)] = x[q];
                }
#pragma unroll
                for (int q = 0; q < kLuCols; ++q) x[q] = xn[q];
            }
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
// Variant of the symmetric-basis FP64 contraction used until the autotuner
// has measured the problem's width (index into contract_sym64_name).
constexpr int kSymDefaultVariant = 2;   // depth 8, upper-triangle diagonal tiles
constexpr int kCircle = 16;               // circle points of a model-det zero search
constexpr int kModelPts = kCircle + 1;    // with the centre: matrices per model LU launch
constexpr int kNorms = kMaxW;             // Frobenius norms of a pass's Taylor terms
struct klsp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int n_cu = 256;
    // problem
    bool have_problem = false;
    bool have_arrays = false;  // DJ, F0l, Omega_res set
    int64_t NPh = 0, N = 0, n_l = 0, K = 0, Kp = 0;
    int Np = 0, m = 0;
    double G = 0.0;
    double *d_DJ = nullptr, *d_F0l = nullptr, *d_Ores = nullptr, *d_Nk = nullptr;
    // W and its low-precision copies, grow-only (in elements): a new problem
    // reuses them when it fits, so a batch of problems allocates its largest
    // W once; klsp_destroy releases them
    double2* d_W64 = nullptr; size_t W64_cap = 0;
    float2* d_W32 = nullptr;  size_t W32_cap = 0;
    uint4* d_Wb = nullptr;    size_t Wb_cap = 0;   // paired bf16 hi/lo rows (BF16x3 builds)
    bool have_W = false;
    bool have_W32 = false, have_Wb = false;   // copies of the current W
    // workspaces
    double* d_om = nullptr;
    double2* d_coef = nullptr;  // [kModelPts][kMaxW] Taylor-model coefficients
    double* d_norm = nullptr;   // [kNorms] sums of squares (k_sumsq) for the mixed search
    double2* d_c = nullptr;   size_t c_cap = 0;
    float2* d_cf = nullptr;   size_t cf_cap = 0;   // FP32 copy of the weights
    double2* d_P = nullptr;   size_t P_cap = 0;
    double2* d_M = nullptr;   size_t M_cap = 0;
    double2* d_A = nullptr;   size_t A_cap = 0;
    double2* d_det = nullptr; size_t det_cap = 0;
    int last_S = 0;           // K-splits of the last contraction launch
    const int* last_count = nullptr;  // per-block piece counts of a balanced launch
    // balanced work tables of the FP64 / FP32 contractions for 1..kMaxBatch
    // omegas, built on first use for the current problem
    struct Balanced {
        bool built = false;
        int wrows = 0, wcols = 0, n_slots = 0, n_units = 0;
        WorkUnit* d_units = nullptr;
        int* d_count = nullptr;
    } bal[2][kMaxBatch + 1];
    Balanced bal_sym;            // ... of the symmetric-basis FP64 contraction
    int bal_sym_variant = -1;    // the kernel variant bal_sym was planned for
    int sym_variant = kSymDefaultVariant;   // kernel variant of symmetric-basis builds
    std::map<int, int> sym_tuned;           // autotuned variant per padded width Np
    // real basis R of a symmetric alpha grid (FP64 one-omega builds): sym is
    // -1 until W has been checked, then 1 (W[:, N-1-a] = conj W[:, a]) or 0;
    // use_sym = 0 turns the real-basis builds off (klsp_set_symmetric_basis)
    double* d_R = nullptr;    size_t R_cap = 0;
    bool have_R = false;
    int sym = -1;
    int use_sym = 1;
    bool last_sym = false;       // the last contraction launch was symmetric-basis
    double2* d_S = nullptr;   size_t S_cap = 0;   // S = P + iQ of that launch
    uint4* d_Rb = nullptr;    size_t Rb_cap = 0;  // bf16 hi/lo planes of R (BF16x3 builds)
    bool have_Rb = false;
    float* d_R32 = nullptr;   size_t R32_cap = 0; // R in FP32 (FP32 builds)
    bool have_R32 = false;
    Balanced bal_sym32;          // work table of the real-basis FP32 contraction
    int* d_flag = nullptr;
    double* d_ones = nullptr; size_t ones_cap = 0;   // unit column scale (linear-form blocks)
    void* d_scratch = nullptr;   // this device's copy of `scratch` (klsp_det_noalloc)
    // the grid scan as one HIP graph: its omegas and dets live on the device,
    // and the instantiated graph is kept for the next scan with the same key
    // (every pointer, size and launch parameter the capture recorded)
    double* d_scan_om = nullptr;   size_t scan_om_cap = 0;
    double2* d_scan_det = nullptr; size_t scan_det_cap = 0;
    hipGraphExec_t scan_exec = nullptr;
    std::vector<int64_t> scan_key;
    int scan_graph = 1;            // klsp_set_scan_graph
    // LU form (klsp_set_lu_form): 0 = step-wise for batches of at most
    // kLuStepMaxMatrices of order <= kLuStepMaxN, one workgroup per matrix
    // otherwise; 1 / 2 force one (2 only up to order kLuStepMaxN)
    void* d_luaux = nullptr;  size_t luaux_cap = 0;   // step-wise LU bookkeeping (bytes)
    int lu_form = 0;
    // real-basis BF16 contraction form (klsp_set_bf16_form): 0 = chosen per
    // problem (bf16_form), kBf16Fold / kBf16Strip force one
    int bf16_form = 0;
    int64_t scan_graph_builds = 0, scan_graph_launches = 0;
    // profiling
    bool prof = false;
    hipEvent_t ev[6] = {};
    klsp_timing last{};
    double last_W_ms = 0.0;
    double last_orbit_ms = 0.0;
};

// Workspace of klsp_det_noalloc: a file-scope device array, so that the
// entry point allocates nothing.  Each GPU the library runs on holds its own
// copy until the process ends.  Host code never passes the array's name to a
// kernel (in host code it is only a stand-in for the symbol): the device
// address of each GPU's copy comes from hipGetSymbolAddress, taken once per
// device at context creation and kept in g_scratch_ptr.
constexpr size_t kScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static std::map<int, void*> g_scratch_ptr;   // device -> address of its `scratch`

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

// K-split geometry of a contraction launch.
struct GangPlan {
    int wrows, wcols, gangs_per_xcd, n_gangs;
    int64_t Kg;
};

// Balanced work table of an FP64 / FP32 contraction shape (see
// klsp_contract.hpp).  A tile of b blocks gets ceil(b * pf / bmax) K pieces,
// so no unit does more than bmax / pf blocks x K of work; pf is the largest
// value whose units fit the resident wave slots of the 8 XCDs in one round
// (with equal gangs, four rounds letting freed slots take the next gang's
// waves measured 7% slower: the gangs lose L2 locality), with at least 256
// rows per unit; piece boundaries are multiples of the kernel's rows per
// loop trip (`quantum`), the last piece ending at Kp.  Units with the same piece count and index
// cover the same K range and are packed onto one XCD where it has room
// (largest groups first, least-loaded XCD).  With `uniform` every tile gets
// the same number of pieces, 8 x (XCD slots / tiles): equal gangs of whole
// tile sets, one K range each.
// Tiles of a launch: work-table index (tile row * wcols + tile column) and
// the number of 16 x 16 output blocks the tile computes.
struct TileDesc {
    int widx, blocks;
};

static bool plan_units(const klsp_ctx* ctx, const std::vector<TileDesc>& tiles, int bmax,
                       int waves_per_cu, bool uniform, int quantum, klsp_ctx::Balanced& out,
                       std::vector<WorkUnit>& units, std::vector<int>& pieces)
{
    const int nt = (int)tiles.size();
    const int slots = std::max(1, ctx->n_cu / 8) * std::max(1, waves_per_cu);
    const int64_t Kq = ctx->Kp / quantum;  // K in units of the loop trip (the last piece ends at Kp)
    pieces.assign(nt, 1);
    const int64_t pf_max = std::max<int64_t>(1, std::min<int64_t>(ctx->Kp / 256, 8 * slots));
    const int64_t pu = std::max<int64_t>(1, std::min<int64_t>(ctx->Kp / 256, 8 * (slots / nt)));
    for (int pf = (int)pf_max; pf >= 1; --pf) {
        int64_t total = 0;
        for (int t = 0; t < nt; ++t) {
            pieces[t] = uniform ? (int)std::min<int64_t>(pu, pf)
                                : (int)std::min<int64_t>(std::max<int64_t>(1, Kq),
                                                         ((int64_t)tiles[t].blocks * pf + bmax - 1) / bmax);
            total += pieces[t];
        }
        if (total > 8 * (int64_t)slots) continue;
        // groups of tiles sharing (pieces, piece index), i.e. one K range
        std::vector<std::pair<std::pair<int, int>, std::vector<int>>> groups;
        {
            std::vector<int> pvals(pieces);
            std::sort(pvals.begin(), pvals.end());
            pvals.erase(std::unique(pvals.begin(), pvals.end()), pvals.end());
            for (int pv : pvals)
                for (int j = 0; j < pv; ++j) {
                    std::vector<int> ts;
                    for (int t = 0; t < nt; ++t)
                        if (pieces[t] == pv) ts.push_back(tiles[t].widx);
                    groups.push_back({{pv, j}, ts});
                }
        }
        std::stable_sort(groups.begin(), groups.end(), [](const auto& x, const auto& y) {
            return x.second.size() > y.second.size();
        });
        std::vector<std::vector<WorkUnit>> bins(8);
        for (const auto& g : groups) {
            const int pv = g.first.first, j = g.first.second;
            const int64_t Kb = (int64_t)j * Kq / pv * quantum;
            const int64_t Ke = j + 1 == pv ? ctx->Kp : (int64_t)(j + 1) * Kq / pv * quantum;
            size_t next = 0;
            while (next < g.second.size()) {
                // least-loaded XCD that holds the rest of the group, else the
                // least-loaded one with any room
                int best = -1;
                for (int x = 0; x < 8; ++x)
                    if ((int)(bins[x].size() + g.second.size() - next) <= slots &&
                        (best < 0 || bins[x].size() < bins[best].size()))
                        best = x;
                if (best < 0)
                    for (int x = 0; x < 8; ++x)
                        if ((int)bins[x].size() < slots &&
                            (best < 0 || bins[x].size() < bins[best].size()))
                            best = x;
                while (next < g.second.size() && (int)bins[best].size() < slots)
                    bins[best].push_back(WorkUnit{Kb, Ke, g.second[next++], j});
            }
        }
        size_t per = 0;
        for (const auto& b : bins) per = std::max(per, b.size());
        units.assign(8 * per, WorkUnit{0, 0, -1, 0});
        for (int x = 0; x < 8; ++x)
            for (size_t k = 0; k < bins[x].size(); ++k) units[8 * k + x] = bins[x][k];
        out.n_slots = *std::max_element(pieces.begin(), pieces.end());
        out.n_units = (int)units.size();
        return true;
    }
    return false;
}

// Balanced work table of an FP64 / FP32 contraction shape over the whole
// output (see klsp_contract.hpp); count[bi + bj nb] = the piece count of the
// tile holding block (bi, bj).
static bool plan_balanced(const klsp_ctx* ctx, const ContractShape& sh, bool uniform, int quantum,
                          klsp_ctx::Balanced& out, std::vector<WorkUnit>& units,
                          std::vector<int>& count)
{
    const int nb = ctx->Np / 16;
    const int wrows = (nb + sh.MB - 1) / sh.MB, wcols = (nb + sh.NB - 1) / sh.NB;
    std::vector<TileDesc> tiles;
    int bmax = 1;
    for (int t = 0; t < wrows * wcols; ++t) {
        const int r = t / wcols, c = t % wcols;
        const int b = std::min(sh.MB, nb - r * sh.MB) * std::min(sh.NB, nb - c * sh.NB);
        tiles.push_back({t, b});
        bmax = std::max(bmax, b);
    }
    std::vector<int> pieces;
    if (!plan_units(ctx, tiles, bmax, sh.waves_per_cu, uniform, quantum, out, units, pieces))
        return false;
    count.assign((size_t)nb * nb, 1);
    for (int bi = 0; bi < nb; ++bi)
        for (int bj = 0; bj < nb; ++bj)
            count[bi + (size_t)bj * nb] = pieces[(bi / sh.MB) * wcols + bj / sh.NB];
    out.wrows = wrows;
    out.wcols = wcols;
    return true;
}

// The same for the symmetric-basis FP64 contraction: square tiles on or above
// the diagonal only, a diagonal tile of m x m blocks computing m (m + 1) / 2
// of them (all m^2 in the full-diagonal variants); count covers blocks
// bi <= bj.
static bool plan_balanced_sym(const klsp_ctx* ctx, int variant, klsp_ctx::Balanced& out,
                              std::vector<WorkUnit>& units, std::vector<int>& count)
{
    const ContractShape sh = contract_sym64_shape(variant);
    const bool full = contract_sym64_full_diag(variant);
    const int nb = ctx->Np / 16;
    const int wt = (nb + sh.MB - 1) / sh.MB;
    std::vector<TileDesc> tiles;
    std::vector<int> index(wt * wt, -1);
    for (int r = 0; r < wt; ++r)
        for (int c = r; c < wt; ++c) {
            const int mv = std::min(sh.MB, nb - r * sh.MB), nv = std::min(sh.MB, nb - c * sh.MB);
            index[r * wt + c] = (int)tiles.size();
            tiles.push_back({r * wt + c, r == c && !full ? mv * (mv + 1) / 2 : mv * nv});
        }
    std::vector<int> pieces;
    if (!plan_units(ctx, tiles, sh.MB * sh.MB, sh.waves_per_cu, false,
                    contract_sym64_quantum(variant), out, units, pieces))
        return false;
    count.assign((size_t)nb * nb, 0);
    for (int bi = 0; bi < nb; ++bi)
        for (int bj = bi; bj < nb; ++bj)
            count[bi + (size_t)bj * nb] = pieces[index[(bi / sh.MB) * wt + bj / sh.MB]];
    out.wrows = wt;
    out.wcols = wt;
    return true;
}

// The same for the real-basis FP32 contraction: 4 x 2 tiles (r, c) with any
// block on or above the diagonal (4r <= 2c + 1), each computing all its
// blocks; count covers blocks bi <= bj.
static bool plan_balanced_sym32(const klsp_ctx* ctx, klsp_ctx::Balanced& out,
                                std::vector<WorkUnit>& units, std::vector<int>& count)
{
    const ContractShape sh = contract_sym32_shape();
    const int nb = ctx->Np / 16;
    const int wr = (nb + sh.MB - 1) / sh.MB, wc = (nb + sh.NB - 1) / sh.NB;
    std::vector<TileDesc> tiles;
    std::vector<int> index(wr * wc, -1);
    for (int r = 0; r < wr; ++r)
        for (int c = 0; c < wc; ++c) {
            if (sh.MB * r > sh.NB * c + sh.NB - 1) continue;
            const int mv = std::min(sh.MB, nb - r * sh.MB), nv = std::min(sh.NB, nb - c * sh.NB);
            index[r * wc + c] = (int)tiles.size();
            tiles.push_back({r * wc + c, mv * nv});
        }
    // every tile cut into the same K ranges: 15.9 -> 13.9 ms at the default
    // grid (pieces in proportion to a tile's blocks give its 8-, 6-, 4- and
    // 3-block tiles four different K partitions); the FP64 real-basis
    // kernel measured 16.0 -> 19.0 ms the same way and keeps proportional
    // pieces
    std::vector<int> pieces;
    if (!plan_units(ctx, tiles, sh.MB * sh.NB, sh.waves_per_cu, true, contract_sym32_quantum(),
                    out, units, pieces))
        return false;
    count.assign((size_t)nb * nb, 0);
    for (int bi = 0; bi < nb; ++bi)
        for (int bj = bi; bj < nb; ++bj)
            count[bi + (size_t)bj * nb] = pieces[index[(bi / sh.MB) * wc + bj / sh.NB]];
    out.wrows = wr;
    out.wcols = wc;
    return true;
}

static ContractShape shape_for(int precision, int nw);

// The balanced table for (precision, nw), built and uploaded on first use.
// Uploads a planned work table into b (stored in b before the copies, so
// that a failed call leaves nothing unowned: a retry or free_balanced
// releases it).
static int store_balanced(klsp_ctx* ctx, klsp_ctx::Balanced& b, const klsp_ctx::Balanced& nbal,
                          const std::vector<WorkUnit>& units, const std::vector<int>& count)
{
    hipFree(b.d_units);
    hipFree(b.d_count);
    b.d_units = nullptr;
    b.d_count = nullptr;
    KLSP_HIP(ctx, hipMalloc((void**)&b.d_units, units.size() * sizeof(WorkUnit)));
    KLSP_HIP(ctx, hipMalloc((void**)&b.d_count, count.size() * sizeof(int)));
    KLSP_HIP(ctx, hipMemcpy(b.d_units, units.data(), units.size() * sizeof(WorkUnit),
                            hipMemcpyHostToDevice));
    KLSP_HIP(ctx, hipMemcpy(b.d_count, count.data(), count.size() * sizeof(int),
                            hipMemcpyHostToDevice));
    b.wrows = nbal.wrows;
    b.wcols = nbal.wcols;
    b.n_slots = nbal.n_slots;
    b.n_units = nbal.n_units;
    b.built = true;
    return KLSP_OK;
}

// The balanced table for (precision, nw), built and uploaded on first use.
static int get_balanced(klsp_ctx* ctx, int precision, int nw, const klsp_ctx::Balanced** out)
{
    klsp_ctx::Balanced& b = ctx->bal[precision == KLSP_PREC_FP64 ? 0 : 1][nw];
    if (!b.built) {
        std::vector<WorkUnit> units;
        std::vector<int> count;
        klsp_ctx::Balanced nbal;
        // the one-omega FP32 kernel (4 x 2 tiles) keeps equal gangs: its
        // balanced table measured 9% slower at the default grid (43.3 vs 39.6
        // ms; more K ranges in flight per XCD), while the FP64 kernels gain 8%
        // (59.5 -> 54.8 ms) and the three-omega FP32 kernel 11%
        const bool uniform = precision == KLSP_PREC_FP32 && nw == 1;
        const int quantum = precision == KLSP_PREC_FP64 ? (nw == 1 ? kRowsPerTrip<true, 1>
                                                                   : kRowsPerTrip<true, 2>)
                                                        : kRowsPerTrip<false, 1>;
        if (!plan_balanced(ctx, shape_for(precision, nw), uniform, quantum, nbal, units, count))
            return fail(ctx, KLSP_ERR_INVALID_ARG, "contraction: no work plan fits the device");
        int rc = store_balanced(ctx, b, nbal, units, count);
        if (rc) return rc;
    }
    *out = &b;
    return KLSP_OK;
}

// The balanced table of the symmetric-basis FP64 contraction.
static int get_balanced_sym(klsp_ctx* ctx, const klsp_ctx::Balanced** out)
{
    klsp_ctx::Balanced& b = ctx->bal_sym;
    if (!b.built || ctx->bal_sym_variant != ctx->sym_variant) {
        std::vector<WorkUnit> units;
        std::vector<int> count;
        klsp_ctx::Balanced nbal;
        if (!plan_balanced_sym(ctx, ctx->sym_variant, nbal, units, count))
            return fail(ctx, KLSP_ERR_INVALID_ARG, "contraction: no work plan fits the device");
        int rc = store_balanced(ctx, b, nbal, units, count);
        if (rc) return rc;
        ctx->bal_sym_variant = ctx->sym_variant;
    }
    *out = &b;
    return KLSP_OK;
}

// The balanced table of the real-basis FP32 contraction.
static int get_balanced_sym32(klsp_ctx* ctx, const klsp_ctx::Balanced** out)
{
    klsp_ctx::Balanced& b = ctx->bal_sym32;
    if (!b.built) {
        std::vector<WorkUnit> units;
        std::vector<int> count;
        klsp_ctx::Balanced nbal;
        if (!plan_balanced_sym32(ctx, nbal, units, count))
            return fail(ctx, KLSP_ERR_INVALID_ARG, "contraction: no work plan fits the device");
        int rc = store_balanced(ctx, b, nbal, units, count);
        if (rc) return rc;
    }
    *out = &b;
    return KLSP_OK;
}

// Gang geometry of the BF16 contraction: workgroup tiles of 320 beta rows x
// 16 alpha columns; gangs per XCD chosen so that the gang workgroups fill the
// XCD's resident workgroup slots in whole rounds where possible.
static GangPlan plan_gangs_bf16(const klsp_ctx* ctx, int nw, int nf)
{
    const Bf16Geometry geo = contract_bf16_geometry(ctx->Np, nw, nf);
    GangPlan gp;
    gp.wrows = geo.wrows;
    gp.wcols = geo.wcols;
    const int gw = gp.wrows * gp.wcols;
    const int slots = std::max(1, ctx->n_cu / 8) * std::max(1, geo.wg_per_cu);
    // gangs per XCD: the fewest from 2 rounds of the XCD's slots on whose
    // last round leaves at most 2% of the slots idle, else the least idle
    // within 12 rounds
    int gpx = 0, xmin = 0;
    double best = 2.0;
    for (int x = std::max(1, (2 * slots) / gw); x <= std::max(1, (12 * slots) / gw); ++x) {
        const double r = (double)x * gw / slots, rounds = std::ceil(r);
        const double tail = (rounds - r) / rounds;
        if (tail <= 0.02) { gpx = x; break; }
        if (tail < best) { best = tail; xmin = x; }
    }
    if (gpx == 0) gpx = std::max(1, xmin);
    const int64_t Kp16 = round_up(ctx->Kp, 16);
    while (gpx > 1 && (int64_t)gpx * 8 * 256 > Kp16) --gpx;
    gp.gangs_per_xcd = gpx;
    gp.n_gangs = 8 * gpx;
    gp.Kg = round_up((Kp16 + gp.n_gangs - 1) / gp.n_gangs, 16);
    return gp;
}

// Work decomposition of the real-basis BF16 contraction (klsp_set_bf16_form):
// the strip form where its block slots fill the triangle nearly as well as
// the fold form's (it reads half the A operands), else the fold form.
constexpr double kBf16StripFill = 0.8;
static int bf16_form(const klsp_ctx* ctx)
{
    if (ctx->bf16_form) return ctx->bf16_form;
    const int Np = (int)ctx->Np;
    return contract_symbf16_slot_use(kBf16Strip, Np, 1) >=
                   kBf16StripFill * contract_symbf16_slot_use(kBf16Fold, Np, 1)
               ? kBf16Strip
               : kBf16Fold;
}

// Whether fold-form passes of three or four weight sets are better run as
// passes of two: their block slots fill the triangle clearly better.
static bool bf16_split(const klsp_ctx* ctx)
{
    const int Np = (int)ctx->Np;
    return contract_symbf16_slot_use(kBf16Fold, Np, 2) >=
           std::max(contract_symbf16_slot_use(kBf16Fold, Np, 3),
                    contract_symbf16_slot_use(kBf16Fold, Np, 4)) + 0.05;
}

// Gang geometry of the real-basis BF16 contraction: column groups (wcols) x
// workgroups per group (wrows) of the given form; gangs per XCD as for the
// complex kernel; K ranges per gang in multiples of 64 rows.
static GangPlan plan_gangs_symbf16(const klsp_ctx* ctx, int form, int nw, int nf, int* mb)
{
    const Bf16Geometry geo = contract_symbf16_geometry(form, ctx->Np, nw, nf);
    GangPlan gp;
    gp.wrows = geo.wrows;
    gp.wcols = geo.wcols;
    *mb = geo.mb;
    const int gw = gp.wrows * gp.wcols;
    const int slots = std::max(1, ctx->n_cu / 8) * std::max(1, geo.wg_per_cu);
    int gpx = 0, xmin = 0;
    double best = 2.0;
    for (int x = std::max(1, (2 * slots) / gw); x <= std::max(1, (12 * slots) / gw); ++x) {
        const double r = (double)x * gw / slots, rounds = std::ceil(r);
        const double tail = (rounds - r) / rounds;
        if (tail <= 0.02) { gpx = x; break; }
        if (tail < best) { best = tail; xmin = x; }
    }
    if (gpx == 0) gpx = std::max(1, xmin);
    // K ranges of whole stage pairs (64 rows: the kernel's loop has no exit
    // between the two stages of a trip)
    const int64_t Kq = round_up(ctx->Kp, 64);
    while (gpx > 1 && (int64_t)gpx * 8 * 256 > Kq) --gpx;
    gp.gangs_per_xcd = gpx;
    gp.n_gangs = 8 * gpx;
    gp.Kg = round_up((Kq + gp.n_gangs - 1) / gp.n_gangs, 64);
    return gp;
}

static ContractShape shape_for(int precision, int nw)
{
    return precision == KLSP_PREC_FP64 ? contract_f64_shape(nw) : contract_f32_shape(nw);
}

// Largest partial-slice count of any launch shape (sizes the partials
// buffer); builds the balanced tables.
static int max_gangs(klsp_ctx* ctx, int* rc)
{
    int g = 8;
    *rc = KLSP_OK;
    for (int prec : {KLSP_PREC_FP64, KLSP_PREC_FP32})
        for (int nw = 1; nw <= kMaxBatch; ++nw) {
            const klsp_ctx::Balanced* b = nullptr;
            if ((*rc = get_balanced(ctx, prec, nw, &b))) return 0;
            g = std::max(g, b->n_slots);
        }
    int mb = 0;
    for (int nw = 1; nw <= kMaxW; ++nw) g = std::max(g, plan_gangs_bf16(ctx, nw, nw).n_gangs);
    for (int form : {kBf16Fold, kBf16Strip})
        for (int nw = 1; nw <= contract_symbf16_max_sets(form); ++nw)
            for (int nf = std::max(0, nw - 1); nf <= nw; ++nf)
                g = std::max(g, plan_gangs_symbf16(ctx, form, nw, nf, &mb).n_gangs);
    g = std::max(g, plan_gangs_bf16(ctx, kMaxW, kMaxW - 1).n_gangs);
    const klsp_ctx::Balanced* b32 = nullptr;
    if ((*rc = get_balanced_sym32(ctx, &b32))) return 0;
    g = std::max(g, b32->n_slots);
    return g;
}

static void free_balanced(klsp_ctx* ctx)
{
    for (auto& row : ctx->bal)
        for (auto& b : row) {
            hipFree(b.d_units);
            hipFree(b.d_count);
            b = klsp_ctx::Balanced();
        }
    hipFree(ctx->bal_sym.d_units);
    hipFree(ctx->bal_sym.d_count);
    ctx->bal_sym = klsp_ctx::Balanced();
    hipFree(ctx->bal_sym32.d_units);
    hipFree(ctx->bal_sym32.d_count);
    ctx->bal_sym32 = klsp_ctx::Balanced();
}

static int free_problem(klsp_ctx* ctx)
{
    free_balanced(ctx);
    hipFree(ctx->d_DJ); hipFree(ctx->d_F0l); hipFree(ctx->d_Ores); hipFree(ctx->d_Nk);
    ctx->d_DJ = ctx->d_F0l = ctx->d_Ores = ctx->d_Nk = nullptr;
    ctx->have_problem = ctx->have_W = ctx->have_arrays = false;
    ctx->have_W32 = ctx->have_Wb = ctx->have_R = ctx->have_Rb = false;
    ctx->sym = -1;
    return KLSP_OK;
}

// The W buffers themselves (kept across problems by free_problem).
static void release_W(klsp_ctx* ctx)
{
    hipFree(ctx->d_W64); hipFree(ctx->d_W32); hipFree(ctx->d_Wb); hipFree(ctx->d_R);
    hipFree(ctx->d_Rb); hipFree(ctx->d_R32);
    ctx->d_R32 = nullptr;
    ctx->R32_cap = 0;
    ctx->have_R32 = false;
    ctx->d_W64 = nullptr;
    ctx->d_W32 = nullptr;
    ctx->d_Wb = nullptr;
    ctx->d_R = nullptr;
    ctx->d_Rb = nullptr;
    ctx->W64_cap = ctx->W32_cap = ctx->Wb_cap = ctx->R_cap = ctx->Rb_cap = 0;
    ctx->have_W = ctx->have_W32 = ctx->have_Wb = ctx->have_R = ctx->have_Rb = ctx->have_R32 = false;
    ctx->sym = -1;
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

int klsp_device_info(int device, char* name_out, int name_cap, int64_t* total_mem_out,
                     int* cu_count_out)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (device < 0 || device >= n) return KLSP_ERR_INVALID_ARG;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return KLSP_ERR_HIP;
    if (name_out && name_cap > 0) {
        // marketing name, then the ISA name (gfx950:sramecc+:xnack-); some
        // drivers report no marketing name
        if (prop.name[0])
            std::snprintf(name_out, (size_t)name_cap, "%s (%s)", prop.name, prop.gcnArchName);
        else
            std::snprintf(name_out, (size_t)name_cap, "%s", prop.gcnArchName);
    }
    if (total_mem_out) *total_mem_out = (int64_t)prop.totalGlobalMem;
    if (cu_count_out) *cu_count_out = prop.multiProcessorCount;
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
    klsp_ctx* ctx = new (std::nothrow) klsp_ctx();
    if (!ctx) return KLSP_ERR_OOM;
    ctx->device = device;
    ctx->n_cu = prop.multiProcessorCount;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return KLSP_ERR_HIP;
    }
    for (auto& e : ctx->ev) hipEventCreate(&e);
    if (hipMalloc((void**)&ctx->d_om, sizeof(double) * 2 * kMaxBatch) != hipSuccess ||
        hipMalloc((void**)&ctx->d_coef, sizeof(double2) * kModelPts * kMaxW) != hipSuccess ||
        hipMalloc((void**)&ctx->d_norm, sizeof(double) * kNorms) != hipSuccess ||
        hipMalloc((void**)&ctx->d_det, sizeof(double2) * kMaxBatch) != hipSuccess ||
        hipMalloc((void**)&ctx->d_flag, sizeof(int)) != hipSuccess) {
        ctx->det_cap = 0;
        klsp_destroy(ctx);
        return KLSP_ERR_OOM;
    }
    ctx->det_cap = kMaxBatch;
    {
        // this device's copy of `scratch`, looked up once per device
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        const auto it = g_scratch_ptr.find(device);
        if (it != g_scratch_ptr.end()) {
            ctx->d_scratch = it->second;
        } else {
            void* ptr = nullptr;
            if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
                klsp_destroy(ctx);
                return KLSP_ERR_HIP;
            }
            g_scratch_ptr[device] = ptr;
            ctx->d_scratch = ptr;
        }
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
    release_W(ctx);
    hipFree(ctx->d_om); hipFree(ctx->d_c); hipFree(ctx->d_P); hipFree(ctx->d_M);
    hipFree(ctx->d_A); hipFree(ctx->d_det); hipFree(ctx->d_coef); hipFree(ctx->d_norm); hipFree(ctx->d_cf);
    hipFree(ctx->d_luaux);
    hipFree(ctx->d_S); hipFree(ctx->d_flag); hipFree(ctx->d_ones);
    if (ctx->scan_exec) hipGraphExecDestroy(ctx->scan_exec);
    hipFree(ctx->d_scan_om); hipFree(ctx->d_scan_det);
    for (auto& e : ctx->ev) hipEventDestroy(e);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return KLSP_OK;
}

const char* klsp_last_error(const klsp_ctx* ctx) { return ctx ? ctx->err.c_str() : g_noctx_err; }

int klsp_upload_problem(klsp_ctx* ctx, const klsp_problem* pr)
try {
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    const bool arrays = pr && pr->DJ && pr->F0l && pr->Omega_res;
    const bool none = pr && !pr->DJ && !pr->F0l && !pr->Omega_res;
    if (!pr || pr->NPh < 1 || pr->N_alpha < 1 || pr->n_l < 1 || !(arrays || none) ||
        !pr->N_kernel)
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
    {
        const auto t = ctx->sym_tuned.find(ctx->Np);
        ctx->sym_variant = t != ctx->sym_tuned.end() ? t->second : kSymDefaultVariant;
    }
    const size_t nK = (size_t)ctx->K;
    KLSP_HIP(ctx, hipMalloc((void**)&ctx->d_DJ, sizeof(double) * ctx->NPh));
    KLSP_HIP(ctx, hipMalloc((void**)&ctx->d_F0l, sizeof(double) * nK));
    KLSP_HIP(ctx, hipMalloc((void**)&ctx->d_Ores, sizeof(double) * nK));
    KLSP_HIP(ctx, hipMalloc((void**)&ctx->d_Nk, sizeof(double) * ctx->N));
    KLSP_HIP(ctx, hipMemcpy(ctx->d_Nk, pr->N_kernel, sizeof(double) * ctx->N, hipMemcpyHostToDevice));
    ctx->have_problem = true;
    if (arrays) return klsp_set_problem_arrays(ctx, pr->DJ, pr->F0l, pr->Omega_res);
    ctx->have_arrays = false;
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

int klsp_set_problem_arrays(klsp_ctx* ctx, const double* DJ, const double* F0l,
                            const double* Omega_res)
try {
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (!ctx->have_problem)
        return fail(ctx, KLSP_ERR_NOT_READY, "klsp_set_problem_arrays: upload the problem first");
    if (!DJ || !F0l || !Omega_res)
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_set_problem_arrays: null array");
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nK = (size_t)ctx->K;
    KLSP_HIP(ctx, hipMemcpy(ctx->d_DJ, DJ, sizeof(double) * ctx->NPh, hipMemcpyHostToDevice));
    KLSP_HIP(ctx, hipMemcpy(ctx->d_F0l, F0l, sizeof(double) * nK, hipMemcpyHostToDevice));
    KLSP_HIP(ctx, hipMemcpy(ctx->d_Ores, Omega_res, sizeof(double) * nK, hipMemcpyHostToDevice));
    ctx->have_arrays = true;
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

// Device W of the current problem, zeroed (the padding rows and columns
// stay zero).  A W that does not fit the buffer releases the low-precision
// copies too before the larger one is allocated.
static int alloc_W64(klsp_ctx* ctx)
{
    const size_t n = (size_t)(ctx->Kp + kRowPad) * ctx->Np;
    ctx->have_W = ctx->have_W32 = ctx->have_Wb = ctx->have_R = ctx->have_Rb = ctx->have_R32 = false;
    ctx->sym = -1;
    if (n > ctx->W64_cap) release_W(ctx);
    int rc = ensure(ctx, ctx->d_W64, ctx->W64_cap, n);
    if (rc) return rc;
    KLSP_HIP(ctx, hipMemsetAsync(ctx->d_W64, 0, sizeof(double2) * n, ctx->stream));
    return KLSP_OK;
}

}  // extern "C"

// Device orbit buffers of one W quadrature ([NPh][nwa] samples, [NPh] frequencies).
struct OrbitBufs {
    double *w1 = nullptr, *ra = nullptr, *pha = nullptr, *o1 = nullptr, *o2 = nullptr;
    ~OrbitBufs() { hipFree(w1); hipFree(ra); hipFree(pha); hipFree(o1); hipFree(o2); }
};

static hipError_t alloc_orbits(OrbitBufs& ob, int64_t NPh, int64_t nwa)
{
    const size_t no = (size_t)NPh * nwa;
    hipError_t e = hipMalloc((void**)&ob.w1, no * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&ob.ra, no * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&ob.pha, no * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&ob.o1, NPh * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&ob.o2, NPh * 8);
    return e;
}

// compute_W_l from orbits already on the device (k_orbit_prep + k_W_f64).
static int W_from_orbits(klsp_ctx* ctx, int64_t nwa, const OrbitBufs& ob, const double* alpha,
                         const int32_t* l_arr, const char* who)
{
    const int64_t NPh = ctx->NPh, nwa_p = round_up(nwa, 4);
    const size_t np = (size_t)NPh * nwa_p;
    double *d_u = nullptr, *d_wt = nullptr, *d_x1 = nullptr, *d_mp = nullptr, *d_al = nullptr;
    int *d_valid = nullptr, *d_l = nullptr;
    auto cleanup = [&]() {
        hipFree(d_u); hipFree(d_wt); hipFree(d_x1); hipFree(d_mp); hipFree(d_al);
        hipFree(d_valid); hipFree(d_l);
    };
    hipError_t e = hipSuccess;
#define KLSP_TRY(call) if (e == hipSuccess) e = (call)
    KLSP_TRY(hipMalloc((void**)&d_u, np * 8));
    KLSP_TRY(hipMalloc((void**)&d_wt, np * 8));
    KLSP_TRY(hipMalloc((void**)&d_x1, np * 8));
    KLSP_TRY(hipMalloc((void**)&d_mp, np * 8));
    KLSP_TRY(hipMalloc((void**)&d_al, ctx->N * 8));
    KLSP_TRY(hipMalloc((void**)&d_valid, NPh * 4));
    KLSP_TRY(hipMalloc((void**)&d_l, ctx->n_l * 4));
    KLSP_TRY(hipMemcpy(d_al, alpha, ctx->N * 8, hipMemcpyHostToDevice));
    KLSP_TRY(hipMemcpy(d_l, l_arr, ctx->n_l * 4, hipMemcpyHostToDevice));
#undef KLSP_TRY
    if (e != hipSuccess) {
        cleanup();
        return fail(ctx, e == hipErrorOutOfMemory ? KLSP_ERR_OOM : KLSP_ERR_HIP,
                    std::string(who) + ": " + hipGetErrorString(e));
    }
    int rc = alloc_W64(ctx);
    if (rc != KLSP_OK) { cleanup(); return rc; }
    hipEventRecord(ctx->ev[0], ctx->stream);
    double lmax = 0.0, amax = 0.0;
    for (int64_t j = 0; j < ctx->n_l; ++j) lmax = std::max(lmax, std::fabs((double)l_arr[j]));
    for (int64_t j = 0; j < ctx->N; ++j) amax = std::max(amax, std::fabs(alpha[j]));
    // alphas evenly spaced to within 8 ulps of the largest (the reference's
    // range(-alpha_max, alpha_max, N)) are generated by rotation in the kernel
    double dalpha = 0.0;
    if (ctx->N >= 2) {
        dalpha = (alpha[ctx->N - 1] - alpha[0]) / (double)(ctx->N - 1);
        const double tol = 8.0 * 2.220446049250313e-16 * amax;
        for (int64_t j = 0; j < ctx->N && dalpha != 0.0; ++j)
            if (!(std::fabs(alpha[j] - (alpha[0] + (double)j * dalpha)) <= tol)) dalpha = 0.0;
    }
    amax = std::max(amax, 16.0 * std::fabs(dalpha));
    hipLaunchKernelGGL(k_orbit_prep, dim3((unsigned)NPh), dim3(256), 0, ctx->stream, ob.w1, ob.ra,
                       ob.pha, ob.o1, ob.o2, (int)nwa, (int)nwa_p, ctx->m, lmax, amax, d_u, d_wt,
                       d_x1, d_mp, d_valid);
    int l_consec = 1;
    for (int64_t j = 1; j < ctx->n_l; ++j)
        if (l_arr[j] != l_arr[0] + j) l_consec = 0;
    // An exactly symmetric alpha grid (the reference's range(-alpha_max,
    // alpha_max, N) is) needs only its upper half: W_l(-alpha) = conj W_l(alpha),
    // since the l-side factor wt cos(l w1 + m phi) / sqrt(r) is real and the
    // non-finite zeroing is the same for both signs.  The kernel computes
    // columns a_first.. and writes each one's conjugate at N - 1 - a too.
    int mirror = ctx->N >= 2;
    for (int64_t j = 0; j < ctx->N && mirror; ++j)
        if (!(alpha[j] == -alpha[ctx->N - 1 - j])) mirror = 0;
    const int a_first = mirror ? (int)(ctx->N / 2) : 0;
    hipLaunchKernelGGL(k_W_f64,
                       dim3((unsigned)NPh, (unsigned)((ctx->N - a_first + kWAc - 1) / kWAc),
                            (unsigned)((ctx->n_l + kWLc - 1) / kWLc)),
                       dim3(64 * kWMmaWaves + 256), 0, ctx->stream, d_u, d_wt, d_x1, d_mp, d_valid, d_al, d_l,
                       l_consec, dalpha, NPh, (int)ctx->N, ctx->Np, (int)ctx->n_l, (int)nwa_p,
                       a_first, mirror, ctx->d_W64);
    hipEventRecord(ctx->ev[1], ctx->stream);
    e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipGetLastError();
    float ms = 0.f;
    hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    ctx->last_W_ms = ms;
    cleanup();
    if (e != hipSuccess) return fail(ctx, KLSP_ERR_HIP, std::string(who) + ": " + hipGetErrorString(e));
    ctx->have_W = true;
    // the mirrored quadrature writes W(-alpha) = conj W(alpha) and a real
    // centre column exactly; other grids are checked when first needed
    if (mirror) ctx->sym = 1;
    return KLSP_OK;
}

// Mirror symmetry of the current W, checked once value by value (W from the
// mirrored quadrature is symmetric by construction and is not checked).
static int check_sym(klsp_ctx* ctx)
{
    if (ctx->sym >= 0 || !ctx->have_W) return KLSP_OK;
    const int64_t rows = ctx->Kp + kRowPad;
    {
        int flag = 1;
        KLSP_HIP(ctx, hipMemsetAsync(ctx->d_flag, 0, sizeof(int), ctx->stream));
        hipLaunchKernelGGL(k_W_mirror_check, dim3(grid_for(rows * (ctx->N - ctx->N / 2))),
                           dim3(256), 0, ctx->stream, ctx->d_W64, rows, (int)ctx->N, ctx->Np,
                           ctx->d_flag);
        KLSP_HIP(ctx, hipGetLastError());
        KLSP_HIP(ctx, hipMemcpyAsync(&flag, ctx->d_flag, sizeof(int), hipMemcpyDeviceToHost,
                                     ctx->stream));
        KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->sym = flag ? 0 : 1;
    }
    return KLSP_OK;
}

// The real basis of the current W for FP64 one-omega builds, made on first
// use where W is symmetric (12.5 GB written, 25 GB read at the default grid:
// a few ms, once per W).
static int prepare_sym(klsp_ctx* ctx)
{
    if (!ctx->use_sym || ctx->sym == 0 || ctx->have_R || !ctx->have_W) return KLSP_OK;
    int rc = check_sym(ctx);
    if (rc || ctx->sym != 1) return rc;
    const int64_t rows = ctx->Kp + kRowPad;
    rc = ensure(ctx, ctx->d_R, ctx->R_cap, (size_t)rows * ctx->Np);
    if (rc) return rc;
    hipLaunchKernelGGL(k_W64_to_R, dim3(grid_for(rows * ctx->Np)), dim3(256), 0, ctx->stream,
                       ctx->d_W64, rows, (int)ctx->N, ctx->Np, ctx->d_R);
    KLSP_HIP(ctx, hipGetLastError());
    KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->have_R = true;
    return KLSP_OK;
}

extern "C" {

int klsp_compute_W(klsp_ctx* ctx, int64_t nwa, const double* w1, const double* ra,
                   const double* pha, const double* Om1, const double* Om2,
                   const double* alpha, const int32_t* l_arr, double* W_out)
try {
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (!ctx->have_problem)
        return fail(ctx, KLSP_ERR_NOT_READY, "klsp_compute_W: upload the problem first");
    if (nwa < 2 || !w1 || !ra || !pha || !Om1 || !Om2 || !alpha || !l_arr)
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_compute_W: bad arguments");
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t NPh = ctx->NPh;
    const size_t no = (size_t)NPh * nwa;
    OrbitBufs ob;
    hipError_t e = alloc_orbits(ob, NPh, nwa);
    if (e == hipSuccess) e = hipMemcpy(ob.w1, w1, no * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ob.ra, ra, no * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ob.pha, pha, no * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ob.o1, Om1, NPh * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ob.o2, Om2, NPh * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess)
        return fail(ctx, e == hipErrorOutOfMemory ? KLSP_ERR_OOM : KLSP_ERR_HIP,
                    std::string("klsp_compute_W: ") + hipGetErrorString(e));
    int rc = W_from_orbits(ctx, nwa, ob, alpha, l_arr, "klsp_compute_W");
    if (rc) return rc;
    if (W_out) return klsp_download_W(ctx, W_out);
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

int klsp_compute_orbits_W(klsp_ctx* ctx, const klsp_params* p, const double* Rc,
                          const double* e, const double* R1, const double* R2, const double* E,
                          const double* L, const double* L2, const double* SGNL,
                          const double* alpha, const int32_t* l_arr, double* Om1, double* Om2,
                          double* w1, double* ra, double* pha)
try {
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (!ctx->have_problem)
        return fail(ctx, KLSP_ERR_NOT_READY, "klsp_compute_orbits_W: upload the problem first");
    if (!p || !Rc || !e || !R1 || !R2 || !E || !L || !L2 || !SGNL || !alpha || !l_arr || !Om1 ||
        !Om2 || p->nw < 3 || p->nwa < 2 || (p->nw - 1) % (p->nwa - 1) != 0 ||
        (int64_t)p->NR * p->Ne != ctx->NPh || p->N_alpha != ctx->N ||
        p->l_max - p->l_min + 1 != ctx->n_l)
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_compute_orbits_W: bad arguments");
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t NPh = ctx->NPh, nwa = p->nwa;
    const int nw = p->nw;
    std::vector<double> wg(nw), cw(nw), sw(nw);
    klsp_orbit_phase_grid(p, wg.data(), cw.data(), sw.data());
    OrbitBufs ob;
    double* d_in = nullptr;     // Rc, e, R1, R2, E, L, L2, SGNL, then wg, cw, sw
    hipError_t st = alloc_orbits(ob, NPh, nwa);
    if (st == hipSuccess) st = hipMalloc((void**)&d_in, sizeof(double) * (8 * NPh + 3 * nw));
    const double* ins[8] = {Rc, e, R1, R2, E, L, L2, SGNL};
    for (int i = 0; i < 8 && st == hipSuccess; ++i)
        st = hipMemcpy(d_in + i * NPh, ins[i], sizeof(double) * NPh, hipMemcpyHostToDevice);
    if (st == hipSuccess) st = hipMemcpy(d_in + 8 * NPh, wg.data(), 8 * nw, hipMemcpyHostToDevice);
    if (st == hipSuccess) st = hipMemcpy(d_in + 8 * NPh + nw, cw.data(), 8 * nw, hipMemcpyHostToDevice);
    if (st == hipSuccess) st = hipMemcpy(d_in + 8 * NPh + 2 * nw, sw.data(), 8 * nw, hipMemcpyHostToDevice);
    if (st == hipSuccess) {
        hipEventRecord(ctx->ev[0], ctx->stream);
        hipLaunchKernelGGL(k_orbits, dim3((unsigned)NPh), dim3(256), 0, ctx->stream, d_in,
                           d_in + NPh, d_in + 2 * NPh, d_in + 3 * NPh, d_in + 4 * NPh,
                           d_in + 5 * NPh, d_in + 6 * NPh, d_in + 7 * NPh, d_in + 8 * NPh,
                           d_in + 8 * NPh + nw, d_in + 8 * NPh + 2 * nw, nw, (int)nwa, ob.w1,
                           ob.ra, ob.pha, ob.o1, ob.o2);
        hipEventRecord(ctx->ev[1], ctx->stream);
        st = hipGetLastError();
    }
    if (st == hipSuccess) st = hipStreamSynchronize(ctx->stream);
    if (st == hipSuccess) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
        ctx->last_orbit_ms = ms;
    }
    hipFree(d_in);
    if (st != hipSuccess)
        return fail(ctx, st == hipErrorOutOfMemory ? KLSP_ERR_OOM : KLSP_ERR_HIP,
                    std::string("klsp_compute_orbits_W: ") + hipGetErrorString(st));
    int rc = W_from_orbits(ctx, nwa, ob, alpha, l_arr, "klsp_compute_orbits_W");
    if (rc) return rc;
    const size_t no = (size_t)NPh * nwa;
    KLSP_HIP(ctx, hipMemcpy(Om1, ob.o1, NPh * 8, hipMemcpyDeviceToHost));
    KLSP_HIP(ctx, hipMemcpy(Om2, ob.o2, NPh * 8, hipMemcpyDeviceToHost));
    if (w1) KLSP_HIP(ctx, hipMemcpy(w1, ob.w1, no * 8, hipMemcpyDeviceToHost));
    if (ra) KLSP_HIP(ctx, hipMemcpy(ra, ob.ra, no * 8, hipMemcpyDeviceToHost));
    if (pha) KLSP_HIP(ctx, hipMemcpy(pha, ob.pha, no * 8, hipMemcpyDeviceToHost));
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

int klsp_upload_W(klsp_ctx* ctx, const double* W)
try {
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
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

int klsp_download_W(klsp_ctx* ctx, double* W_out)
try {
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
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

int klsp_prepare(klsp_ctx* ctx, int precision)
try {
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (precision != KLSP_PREC_FP64 && precision != KLSP_PREC_FP32 &&
        precision != KLSP_PREC_BF16X3)
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_prepare: unknown precision");
    if (!ctx->have_W) return fail(ctx, KLSP_ERR_NOT_READY, "klsp_prepare: no W on device");
    int rc;
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    if (precision == KLSP_PREC_FP64) {
        // the real basis where W is symmetric, and the work table of one-omega
        // builds (klsp_det_noalloc needs both made here)
        if ((rc = prepare_sym(ctx)) || (rc = check_sym(ctx))) return rc;
        const klsp_ctx::Balanced* b = nullptr;
        if (ctx->use_sym && ctx->sym == 1) rc = get_balanced_sym(ctx, &b);
        else rc = get_balanced(ctx, KLSP_PREC_FP64, 1, &b);
        if (rc) return rc;
    }
    if (precision == KLSP_PREC_FP32 && (rc = check_sym(ctx))) return rc;
    if (precision == KLSP_PREC_FP32 && ctx->use_sym && ctx->sym == 1 && !ctx->have_R32) {
        // the real basis in FP32
        const int64_t rows = ctx->Kp + kRowPad;
        if ((rc = ensure(ctx, ctx->d_R32, ctx->R32_cap, (size_t)(rows * ctx->Np)))) return rc;
        KLSP_HIP(ctx, split_R_f32(ctx->d_W64, rows, (int)ctx->N, ctx->Np, ctx->d_R32,
                                  ctx->stream));
        KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->have_R32 = true;
    }
    if (precision == KLSP_PREC_FP32 && !(ctx->use_sym && ctx->sym == 1) && !ctx->have_W32) {
        KLSP_HIP(ctx, hipSetDevice(ctx->device));
        const int64_t n = (ctx->Kp + kRowPad) * ctx->Np;
        if ((rc = ensure(ctx, ctx->d_W32, ctx->W32_cap, (size_t)n))) return rc;
        hipLaunchKernelGGL(k_W64_to_32, dim3(grid_for(n)), dim3(256), 0, ctx->stream,
                           ctx->d_W64, n, ctx->d_W32);
        KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->have_W32 = true;
    }
    if (precision == KLSP_PREC_BF16X3 && (rc = check_sym(ctx))) return rc;
    if (precision == KLSP_PREC_BF16X3 && ctx->use_sym && ctx->sym == 1 && !ctx->have_Rb) {
        // bf16 hi and lo planes of the real basis
        const int64_t nq = symbf16_records(ctx->Kp);
        if ((rc = ensure(ctx, ctx->d_Rb, ctx->Rb_cap, (size_t)2 * nq * ctx->Np))) return rc;
        KLSP_HIP(ctx, split_R_bf16(ctx->d_W64, ctx->Kp + kRowPad, (int)ctx->N, ctx->Np, nq,
                                   ctx->d_Rb, ctx->stream));
        KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->have_Rb = true;
    }
    if (precision == KLSP_PREC_BF16X3 && !(ctx->use_sym && ctx->sym == 1) && !ctx->have_Wb) {
        KLSP_HIP(ctx, hipSetDevice(ctx->device));
        const int64_t nq = round_up(ctx->Kp, 16) / 4;
        // two planes (hi, lo) of nq x Np 16-byte groups of four K rows
        if ((rc = ensure(ctx, ctx->d_Wb, ctx->Wb_cap, (size_t)2 * nq * ctx->Np))) return rc;
        KLSP_HIP(ctx, split_W_bf16(ctx->d_W64, ctx->Kp + kRowPad, ctx->Np, nq, ctx->d_Wb,
                                   ctx->stream));
        KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->have_Wb = true;
    }
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

}  // extern "C"

// Launches the contraction for nb omegas (weights already in ctx->d_c); in
// BF16x3 the weight sets from nf on take the single hi x hi product.
static int launch_contract(klsp_ctx* ctx, int precision, int nb, int nf)
{
    const bool bf = precision == KLSP_PREC_BF16X3;
    GangPlan gp{};
    const klsp_ctx::Balanced* bal = nullptr;
    int rc;
    if (precision == KLSP_PREC_FP64 && nb == 1 && (rc = prepare_sym(ctx))) return rc;
    const bool symb = precision == KLSP_PREC_FP64 && nb == 1 && ctx->use_sym && ctx->sym == 1 &&
                      ctx->have_R;
    if (symb) {
        if ((rc = get_balanced_sym(ctx, &bal))) return rc;
        ContractArgs a{};
        a.units = bal->d_units;
        a.n_units = bal->n_units;
        a.W = ctx->d_R;
        a.cw = ctx->d_c;
        a.P = ctx->d_P;
        a.Kp = ctx->Kp;
        a.Np = ctx->Np;
        a.n_gangs = bal->n_slots;
        a.wrows = bal->wrows;
        a.wcols = bal->wcols;
        const hipError_t e = contract_sym64_launch(ctx->sym_variant, a, ctx->stream);
        if (e != hipSuccess)
            return fail(ctx, KLSP_ERR_HIP, std::string("contraction launch: ") + hipGetErrorString(e));
        ctx->last_S = bal->n_slots;
        ctx->last_count = bal->d_count;
        ctx->last_sym = true;
        return KLSP_OK;
    }
    if (precision == KLSP_PREC_FP32 && ctx->use_sym && ctx->sym == 1 && ctx->have_R32) {
        // one launch per weight set: set w reads its weights at d_cf + w Kpc
        // and writes its slices at d_P + w nsl Np^2 (the reduce's layout)
        const klsp_ctx::Balanced* b = nullptr;
        if ((rc = get_balanced_sym32(ctx, &b))) return rc;
        const int64_t Kpc = ctx->Kp + kRowPad;
        for (int w = 0; w < nb; ++w) {
            ContractArgs a{};
            a.units = b->d_units;
            a.n_units = b->n_units;
            a.W = ctx->d_R32;
            a.cwf = ctx->d_cf + w * Kpc;
            a.P = ctx->d_P + (int64_t)w * b->n_slots * ctx->Np * ctx->Np;
            a.Kp = ctx->Kp;
            a.Np = ctx->Np;
            a.n_gangs = b->n_slots;
            a.wrows = b->wrows;
            a.wcols = b->wcols;
            const hipError_t e = contract_sym32_launch(a, ctx->stream);
            if (e != hipSuccess)
                return fail(ctx, KLSP_ERR_HIP,
                            std::string("contraction launch: ") + hipGetErrorString(e));
        }
        ctx->last_S = b->n_slots;
        ctx->last_count = b->d_count;
        ctx->last_sym = true;
        return KLSP_OK;
    }
    if (bf && ctx->use_sym && ctx->sym == 1 && ctx->have_Rb) {
        // weight sets in launches of at most `per`: two in the strip form
        // (its register budget); in the fold form all of them, or four as
        // two of two where that fills the block slots better (N = 301: 95%
        // against 79%).  Launch h takes sets from per * h on, with their
        // weights and partial slices, so the reduce sees one pass of nb
        // sets; all launches use the first one's gang plan (the partial
        // slices of a pass share one gang count)
        const int form = bf16_form(ctx);
        const int per = form == kBf16Strip ? contract_symbf16_max_sets(form)
                        : nb == 4 && bf16_split(ctx) ? 2
                                                     : nb;
        const int nl = (nb + per - 1) / per;
        int mb = 0;
        gp = plan_gangs_symbf16(ctx, form, std::min(nb, per), nl > 1 ? std::min(nf, per) : nf, &mb);
        for (int h = 0; h < nl; ++h) {
            const int nw = std::min(per, nb - h * per);
            ContractArgs a{};
            a.W = ctx->d_Rb;
            a.cw = ctx->d_c;
            a.cwf = ctx->d_cf + (size_t)h * per * (ctx->Kp + kRowPad);
            a.P = ctx->d_P + (size_t)h * per * gp.n_gangs * ctx->Np * ctx->Np;
            a.Kp = ctx->Kp;
            a.Np = ctx->Np;
            a.Kg = gp.Kg;
            a.n_gangs = gp.n_gangs;
            a.wrows = gp.wrows;
            a.wcols = gp.wcols;
            const int nfh = std::min(nw, std::max(0, nf - h * per));
            const hipError_t e = contract_symbf16_launch(form, nw, nfh, mb, a, ctx->stream);
            if (e != hipSuccess)
                return fail(ctx, KLSP_ERR_HIP,
                            std::string("contraction launch: ") + hipGetErrorString(e));
        }
        ctx->last_S = gp.n_gangs;
        ctx->last_count = nullptr;
        ctx->last_sym = true;
        return KLSP_OK;
    }
    ctx->last_sym = false;
    if (bf) {
        gp = plan_gangs_bf16(ctx, nb, nf);
    } else {
        if ((rc = get_balanced(ctx, precision, nb, &bal))) return rc;
        gp.wrows = bal->wrows;
        gp.wcols = bal->wcols;
        gp.n_gangs = bal->n_slots;
        gp.Kg = 0;
    }
    ContractArgs a;
    a.units = bal ? bal->d_units : nullptr;
    a.n_units = bal ? bal->n_units : 0;
    a.W = precision == KLSP_PREC_FP64 ? (const void*)ctx->d_W64
          : bf                        ? (const void*)ctx->d_Wb
                                      : (const void*)ctx->d_W32;
    a.cw = ctx->d_c;
    a.cwf = ctx->d_cf;
    a.P = ctx->d_P;
    a.Kp = ctx->Kp;
    a.Np = ctx->Np;
    a.Kg = gp.Kg;
    a.n_gangs = gp.n_gangs;
    a.wrows = gp.wrows;
    a.wcols = gp.wcols;
    const hipError_t e = precision == KLSP_PREC_FP64 ? contract_f64_launch(nb, a, ctx->stream)
                         : bf                        ? contract_bf16_launch(nb, nf, a, ctx->stream)
                                                     : contract_f32_launch(nb, a, ctx->stream);
    if (e != hipSuccess)
        return fail(ctx, KLSP_ERR_HIP, std::string("contraction launch: ") + hipGetErrorString(e));
    ctx->last_S = gp.n_gangs;
    ctx->last_count = bal ? bal->d_count : nullptr;
    return KLSP_OK;
}

// Builds M for nb <= kMaxBatch omegas into ctx->d_M ([nb][N,N]).  With
// taylor set, om is one centre and the nb matrices are its Taylor terms.
// nf < nb (BF16x3 Taylor terms only): terms nf.. take the single product.
static void add_build_timing(klsp_ctx* ctx, int nb);

// ... into matrix slots slot..slot+nb-1 of ctx->d_M.
// With om_dev set, the omegas are read from that device array (no host copy:
// the form a captured graph needs) and om is not used.
static int build_batch(klsp_ctx* ctx, const double* om, int nb, int precision,
                       bool taylor = false, int slot = 0, int nf = 0,
                       const double* om_dev = nullptr)
{
    if (nf <= 0 || nf > nb || precision != KLSP_PREC_BF16X3 || nb != kMaxW) nf = nb;
    const int64_t Kpc = ctx->Kp + kRowPad;
    int rc;
    if ((rc = ensure(ctx, ctx->d_c, ctx->c_cap, (size_t)kMaxW * Kpc))) return rc;
    float2* cf = nullptr;   // the complex64 copy the BF16x3 and FP32 builds read
    if (precision == KLSP_PREC_BF16X3 || precision == KLSP_PREC_FP32) {
        if ((rc = ensure(ctx, ctx->d_cf, ctx->cf_cap, (size_t)kMaxW * Kpc))) return rc;
        cf = ctx->d_cf;
    }
    int mg = max_gangs(ctx, &rc);
    if (rc) return rc;
    if (precision == KLSP_PREC_FP64 && nb == 1) {
        if ((rc = prepare_sym(ctx))) return rc;
        if (ctx->use_sym && ctx->sym == 1) {
            const klsp_ctx::Balanced* bs = nullptr;
            if ((rc = get_balanced_sym(ctx, &bs))) return rc;
            mg = std::max(mg, bs->n_slots);
        }
    }
    if (ctx->use_sym && ctx->sym == 1 &&
        (rc = ensure(ctx, ctx->d_S, ctx->S_cap, (size_t)kMaxW * ctx->N * ctx->N)))
        return rc;
    if ((rc = ensure(ctx, ctx->d_P, ctx->P_cap, (size_t)kMaxW * mg * ctx->Np * ctx->Np)))
        return rc;
    if ((rc = ensure(ctx, ctx->d_M, ctx->M_cap,
                     (size_t)std::max(kMaxW, slot + nb) * ctx->N * ctx->N)))
        return rc;
    if (!taylor && !om_dev)
        KLSP_HIP(ctx, hipMemcpyAsync(ctx->d_om, om, sizeof(double) * 2 * nb,
                                     hipMemcpyHostToDevice, ctx->stream));
    if (ctx->prof) hipEventRecord(ctx->ev[0], ctx->stream);
    if (taylor)
        hipLaunchKernelGGL(k_weights_taylor, dim3(grid_for(Kpc)), dim3(256), 0, ctx->stream,
                           ctx->d_DJ, ctx->d_F0l, ctx->d_Ores, ctx->NPh, ctx->K, Kpc, om[0],
                           om[1], nb, ctx->d_c, cf);
    else
        hipLaunchKernelGGL(k_weights, dim3(grid_for((int64_t)nb * Kpc)), dim3(256), 0,
                           ctx->stream, ctx->d_DJ, ctx->d_F0l, ctx->d_Ores, ctx->NPh, ctx->K,
                           Kpc, om_dev ? om_dev : ctx->d_om, nb, ctx->d_c, cf);
    if (ctx->prof) hipEventRecord(ctx->ev[1], ctx->stream);
    if ((rc = launch_contract(ctx, precision, nb, nf))) return rc;
    if (ctx->prof) hipEventRecord(ctx->ev[2], ctx->stream);
    if (ctx->last_sym) {
        const int64_t NN = ctx->N * ctx->N;
        hipLaunchKernelGGL(k_reduce_sym, dim3(grid_for(nb * NN, kRedElems)),
                           dim3(kRedElems * kRedParts), 0, ctx->stream, ctx->d_P, ctx->last_S,
                           ctx->last_count, ctx->Np, (int)ctx->N, nb, ctx->d_S);
        hipLaunchKernelGGL(k_expand_sym, dim3(grid_for(nb * NN)), dim3(256), 0, ctx->stream,
                           ctx->d_S, (int)ctx->N, nb, ctx->G, ctx->d_Nk,
                           ctx->d_M + (size_t)slot * NN);
    } else {
        hipLaunchKernelGGL(k_reduce, dim3(grid_for((int64_t)nb * ctx->N * ctx->N, kRedElems)),
                           dim3(kRedElems * kRedParts), 0,
                           ctx->stream, ctx->d_P, ctx->last_S, ctx->last_count, ctx->Np,
                           (int)ctx->N, nb, ctx->G, ctx->d_Nk,
                           ctx->d_M + (size_t)slot * ctx->N * ctx->N);
    }
    if (ctx->prof) hipEventRecord(ctx->ev[3], ctx->stream);
    KLSP_HIP(ctx, hipGetLastError());
    add_build_timing(ctx, nb);
    return KLSP_OK;
}

// With profiling on: waits for the build just enqueued and adds its phases.
static void add_build_timing(klsp_ctx* ctx, int nb)
{
    if (!ctx->prof) return;
    hipEventSynchronize(ctx->ev[3]);
    float a = 0, b = 0, c = 0;
    hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]);
    hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]);
    hipEventElapsedTime(&c, ctx->ev[2], ctx->ev[3]);
    ctx->last.weights_ms += a;
    ctx->last.contract_ms += b;
    ctx->last.reduce_ms += c;
    ctx->last.n_omega += nb;
    ctx->last.contract_launches += 1;
}

// Whether an LU launch of nd matrices takes the step-wise LU (klsp_lu.hip).
static bool lu_steps(const klsp_ctx* ctx, int nd)
{
    return ctx->N <= kLuStepMaxN &&
           (ctx->lu_form == 2 || (ctx->lu_form == 0 && nd <= kLuStepMaxMatrices));
}

// The step-wise LU's bookkeeping for nd matrices (grow-only; callers that
// capture a graph make sure of it before the capture starts).
static int ensure_lu_aux(klsp_ctx* ctx, int nd)
{
    unsigned char* aux = static_cast<unsigned char*>(ctx->d_luaux);
    const int rc = ensure(ctx, aux, ctx->luaux_cap, (size_t)nd * lu_step_aux_bytes());
    ctx->d_luaux = aux;
    return rc;
}

// LU launch over matrices [0, nd) of ctx->d_M (coef as k_lu_det), dets to
// ctx->d_det, with its device time added to the profile.
static int launch_lu(klsp_ctx* ctx, int nd, const double2* coef, int nt,
                     double2* det_out = nullptr)
{
    const bool steps = lu_steps(ctx, nd);
    if (steps) {
        int rc;
        if ((rc = ensure_lu_aux(ctx, nd))) return rc;
    }
    if (ctx->prof) hipEventRecord(ctx->ev[4], ctx->stream);
    if (steps)
        KLSP_HIP(ctx, launch_lu_steps(ctx->stream, ctx->d_M, coef, nt, ctx->d_A, (int)ctx->N, nd,
                                      ctx->d_luaux, det_out ? det_out : ctx->d_det));
    else
        hipLaunchKernelGGL(k_lu_det, dim3(nd), dim3(kLuThreads), 0, ctx->stream, ctx->d_M, coef,
                           nt, ctx->d_A, (int)ctx->N, det_out ? det_out : ctx->d_det);
    if (ctx->prof) hipEventRecord(ctx->ev[5], ctx->stream);
    KLSP_HIP(ctx, hipGetLastError());
    return KLSP_OK;
}

static void add_lu_timing(klsp_ctx* ctx)
{
    if (!ctx->prof) return;
    hipEventSynchronize(ctx->ev[5]);
    float d = 0;
    hipEventElapsedTime(&d, ctx->ev[4], ctx->ev[5]);
    ctx->last.lu_ms += d;
}

// Matrices built before one LU launch (the launch runs one workgroup per
// matrix, so up to a few dozen cost the same time as one).
static int det_slots(const klsp_ctx* ctx)
{
    const int64_t fit = ((int64_t)256 << 20) / std::max<int64_t>(1, 16 * ctx->N * ctx->N);
    return (int)std::max<int64_t>(kMaxBatch, std::min<int64_t>(64, fit));
}

// Matrices per LU launch for n omegas built maxb at a time: whole batches
// only (a launch never ends on a narrower contraction pass unless the omegas
// run out), so a scan groups its omegas the same way in every form.
static int batch_slots(const klsp_ctx* ctx, int64_t n, int maxb)
{
    const int fit = std::max(maxb, det_slots(ctx) / maxb * maxb);
    return (int)std::min<int64_t>(fit, round_up(n, maxb));
}

static int check_ready(klsp_ctx* ctx, int precision, const char* who)
{
    if (!ctx->have_problem || !ctx->have_W || !ctx->have_arrays)
        return fail(ctx, KLSP_ERR_NOT_READY,
                    std::string(who) + ": problem arrays and W must be on device");
    if (precision != KLSP_PREC_FP64 && precision != KLSP_PREC_FP32 &&
        precision != KLSP_PREC_BF16X3)
        return fail(ctx, KLSP_ERR_INVALID_ARG, std::string(who) + ": unknown precision");
    if (precision == KLSP_PREC_FP32 && !ctx->have_W32 && !ctx->have_R32)
        return klsp_prepare(ctx, precision);
    if (precision == KLSP_PREC_FP32 && ctx->use_sym && ctx->sym == 1 && !ctx->have_R32)
        return klsp_prepare(ctx, precision);
    if (precision == KLSP_PREC_FP32 && !(ctx->use_sym && ctx->sym == 1) && !ctx->have_W32)
        return klsp_prepare(ctx, precision);
    if (precision == KLSP_PREC_BF16X3 && !ctx->have_Wb && !ctx->have_Rb)
        return klsp_prepare(ctx, precision);
    if (precision == KLSP_PREC_BF16X3 && ctx->use_sym && ctx->sym == 1 && !ctx->have_Rb)
        return klsp_prepare(ctx, precision);
    if (precision == KLSP_PREC_BF16X3 && !(ctx->use_sym && ctx->sym == 1) && !ctx->have_Wb)
        return klsp_prepare(ctx, precision);
    return KLSP_OK;
}

// Omegas per contraction pass: FP64 is MFMA-bound at one omega per pass
// (larger batches only lower the occupancy); the low-precision passes share
// W loads among up to three, or in the real basis in BF16x3 two where one
// launch takes no more (strip form) or passes of two fill the block slots
// better (fold form, bf16_split).
static int omega_batch(const klsp_ctx* ctx, int precision)
{
    if (precision == KLSP_PREC_FP64) return 1;
    if (precision == KLSP_PREC_BF16X3 && ctx->use_sym && ctx->sym == 1 &&
        (bf16_form(ctx) == kBf16Strip || bf16_split(ctx)))
        return 2;
    return kMaxBatch;
}

// det(I - M(omega_i)) for n omegas; the workhorse of klsp_det/newton/scan.
static int det_many(klsp_ctx* ctx, const double* om, int64_t n, int precision, double* out)
{
    int rc = check_ready(ctx, precision, "klsp_det");
    if (rc) return rc;
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    const int maxb = omega_batch(ctx, precision);
    const int slots = batch_slots(ctx, n, maxb);
    const size_t NN = (size_t)ctx->N * ctx->N;
    if ((rc = ensure(ctx, ctx->d_M, ctx->M_cap, (size_t)std::max(slots, kMaxW) * NN))) return rc;
    if ((rc = ensure(ctx, ctx->d_A, ctx->A_cap, (size_t)std::max(slots, kMaxBatch) * NN))) return rc;
    if ((rc = ensure(ctx, ctx->d_det, ctx->det_cap, (size_t)std::max(slots, kMaxBatch)))) return rc;
    for (int64_t i0 = 0; i0 < n; i0 += slots) {
        const int nd = (int)std::min<int64_t>(slots, n - i0);
        for (int j0 = 0; j0 < nd; j0 += maxb) {
            const int nb = std::min(maxb, nd - j0);
            if ((rc = build_batch(ctx, om + 2 * (i0 + j0), nb, precision, false, j0))) return rc;
        }
        if ((rc = launch_lu(ctx, nd, nullptr, 0))) return rc;
        KLSP_HIP(ctx, hipMemcpyAsync(out + 2 * i0, ctx->d_det, sizeof(double2) * nd,
                                     hipMemcpyDeviceToHost, ctx->stream));
        KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        add_lu_timing(ctx);
    }
    return KLSP_OK;
}

// Everything a captured scan recorded that can change between calls: buffer
// addresses, sizes, launch parameters, the contraction path and its variant.
static std::vector<int64_t> scan_graph_key(const klsp_ctx* ctx, int64_t n, int precision,
                                           int slots, int maxb)
{
    int64_t g = 0;
    std::memcpy(&g, &ctx->G, sizeof g);
    auto P = [](const void* q) { return (int64_t)(uintptr_t)q; };
    std::vector<int64_t> k = {
        n, precision, slots, maxb, ctx->N, ctx->Np, ctx->Kp, ctx->K, ctx->NPh, ctx->n_cu, g,
        ctx->use_sym, ctx->sym, ctx->have_R, ctx->have_Rb, ctx->have_Wb, ctx->have_W32,
        ctx->sym_variant, P(ctx->d_W64), P(ctx->d_W32), P(ctx->d_Wb), P(ctx->d_R), P(ctx->d_Rb),
        P(ctx->d_c), P(ctx->d_cf), P(ctx->d_P), P(ctx->d_S), P(ctx->d_M), P(ctx->d_A),
        P(ctx->d_DJ), P(ctx->d_F0l), P(ctx->d_Ores), P(ctx->d_Nk), P(ctx->d_scan_om),
        P(ctx->d_scan_det), P(ctx->bal_sym.d_units), P(ctx->bal_sym.d_count),
        ctx->bal_sym.n_units, ctx->bal_sym.n_slots, P(ctx->d_R32), ctx->have_R32,
        P(ctx->bal_sym32.d_units), P(ctx->bal_sym32.d_count), ctx->bal_sym32.n_units,
        ctx->bal_sym32.n_slots, P(ctx->d_luaux), ctx->lu_form, ctx->bf16_form};
    for (const auto& row : ctx->bal)
        for (const auto& b : row) {
            k.push_back(P(b.d_units));
            k.push_back(P(b.d_count));
            k.push_back(b.n_units);
            k.push_back(b.n_slots);
        }
    return k;
}

// Profiling off for the scope (events cannot be recorded into a captured
// graph), restored on every return.
struct ProfilingOff {
    klsp_ctx* c;
    bool was;
    explicit ProfilingOff(klsp_ctx* ctx) : c(ctx), was(ctx->prof) { c->prof = false; }
    ~ProfilingOff() { c->prof = was; }
};

// det(I - M(omega_i)) for the n omegas of a grid scan as one HIP graph: the
// omegas are uploaded once, each batch's weights read them on the device,
// and the LU launches write their dets into a device array, so the graph
// holds kernel nodes only (no host copies, no memory sets).  A graph with
// the same key as the last one is launched again (its omegas are read from
// the device array at launch).  Batches are built eagerly once before a
// capture, so that every workspace and work table exists (nothing is
// allocated while capturing).
static int scan_by_graph(klsp_ctx* ctx, const double* om, int64_t n, int precision, double* out)
{
    int rc = check_ready(ctx, precision, "klsp_grid_scan");
    if (rc) return rc;
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    const int maxb = omega_batch(ctx, precision);
    const int slots = batch_slots(ctx, n, maxb);
    const size_t NN = (size_t)ctx->N * ctx->N;
    if ((rc = ensure(ctx, ctx->d_M, ctx->M_cap, (size_t)std::max(slots, kMaxW) * NN))) return rc;
    if ((rc = ensure(ctx, ctx->d_A, ctx->A_cap, (size_t)std::max(slots, kMaxBatch) * NN))) return rc;
    if ((rc = ensure(ctx, ctx->d_det, ctx->det_cap, (size_t)std::max(slots, kMaxBatch)))) return rc;
    if ((rc = ensure(ctx, ctx->d_scan_om, ctx->scan_om_cap, (size_t)2 * n))) return rc;
    if ((rc = ensure(ctx, ctx->d_scan_det, ctx->scan_det_cap, (size_t)n))) return rc;
    if ((rc = ensure_lu_aux(ctx, std::max(slots, kMaxBatch)))) return rc;
    KLSP_HIP(ctx, hipMemcpyAsync(ctx->d_scan_om, om, sizeof(double) * 2 * n,
                                 hipMemcpyHostToDevice, ctx->stream));
    const bool prof = ctx->prof;
    ProfilingOff off(ctx);
    const int64_t builds = (n + maxb - 1) / maxb + (n / slots);   // upper bound for the timing
    std::vector<int64_t> key = scan_graph_key(ctx, n, precision, slots, maxb);
    if (!ctx->scan_exec || key != ctx->scan_key) {
        // every batch width the graph uses (full batches and the last one's
        // remainder; slots hold whole batches), built once outside the
        // capture from the first omegas
        const int widths[2] = {(int)std::min<int64_t>(maxb, n), (int)(n % maxb)};
        for (int nb : widths)
            if (nb > 0 &&
                (rc = build_batch(ctx, nullptr, nb, precision, false, 0, 0, ctx->d_scan_om)))
                return rc;
        key = scan_graph_key(ctx, n, precision, slots, maxb);
        if (ctx->scan_exec) {
            hipGraphExecDestroy(ctx->scan_exec);
            ctx->scan_exec = nullptr;
        }
        KLSP_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int crc = KLSP_OK;
        for (int64_t i0 = 0; i0 < n && crc == KLSP_OK; i0 += slots) {
            const int nd = (int)std::min<int64_t>(slots, n - i0);
            for (int j0 = 0; j0 < nd && crc == KLSP_OK; j0 += maxb) {
                const int nb = std::min(maxb, nd - j0);
                crc = build_batch(ctx, nullptr, nb, precision, false, j0, 0,
                                  ctx->d_scan_om + 2 * (i0 + j0));
            }
            if (crc == KLSP_OK) crc = launch_lu(ctx, nd, nullptr, 0, ctx->d_scan_det + i0);
        }
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
        if (crc != KLSP_OK || ec != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            if (crc != KLSP_OK) return crc;
            return fail(ctx, KLSP_ERR_HIP,
                        std::string("klsp_grid_scan: graph capture: ") + hipGetErrorString(ec));
        }
        hipGraphExec_t ex = nullptr;
        const hipError_t ei = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ei != hipSuccess)
            return fail(ctx, KLSP_ERR_HIP,
                        std::string("klsp_grid_scan: graph instantiation: ") + hipGetErrorString(ei));
        ctx->scan_exec = ex;
        ctx->scan_key = key;
        ++ctx->scan_graph_builds;
    }
    if (prof) hipEventRecord(ctx->ev[4], ctx->stream);
    KLSP_HIP(ctx, hipGraphLaunch(ctx->scan_exec, ctx->stream));
    if (prof) hipEventRecord(ctx->ev[5], ctx->stream);
    ++ctx->scan_graph_launches;
    KLSP_HIP(ctx, hipMemcpyAsync(out, ctx->d_scan_det, sizeof(double2) * n,
                                 hipMemcpyDeviceToHost, ctx->stream));
    KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (prof) {
        float d = 0;
        hipEventElapsedTime(&d, ctx->ev[4], ctx->ev[5]);
        ctx->last.contract_ms += d;   // the whole graph (no per-phase split)
    }
    ctx->last.n_omega += n;
    ctx->last.contract_launches += builds;
    return KLSP_OK;
}

static void reset_timing(klsp_ctx* ctx) { ctx->last = klsp_timing{}; }

// Timings of the inner calls of a driver, summed.
static void acc_timing(klsp_timing& acc, klsp_ctx* ctx)
{
    acc.weights_ms += ctx->last.weights_ms;
    acc.contract_ms += ctx->last.contract_ms;
    acc.reduce_ms += ctx->last.reduce_ms;
    acc.lu_ms += ctx->last.lu_ms;
    acc.n_omega += ctx->last.n_omega;
    acc.contract_launches += ctx->last.contract_launches;
    ctx->last = klsp_timing{};
}

// dets of I - sum_j cf[i*nt + j] T_j for nd coefficient sets over the nt
// matrices T_j in ctx->d_M (nd <= 3, nt <= kMaxW).
static int model_dets(klsp_ctx* ctx, int nt, const std::complex<double>* cf, int nd,
                      std::complex<double>* dets)
{
    int rc;
    if (nd > kModelPts || nt > kMaxW) return fail(ctx, KLSP_ERR_INVALID_ARG, "model_dets: too many points");
    if ((rc = ensure(ctx, ctx->d_A, ctx->A_cap,
                     (size_t)std::max(kMaxBatch, nd) * ctx->N * ctx->N)))
        return rc;
    if ((rc = ensure(ctx, ctx->d_det, ctx->det_cap, (size_t)std::max(kMaxBatch, nd)))) return rc;
    double2 coef[kModelPts * kMaxW];
    for (int i = 0; i < nd * nt; ++i) coef[i] = make_double2(cf[i].real(), cf[i].imag());
    KLSP_HIP(ctx, hipMemcpyAsync(ctx->d_coef, coef, sizeof(double2) * nd * nt,
                                 hipMemcpyHostToDevice, ctx->stream));
    if ((rc = launch_lu(ctx, nd, ctx->d_coef, nt))) return rc;
    KLSP_HIP(ctx, hipMemcpyAsync(dets, ctx->d_det, sizeof(double2) * nd, hipMemcpyDeviceToHost,
                                 ctx->stream));
    KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    add_lu_timing(ctx);
    return KLSP_OK;
}

// Zero of a model det by one LU launch: its values at the kCircle points
// s0 + r e^(2 pi i k / kCircle) and at s0 give, by a discrete Fourier
// transform, its Taylor coefficients around s0,
//   a_n = (1 / kCircle) sum_k f_k e^(-2 pi i n k / kCircle) / r^n,
// exact up to the aliased terms ~(r / R)^kCircle (R >= Im w, the distance to
// the det's nearest singularity); Newton on that polynomial from -a_0 / a_1
// gives the zero t nearest s0 (t relative to s0).  coef_at(s, cf) fills the
// nt model coefficients of the point s.  ok is false where the polynomial is
// not to be trusted: a_0 differs from the centre value by more than 1e-8 of
// the circle's values (aliasing), or no zero within the circle was found.
template <typename CoefAt>
static int circle_zero(klsp_ctx* ctx, int nt, CoefAt coef_at, std::complex<double> s0, double r,
                       std::complex<double>* t_out, std::complex<double>* f0_out, bool* ok)
{
    typedef std::complex<double> cd;
    const int P = kCircle;
    cd cf[kModelPts * kMaxW], f[kModelPts];
    cd e[kCircle];
    const double two_pi = 6.283185307179586;
    for (int k = 0; k < P; ++k) e[k] = std::polar(1.0, two_pi * k / P);
    for (int k = 0; k < P; ++k) coef_at(s0 + r * e[k], cf + k * nt);
    coef_at(s0, cf + P * nt);
    int rc = model_dets(ctx, nt, cf, P + 1, f);
    if (rc) return rc;
    *f0_out = f[P];
    *ok = false;
    double fmax = 0.0;
    for (int k = 0; k < P; ++k) fmax = std::max(fmax, std::abs(f[k]));
    cd a[kCircle];
    for (int n = 0; n < P; ++n) {
        cd sum = 0.0;
        for (int k = 0; k < P; ++k) sum += f[k] * std::conj(e[(n * k) % P]);
        a[n] = sum / (double)P / std::pow(r, n);
    }
    if (!(fmax > 0.0) || !std::isfinite(fmax) || std::abs(a[0] - f[P]) > 1e-8 * fmax ||
        !(std::abs(a[1]) > 0.0))
        return KLSP_OK;
    cd t = -a[0] / a[1];
    for (int it = 0; it < 60; ++it) {
        cd p = a[P - 1], dp = 0.0;
        for (int n = P - 2; n >= 0; --n) {
            dp = dp * t + p;
            p = p * t + a[n];
        }
        if (!(std::abs(dp) > 0.0)) return KLSP_OK;
        const cd dt = p / dp;
        t -= dt;
        if (!(std::abs(t) <= 2.0 * r)) return KLSP_OK;
        if (std::abs(dt) <= 1e-15 * r) break;
    }
    if (std::abs(t) <= r) {
        *t_out = t;
        *ok = true;
    }
    return KLSP_OK;
}

// Mixed-precision Newton (KLSP_PREC_MIXED; documented at klsp_newton in
// klsp.h).  omega = m*Omega_p + i*gamma; det(I - M(omega)) is holomorphic in
// omega, so the step is the complex Newton step -D/D'.
static int newton_mixed(klsp_ctx* ctx, const klsp_newton_params* np, int lp_prec,
                        klsp_newton_result* res, double* history)
{
    typedef std::complex<double> cd;
    const double m = (double)ctx->m;
    const double switch_step = 1e-6;    // leave low precision below this move
    const double accept_step = 1e-5;    // a model Newton step below this ends a pass
    const double h = 1e-6;              // central-difference step on the model
    klsp_timing acc{};
    int rows = 0, n_lp = 0, n_64 = 0, rc;
    auto record = [&](cd w, double absdet) {
        if (history && rows < np->max_iter) {
            history[rows] = w.real() / m;
            history[rows + np->max_iter] = w.imag();
            history[rows + 2 * np->max_iter] = absdet;
        }
        ++rows;
    };
    auto bounded = [&](cd w) {
        double Op = w.real() / m, g = w.imag();
        if (!std::isnan(g) && g < 1e-4) g = 1e-4;
        if (Op > 1.0) Op = 1.0; else if (Op < 0.0) Op = 0.0;
        return cd(m * Op, g);
    };
    cd w(m * np->Omega_p_init, np->gamma_init);
    w = bounded(w);
    cd wc = w;                          // centre of the last low-precision pass
    // 1. low-precision passes on the Taylor model.  A pass with nt terms that
    // moved by m leaves the root off by e ~ r^nt m and its slope off by
    // ~nt r^(nt-1) (relative), r = m / Im w (the series converges for
    // |s| < Im w).  The anchored finish below corrects e with that slope, so
    // the passes stop once e < 1e-6 and e nt r^(nt-1) < 1e-12 (or m < 1e-6);
    // from the default start (m = 4e-3, r = 0.03) one 4-term BF16x3 pass
    // suffices.  When e < 1e-4 the next pass builds three terms: the
    // anchored finish takes its slope between the anchor and the root from
    // T_2 and T_3, and along the det's null vectors T_3 can be ~150 times T_2
    // per unit of omega (Im w ~ 0.017), so a two-term model's slope would be
    // off by ~300 |s_a| ~ 1e-4 there.  A pass ends once a model Newton step
    // below accept_step is applied: from there the step error is
    // ~|step|^2 / Im w.
    const int nt_max = lp_prec == KLSP_PREC_BF16X3 ? kMaxTerms : kMaxBatch;
    const int64_t NN = ctx->N * ctx->N;
    double tnorm[kMaxW] = {};           // Frobenius norms of the last pass's terms
    int nt = nt_max;                    // Taylor terms of the next pass
    int nt_last = nt;                   // ... of the last pass built
    double slope_last = 1.0;            // predicted relative slope error of its model
    for (int pass = 0; pass < np->max_iter; ++pass) {
        wc = w;
        const double om[2] = {w.real(), w.imag()};
        // four BF16x3 terms: T_4 takes the single hi x hi product (it enters the
        // model times s^3 and the anchored finish times 3 s_a^2 (s - s_a))
        if ((rc = build_batch(ctx, om, nt, lp_prec, true, 0, nt == kMaxW ? kMaxW - 1 : nt)))
            return rc;
        hipLaunchKernelGGL(k_sumsq, dim3(nt), dim3(256), 0, ctx->stream, ctx->d_M,
                           (const double2*)nullptr, NN, ctx->d_norm);
        KLSP_HIP(ctx, hipGetLastError());
        double tn2[kMaxW];
        KLSP_HIP(ctx, hipMemcpyAsync(tn2, ctx->d_norm, sizeof(double) * nt, hipMemcpyDeviceToHost,
                                     ctx->stream));
        KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int j = 0; j < nt; ++j) tnorm[j] = std::sqrt(tn2[j]);
        acc_timing(acc, ctx);
        ++n_lp;
        nt_last = nt;
        const double rho = 0.25 * w.imag();
        cd s = 0.0, D0 = 0.0;
        auto model_at = [&](cd si, cd* c) {   // M(w_c + s) = sum_j (-s)^j T_{j+1}
            cd p = 1.0;
            for (int j = 0; j < nt; ++j, p *= -si) c[j] = p;
        };
        // the model's zero within rho from one launch of kCircle + 1 LUs;
        // model Newton steps (three LUs a launch) where that fails
        bool found = false;
        {
            cd t;
            if ((rc = circle_zero(ctx, nt, model_at, 0.0, rho, &t, &D0, &found))) return rc;
            acc_timing(acc, ctx);
            if (found) s = t;
        }
        for (int it = 0; it < 12 && !found; ++it) {
            const cd pts[3] = {s, s + h, s - h};
            cd cf[3 * kMaxW], d[3];
            for (int i = 0; i < 3; ++i) {     // M(w_c + s) = sum_j (-s)^j T_{j+1}
                cd p = 1.0;
                for (int j = 0; j < nt; ++j, p *= -pts[i]) cf[i * nt + j] = p;
            }
            if ((rc = model_dets(ctx, nt, cf, 3, d))) return rc;
            acc_timing(acc, ctx);
            const cd Dp = (d[1] - d[2]) / (2.0 * h);
            if (!(std::abs(Dp) > 0.0) || !std::isfinite(std::abs(Dp))) break;
            cd sn = s - d[0] / Dp;
            // a zero beyond the circle: a step of up to 2 rho (the 4-term
            // model's relative error there is ~(1/2)^4), then a new pass
            const bool clipped = std::abs(sn) > 2.0 * rho;
            if (clipped) sn *= 2.0 * rho / std::abs(sn);
            const double step = std::abs(sn - s);
            s = sn;
            if (clipped || step <= accept_step * std::max(1.0, std::abs(w))) break;
        }
        record(w, std::abs(D0));
        const cd wn = bounded(w + s);
        const double moved = std::abs(wn - w);
        w = wn;
        if (!(moved >= switch_step)) break;
        const double r = moved / std::max(wn.imag(), 1e-300);
        const double e = std::pow(r, nt) * moved, slope_err = nt * std::pow(r, nt - 1);
        slope_last = slope_err;
        if (e < switch_step && e * slope_err < 1e-12) break;
        nt = e < 1e-4 ? 3 : nt_max;
    }
    // 2. FP64 finish on the anchored model.  One FP64 build at the
    // low-precision root w_a gives M64(w_a); with s = w - w_c and s_a = w_a - w_c
    //   M(w) = M64(w_a) + [M_lp(w) - M_lp(w_a)]
    //        = M64(w_a) + sum_{j>=1} ((-s)^j - (-s_a)^j) T_{j+1}
    //        = M64(w_a) - (s - s_a) T_2 + (s - s_a)(s + s_a) T_3 - ...,
    // so the low-precision terms enter only multiplied by |s - s_a|.  (|det| <
    // tol alone is no test of the root: where |det'| is small, ~1e-4 at the
    // default grid, it admits |d omega| ~ 1e-8.)  Newton on the model needs
    // FP64 LUs only; the anchor is rebuilt where the root moves too far.
    const double h64 = 1e-5;            // central-difference step on the model
    // The anchored model's error moves its zero by ~sigma |s - s_a|, sigma the
    // error of the model's slope relative to the slope itself, both taken
    // along the null vectors u, v of I - M at the zero: a term E of M moves
    // the zero by v^H E u / v^H T_2 u.  Norms do not bound these ratios
    // (v^H T_2 u is small where the resonant parts of T_2 cancel; at
    // Im w ~ 0.017 |v^H T_3 u| / |v^H T_2 u| ~ 150 where |T_3| / |T_2| ~ 15),
    // so they are measured: near the zero adj(I - M) ~ kappa u v^H, and the
    // partial derivatives of the anchored model's det in its coefficients,
    //   g_j = |d det / d c_j| ~ |kappa| |v^H T_{j+1} u|   (T_1 = the anchor),
    // come from one LU launch of central differences at s = s_a.  Then
    //  - truncation of the pass's series: the first term left out enters the
    //    slope as nt s_a^(nt-1) T_{nt+1}; its ratio q = g_nt / g_(nt-1) is
    //    continued from the last built ones as max(1 / Im w_c, q_last,
    //    q_last^2 / q_prev), and taken four times: ratios along u, v can grow
    //    faster than geometrically (1.7, 3.4, then ~24 where 1/Im w = 8.1, on
    //    the small grid's published root after a three-term FP32 pass);
    //  - rounding of the low-precision terms.  The rounding E_1 of T_1 moved
    //    the pass's zero by delta_b = |s - s_a| = |v^H E_1 u| / |v^H T_2 u| at
    //    the first anchor, and T_2's relative error along u, v is
    //    |v^H E_2 u| / |v^H T_2 u| = delta_b |v^H E_2 u| / |v^H E_1 u|: a ratio
    //    of two projections of rounding matrices, which follows |E_2| / |E_1|
    //    ~ |T_2| / |T_1| up to a heavy-tailed factor (0.5 to 12.7 where
    //    delta_b >= 1e-7 over 48 measured anchors; kRoundRatio = 16 covers
    //    it: the largest, 9.6e-5 of BF16x3 rounding where |v^H u| = 0.016,
    //    came with delta_b = 1e-6).  Where delta_b happens to be small the
    //    factor can be larger (20 at delta_b = 5e-8), so the error is also
    //    held to at least eps_lp (about twice the largest seen at such
    //    anchors: 2.1e-5 BF16x3, 5.8e-7 FP32).  T_{j+1}, j >= 2, adds
    //    j |s_a|^(j-1) times the larger of
    //    g_j / g_1 eps_lp (BF16x3's hi x hi fourth term: 1e-2) and
    //    kRoundRatio delta_b |T_{j+1}| / |T_1|.
    // A zero is accepted where
    //   |s - s_a| (4 trunc + max(eps_lp, 16 delta_b |T_2| / |T_1|) + higher) <= 5e-11,
    // half the 1e-10 the roots are held to; further out a new anchor is built
    // at the model's zero.  (Default grid: the BF16x3 zero 5.8e-7 off,
    // estimate about 4e-11, one FP64 build.)  Without finite partials the bound
    // falls back to nt r^(nt-1) + 1e-4 (BF16x3) or 1e-5 (FP32).
    const double eps_fallback = lp_prec == KLSP_PREC_BF16X3 ? 1e-4 : 1e-5;
    const double eps_lp = lp_prec == KLSP_PREC_BF16X3 ? 4e-5 : 2e-6;
    const double kRoundRatio = 16.0;
    double delta_b = NAN;               // |s - s_a| at the first anchor
    cd D(NAN, NAN);
    int conv = 0;
    for (int anchor = 0; anchor < std::max(1, np->max_iter); ++anchor) {
        // converged is decided by the last anchor alone: its Newton loop must
        // end on the step test with the zero inside anchor_tol of the anchor
        conv = 0;
        bool stepped = false;
        {
            const double om[2] = {w.real(), w.imag()};
            if ((rc = build_batch(ctx, om, 1, KLSP_PREC_FP64))) return rc;   // d_M[0]
            KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
            acc_timing(acc, ctx);
            ++n_64;
        }
        const cd sa = w - wc;
        // g_j: the anchored model's det at s = s_a (c_0 = 1, c_j = 0) with c_j
        // moved by +-h_j, h_j |T_{j+1}| = 1e-6 |T_1|, in one LU launch
        double g[kMaxW] = {};
        bool have_g = nt_last >= 3 && 2 * nt_last <= kModelPts && tnorm[0] > 0.0;
        if (have_g) {
            double hj[kMaxW];
            cd cf[2 * kMaxW * kMaxW], d[2 * kMaxW];
            for (int j = 0; j < nt_last; ++j) {
                hj[j] = 1e-6 * tnorm[0] / tnorm[j];
                if (!(hj[j] > 0.0) || !std::isfinite(hj[j])) have_g = false;
                for (int e = 0; e < 2; ++e) {
                    cd* c = cf + (2 * j + e) * nt_last;
                    for (int i = 0; i < nt_last; ++i) c[i] = i == 0 ? 1.0 : 0.0;
                    c[j] += e ? -hj[j] : hj[j];
                }
            }
            if (have_g) {
                if ((rc = model_dets(ctx, nt_last, cf, 2 * nt_last, d))) return rc;
                acc_timing(acc, ctx);
                for (int j = 0; j < nt_last; ++j) {
                    g[j] = std::abs(d[2 * j] - d[2 * j + 1]) / (2.0 * hj[j]);
                    if (!(g[j] > 0.0) || !std::isfinite(g[j])) have_g = false;
                }
            }
        }
        // anchor_tol: the largest |s - s_a| the bound above accepts (at the
        // first anchor delta_b is that same distance: a quadratic)
        double anchor_tol = NAN;
        if (have_g) {
            const double asa = std::abs(sa);
            const double q_last = g[nt_last - 1] / g[nt_last - 2];
            const double q_prev = g[nt_last - 2] / g[nt_last - 3];
            const double q_next = std::max(std::max(1.0 / std::max(wc.imag(), 1e-300), q_last),
                                           q_last * q_last / q_prev);
            const double trunc =
                nt_last * std::pow(asa, nt_last - 1) * g[nt_last - 1] * q_next / g[1];
            // sigma for a pass zero that lay x from the anchored one
            auto sigma = [&](double x) {
                double sg = 4.0 * trunc + std::max(eps_lp, kRoundRatio * x * tnorm[1] / tnorm[0]);
                for (int j = 2; j < nt_last; ++j) {
                    // a pass of kMaxW terms builds the last from hi x hi alone
                    const double eps_j = nt_last == kMaxW && j == kMaxW - 1 ? 1e-2 : eps_lp;
                    sg += j * std::pow(asa, j - 1) *
                          std::max(g[j] / g[1] * eps_j, kRoundRatio * x * tnorm[j] / tnorm[0]);
                }
                return sg;
            };
            if (anchor == 0) {
                // x sigma(x) increases with x: the largest x it keeps <= 5e-11
                double lo = 0.0, hi = 5e-11 / sigma(0.0);
                for (int it = 0; it < 60; ++it) {
                    const double mid = 0.5 * (lo + hi);
                    (mid * sigma(mid) <= 5e-11 ? lo : hi) = mid;
                }
                anchor_tol = lo;
            } else {
                anchor_tol = 5e-11 / sigma(delta_b);
            }
        }
        if (!(anchor_tol > 0.0) || !std::isfinite(anchor_tol))
            anchor_tol = 5e-11 / (std::min(slope_last, 1.0) + eps_fallback);
        cd s = sa;
        double prev_step = INFINITY;
        bool far = false;
        auto anchored_at = [&](cd si, cd* c) {
            // (-s)^j - (-s_a)^j = (-1)^j (s - s_a) sum_{q<j} s^q s_a^(j-1-q), free
            // of the cancellation of the difference of powers
            const cd ds = si - sa;
            c[0] = 1.0;
            for (int j = 1; j < nt_last; ++j) {
                cd sum = 0.0, sp = 1.0;
                for (int q = 0; q < j; ++q, sp *= si) sum += sp * std::pow(sa, j - 1 - q);
                c[j] = ((j & 1) ? -1.0 : 1.0) * ds * sum;
            }
        };
        // the anchored model's zero within anchor_tol of s_a from one launch;
        // the Newton steps below then confirm it (normally in one launch)
        {
            cd t, D_a;
            bool found = false;
            if ((rc = circle_zero(ctx, nt_last, anchored_at, sa, anchor_tol, &t, &D_a, &found)))
                return rc;
            acc_timing(acc, ctx);
            record(w, std::abs(D_a));
            if (found) s = sa + t;
        }
        for (int it = 0; it < 30; ++it) {
            const cd pts[3] = {s, s + h64, s - h64};
            cd cf[3 * kMaxW], d[3];
            for (int i = 0; i < 3; ++i) anchored_at(pts[i], cf + nt_last * i);
            if ((rc = model_dets(ctx, nt_last, cf, 3, d))) return rc;
            acc_timing(acc, ctx);
            D = d[0];
            const cd Dp = (d[1] - d[2]) / (2.0 * h64);
            if (!(std::abs(Dp) > 0.0) || !std::isfinite(std::abs(Dp))) break;
            const cd wn = bounded(wc + s - D / Dp);
            const double step = std::abs(wn - (wc + s));
            // converged: a step below 1e-12 |w| (a hundredth of the 1e-10 the
            // roots are held to; the point kept is the one the step starts
            // from), or no longer shrinking (the rounding floor of the FP64
            // dets)
            if (step <= 1e-12 * std::max(1.0, std::abs(wn)) || (it >= 2 && step > 0.5 * prev_step)) {
                stepped = true;
                break;
            }
            prev_step = step;
            s = wn - wc;
            // beyond anchor_tol the anchored model's zero may be off by more
            // than 5e-11, but it is still far better than the point where it
            // left: converge on it, then build the next anchor there (unless
            // it strays so far that the model itself is doubtful)
            if (std::abs(s - sa) > anchor_tol) far = true;
            if (std::abs(s - sa) > 0.25 * wc.imag()) break;
            // the next anchor is built here: once Newton's next step (~ step^2
            // / Im w near a simple zero) is below a thousandth of anchor_tol,
            // further steps on this anchor's model change nothing it needs
            if (far && step * step < 1e-3 * anchor_tol * wc.imag()) break;
        }
        if (anchor == 0) delta_b = std::abs(s - sa);
        w = wc + s;
        if (!far) {
            conv = stepped && std::abs(D) < np->tol;
            break;
        }
    }
    record(w, std::abs(D));
    res->Omega_p = w.real() / m;
    res->gamma = w.imag();
    res->omega[0] = w.real();
    res->omega[1] = w.imag();
    res->det[0] = D.real();
    res->det[1] = D.imag();
    res->iterations = rows;
    res->converged = conv;
    res->n_det_evals = n_lp + n_64;
    res->n_fp64_builds = n_64;
    res->n_lp_builds = n_lp;
    ctx->last = acc;
    return KLSP_OK;
}

extern "C" {

int klsp_build_M(klsp_ctx* ctx, double om_re, double om_im, int precision, double* M_out)
try {
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
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

int klsp_build_taylor(klsp_ctx* ctx, double om_re, double om_im, int n_terms, int precision,
                      double* T_out)
try {
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (!T_out || n_terms < 1 ||
        n_terms > (precision == KLSP_PREC_BF16X3 ? kMaxTerms : kMaxBatch))
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_build_taylor: bad arguments");
    int rc = check_ready(ctx, precision, "klsp_build_taylor");
    if (rc) return rc;
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    reset_timing(ctx);
    const double om[2] = {om_re, om_im};
    if ((rc = build_batch(ctx, om, n_terms, precision, true))) return rc;
    KLSP_HIP(ctx, hipMemcpyAsync(T_out, ctx->d_M, sizeof(double2) * n_terms * ctx->N * ctx->N,
                                 hipMemcpyDeviceToHost, ctx->stream));
    KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

int klsp_build_pencil(klsp_ctx* ctx, double* D_out, double* E_out, double* F_out)
try {
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (!D_out && !E_out && !F_out) return KLSP_OK;
    int rc = check_ready(ctx, KLSP_PREC_FP64, "klsp_build_pencil");
    if (rc) return rc;
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    reset_timing(ctx);
    const int64_t NPh = ctx->NPh, N = ctx->N, NN = N * N;
    const int64_t Kps = round_up(NPh, 8), kstride = Kps + kRowPad;   // one slab of K
    // the three weight sets share each W pass: the FP64 three-omega kernel
    // over one slab's rows, in gangs of at least 256 rows filling the device
    const ContractShape sh = contract_f64_shape(3);
    const int nbk = ctx->Np / 16;
    const int wrows = (nbk + sh.MB - 1) / sh.MB, wcols = (nbk + sh.NB - 1) / sh.NB;
    const int gw = wrows * wcols;
    const int slots = std::max(1, ctx->n_cu / 8) * std::max(1, sh.waves_per_cu);
    int gpx = (int)std::max<int64_t>(1, std::min<int64_t>(Kps / (8 * 256), std::max(1, slots / gw)));
    const int n_gangs = 8 * gpx;
    const int64_t Kg = round_up((Kps + n_gangs - 1) / n_gangs, 8);
    if ((rc = ensure(ctx, ctx->d_c, ctx->c_cap, (size_t)3 * kstride))) return rc;
    if ((rc = ensure(ctx, ctx->d_P, ctx->P_cap, (size_t)3 * n_gangs * ctx->Np * ctx->Np))) return rc;
    if ((rc = ensure(ctx, ctx->d_M, ctx->M_cap, (size_t)std::max<int64_t>(kMaxW, 3) * NN))) return rc;
    if ((rc = ensure(ctx, ctx->d_ones, ctx->ones_cap, (size_t)N))) return rc;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(N)), dim3(256), 0, ctx->stream, ctx->d_ones, N, 1.0);
    double* outs[3] = {D_out, E_out, F_out};
    for (int64_t li = 0; li < ctx->n_l; ++li) {
        hipLaunchKernelGGL(k_weights_pencil, dim3(grid_for(kstride)), dim3(256), 0, ctx->stream,
                           ctx->d_DJ, ctx->d_F0l, ctx->d_Ores, NPh, (int)li, kstride, ctx->d_c);
        ContractArgs a{};
        a.W = ctx->d_W64 + (size_t)li * NPh * ctx->Np;   // the slab's first row
        a.cw = ctx->d_c;
        a.P = ctx->d_P;
        a.Kp = Kps;   // rows past the slab meet zero weights
        a.Np = ctx->Np;
        a.Kg = Kg;
        a.n_gangs = n_gangs;
        a.wrows = wrows;
        a.wcols = wcols;
        a.units = nullptr;
        // the kernel reads weights at stride Kp + kRowPad = kstride
        KLSP_HIP(ctx, contract_f64_launch(3, a, ctx->stream));
        hipLaunchKernelGGL(k_reduce, dim3(grid_for(3 * NN, kRedElems)), dim3(kRedElems * kRedParts),
                           0, ctx->stream, ctx->d_P, n_gangs, (const int*)nullptr, ctx->Np, (int)N,
                           3, 1.0, ctx->d_ones, ctx->d_M);
        KLSP_HIP(ctx, hipGetLastError());
        for (int w = 0; w < 3; ++w)
            if (outs[w])
                KLSP_HIP(ctx, hipMemcpyAsync(outs[w] + 2 * NN * li, ctx->d_M + (size_t)w * NN,
                                             sizeof(double2) * NN, hipMemcpyDeviceToHost,
                                             ctx->stream));
        KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

int klsp_det(klsp_ctx* ctx, const double* omegas, int64_t n, int precision, double* dets_out)
try {
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (n < 0 || (n > 0 && (!omegas || !dets_out)))
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_det: bad arguments");
    reset_timing(ctx);
    if (n == 0) return KLSP_OK;
    return det_many(ctx, omegas, n, precision, dets_out);
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

// klsp_det in FP64 with every workspace carved from this device's `scratch`
// (documented in klsp.h).  Layout: weights c [Kp + kRowPad] | K-piece
// partials [n_slots][Np][Np] | S [N][N] | M, LU work [slots][N][N] | dets,
// omegas [slots]; nothing needs clearing (every region is written before it
// is read).
int klsp_det_noalloc(klsp_ctx* ctx, const double* omegas, int64_t n, double* dets_out)
try {
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (n < 0 || (n > 0 && (!omegas || !dets_out)))
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_det_noalloc: bad arguments");
    reset_timing(ctx);
    if (n == 0) return KLSP_OK;
    if (!ctx->have_problem || !ctx->have_W || !ctx->have_arrays)
        return fail(ctx, KLSP_ERR_NOT_READY,
                    "klsp_det_noalloc: problem arrays and W must be on device");
    if (!ctx->d_scratch)
        return fail(ctx, KLSP_ERR_NOT_READY, "klsp_det_noalloc: no scratch on this device");
    const bool symb = ctx->use_sym && ctx->sym == 1;
    const klsp_ctx::Balanced* bal = symb ? &ctx->bal_sym : &ctx->bal[0][1];
    if (ctx->sym < 0 || !bal->built || (symb && (!ctx->have_R || ctx->bal_sym_variant != ctx->sym_variant)))
        return fail(ctx, KLSP_ERR_NOT_READY,
                    "klsp_det_noalloc: call klsp_prepare(ctx, KLSP_PREC_FP64) first");
    const int64_t Kpc = ctx->Kp + kRowPad, NN = ctx->N * ctx->N;
    const size_t Npp = (size_t)ctx->Np * ctx->Np;
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t b_c = al(sizeof(double2) * (size_t)Kpc);
    const size_t b_P = al(sizeof(double2) * Npp * (size_t)bal->n_slots);
    const size_t b_S = al(sizeof(double2) * (size_t)NN);
    const size_t b_M = al(sizeof(double2) * (size_t)NN);
    const size_t fixed = b_c + b_P + b_S;
    const size_t per_slot = 2 * b_M + 2 * 256;
    int64_t slots = 0;
    if (fixed + per_slot <= kScratchBytes)
        slots = std::min<int64_t>({(int64_t)16, n, (int64_t)((kScratchBytes - fixed) / per_slot)});
    if (slots < 1)
        return fail(ctx, KLSP_ERR_INVALID_ARG,
                    "klsp_det_noalloc: the problem needs " + std::to_string(fixed + per_slot) +
                        " bytes of workspace, more than the " + std::to_string(kScratchBytes) +
                        "-byte scratch");
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    char* base = (char*)ctx->d_scratch;
    double2* c_s = (double2*)base;
    double2* P_s = (double2*)(base + b_c);
    double2* S_s = (double2*)(base + b_c + b_P);
    double2* M_s = (double2*)(base + fixed);
    double2* A_s = (double2*)(base + fixed + slots * b_M);
    double2* det_s = (double2*)(base + fixed + 2 * slots * b_M);
    double* om_s = (double*)(base + fixed + 2 * slots * b_M + al(sizeof(double2) * slots));
    for (int64_t i0 = 0; i0 < n; i0 += slots) {
        const int nd = (int)std::min<int64_t>(slots, n - i0);
        KLSP_HIP(ctx, hipMemcpyAsync(om_s, omegas + 2 * i0, sizeof(double) * 2 * nd,
                                     hipMemcpyHostToDevice, ctx->stream));
        for (int j = 0; j < nd; ++j) {
            hipLaunchKernelGGL(k_weights, dim3(grid_for(Kpc)), dim3(256), 0, ctx->stream,
                               ctx->d_DJ, ctx->d_F0l, ctx->d_Ores, ctx->NPh, ctx->K, Kpc,
                               om_s + 2 * j, 1, c_s, (float2*)nullptr);
            ContractArgs a{};
            a.units = bal->d_units;
            a.n_units = bal->n_units;
            a.W = symb ? (const void*)ctx->d_R : (const void*)ctx->d_W64;
            a.cw = c_s;
            a.P = P_s;
            a.Kp = ctx->Kp;
            a.Np = ctx->Np;
            a.n_gangs = bal->n_slots;
            a.wrows = bal->wrows;
            a.wcols = bal->wcols;
            const hipError_t e = symb ? contract_sym64_launch(ctx->sym_variant, a, ctx->stream)
                                      : contract_f64_launch(1, a, ctx->stream);
            if (e != hipSuccess)
                return fail(ctx, KLSP_ERR_HIP, std::string("contraction launch: ") + hipGetErrorString(e));
            if (symb) {
                hipLaunchKernelGGL(k_reduce_sym, dim3(grid_for(NN, kRedElems)),
                                   dim3(kRedElems * kRedParts), 0, ctx->stream, P_s, bal->n_slots,
                                   bal->d_count, ctx->Np, (int)ctx->N, 1, S_s);
                hipLaunchKernelGGL(k_expand_sym, dim3(grid_for(NN)), dim3(256), 0, ctx->stream,
                                   S_s, (int)ctx->N, 1, ctx->G, ctx->d_Nk, M_s + (size_t)j * NN);
            } else {
                hipLaunchKernelGGL(k_reduce, dim3(grid_for(NN, kRedElems)),
                                   dim3(kRedElems * kRedParts), 0, ctx->stream, P_s, bal->n_slots,
                                   bal->d_count, ctx->Np, (int)ctx->N, 1, ctx->G, ctx->d_Nk,
                                   M_s + (size_t)j * NN);
            }
        }
        hipLaunchKernelGGL(k_lu_det, dim3(nd), dim3(kLuThreads), 0, ctx->stream, M_s,
                           (const double2*)nullptr, 0, A_s, (int)ctx->N, det_s);
        KLSP_HIP(ctx, hipGetLastError());
        KLSP_HIP(ctx, hipMemcpyAsync(dets_out + 2 * i0, det_s, sizeof(double2) * nd,
                                     hipMemcpyDeviceToHost, ctx->stream));
        KLSP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

// newton_search, NewtonSolver.jl:36-172, with the forward-difference pair
// (omega + m*delta, omega + i*delta) evaluated as one two-omega batch.
int klsp_newton(klsp_ctx* ctx, const klsp_newton_params* np, klsp_newton_result* res,
                double* history)
try {
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (!np || !res || np->max_iter < 0 || np->max_line_search < 0)
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_newton: bad arguments");
    if (np->precision == KLSP_PREC_MIXED || np->precision == KLSP_PREC_MIXED_FP32) {
        const int lp = np->precision == KLSP_PREC_MIXED ? KLSP_PREC_BF16X3 : KLSP_PREC_FP32;
        int rc = check_ready(ctx, lp, "klsp_newton");
        if (rc) return rc;
        reset_timing(ctx);
        return newton_mixed(ctx, np, lp, res, history);
    }
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
        acc_timing(acc, ctx);
        return r;
    };
    auto det2 = [&](const double* om4, double* d4) -> int {
        n_eval += 2;
        int r = det_many(ctx, om4, 2, prec, d4);
        acc_timing(acc, ctx);
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
    res->n_fp64_builds = prec == KLSP_PREC_FP64 ? n_eval : 0;
    res->n_lp_builds = prec == KLSP_PREC_FP64 ? 0 : n_eval;
    ctx->last = acc;
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

int klsp_grid_scan(klsp_ctx* ctx, const klsp_scan_params* sp, double* det_grid,
                   double* Op_init, double* g_init)
try {
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
    if ((rc = ctx->scan_graph ? scan_by_graph(ctx, om.data(), n, sp->precision, det_grid)
                              : det_many(ctx, om.data(), n, sp->precision, det_grid)))
        return rc;
    // find_minimum (GridScan.jl:234-246): argmin |det| in column-major
    // order, first occurrence; Julia's argmin takes NaN as the minimum
    int64_t best = 0;
    double bv = INFINITY;
    for (int64_t x = 0; x < n; ++x) {
        const double v = std::hypot(det_grid[2 * x], det_grid[2 * x + 1]);
        if (std::isnan(v)) { best = x; break; }
        if (v < bv) { bv = v; best = x; }
    }
    if (Op_init) *Op_init = Opg[best / sp->n_gamma];
    if (g_init) *g_init = gg[best % sp->n_gamma];
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

int klsp_device_bytes(klsp_ctx* ctx, int64_t* bytes_out)
{
    if (!ctx || !bytes_out) return KLSP_ERR_INVALID_ARG;
    int64_t b = 0;
    b += (int64_t)ctx->W64_cap * sizeof(double2) + (int64_t)ctx->W32_cap * sizeof(float2);
    b += (int64_t)ctx->Wb_cap * sizeof(uint4) + (int64_t)ctx->R_cap * sizeof(double);
    b += (int64_t)ctx->Rb_cap * sizeof(uint4) + (int64_t)ctx->c_cap * sizeof(double2);
    b += (int64_t)ctx->cf_cap * sizeof(float2) + (int64_t)ctx->P_cap * sizeof(double2);
    b += (int64_t)ctx->M_cap * sizeof(double2) + (int64_t)ctx->A_cap * sizeof(double2);
    b += (int64_t)ctx->det_cap * sizeof(double2) + (int64_t)ctx->S_cap * sizeof(double2);
    b += (int64_t)ctx->ones_cap * sizeof(double) + (int64_t)ctx->scan_om_cap * sizeof(double);
    b += (int64_t)ctx->scan_det_cap * sizeof(double2) + (int64_t)ctx->R32_cap * sizeof(float);
    if (ctx->have_problem)   // DJ, F0l, Omega_res, N_kernel
        b += (ctx->NPh + 2 * ctx->K + ctx->N) * (int64_t)sizeof(double);
    for (const auto& row : ctx->bal)
        for (const auto& t : row) b += (int64_t)t.n_units * sizeof(WorkUnit) + 4096;
    b += (int64_t)ctx->bal_sym.n_units * sizeof(WorkUnit) + 4096;
    b += (int64_t)ctx->bal_sym32.n_units * sizeof(WorkUnit) + 4096;
    *bytes_out = b;
    return KLSP_OK;
}

int klsp_set_scan_graph(klsp_ctx* ctx, int on)
{
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    ctx->scan_graph = on ? 1 : 0;
    return KLSP_OK;
}

int klsp_set_lu_form(klsp_ctx* ctx, int form)
{
    if (!ctx || form < 0 || form > 2) return KLSP_ERR_INVALID_ARG;
    ctx->lu_form = form;
    return KLSP_OK;
}

int klsp_set_bf16_form(klsp_ctx* ctx, int form)
{
    if (!ctx || form < 0 || form > 2) return KLSP_ERR_INVALID_ARG;
    ctx->bf16_form = form;
    return KLSP_OK;
}

int klsp_scan_graph_stats(klsp_ctx* ctx, int64_t* builds_out, int64_t* launches_out)
{
    if (!ctx || !builds_out || !launches_out) return KLSP_ERR_INVALID_ARG;
    *builds_out = ctx->scan_graph_builds;
    *launches_out = ctx->scan_graph_launches;
    return KLSP_OK;
}

int klsp_autotune(klsp_ctx* ctx, int precision, int reps, int* n_variants, double* ms_out,
                  int n_max, int* chosen)
try {
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (reps < 1 || !n_variants || (n_max > 0 && !ms_out))
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_autotune: bad arguments");
    int rc = check_ready(ctx, precision, "klsp_autotune");
    if (rc) return rc;
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    *n_variants = 1;
    if (chosen) *chosen = 0;
    if (precision != KLSP_PREC_FP64 || (rc = prepare_sym(ctx)) || !(ctx->use_sym && ctx->sym == 1)) {
        // one form only: its time
        if (rc) return rc;
        const bool prof = ctx->prof;
        ctx->prof = true;
        const double om[2] = {1.0, 0.1};
        double t = 0.0;
        for (int r = 0; r <= reps && rc == KLSP_OK; ++r) {
            reset_timing(ctx);
            rc = build_batch(ctx, om, 1, precision);
            if (r > 0) t += ctx->last.contract_ms;
        }
        ctx->prof = prof;
        reset_timing(ctx);
        if (rc) return rc;
        if (n_max > 0) ms_out[0] = t / reps;
        return KLSP_OK;
    }
    // FP64 on a symmetric grid: every compiled variant of the real-basis
    // contraction, each with one untimed launch and `reps` timed ones
    const int nv = contract_sym64_variants();
    const bool prof = ctx->prof;
    const int keep = ctx->sym_variant;
    ctx->prof = true;
    const double om[2] = {1.0, 0.1};
    int best = keep;
    double best_ms = INFINITY;
    for (int v = 0; v < nv && rc == KLSP_OK; ++v) {
        ctx->sym_variant = v;
        double t = 0.0;
        for (int r = 0; r <= reps && rc == KLSP_OK; ++r) {
            reset_timing(ctx);
            rc = build_batch(ctx, om, 1, KLSP_PREC_FP64);
            if (r > 0) t += ctx->last.contract_ms;
        }
        t /= reps;
        if (v < n_max) ms_out[v] = t;
        if (t < best_ms) { best_ms = t; best = v; }
    }
    ctx->prof = prof;
    reset_timing(ctx);
    if (rc) {
        ctx->sym_variant = keep;
        return rc;
    }
    ctx->sym_variant = best;
    ctx->sym_tuned[ctx->Np] = best;
    {
        const klsp_ctx::Balanced* b = nullptr;
        if ((rc = get_balanced_sym(ctx, &b))) return rc;
    }
    *n_variants = nv;
    if (chosen) *chosen = best;
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
}

int klsp_set_variant(klsp_ctx* ctx, int precision, int variant)
{
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    if (precision != KLSP_PREC_FP64 || variant < 0 || variant >= contract_sym64_variants())
        return fail(ctx, KLSP_ERR_INVALID_ARG, "klsp_set_variant: no such variant");
    ctx->sym_variant = variant;
    return KLSP_OK;
}

const char* klsp_variant_name(int precision, int variant)
{
    if (precision == KLSP_PREC_FP64 && variant >= 0 && variant < contract_sym64_variants())
        return contract_sym64_name(variant);
    return "";
}

int klsp_set_symmetric_basis(klsp_ctx* ctx, int on)
{
    if (!ctx) return KLSP_ERR_INVALID_ARG;
    ctx->use_sym = on != 0;
    return KLSP_OK;
}

int klsp_symmetric_basis(klsp_ctx* ctx, int* active_out)
try {
    if (!ctx || !active_out) return KLSP_ERR_INVALID_ARG;
    if (!ctx->have_W) return fail(ctx, KLSP_ERR_NOT_READY, "klsp_symmetric_basis: no W on device");
    KLSP_HIP(ctx, hipSetDevice(ctx->device));
    int rc = prepare_sym(ctx);
    if (rc) return rc;
    *active_out = ctx->use_sym && ctx->sym == 1 && ctx->have_R;
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(ctx ? &ctx->err : nullptr);
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

int klsp_last_orbit_ms(klsp_ctx* ctx, double* ms)
{
    if (!ctx || !ms) return KLSP_ERR_INVALID_ARG;
    *ms = ctx->last_orbit_ms;
    return KLSP_OK;
}

}  // extern "C"
