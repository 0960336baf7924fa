// klsp_host.cpp -- host (CPU, FP64) half of the precompute: quadrature
// weights, the complex Gamma ratio N_m(alpha), the Toomre-Zang model, the
// (Rc, v) grids, orbit sampling, Jacobian and DF.  Everything here is cheap
// scalar work except compute_orbits, which is spread over host threads (one
// orbit per task; each orbit's cumulative sums run in order, so results do
// not depend on the thread count).
//
// Reference: src/KalnajsLogSpiral/{Models,OrbitIntegration,BasisFunctions}.jl.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstring>
#include <thread>
#include <vector>

#include "klsp.h"
#include "klsp_internal.hpp"

namespace klsp {

namespace {

constexpr double kPi = 3.14159265358979323846;

// Julia range(a, b, length=n) element i: one rounding of the exact lerp.
inline double range_at(double a, double b, int64_t n, int64_t i)
{
    if (i == 0) return a;
    if (i == n - 1) return b;
    long double num = (long double)a * (long double)(n - 1 - i) + (long double)b * (long double)i;
    return (double)(num / (long double)(n - 1));
}

// Julia inv(::ComplexF64): Smith's algorithm (range scaling omitted where the
// arguments are O(1); it is kept for the general entry points).
std::complex<double> jl_inv(std::complex<double> w)
{
    double c = w.real(), d = w.imag();
    if (std::isinf(c) || std::isinf(d))
        return {std::copysign(0.0, c), d < 0 ? 0.0 : -0.0};
    double s = 1.0;
    const double cd = std::max(std::fabs(c), std::fabs(d));
    const double eps = 2.220446049250313e-16;
    if (cd >= 0.5 * 1.7976931348623157e308) { c *= 0.5; d *= 0.5; s *= 0.5; }
    if (cd <= 2.2250738585072014e-308 * 2.0 / eps) {
        const double bs = 2.0 / (eps * eps);
        c *= bs; d *= bs; s *= bs;
    }
    double p, q;
    if (std::fabs(d) <= std::fabs(c)) {
        const double r = d / c, t = 1.0 / (c + d * r);
        p = t; q = -r * t;
    } else {
        const double r = c / d, t = 1.0 / (d + c * r);
        p = r * t; q = -t;
    }
    return {p * s, q * s};
}

std::complex<double> cmul(std::complex<double> a, std::complex<double> b)
{
    return {a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real()};
}

std::complex<double> cdiv(std::complex<double> a, std::complex<double> b)
{
    const double c = b.real(), d = b.imag();
    if (std::fabs(d) <= std::fabs(c)) {
        const double r = d / c, t = 1.0 / (c + d * r);
        return {(a.real() + a.imag() * r) * t, (a.imag() - a.real() * r) * t};
    }
    const double r = c / d, t = 1.0 / (c * r + d);
    return {(a.real() * r + a.imag()) * t, (a.imag() * r - a.real()) * t};
}

std::complex<double> cexp_(std::complex<double> z)
{
    const double er = std::exp(z.real());
    return {er * std::cos(z.imag()), er * std::sin(z.imag())};
}

std::complex<double> cpow_(std::complex<double> z, std::complex<double> p)
{
    const std::complex<double> lz{std::log(std::hypot(z.real(), z.imag())),
                                  std::atan2(z.imag(), z.real())};
    return cexp_(cmul(p, lz));
}

// Toomre-Zang model, Models.jl:24-169 (V = ln r, flat rotation curve).
struct Model {
    double L0, sigma_r0, s2, DF_const;
    int n_zang, q_zang;

    explicit Model(const klsp_params& p)
    {
        L0 = p.L0;
        n_zang = p.n_zang;
        q_zang = p.q1 - 1;
        sigma_r0 = 1.0 / std::sqrt((double)p.q1);
        s2 = sigma_r0 * sigma_r0;
        DF_const = 1.0 / (2.0 * kPi * std::pow(2.0, q_zang / 2.0) * std::sqrt(kPi) *
                          std::tgamma((double)(q_zang + 1) / 2) *
                          std::pow(sigma_r0, (double)(q_zang + 2)));
    }
    static double V(double r) { return std::log(r); }
    static double dV(double r) { return 1.0 / r; }
    double taper(double L) const
    {
        if (n_zang == 0) return 1.0;
        const double La = std::fabs(L);
        if (La < 1e-12) return 0.0;
        return 1.0 / (1.0 + std::pow(L0 / La, (double)n_zang));
    }
    double F(double E, double L) const
    {
        const double La = std::fabs(L);
        if (La < 1e-12) return 0.0;
        const double r = DF_const * std::exp(-E / s2) * std::pow(La, (double)q_zang) * taper(L);
        return std::isfinite(r) ? r : 0.0;
    }
    double FE(double E, double L) const { return -F(E, L) / s2; }
    double FL(double E, double L) const
    {
        const double La = std::fabs(L);
        if (La < 1e-12) return 0.0;
        return F(E, L) * ((double)(q_zang + n_zang) - n_zang * taper(L)) / La;
    }
};

bool params_ok(const klsp_params* p)
{
    return p && p->NR >= 1 && p->Ne >= 1 && p->nw >= 3 && p->nwa >= 2 &&
           (p->nw - 1) % (p->nwa - 1) == 0 && p->Rc_min > 0 && p->Rc_max > 0 &&
           p->l_max >= p->l_min && p->N_alpha >= 1;
}

}  // namespace

void host_gamma(double re, double im, double* out)
{
    static const double g = 607.0 / 128.0;
    static const double c[15] = {
        0.99999999999999709182,    57.156235665862923517,    -59.597960355475491248,
        14.136097974741747174,     -0.49191381609762019978,  0.33994649984811888699e-4,
        0.46523628927048575665e-4, -0.98374475304879564677e-4, 0.15808870322491248884e-3,
        -0.21026444172410488319e-3, 0.21743961811521264320e-3, -0.16431810653676389022e-3,
        0.84418223983852743293e-4, -0.26190838401581408670e-4, 0.36899182659531622704e-5};
    const std::complex<double> zz{re, im};
    std::complex<double> z = zz.real() < 0 ? -zz : zz;
    z -= 1.0;
    const std::complex<double> zh = z + 0.5, zgh = zh + g;
    const std::complex<double> zp = cpow_(zgh, zh * 0.5);
    std::complex<double> ss{0.0, 0.0};
    for (int k = 14; k >= 1; --k) ss += c[k] * jl_inv(z + (double)k);
    const double sq2pi = 2.5066282746310005024157652848110;
    std::complex<double> f =
        cmul(cmul(sq2pi * (c[0] + ss), cmul(zp, cexp_(-zgh))), zp);
    if (z == std::complex<double>(0.0, 0.0) || z == std::complex<double>(1.0, 0.0)) f = 1.0;
    if (zz.real() < 0) {
        const std::complex<double> pz = kPi * zz;
        const std::complex<double> s{std::sin(pz.real()) * std::cosh(pz.imag()),
                                     std::cos(pz.real()) * std::sinh(pz.imag())};
        f = -kPi * jl_inv(cmul(cmul(zz, f), s));
    }
    if (std::round(zz.real()) == zz.real() && zz.imag() == 0 && zz.real() <= 0)
        f = {INFINITY, 0.0};
    out[0] = f.real();
    out[1] = f.imag();
}

}  // namespace klsp

using namespace klsp;

extern "C" {

int klsp_linrange(double a, double b, int64_t n, double* out)
{
    if (n < 1 || !out) return KLSP_ERR_INVALID_ARG;
    for (int64_t i = 0; i < n; ++i) out[i] = (n == 1) ? a : range_at(a, b, n, i);
    return KLSP_OK;
}

int klsp_trapezoid_coef(const double* x, int64_t n, double* S)
{
    if (n < 0 || (n > 0 && (!x || !S))) return KLSP_ERR_INVALID_ARG;
    std::fill(S, S + n, 0.0);
    if (n >= 2) {
        S[0] = (x[1] - x[0]) / 2;
        S[n - 1] = (x[n - 1] - x[n - 2]) / 2;
        for (int64_t i = 1; i + 1 < n; ++i) S[i] = (x[i + 1] - x[i - 1]) / 2;
    }
    return KLSP_OK;
}

int klsp_simpson_coef(const double* x, int64_t n, double* S)
{
    if (n < 3 || n % 2 == 0) return klsp_trapezoid_coef(x, n, S);
    const double dx = x[1] - x[0];
    std::fill(S, S + n, 0.0);
    S[0] = S[n - 1] = dx / 3;
    for (int64_t i = 1; i + 1 < n; i += 2) S[i] = 4 * dx / 3;
    for (int64_t i = 2; i + 2 < n; i += 2) S[i] = 2 * dx / 3;
    return KLSP_OK;
}

int klsp_gamma_complex(double re, double im, double* out)
{
    if (!out) return KLSP_ERR_INVALID_ARG;
    host_gamma(re, im, out);
    return KLSP_OK;
}

int klsp_compute_N_kernel(int m, const double* alpha, const double* d_alpha, int64_t N,
                          double* out)
{
    if (N < 0 || (N > 0 && (!alpha || !d_alpha || !out))) return KLSP_ERR_INVALID_ARG;
    for (int64_t i = 0; i < N; ++i) {
        const double a = alpha[i];
        double g[4][2];
        host_gamma((m + 0.5) / 2, a / 2, g[0]);
        host_gamma((m + 0.5) / 2, -a / 2, g[1]);
        host_gamma((m + 1.5) / 2, a / 2, g[2]);
        host_gamma((m + 1.5) / 2, -a / 2, g[3]);
        const std::complex<double> num = cmul({g[0][0], g[0][1]}, {g[1][0], g[1][1]});
        const std::complex<double> den = cmul({g[2][0], g[2][1]}, {g[3][0], g[3][1]});
        out[i] = cdiv(num, den).real() * kPi * d_alpha[i];
    }
    return KLSP_OK;
}

int klsp_compute_grids(const klsp_params* p, double* Rc, double* e, double* v, double* R1,
                       double* R2, double* E, double* L, double* L2, double* SGNL,
                       double* S_RC, double* S_e)
try {
    if (!params_ok(p) || !Rc || !e || !v || !R1 || !R2 || !E || !L || !L2 || !SGNL || !S_RC ||
        !S_e)
        return KLSP_ERR_INVALID_ARG;
    const int NR = p->NR, Ne = p->Ne;
    std::vector<double> Rc1(NR), uRc(NR), tr(NR), va(Ne), se(Ne);
    const double l0 = std::log10(p->Rc_min), l1 = std::log10(p->Rc_max);
    for (int i = 0; i < NR; ++i) {
        Rc1[i] = std::pow(10.0, NR == 1 ? l0 : range_at(l0, l1, NR, i));
        uRc[i] = std::log(Rc1[i]);
    }
    klsp_trapezoid_coef(uRc.data(), NR, tr.data());
    for (int i = 0; i < NR; ++i) S_RC[i] = Rc1[i] * tr[i];
    for (int j = 0; j < Ne; ++j) va[j] = Ne == 1 ? p->v_min : range_at(p->v_min, 1.0, Ne, j);
    klsp_simpson_coef(va.data(), Ne, se.data());
    for (int j = 0; j < Ne; ++j)
        for (int i = 0; i < NR; ++i) {
            const int64_t k = i + (int64_t)NR * j;
            double vv = va[j];
            if (std::fabs(vv) < 1e-14) vv = 1e-12;
            const double rc = Rc1[i];
            const double sg = vv > 0 ? 1.0 : (vv < 0 ? -1.0 : 0.0);
            const double ec = 1.0 - std::fabs(vv);
            const double r1 = rc * (1.0 - ec), r2 = rc * (1.0 + ec);
            double En;
            if (std::fabs(r2 * r2 - r1 * r1) > 1e-14)
                En = (Model::V(r2) * (r2 * r2) - Model::V(r1) * (r1 * r1)) / (r2 * r2 - r1 * r1);
            else
                En = rc * Model::dV(rc) / 2 + Model::V(rc);
            if (!std::isfinite(En)) En = Model::V(r2);
            double l2 = 2.0 * (En - Model::V(r2)) * (r2 * r2);
            if (l2 < 1e-14) l2 = 0.0;
            Rc[k] = rc; v[k] = vv; SGNL[k] = sg; e[k] = ec; R1[k] = r1; R2[k] = r2;
            S_e[k] = se[j]; E[k] = En; L2[k] = l2; L[k] = std::sqrt(l2) * sg;
        }
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(nullptr);
}

int klsp_compute_orbits(const klsp_params* p, const double* Rc, const double* e,
                        const double* R1, const double* R2, const double* E, const double* L,
                        const double* L2, const double* SGNL, double* Om1, double* Om2,
                        double* w1, double* ra, double* pha, double* jac, double* F0,
                        double* FE, double* FL, int n_threads)
try {
    if (!params_ok(p) || !Rc || !e || !R1 || !R2 || !E || !L || !L2 || !SGNL || !Om1 || !Om2 ||
        !w1 || !ra || !pha || !jac || !F0 || !FE || !FL)
        return KLSP_ERR_INVALID_ARG;
    const int NR = p->NR, nw = p->nw, nwa = p->nwa;
    const int64_t NPh = (int64_t)NR * p->Ne;
    const int stride = (nw - 1) / (nwa - 1);
    std::vector<double> w(nw), cw(nw), sw(nw);
    for (int i = 0; i < nw; ++i) {
        w[i] = range_at(1e-8, kPi - 1e-8, nw, i);
        cw[i] = std::cos(w[i]);
        sw[i] = std::sin(w[i]);
    }
    const double dw = w[1] - w[0];
    const Model md(*p);

    // One orbit: OrbitIntegration.jl:209-302.
    auto orbit = [&](int64_t k, std::vector<double>& xs, std::vector<double>& rv,
                     std::vector<double>& sv, std::vector<double>& cum) {
        const double Ri = Rc[k % NR];
        const double kap = std::sqrt(2.0) / Ri, Om = 1.0 / Ri;
        const double ecc = e[k], r1 = R1[k], r2 = R2[k], Ek = E[k], l2 = L2[k];
        double* W1 = w1 + k * nwa;
        double* RA = ra + k * nwa;
        double* PH = pha + k * nwa;
        if (std::fabs(ecc) < 1e-10 || ecc < 0.01) {
            // circular (:221-229) and epicyclic (:231-241) branches
            const bool circ = std::fabs(ecc) < 1e-10;
            const double aj = (r2 - r1) / 2;
            Om1[k] = kap;
            Om2[k] = Om * SGNL[k];
            const double c2 = 2 * Om / kap * aj / Ri;
            for (int a = 0; a < nwa; ++a) {
                const int i = a * stride;
                W1[a] = w[i];
                RA[a] = circ ? Ri : Ri - aj * cw[i];
                PH[a] = w[i] * Om2[k] / Om1[k] + (circ ? 0.0 : c2 * sw[i] * SGNL[k]);
            }
            return;
        }
        const double rs = (r1 + r2) / 2, drs = (r2 - r1) / 2;
        const bool nonradial = l2 > 1e-12;
        for (int i = 0; i < nw; ++i) {
            const double x = rs - drs * cw[i];
            xs[i] = x;
            const double q = nonradial ? 2.0 * (Ek - Model::V(x)) * (x * x) - l2
                                       : 2.0 * (Ek - Model::V(x));
            rv[i] = std::sqrt(std::max(0.0, q));
        }
        for (int i = 0; i < nw; ++i) sv[i] = nonradial ? sw[i] * xs[i] / rv[i] : sw[i] / rv[i];
        if (nonradial) sv[0] = 2.0 * sv[1] - sv[2];
        sv[nw - 1] = 2.0 * sv[nw - 2] - sv[nw - 3];
        // t = cumsum of trapezoid midpoints of drs*dw*svr (:277-281)
        double prev = drs * dw * sv[0], acc = 0.0;
        cum[0] = 0.0;
        for (int i = 1; i < nw; ++i) {
            const double cur = drs * dw * sv[i];
            acc += (prev + cur) / 2;
            cum[i] = acc;
            prev = cur;
        }
        const double o1 = kPi / cum[nw - 1];
        Om1[k] = o1;
        for (int a = 0; a < nwa; ++a) {
            W1[a] = cum[a * stride] * o1;
            RA[a] = xs[a * stride];
        }
        if (std::fabs(1.0 - ecc) > 1e-10 && nonradial) {   // :287-295
            for (int i = 0; i < nw; ++i) sv[i] = sw[i] / rv[i];
            sv[0] = 2.0 * sv[1] - sv[2];
            sv[nw - 1] = 2.0 * sv[nw - 2] - sv[nw - 3];
            prev = drs * dw * sv[0] / xs[0];
            acc = 0.0;
            PH[0] = L[k] * 0.0;
            for (int i = 1; i < nw; ++i) {
                const double cur = drs * dw * sv[i] / xs[i];
                acc += (prev + cur) / 2;
                prev = cur;
                if (i % stride == 0) PH[i / stride] = L[k] * acc;
            }
        } else {                                              // :296-299
            for (int a = 0; a < nwa; ++a) PH[a] = kPi / 2 * 1.0 * SGNL[k];
            PH[0] = 0.0;
        }
        Om2[k] = Om1[k] * PH[nwa - 1] / kPi;
    };

    unsigned nt = n_threads > 0 ? (unsigned)n_threads : std::thread::hardware_concurrency();
    nt = std::max(1u, std::min<unsigned>(nt, 64u));
    std::atomic<int64_t> next{0};
    std::atomic<bool> failed{false};
    // Orbits are taken 16 at a time from one shared counter.  A worker that
    // cannot allocate its scratch drains the counter and flags the call,
    // which fails once every worker has been joined (an exception must not
    // leave a thread function, nor unwind past joinable threads).
    auto worker = [&]() noexcept {
        try {
            std::vector<double> xs(nw), rv(nw), sv(nw), cum(nw);
            for (;;) {
                const int64_t k0 = next.fetch_add(16);
                if (k0 >= NPh) break;
                for (int64_t k = k0; k < std::min<int64_t>(k0 + 16, NPh); ++k)
                    orbit(k, xs, rv, sv, cum);
            }
        } catch (...) {
            failed = true;
            next = NPh;
        }
    };
    std::vector<std::thread> pool;
    try {
        pool.reserve(nt - 1);
        for (unsigned t = 1; t < nt; ++t) pool.emplace_back(worker);
    } catch (...) {
        // fewer threads than asked for: the counter is shared, so the ones
        // that started (and this thread) still cover every orbit
    }
    worker();
    for (auto& th : pool) th.join();
    if (failed) return KLSP_ERR_OOM;

    return klsp_compute_jacobian_df(p, Rc, R1, R2, E, L, Om1, jac, F0, FE, FL);
} catch (...) {
    return klsp::exception_status(nullptr);
}

int klsp_orbit_phase_grid(const klsp_params* p, double* w, double* cw, double* sw)
{
    if (!params_ok(p) || !w || !cw || !sw) return KLSP_ERR_INVALID_ARG;
    for (int i = 0; i < p->nw; ++i) {
        w[i] = range_at(1e-8, kPi - 1e-8, p->nw, i);
        cw[i] = std::cos(w[i]);
        sw[i] = std::sin(w[i]);
    }
    return KLSP_OK;
}

int klsp_compute_jacobian_df(const klsp_params* p, const double* Rc, const double* R1,
                             const double* R2, const double* E, const double* L,
                             const double* Om1, double* jac, double* F0, double* FE, double* FL)
try {
    if (!params_ok(p) || !Rc || !R1 || !R2 || !E || !L || !Om1 || !jac || !F0 || !FE || !FL)
        return KLSP_ERR_INVALID_ARG;
    const int64_t NPh = (int64_t)p->NR * p->Ne;
    const Model md(*p);
    // Jacobian (:341-382) and DF on the grid (:317-333).
    for (int64_t k = 0; k < NPh; ++k) {
        const double r1 = R1[k], r2 = R2[k], rc = Rc[k], Ek = E[k], Lk = L[k], o1 = Om1[k];
        double J = 0.0;
        if (std::fabs(r2 - r1) > 1e-12 && std::fabs(Lk) > 1e-12) {
            const double t1 = 2 * (Ek - Model::V(r1)) * r1 - Model::dV(r1) * (r1 * r1);
            const double t2 = 2 * (Ek - Model::V(r2)) * r2 - Model::dV(r2) * (r2 * r2);
            J = 2 * std::fabs(t1 * t2 / (r2 * r2 - r1 * r1)) * rc / o1 / std::fabs(Lk);
        }
        if (std::fabs(Lk) <= 1e-12) {
            J = std::sqrt(std::max(2.0 * (Ek - Model::V(r1)), 0.0)) * Model::dV(r2) * 2 * rc / o1;
            if (!std::isfinite(J)) J = 0.0;
        }
        jac[k] = J;
        const double f = md.F(Ek, Lk), fe = md.FE(Ek, Lk), fl = md.FL(Ek, Lk);
        F0[k] = std::isfinite(f) ? f : 0.0;
        FE[k] = std::isfinite(fe) ? fe : 0.0;
        FL[k] = std::isfinite(fl) ? fl : 0.0;
    }
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(nullptr);
}

int klsp_precompute_arrays(const klsp_params* p, const double* S_RC, const double* S_e,
                           const double* jac, const double* Om1, const double* Om2,
                           const double* FE, const double* FL, double* DJ, double* F0l,
                           double* Ores)
try {
    if (!p || p->NR < 1 || p->Ne < 1 || p->l_max < p->l_min || !S_RC || !S_e || !jac || !Om1 ||
        !Om2 || !FE || !FL || !DJ || !F0l || !Ores)
        return KLSP_ERR_INVALID_ARG;
    const int64_t NPh = (int64_t)p->NR * p->Ne;
    for (int64_t k = 0; k < NPh; ++k) DJ[k] = S_RC[k % p->NR] * S_e[k] * jac[k];
    for (int l = p->l_min; l <= p->l_max; ++l) {
        const int64_t off = NPh * (int64_t)(l - p->l_min);
        for (int64_t k = 0; k < NPh; ++k) {
            const double f = l * Om1[k] + p->m * Om2[k];
            F0l[off + k] = f * FE[k] + p->m * FL[k];
            Ores[off + k] = f;
        }
    }
    return KLSP_OK;
} catch (...) {
    return klsp::exception_status(nullptr);
}

}  // extern "C"
