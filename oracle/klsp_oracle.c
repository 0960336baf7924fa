/*
 * klsp_oracle.c -- FP64 CPU oracle for the Kalnajs log-spiral hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference Julia package (epolyach/Kalnajs_log_spirals, src/KalnajsLogSpiral/)
 * used as the checker for the HIP library.  Only tests/, the smoke() entry and
 * the CPU-baseline leg of bench.py may load it; the product library never
 * links or calls it.
 *
 * Every function cites the reference lines it follows.  Arithmetic is written
 * out in real/imaginary parts in the reference's operation order (Julia
 * evaluates a .* b ./ c left to right), compiled with -ffp-contract=off so no
 * fused multiply-adds change the rounding.
 *
 * Parity pins (see tests/test_oracle_*.py): quadrature weights are checked
 * against polynomial exactness, the Lanczos gamma against mpmath, orbits
 * against scipy quadrature of the radial/azimuthal periods and the circular
 * limit, W against the circular-orbit closed form and scale invariance, M
 * against an independent numpy BLAS-form build and the det(conj w) =
 * conj det(w) symmetry, det against LAPACK, Newton against a synthetic
 * holomorphic function, and the whole chain against the published
 * small-grid root (convergence_table.txt:13).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* Julia-style range(a, b, length=n): value i (0-based) of the evenly spaced */
/* grid, computed in extended precision and rounded once, which is what      */
/* Julia's TwicePrecision linspace returns to within one ulp.                */
/* ------------------------------------------------------------------------ */
void or_linrange(double a, double b, long n, double* out)
{
    if (n == 1) { out[0] = a; return; }
    long double la = a, lb = b, d = (long double)(n - 1);
    for (long i = 0; i < n; ++i) {
        long double v = (la * (long double)(n - 1 - i) + lb * (long double)i) / d;
        out[i] = (double)v;
    }
    out[0] = a;
    out[n - 1] = b;
}

/* trapezoid_coef -- BasisFunctions.jl:174-188, OrbitIntegration.jl:63-74 */
void or_trapezoid_coef(const double* x, long n, double* S)
{
    for (long i = 0; i < n; ++i) S[i] = 0.0;
    if (n >= 2) {
        S[0] = (x[1] - x[0]) / 2;
        S[n - 1] = (x[n - 1] - x[n - 2]) / 2;
        for (long i = 1; i < n - 1; ++i) S[i] = (x[i + 1] - x[i - 1]) / 2;
    }
}

/* simpson_coef -- OrbitIntegration.jl:79-95 (trapezoid if n<3 or n even) */
void or_simpson_coef(const double* x, long n, double* S)
{
    if (n < 3 || (n % 2) == 0) { or_trapezoid_coef(x, n, S); return; }
    double dx = x[1] - x[0];
    for (long i = 0; i < n; ++i) S[i] = 0.0;
    S[0] = dx / 3;
    S[n - 1] = dx / 3;
    for (long i = 1; i < n - 1; i += 2) S[i] = 4 * dx / 3;  /* Julia 2:2:n-1 */
    for (long i = 2; i < n - 2; i += 2) S[i] = 2 * dx / 3;  /* Julia 3:2:n-2 */
}

/* ------------------------------------------------------------------------ */
/* Complex helpers (explicit re/im, Julia Base semantics)                    */
/* ------------------------------------------------------------------------ */
typedef struct { double re, im; } cplx;

static cplx c_mk(double r, double i) { cplx z = {r, i}; return z; }
static cplx c_add(cplx a, cplx b) { return c_mk(a.re + b.re, a.im + b.im); }
static cplx c_sub(cplx a, cplx b) { return c_mk(a.re - b.re, a.im - b.im); }
static cplx c_mul(cplx a, cplx b)
{
    return c_mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static cplx c_scale(cplx a, double s) { return c_mk(a.re * s, a.im * s); }

/* Julia inv(::ComplexF64) (Smith's algorithm with range scaling). */
static cplx c_inv(cplx w)
{
    double c = w.re, d = w.im;
    if (isinf(c) || isinf(d)) return c_mk(copysign(0.0, c), -copysign(0.0, d) * 1.0);
    const double half = 0.5, two = 2.0;
    double cd = fmax(fabs(c), fabs(d));
    const double ov = 1.7976931348623157e308, un = 2.2250738585072014e-308;
    const double eps = 2.220446049250313e-16;
    double bs = two / (eps * eps), s = 1.0;
    if (cd >= half * ov) { c = half * c; d = half * d; s = s * half; }
    if (cd <= un * two / eps) { c = bs * c; d = bs * d; s = s * bs; }
    double p, q;
    if (fabs(d) <= fabs(c)) {
        double r = d / c, t = 1.0 / (c + d * r);
        p = t; q = -r * t;
    } else {
        double tmp = c; c = d; d = tmp;
        double r = d / c, t = 1.0 / (c + d * r);
        p = r * t; q = -t;
    }
    return c_mk(p * s, q * s);
}

/* Julia complex division a / b (robust Baudin-Smith); used only in gamma. */
static cplx c_div(cplx a, cplx b)
{
    /* a/b = a * conj(b) / |b|^2 with Smith scaling; accuracy ~1 ulp. */
    double c = b.re, d = b.im;
    if (fabs(d) <= fabs(c)) {
        double r = d / c, t = 1.0 / (c + d * r);
        return c_mk((a.re + a.im * r) * t, (a.im - a.re * r) * t);
    } else {
        double r = c / d, t = 1.0 / (c * r + d);
        return c_mk((a.re * r + a.im) * t, (a.im * r - a.re) * t);
    }
}

static cplx c_exp(cplx z)
{
    double er = exp(z.re);
    return c_mk(er * cos(z.im), er * sin(z.im));
}

static cplx c_log(cplx z) { return c_mk(log(hypot(z.re, z.im)), atan2(z.im, z.re)); }

static cplx c_pow(cplx z, cplx p) { return c_exp(c_mul(p, c_log(z))); }

/* gamma_complex -- BasisFunctions.jl:309-366 (Godfrey's Lanczos, g=607/128,
 * 15 coefficients, reflection for Re z < 0; orig. Matlab/gamma_complex.m). */
void or_gamma_complex(double zr, double zi, double* out_re, double* out_im)
{
    static const double g = 607.0 / 128.0;
    static const double c[15] = {
        0.99999999999999709182,     57.156235665862923517,
        -59.597960355475491248,     14.136097974741747174,
        -0.49191381609762019978,    0.33994649984811888699e-4,
        0.46523628927048575665e-4,  -0.98374475304879564677e-4,
        0.15808870322491248884e-3,  -0.21026444172410488319e-3,
        0.21743961811521264320e-3,  -0.16431810653676389022e-3,
        0.84418223983852743293e-4,  -0.26190838401581408670e-4,
        0.36899182659531622704e-5};
    cplx zz = c_mk(zr, zi), z = zz;
    if (z.re < 0) z = c_mk(-z.re, -z.im);
    z = c_mk(z.re - 1, z.im);
    cplx zh = c_mk(z.re + 0.5, z.im);
    cplx zgh = c_mk(zh.re + g, zh.im);
    cplx zp = c_pow(zgh, c_scale(zh, 0.5));
    cplx ss = c_mk(0.0, 0.0);
    for (int pp = 14; pp >= 1; --pp) {
        /* Julia Real / Complex is a * inv(z) */
        ss = c_add(ss, c_scale(c_inv(c_mk(z.re + pp, z.im)), c[pp]));
    }
    const double sq2pi = 2.5066282746310005024157652848110;
    cplx t1 = c_scale(c_mk(c[0] + ss.re, ss.im), sq2pi);
    cplx t2 = c_mul(zp, c_exp(c_mk(-zgh.re, -zgh.im)));
    cplx f = c_mul(c_mul(t1, t2), zp);
    if ((z.re == 0 && z.im == 0) || (z.re == 1 && z.im == 0)) f = c_mk(1.0, 0.0);
    if (zz.re < 0) {
        /* f = -pi / (zz * f * sin(pi*zz)) */
        cplx pz = c_mk(OR_PI * zz.re, OR_PI * zz.im);
        cplx s = c_mk(sin(pz.re) * cosh(pz.im), cos(pz.re) * sinh(pz.im));
        cplx den = c_mul(c_mul(zz, f), s);
        f = c_scale(c_inv(den), -OR_PI);
    }
    if (round(zz.re) == zz.re && zz.im == 0 && zz.re <= 0) f = c_mk(INFINITY, 0.0);
    *out_re = f.re;
    *out_im = f.im;
}

/* compute_N_kernel -- BasisFunctions.jl:281-304:
 * N[a] = Re[G(z1)G(z2)/(G(z3)G(z4))] * pi * d_alpha[a]. */
void or_N_kernel(int m, const double* alpha, const double* d_alpha, long N, double* out)
{
    for (long i = 0; i < N; ++i) {
        double a = alpha[i];
        cplx z1 = c_mk((m + 0.5) / 2, a / 2), z2 = c_mk((m + 0.5) / 2, -a / 2);
        cplx z3 = c_mk((m + 1.5) / 2, a / 2), z4 = c_mk((m + 1.5) / 2, -a / 2);
        cplx g1, g2, g3, g4;
        or_gamma_complex(z1.re, z1.im, &g1.re, &g1.im);
        or_gamma_complex(z2.re, z2.im, &g2.re, &g2.im);
        or_gamma_complex(z3.re, z3.im, &g3.re, &g3.im);
        or_gamma_complex(z4.re, z4.im, &g4.re, &g4.im);
        cplx ratio = c_div(c_mul(g1, g2), c_mul(g3, g4));
        out[i] = ratio.re * OR_PI * d_alpha[i];
    }
}

/* ------------------------------------------------------------------------ */
/* Toomre-Zang model -- Models.jl:24-169                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    double L0;
    int n_zang, q1, q_zang;
    double sigma_r0, G, DF_const;
} or_model;

/* create_toomre_zang_model -- Models.jl:40-56 */
static or_model model_make(double L0, int n_zang, int q1, double G)
{
    or_model md;
    md.L0 = L0; md.n_zang = n_zang; md.q1 = q1; md.G = G;
    md.sigma_r0 = 1.0 / sqrt((double)q1);
    md.q_zang = q1 - 1;
    md.DF_const = 1.0 / (2.0 * OR_PI * pow(2.0, md.q_zang / 2.0) * sqrt(OR_PI) *
                         tgamma((double)(md.q_zang + 1) / 2) *
                         pow(md.sigma_r0, (double)(md.q_zang + 2)));
    return md;
}

double or_df_const(double L0, int n_zang, int q1, double G)
{
    return model_make(L0, n_zang, q1, G).DF_const;
}

static double V_(double r) { return log(r); }        /* Models.jl:64 */
static double dV_(double r) { return 1.0 / r; }      /* Models.jl:67 */
static double Om_(double r) { return 1.0 / r; }      /* Models.jl:70 */
static double kap_(double r) { return sqrt(2.0) / r; } /* Models.jl:73 */

/* taper -- Models.jl:91-102 */
static double taper_(const or_model* md, double L)
{
    if (md->n_zang == 0) return 1.0;
    double La = fabs(L);
    if (La < 1e-12) return 0.0;
    return 1.0 / (1.0 + pow(md->L0 / La, (double)md->n_zang));
}

/* distribution_function -- Models.jl:137-150 */
static double DF_(const or_model* md, double E, double L)
{
    double La = fabs(L);
    if (La < 1e-12) return 0.0;
    double s2 = md->sigma_r0 * md->sigma_r0;
    double tp = taper_(md, L);
    double r = md->DF_const * exp(-E / s2) * pow(La, (double)md->q_zang) * tp;
    return isfinite(r) ? r : 0.0;
}

/* DF_E -- Models.jl:153-155 */
static double DFE_(const or_model* md, double E, double L)
{
    return -DF_(md, E, L) / (md->sigma_r0 * md->sigma_r0);
}

/* DF_L -- Models.jl:158-169 */
static double DFL_(const or_model* md, double E, double L)
{
    double La = fabs(L);
    if (La < 1e-12) return 0.0;
    double f0 = DF_(md, E, L), tp = taper_(md, L);
    return f0 * ((double)(md->q_zang + md->n_zang) - md->n_zang * tp) / La;
}

void or_df_eval(double L0, int n_zang, int q1, double G, double E, double L,
                double* F, double* FE, double* FL)
{
    or_model md = model_make(L0, n_zang, q1, G);
    *F = DF_(&md, E, L);
    *FE = DFE_(&md, E, L);
    *FL = DFL_(&md, E, L);
}

/* ------------------------------------------------------------------------ */
/* compute_grids -- OrbitIntegration.jl:102-170                              */
/* Outputs are [NR*Ne] with iR fastest (Julia column-major [NR, Ne]).         */
/* ------------------------------------------------------------------------ */
int or_compute_grids(int NR, int Ne, double Rc_min, double Rc_max, double v_min,
                     double* Rc, double* e, double* v, double* R1, double* R2,
                     double* E, double* L, double* L2, double* SGNL,
                     double* S_RC, double* S_e)
{
    if (NR < 1 || Ne < 1) return -1;
    double* lg = (double*)malloc(sizeof(double) * NR);
    double* Rc1 = (double*)malloc(sizeof(double) * NR);
    double* uRc = (double*)malloc(sizeof(double) * NR);
    double* tr = (double*)malloc(sizeof(double) * NR);
    double* va = (double*)malloc(sizeof(double) * Ne);
    double* se = (double*)malloc(sizeof(double) * Ne);
    or_linrange(log10(Rc_min), log10(Rc_max), NR, lg);
    for (int i = 0; i < NR; ++i) { Rc1[i] = pow(10.0, lg[i]); uRc[i] = log(Rc1[i]); }
    or_trapezoid_coef(uRc, NR, tr);
    for (int i = 0; i < NR; ++i) S_RC[i] = Rc1[i] * tr[i];
    or_linrange(v_min, 1.0, Ne, va);
    or_simpson_coef(va, Ne, se);
    for (int ie = 0; ie < Ne; ++ie)
        for (int iR = 0; iR < NR; ++iR) {
            long k = iR + (long)NR * ie;
            S_e[k] = se[ie];
            v[k] = va[ie];
            Rc[k] = Rc1[iR];
        }
    long NPh = (long)NR * Ne;
    for (long k = 0; k < NPh; ++k) {
        if (fabs(v[k]) < 1e-14) v[k] = 1e-12;
        SGNL[k] = (v[k] > 0) ? 1.0 : ((v[k] < 0) ? -1.0 : 0.0);
        e[k] = 1.0 - fabs(v[k]);
        R1[k] = Rc[k] * (1.0 - e[k]);
        R2[k] = Rc[k] * (1.0 + e[k]);
    }
    for (long k = 0; k < NPh; ++k) {
        double r1 = R1[k], r2 = R2[k];
        double V1 = V_(r1), V2 = V_(r2);
        double Ev;
        if (fabs(r2 * r2 - r1 * r1) > 1e-14)
            Ev = (V2 * (r2 * r2) - V1 * (r1 * r1)) / (r2 * r2 - r1 * r1);
        else {
            double r = Rc[k];
            Ev = r * dV_(r) / 2 + V_(r);
        }
        if (!isfinite(Ev)) Ev = V_(R2[k]);
        E[k] = Ev;
    }
    for (long k = 0; k < NPh; ++k) {
        double r2 = R2[k];
        L2[k] = 2.0 * (E[k] - V_(r2)) * (r2 * r2);
        if (L2[k] < 1e-14) L2[k] = 0.0;
        L[k] = sqrt(L2[k]) * SGNL[k];
    }
    free(lg); free(Rc1); free(uRc); free(tr); free(va); free(se);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* compute_orbits -- OrbitIntegration.jl:181-336, compute_jacobian!         */
/* :341-382.  w1/ra/pha are [nwa, NR, Ne] (sample fastest).                  */
/* ------------------------------------------------------------------------ */
int or_compute_orbits(int NR, int Ne, int nw, int nwa,
                      double L0, int n_zang, int q1, double G,
                      const double* Rc, const double* e, const double* R1,
                      const double* R2, const double* E, const double* L,
                      const double* L2, const double* SGNL,
                      double* Om1, double* Om2, double* w1, double* ra, double* pha,
                      double* jac, double* F0, double* FE, double* FL)
{
    if (nw < 3 || nwa < 2 || (nw - 1) % (nwa - 1) != 0) return -1;
    or_model md = model_make(L0, n_zang, q1, G);
    const double eps_w = 1e-8;
    double* w = (double*)malloc(sizeof(double) * nw);
    double* cosw = (double*)malloc(sizeof(double) * nw);
    double* sinw = (double*)malloc(sizeof(double) * nw);
    or_linrange(eps_w, OR_PI - eps_w, nw, w);
    double dw = w[1] - w[0];
    for (int i = 0; i < nw; ++i) { cosw[i] = cos(w[i]); sinw[i] = sin(w[i]); }
    int nwai = (nw - 1) / (nwa - 1);
    long NPh = (long)NR * Ne;
    for (long x = 0; x < NPh * nwa; ++x) { w1[x] = NAN; ra[x] = NAN; pha[x] = NAN; }

#pragma omp parallel
    {
        double* xs = (double*)malloc(sizeof(double) * nw);
        double* rvr = (double*)malloc(sizeof(double) * nw);
        double* svr = (double*)malloc(sizeof(double) * nw);
        double* dt = (double*)malloc(sizeof(double) * nw);
        double* t = (double*)malloc(sizeof(double) * nw);
#pragma omp for schedule(dynamic, 1)
        for (long k = 0; k < NPh; ++k) {
            int iR = (int)(k % NR);
            double R_i = Rc[iR];
            double kap = kap_(R_i), Om = Om_(R_i);
            double E_i = E[k], L_j = L[k], L2_j = L2[k], ecc = e[k];
            double r1 = R1[k], r2 = R2[k];
            double* W1 = w1 + k * nwa;
            double* RA = ra + k * nwa;
            double* PH = pha + k * nwa;
            if (fabs(ecc) < 1e-10) {           /* circular :221-229 */
                Om1[k] = kap;
                Om2[k] = Om * SGNL[k];
                for (int a = 0; a < nwa; ++a) {
                    double wa = w[a * nwai];
                    W1[a] = wa; RA[a] = R_i; PH[a] = wa * Om2[k] / Om1[k];
                }
                continue;
            }
            if (ecc < 0.01) {                  /* epicyclic :231-241 */
                double a_j = (r2 - r1) / 2;
                Om1[k] = kap;
                Om2[k] = Om * SGNL[k];
                double coef = 2 * Om / kap * a_j / R_i;
                for (int a = 0; a < nwa; ++a) {
                    int i = a * nwai;
                    double wa = w[i];
                    W1[a] = wa;
                    RA[a] = R_i - a_j * cosw[i];
                    PH[a] = wa * Om2[k] / Om1[k] + coef * sinw[i] * SGNL[k];
                }
                continue;
            }
            /* general :243-302 */
            double rs = (r1 + r2) / 2, drs = (r2 - r1) / 2;
            for (int i = 0; i < nw; ++i) xs[i] = rs - drs * cosw[i];
            if (L2_j > 1e-12) {
                for (int i = 0; i < nw; ++i) {
                    double x = xs[i];
                    double vr_sq = 2.0 * (E_i - V_(x)) * (x * x) - L2_j;
                    rvr[i] = sqrt(fmax(0.0, vr_sq));
                }
                for (int i = 0; i < nw; ++i) svr[i] = sinw[i] * xs[i] / rvr[i];
                svr[0] = 2.0 * svr[1] - svr[2];
                svr[nw - 1] = 2.0 * svr[nw - 2] - svr[nw - 3];
            } else {
                for (int i = 0; i < nw; ++i) {
                    double x = xs[i];
                    double vr_sq = 2.0 * (E_i - V_(x));
                    rvr[i] = sqrt(fmax(0.0, vr_sq));
                }
                for (int i = 0; i < nw; ++i) svr[i] = sinw[i] / rvr[i];
                svr[nw - 1] = 2.0 * svr[nw - 2] - svr[nw - 3];
            }
            /* radial period :277-281 */
            for (int i = 0; i < nw; ++i) dt[i] = drs * dw * svr[i];
            t[0] = 0.0;
            for (int i = 1; i < nw; ++i) t[i] = t[i - 1] + (dt[i - 1] + dt[i]) / 2;
            Om1[k] = OR_PI / t[nw - 1];
            for (int a = 0; a < nwa; ++a) {
                int i = a * nwai;
                W1[a] = t[i] * Om1[k];
                RA[a] = xs[i];
            }
            /* azimuthal phase :287-299 */
            if (fabs(1.0 - ecc) > 1e-10 && L2_j > 1e-12) {
                for (int i = 0; i < nw; ++i) svr[i] = sinw[i] / rvr[i];
                svr[0] = 2.0 * svr[1] - svr[2];
                svr[nw - 1] = 2.0 * svr[nw - 2] - svr[nw - 3];
                for (int i = 0; i < nw; ++i) dt[i] = drs * dw * svr[i] / xs[i];
                t[0] = 0.0;
                for (int i = 1; i < nw; ++i) t[i] = t[i - 1] + (dt[i - 1] + dt[i]) / 2;
                for (int a = 0; a < nwa; ++a) PH[a] = L_j * t[a * nwai];
            } else {
                for (int a = 0; a < nwa; ++a) PH[a] = OR_PI / 2 * 1.0 * SGNL[k];
                PH[0] = 0.0;
            }
            Om2[k] = Om1[k] * PH[nwa - 1] / OR_PI;
        }
        free(xs); free(rvr); free(svr); free(dt); free(t);
    }

    /* compute_jacobian! :341-382 */
    for (long k = 0; k < NPh; ++k) {
        double r1 = R1[k], r2 = R2[k], rc = Rc[k], Ev = E[k], Lv = L[k], o1 = Om1[k];
        jac[k] = 0.0;
        if (fabs(r2 - r1) > 1e-12 && fabs(Lv) > 1e-12) {
            double V1 = V_(r1), V2 = V_(r2), dV1 = dV_(r1), dV2 = dV_(r2);
            double t1 = 2 * (Ev - V1) * r1 - dV1 * (r1 * r1);
            double t2 = 2 * (Ev - V2) * r2 - dV2 * (r2 * r2);
            double den = r2 * r2 - r1 * r1;
            double dj = fabs(t1 * t2 / den);
            jac[k] = 2 * dj * rc / o1 / fabs(Lv);
        }
        if (fabs(Lv) <= 1e-12) {
            double DelE = 2.0 * (Ev - V_(r1));
            jac[k] = sqrt(fmax(DelE, 0.0)) * dV_(r2) * 2 * rc / o1;
            if (!isfinite(jac[k])) jac[k] = 0.0;
        }
    }
    /* DF on the grid :317-333 */
    for (long k = 0; k < NPh; ++k) {
        F0[k] = DF_(&md, E[k], L[k]);
        FE[k] = DFE_(&md, E[k], L[k]);
        FL[k] = DFL_(&md, E[k], L[k]);
        if (!isfinite(F0[k])) F0[k] = 0.0;
        if (!isfinite(FE[k])) FE[k] = 0.0;
        if (!isfinite(FL[k])) FL[k] = 0.0;
    }
    free(w); free(cosw); free(sinw);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* compute_W_l -- BasisFunctions.jl:199-266.  W is [NPh, N, n_l] complex,    */
/* interleaved re/im, Julia column-major.                                    */
/* The reference's per-(l,alpha) inner loop is evaluated here with the      */
/* l-only factor cos(l*w1+m*phi) and the alpha-only factor exp(-i alpha      */
/* log r) hoisted (same values, same operation order per term).              */
/* ------------------------------------------------------------------------ */
int or_compute_W(int NR, int Ne, int nwa, int m,
                 const double* w1, const double* ra, const double* pha,
                 const double* Om1, const double* Om2,
                 const double* alpha, long N, const int* l_arr, long n_l, double* W)
{
    long NPh = (long)NR * Ne;
    memset(W, 0, sizeof(double) * 2 * NPh * N * n_l);
#pragma omp parallel
    {
        double* Sw1 = (double*)malloc(sizeof(double) * nwa);
        double* phi = (double*)malloc(sizeof(double) * nwa);
        double* logr = (double*)malloc(sizeof(double) * nwa);
        double* sqr = (double*)malloc(sizeof(double) * nwa);
        double* cl = (double*)malloc(sizeof(double) * nwa);
        double* er = (double*)malloc(sizeof(double) * nwa * N);
        double* ei = (double*)malloc(sizeof(double) * nwa * N);
#pragma omp for schedule(dynamic, 4)
        for (long k = 0; k < NPh; ++k) {
            const double* W1 = w1 + k * nwa;
            const double* RA = ra + k * nwa;
            const double* TH = pha + k * nwa;
            int ok = 1;
            for (int t = 0; t < nwa; ++t)
                if (!isfinite(RA[t]) || !isfinite(TH[t]) || !isfinite(W1[t])) { ok = 0; break; }
            if (!ok) continue;                                   /* :224-226 */
            or_trapezoid_coef(W1, nwa, Sw1);                      /* :229 */
            double r21 = Om2[k] / Om1[k];                         /* :232 */
            for (int t = 0; t < nwa; ++t) {
                phi[t] = r21 * W1[t] - TH[t];                     /* :235 */
                logr[t] = log(RA[t]);
                sqr[t] = sqrt(RA[t]);
            }
            for (long a = 0; a < N; ++a)
                for (int t = 0; t < nwa; ++t) {
                    double y = -alpha[a] * logr[t];
                    er[a * nwa + t] = cos(y);
                    ei[a * nwa + t] = sin(y);
                }
            for (long il = 0; il < n_l; ++il) {
                double l = (double)l_arr[il];
                for (int t = 0; t < nwa; ++t) cl[t] = cos(l * W1[t] + m * phi[t]);  /* :240,245 */
                for (long a = 0; a < N; ++a) {
                    double sr = 0.0, si = 0.0;
                    const double* ER = er + a * nwa;
                    const double* EI = ei + a * nwa;
                    for (int t = 0; t < nwa; ++t) {
                        double ir = cl[t] * ER[t] / sqr[t];
                        double ii = cl[t] * EI[t] / sqr[t];
                        if (!isfinite(ir) || !isfinite(ii)) { ir = 0.0; ii = 0.0; }  /* :248 */
                        sr += Sw1[t] * ir;
                        si += Sw1[t] * ii;
                    }
                    long idx = k + NPh * (a + N * il);
                    W[2 * idx] = sr / OR_PI;                       /* :250 */
                    W[2 * idx + 1] = si / OR_PI;
                }
            }
        }
        free(Sw1); free(phi); free(logr); free(sqr); free(cl); free(er); free(ei);
    }
    return 0;
}

/* omega-independent arrays -- BasisFunctions.jl:127-153 */
void or_precompute_arrays(int NR, int Ne, int m, const int* l_arr, long n_l,
                          const double* S_RC, const double* S_e, const double* jac,
                          const double* Om1, const double* Om2,
                          const double* FE, const double* FL,
                          double* DJ, double* F0l, double* Ores)
{
    long NPh = (long)NR * Ne;
    for (long k = 0; k < NPh; ++k) DJ[k] = S_RC[k % NR] * S_e[k] * jac[k];
    for (long il = 0; il < n_l; ++il) {
        double l = (double)l_arr[il];
        for (long k = 0; k < NPh; ++k) {
            double f = l * Om1[k] + m * Om2[k];
            F0l[k + NPh * il] = f * FE[k] + m * FL[k];
            Ores[k + NPh * il] = f;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* build_M_matrix_cpu -- MatrixBuilder.jl:27-69.  M is complex [N, N]        */
/* column-major, interleaved.                                                */
/* ------------------------------------------------------------------------ */
int or_build_M(long NPh, long N, long n_l, const double* W, const double* DJ,
               const double* F0l, const double* Ores, const double* Nk, double G,
               double om_re, double om_im, double* M)
{
    memset(M, 0, sizeof(double) * 2 * N * N);
    double* wr = (double*)malloc(sizeof(double) * NPh);
    double* wi = (double*)malloc(sizeof(double) * NPh);
    double* Pr = (double*)malloc(sizeof(double) * NPh * N);
    double* Pi = (double*)malloc(sizeof(double) * NPh * N);
    for (long il = 0; il < n_l; ++il) {
        const double* Wl = W + 2 * NPh * N * il;
        for (long k = 0; k < NPh; ++k) {
            cplx d = c_mk(om_re - Ores[k + NPh * il], om_im - 0.0);    /* :44 */
            if (hypot(d.re, d.im) < 1e-12) d = c_mk(1e-12, 1e-12);     /* :47-48 */
            double x = DJ[k] * F0l[k + NPh * il];                        /* :51 */
            cplx iv = c_inv(d);
            wr[k] = x * iv.re;
            wi[k] = x * iv.im;
        }
        /* P[:, beta] = W[:, beta] .* weight */
#pragma omp parallel for schedule(static)
        for (long b = 0; b < N; ++b)
            for (long k = 0; k < NPh; ++k) {
                cplx wv = c_mk(Wl[2 * (k + NPh * b)], Wl[2 * (k + NPh * b) + 1]);
                cplx p = c_mul(wv, c_mk(wr[k], wi[k]));
                Pr[k + NPh * b] = p.re;
                Pi[k + NPh * b] = p.im;
            }
        /* M[beta, alpha] += sum(P[:, beta] .* conj(W[:, alpha])) :55-59 */
#pragma omp parallel for schedule(static)
        for (long a = 0; a < N; ++a)
            for (long b = 0; b < N; ++b) {
                double sr = 0.0, si = 0.0;
                const double* PR = Pr + NPh * b;
                const double* PI = Pi + NPh * b;
                const double* WA = Wl + 2 * NPh * a;
                for (long k = 0; k < NPh; ++k) {
                    double cr = WA[2 * k], ci = -WA[2 * k + 1];
                    sr += PR[k] * cr - PI[k] * ci;
                    si += PR[k] * ci + PI[k] * cr;
                }
                M[2 * (b + N * a)] += sr;
                M[2 * (b + N * a) + 1] += si;
            }
    }
    for (long a = 0; a < N; ++a)                                      /* :63-66 */
        for (long b = 0; b < N; ++b) {
            long x = b + N * a;
            M[2 * x] = M[2 * x] * G * Nk[a];
            M[2 * x + 1] = M[2 * x + 1] * G * Nk[a];
        }
    free(wr); free(wi); free(Pr); free(Pi);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* det(A) by LU with partial pivoting, LAPACK zgetf2 semantics: pivot =      */
/* first row maximising |re|+|im| (izamax), reciprocal-scaled column,        */
/* rank-1 update; det = prod(diag U) * (-1)^swaps in Julia's order           */
/* (LinearAlgebra det(::LU)), and exactly 0 when a pivot is exactly zero     */
/* (zgetf2's info > 0, for which det(::LU) returns zero(T)).  A NaN on the   */
/* diagonal is the pivot (izamax seeds its search with the first entry and   */
/* takes a later one only if strictly larger), so it reaches det.            */
/* A is complex [N, N] column-major, in place.                               */
/* ------------------------------------------------------------------------ */
int or_det_inplace(long N, double* A, double* det_re, double* det_im)
{
    int swaps = 0, singular = 0;
#define AR(i, j) A[2 * ((i) + N * (j))]
#define AI(i, j) A[2 * ((i) + N * (j)) + 1]
    for (long j = 0; j < N; ++j) {
        long p = j;
        double best = fabs(AR(j, j)) + fabs(AI(j, j));
        for (long i = j + 1; i < N; ++i) {
            double v = fabs(AR(i, j)) + fabs(AI(i, j));
            if (v > best) { best = v; p = i; }
        }
        if (AR(p, j) == 0.0 && AI(p, j) == 0.0) { singular = 1; continue; }
        if (p != j) {
            ++swaps;
            for (long c = 0; c < N; ++c) {
                double tr = AR(j, c), ti = AI(j, c);
                AR(j, c) = AR(p, c); AI(j, c) = AI(p, c);
                AR(p, c) = tr; AI(p, c) = ti;
            }
        }
        cplx rp = c_inv(c_mk(AR(j, j), AI(j, j)));
        for (long i = j + 1; i < N; ++i) {
            cplx x = c_mul(c_mk(AR(i, j), AI(i, j)), rp);
            AR(i, j) = x.re; AI(i, j) = x.im;
        }
        for (long c = j + 1; c < N; ++c) {
            cplx u = c_mk(AR(j, c), AI(j, c));
            if (u.re == 0.0 && u.im == 0.0) continue;
            for (long i = j + 1; i < N; ++i) {
                cplx lu = c_mul(c_mk(AR(i, j), AI(i, j)), u);
                AR(i, c) -= lu.re; AI(i, c) -= lu.im;
            }
        }
    }
    cplx d = c_mk(1.0, 0.0);
    for (long i = 0; i < N; ++i) d = c_mul(d, c_mk(AR(i, i), AI(i, i)));
    if (swaps & 1) d = c_mk(-d.re, -d.im);
    if (singular) d = c_mk(0.0, 0.0);
#undef AR
#undef AI
    *det_re = d.re;
    *det_im = d.im;
    return 0;
}

/* compute_determinant_cpu -- MatrixBuilder.jl:76-87: det(I - M) */
int or_compute_det(long NPh, long N, long n_l, const double* W, const double* DJ,
                   const double* F0l, const double* Ores, const double* Nk, double G,
                   double om_re, double om_im, double* det_re, double* det_im)
{
    double* M = (double*)malloc(sizeof(double) * 2 * N * N);
    or_build_M(NPh, N, n_l, W, DJ, F0l, Ores, Nk, G, om_re, om_im, M);
    for (long x = 0; x < N * N; ++x) { M[2 * x] = -M[2 * x]; M[2 * x + 1] = -M[2 * x + 1]; }
    for (long i = 0; i < N; ++i) M[2 * (i + N * i)] += 1.0;
    or_det_inplace(N, M, det_re, det_im);
    free(M);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* newton_search -- NewtonSolver.jl:36-172.  det is supplied by a callback  */
/* so the driver can be pinned on a synthetic holomorphic function.         */
/* history is [max_iter, 3] column-major (Omega_p, gamma, |det|).           */
/* ------------------------------------------------------------------------ */
typedef void (*or_det_fn)(double om_re, double om_im, double* det_re, double* det_im, void* user);

int or_newton(or_det_fn f, void* user, int m, int max_iter, double tol, double delta,
              int use_line_search, int max_line_search, double Op0, double g0,
              double* out /* Op, gamma, det_re, det_im */, int* iterations, int* converged,
              double* history)
{
    double Op = Op0, gv = g0;
    int conv = 0, iter_done = 0;
    for (int iter = 1; iter <= max_iter; ++iter) {
        iter_done = iter;
        double dr, di;
        f(m * Op, gv, &dr, &di, user);                           /* :74 */
        double Fr = dr, Fi = di;
        double res = hypot(Fr, Fi);                               /* :78 */
        history[(iter - 1)] = Op;
        history[(iter - 1) + max_iter] = gv;
        history[(iter - 1) + 2 * max_iter] = hypot(dr, di);
        if (res < tol) { conv = 1; break; }                       /* :89 */
        double aOr, aOi, agr, agi;
        f(m * (Op + delta), gv, &aOr, &aOi, user);                /* :97-101 */
        f(m * Op, gv + delta, &agr, &agi, user);                  /* :98-102 */
        double J11 = (aOr - dr) / delta, J12 = (agr - dr) / delta;
        double J21 = (aOi - di) / delta, J22 = (agi - di) / delta;
        /* delta_x = -J \ F  (2x2 solve, LAPACK-style partial pivoting) */
        double x1, x2;
        {
            double a11 = J11, a12 = J12, a21 = J21, a22 = J22, b1 = Fr, b2 = Fi;
            if (fabs(a21) > fabs(a11)) {
                double t;
                t = a11; a11 = a21; a21 = t;
                t = a12; a12 = a22; a22 = t;
                t = b1; b1 = b2; b2 = t;
            }
            double l21 = a21 * (1.0 / a11);   /* dgetf2 reciprocal scaling */
            double u22 = a22 - l21 * a12;
            double y2 = b2 - l21 * b1;
            x2 = y2 / u22;
            x1 = (b1 - a12 * x2) / a11;
            x1 = -x1; x2 = -x2;
        }
        if (use_line_search) {                                    /* :115-147 */
            double al = 1.0;
            for (int ls = 1; ls <= max_line_search; ++ls) {
                double Opn = Op + al * x1, gn = gv + al * x2;
                if (gn < 1e-4 || Opn < 0.0 || Opn > 1.0) { al *= 0.5; continue; }
                double nr, ni;
                f(m * Opn, gn, &nr, &ni, user);
                if (hypot(nr, ni) < res) { Op = Opn; gv = gn; break; }
                al *= 0.5;
                if (ls == max_line_search) { Op += x1; gv += x2; }
            }
        } else {
            Op += x1;                                             /* :142-143 */
            gv += x2;
        }
        /* :155-156, Julia max/clamp (NaN propagates) */
        if (!isnan(gv) && gv < 1e-4) gv = 1e-4;
        if (Op > 1.0) Op = 1.0; else if (Op < 0.0) Op = 0.0;
    }
    double fr, fi;
    f(m * Op, gv, &fr, &fi, user);                                /* :159-161 */
    out[0] = Op; out[1] = gv; out[2] = fr; out[3] = fi;
    *iterations = iter_done;
    *converged = conv;
    return 0;
}

/* Precomputed-data det callback for or_newton / or_grid_scan. */
typedef struct {
    long NPh, N, n_l;
    const double *W, *DJ, *F0l, *Ores, *Nk;
    double G;
} or_problem;

void or_problem_det(double om_re, double om_im, double* dr, double* di, void* user)
{
    const or_problem* p = (const or_problem*)user;
    or_compute_det(p->NPh, p->N, p->n_l, p->W, p->DJ, p->F0l, p->Ores, p->Nk, p->G,
                   om_re, om_im, dr, di);
}

int or_newton_problem(long NPh, long N, long n_l, const double* W, const double* DJ,
                      const double* F0l, const double* Ores, const double* Nk, double G,
                      int m, int max_iter, double tol, double delta, int use_line_search,
                      int max_line_search, double Op0, double g0, double* out,
                      int* iterations, int* converged, double* history)
{
    or_problem p = {NPh, N, n_l, W, DJ, F0l, Ores, Nk, G};
    return or_newton(or_problem_det, &p, m, max_iter, tol, delta, use_line_search,
                     max_line_search, Op0, g0, out, iterations, converged, history);
}

/* grid_scan + find_minimum -- GridScan.jl:25-82, :202-246.  det_grid is    */
/* complex [n_gamma, n_Op] column-major.                                     */
int or_grid_scan(long NPh, long N, long n_l, const double* W, const double* DJ,
                 const double* F0l, const double* Ores, const double* Nk, double G, int m,
                 int n_Op, int n_gamma, double Op_min, double Op_max, double g_min,
                 double g_max, double* det_grid, double* Op_init, double* g_init)
{
    double* Opg = (double*)malloc(sizeof(double) * n_Op);
    double* gg = (double*)malloc(sizeof(double) * n_gamma);
    or_linrange(Op_min, Op_max, n_Op, Opg);
    or_linrange(g_min, g_max, n_gamma, gg);
    for (int j = 0; j < n_Op; ++j)
        for (int i = 0; i < n_gamma; ++i) {
            double dr, di;
            or_compute_det(NPh, N, n_l, W, DJ, F0l, Ores, Nk, G, m * Opg[j], gg[i], &dr, &di);
            det_grid[2 * (i + (long)n_gamma * j)] = dr;
            det_grid[2 * (i + (long)n_gamma * j) + 1] = di;
        }
    long best = 0;
    double bv = INFINITY;
    for (long x = 0; x < (long)n_Op * n_gamma; ++x) {
        double v = hypot(det_grid[2 * x], det_grid[2 * x + 1]);
        if (v < bv) { bv = v; best = x; }
    }
    *Op_init = Opg[best / n_gamma];
    *g_init = gg[best % n_gamma];
    free(Opg); free(gg);
    return 0;
}
