/*
 * klsp.h -- C ABI of the MI355X-native Kalnajs log-spiral hot path.
 *
 * One shared library (libklsp.so) holds hand-written CDNA4 (gfx950) HIP
 * kernels for the three stages of the reference's det(I - M(omega)) path and
 * the host drivers above them.  Plain pointers and sizes only; no C++ or torch
 * types cross this boundary.  Every function returns an int status (0 = OK);
 * klsp_last_error() gives the message of the last failure on a context.
 *
 * Layouts are the reference's Julia column-major layouts, complex numbers are
 * interleaved (re, im) doubles:
 *   orbit arrays w1/ra/pha  [nwa, NR, Ne]   sample fastest, then iR, then ie
 *   per-orbit arrays        [NR, Ne]        k = iR + NR*ie  ("NPh" = NR*Ne)
 *   W_l_mat                 [NPh, N, n_l]   complex
 *   F0l_all, Omega_res      [NPh, n_l]
 *   M                       [N, N]          complex
 *   det_grid                [n_gamma, n_Op] complex
 *   Newton history          [iterations, 3] (Omega_p, gamma, |det|)
 *
 * Ownership: the caller owns every host buffer; the library never keeps a
 * host pointer after a call returns.  A context owns its device memory until
 * klsp_destroy().  Calls on one context are serialised on one HIP stream and
 * are blocking; contexts are independent.  The library has no CPU fallback:
 * every compute entry point runs on the GPU or fails with KLSP_ERR_NO_DEVICE.
 */
#ifndef KLSP_H
#define KLSP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KLSP_ABI_VERSION 1

/* Status codes. */
enum klsp_status {
    KLSP_OK = 0,
    KLSP_ERR_INVALID_ARG = 1,  /* bad size, null pointer, inconsistent dims */
    KLSP_ERR_OOM = 2,          /* device or host allocation failed          */
    KLSP_ERR_HIP = 3,          /* a HIP runtime call failed                 */
    KLSP_ERR_NOT_READY = 4,    /* e.g. det before the problem/W is on device */
    KLSP_ERR_NO_DEVICE = 5,    /* no gfx950 device visible                  */
    KLSP_ERR_HOST = 6          /* another host-side failure (e.g. a worker
                                  thread could not start)                  */
};

/* Arithmetic of an M(omega) build.  The determinant (LU) is FP64 always. */
enum klsp_precision {
    KLSP_PREC_FP64 = 0,   /* W complex128, FP64 MFMA contraction            */
    KLSP_PREC_FP32 = 1,   /* W complex64, FP32 MFMA with FP64 K-tile sums   */
    KLSP_PREC_MIXED = 2,  /* klsp_newton only: BF16x3 Taylor-model passes
                             until the step is below 1e-6, then an FP64
                             anchored finish (see klsp_newton)             */
    KLSP_PREC_BF16X3 = 3, /* W as bf16 hi + lo, BF16 MFMA with three-term
                             split products, FP32 accumulation, FP64 gang
                             sums (~1e-5 relative)                         */
    KLSP_PREC_MIXED_FP32 = 4 /* klsp_newton only: as MIXED with FP32 passes */
};

typedef struct klsp_ctx klsp_ctx;

/* ------------------------------------------------------------------ */
/* Library and context                                                 */
/* ------------------------------------------------------------------ */

/* Version string of the library ("klsp <semver> gfx950"). */
const char* klsp_version(void);

/* Number of visible HIP devices (0 if none); never fails on a CPU host. */
int klsp_device_count(int* n_out);

/* Name ("<marketing name> (<gfx ISA>)", or the ISA name alone where the
 * driver reports no marketing name; truncated to name_cap bytes with the
 * terminating NUL), total device memory in bytes and compute-unit count of
 * HIP device `device`; each output may be NULL.  Replaces
 * GPUBackend.gpu_device_name (GPUBackend.jl:246) and the runner's
 * gpu_device_name / gpu_total_memory (run_kalnajs_gpu.jl:62-79). */
int klsp_device_info(int device, char* name_out, int name_cap, int64_t* total_mem_out,
                     int* cu_count_out);

/* Create a context on HIP device `device`.  Replaces the device side of
 * GPUBackend.gpu_device! + to_device (GPUBackend.jl:147, BasisFunctions.jl:397). */
int klsp_create(int device, klsp_ctx** ctx_out);

/* Free every device buffer of the context (the Julia finalizer calls this). */
int klsp_destroy(klsp_ctx* ctx);

/* Message of the last failed call on `ctx` (static string if ctx is NULL). */
const char* klsp_last_error(const klsp_ctx* ctx);

/* ------------------------------------------------------------------ */
/* Host precompute (CPU, FP64; no device needed)                       */
/* ------------------------------------------------------------------ */

/* Model and grid parameters: Configuration.jl:31-67 (PhysicsConfig,
 * ModelConfig, GridConfig); same keys as the TOML config files. */
typedef struct klsp_params {
    int m;            /* azimuthal mode number                          */
    double G;         /* gravitational constant                         */
    double L0;        /* taper angular-momentum scale                   */
    int n_zang;       /* Zang taper exponent                            */
    int q1;           /* velocity-dispersion index (q = q1 - 1)         */
    int NR, Ne;       /* guiding-centre and circulation grid sizes      */
    double Rc_min, Rc_max, v_min;
    int N_alpha;      /* log-spiral wavenumbers alpha in [-amax, amax]   */
    double alpha_max;
    int l_min, l_max; /* radial harmonics                               */
    int nw, nwa;      /* orbit integration points / stored samples      */
} klsp_params;

/* trapezoid_coef, BasisFunctions.jl:174 (= OrbitIntegration.jl:63). */
int klsp_trapezoid_coef(const double* x, int64_t n, double* S_out);

/* simpson_coef, OrbitIntegration.jl:79 (falls back to trapezoid if n even). */
int klsp_simpson_coef(const double* x, int64_t n, double* S_out);

/* Evenly spaced grid, Julia range(a, b, length=n). */
int klsp_linrange(double a, double b, int64_t n, double* out);

/* gamma_complex, BasisFunctions.jl:309 (Godfrey-Lanczos). out = {re, im}. */
int klsp_gamma_complex(double re, double im, double* out);

/* compute_N_kernel, BasisFunctions.jl:281: N[a] = Re ratio * pi * d_alpha[a]. */
int klsp_compute_N_kernel(int m, const double* alpha, const double* d_alpha,
                          int64_t N, double* N_out);

/* compute_grids, OrbitIntegration.jl:102.  Per-orbit outputs are [NR*Ne]
 * (k = iR + NR*ie); S_RC is [NR]. */
int klsp_compute_grids(const klsp_params* p, double* Rc, double* e, double* v,
                       double* R1, double* R2, double* E, double* L, double* L2,
                       double* SGNL, double* S_RC, double* S_e);

/* compute_orbits + compute_jacobian! + DF fill, OrbitIntegration.jl:181,
 * :341, :317.  Inputs are the compute_grids outputs; w1/ra/pha are
 * [nwa, NR, Ne].  Runs on `n_threads` host threads (<= 0: all cores). */
int klsp_compute_orbits(const klsp_params* p, const double* Rc, const double* e,
                        const double* R1, const double* R2, const double* E,
                        const double* L, const double* L2, const double* SGNL,
                        double* Omega_1, double* Omega_2, double* w1, double* ra,
                        double* pha, double* jacobian, double* F0, double* FE,
                        double* FL, int n_threads);

/* The Jacobian |d(E,L)/d(Rc,v)| / Omega_1 and the DF on the grid, the tail of
 * compute_orbits (compute_jacobian!, OrbitIntegration.jl:341-382; DF fill
 * :317-333), for orbits integrated on the device (klsp_compute_orbits_W). */
int klsp_compute_jacobian_df(const klsp_params* p, const double* Rc, const double* R1,
                             const double* R2, const double* E, const double* L,
                             const double* Omega_1, double* jacobian, double* F0, double* FE,
                             double* FL);

/* The radial-phase samples of compute_orbits (OrbitIntegration.jl:190-193):
 * w[nw] = range(1e-8, pi - 1e-8, length = nw), cw = cos(w), sw = sin(w). */
int klsp_orbit_phase_grid(const klsp_params* p, double* w, double* cw, double* sw);

/* omega-independent arrays of precompute_all, BasisFunctions.jl:127-153:
 * DJ[NPh], F0l[NPh, n_l], Omega_res[NPh, n_l]. */
int klsp_precompute_arrays(const klsp_params* p, const double* S_RC, const double* S_e,
                           const double* jacobian, const double* Omega_1,
                           const double* Omega_2, const double* FE, const double* FL,
                           double* DJ, double* F0l, double* Omega_res);

/* ------------------------------------------------------------------ */
/* Device residency (PrecomputedData -> DevicePrecomputed)             */
/* ------------------------------------------------------------------ */

/* omega-independent arrays of PrecomputedData (BasisFunctions.jl:30-60):
 * sets the problem dimensions of the context and copies the small arrays to
 * the device (the reference keeps them on the host, BasisFunctions.jl:390). */
typedef struct klsp_problem {
    int64_t NPh;            /* number of orbits (NR*Ne)                     */
    int64_t N_alpha;        /* N                                            */
    int64_t n_l;            /* number of radial harmonics                   */
    int m;
    double G;
    const double* DJ;       /* [NPh]  (DJ, F0l, Omega_res may all be NULL:   */
    const double* F0l;      /* [NPh, n_l]   set them later with             */
    const double* Omega_res;/* [NPh, n_l]   klsp_set_problem_arrays)        */
    const double* N_kernel; /* [N]                                          */
} klsp_problem;

/* Replaces the context's problem (dimensions, arrays; W must be given
 * again).  The W buffers are kept and reused when the new W fits (a batch of
 * problems then allocates its largest W once); klsp_destroy releases them. */
int klsp_upload_problem(klsp_ctx* ctx, const klsp_problem* prob);

/* DJ [NPh], F0l and Omega_res [NPh, n_l] of an uploaded problem (dimensions
 * unchanged; W is kept).  Needed before any build when the problem was
 * uploaded without them. */
int klsp_set_problem_arrays(klsp_ctx* ctx, const double* DJ, const double* F0l,
                            const double* Omega_res);

/* compute_orbits (OrbitIntegration.jl:181-302) followed by compute_W_l
 * (BasisFunctions.jl:199), both on the device: one workgroup per orbit
 * integrates its nw phase samples (the cumulative trapezoid sums as block
 * scans, so the sums are grouped differently from the sequential host loop
 * at the 1e-16 level); the nwa stored samples feed the W quadrature without
 * leaving the device.  Grid inputs as from klsp_compute_grids; the problem
 * (its dimensions at least) must be uploaded.  Omega_1, Omega_2 [NPh] are
 * returned; w1, ra, pha [nwa, NPh] are copied back if non-NULL. */
int klsp_compute_orbits_W(klsp_ctx* ctx, const klsp_params* p, const double* Rc,
                          const double* e, const double* R1, const double* R2, const double* E,
                          const double* L, const double* L2, const double* SGNL,
                          const double* alpha, const int32_t* l_arr, double* Omega_1,
                          double* Omega_2, double* w1, double* ra, double* pha);

/* compute_W_l, BasisFunctions.jl:199 (W_l(alpha) orbit quadrature, Eq. 41),
 * evaluated on the device as one GEMM per orbit and kept device-resident in
 * FP64.  Orbit arrays are [nwa, NPh]; alpha [N]; l_arr [n_l] (must match the
 * uploaded problem).  If W_out is non-NULL the [NPh, N, n_l] complex array
 * is also copied back (24.7 GB at the default grid: pass NULL normally). */
int klsp_compute_W(klsp_ctx* ctx, int64_t nwa, const double* w1, const double* ra,
                   const double* pha, const double* Omega_1, const double* Omega_2,
                   const double* alpha, const int32_t* l_arr, double* W_out);

/* Upload a caller-supplied W_l_mat [NPh, N, n_l] complex128 (to_device,
 * BasisFunctions.jl:421). */
int klsp_upload_W(klsp_ctx* ctx, const double* W);

/* Copy the device W back in the reference layout [NPh, N, n_l]. */
int klsp_download_W(klsp_ctx* ctx, double* W_out);

/* Materialise the device copy of W used by `precision` (the complex64 copy
 * for KLSP_PREC_FP32).  Called implicitly by the first build that needs it. */
int klsp_prepare(klsp_ctx* ctx, int precision);

/* ------------------------------------------------------------------ */
/* The hot path                                                        */
/* ------------------------------------------------------------------ */

/* build_M_matrix, MatrixBuilder.jl:242 (= build_M_matrix_cpu :27 /
 * build_M_matrix_gpu :100): M[b,a] = G*N[a]*sum_l sum_k W[k,b,l] c[k,l]
 * conj(W[k,a,l]), c = DJ*F0l/(omega - Omega_res) with the 1e-12 clamp.
 * M_out is [N, N] complex, column-major. */
int klsp_build_M(klsp_ctx* ctx, double omega_re, double omega_im, int precision,
                 double* M_out);

/* Taylor terms of M around a centre omega_c (no counterpart in the
 * reference; the building block of the KLSP_PREC_MIXED Newton search):
 * T_out is [n_terms][N, N] complex, T_j = G*N[a]*sum W c/d^j conj W with
 * c = DJ*F0l and d = omega_c - Omega_res (j = 1..n_terms, 1e-12 clamp), so
 * M(omega_c + s) = sum_j (-s)^(j-1) T_j for |s| < Im omega_c and T_1 =
 * M(omega_c).  1 <= n_terms <= 3. */
int klsp_build_taylor(klsp_ctx* ctx, double omega_re, double omega_im, int n_terms,
                      int precision, double* T_out);

/* The blocks of the linear generalized eigenproblem omega D C = (E + B) C,
 * Matlab/Matrix_LogExp.m:373-403, for every radial harmonic l of the
 * uploaded problem:
 *   D_l = W_l^T diag(DJ) conj W_l,   E_l = W_l^T diag(DJ Omega_res_l) conj W_l,
 *   F_l = W_l^T diag(DJ F0l_l) conj W_l
 * (the contraction of klsp_build_M restricted to one l, with omega-independent
 * weights; FP64 on the device).  Each output is [N, N, n_l] complex,
 * column-major (block l at offset N*N*l); any may be NULL.  The global
 * matrices and B_{l,l'} = G F_l diag(N_kernel) D_{l'} (Matrix_LogExp.m:
 * 414-449) are small dense algebra on these blocks. */
int klsp_build_pencil(klsp_ctx* ctx, double* D_out, double* E_out, double* F_out);

/* compute_determinant / compute_determinant_batched, MatrixBuilder.jl:197,
 * :215: det(I - M(omega_i)) for n values of omega (interleaved complex),
 * FP64 complex LU with partial pivoting on the device. */
int klsp_det(klsp_ctx* ctx, const double* omegas, int64_t n, int precision,
             double* dets_out);

/* det(I - M(omega_i)) as klsp_det with KLSP_PREC_FP64, allocating nothing.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its workspaces -- the weights, the K-piece partials and
 * the N x N matrices, about 16 (Kp + 64) + 16 Np^2 P + 48 N^2 bytes with P
 * the work table's K pieces (~211 MB at the default grid) -- are carved from
 * one 256 MiB array per GPU that the library holds until the process ends;
 * larger problems fail with KLSP_ERR_INVALID_ARG.  Needs
 * klsp_prepare(ctx, KLSP_PREC_FP64) after the problem and W are on the device
 * (which builds the real basis and the work table; KLSP_ERR_NOT_READY
 * otherwise).  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (from several contexts or threads) race. */
int klsp_det_noalloc(klsp_ctx* ctx, const double* omegas, int64_t n, double* dets_out);

/* newton_search, NewtonSolver.jl:36 (NewtonResult :19). */
typedef struct klsp_newton_params {
    int max_iter;          /* NewtonConfig.max_iter                         */
    double tol;            /* converged when |det| < tol                    */
    double delta;          /* forward-difference step in (Omega_p, gamma)   */
    int use_line_search;   /* halving line search with bounds               */
    int max_line_search;
    double Omega_p_init;
    double gamma_init;
    int precision;         /* klsp_precision of every build                 */
} klsp_newton_params;

typedef struct klsp_newton_result {
    double Omega_p, gamma;
    double omega[2];       /* m*Omega_p + i*gamma                           */
    double det[2];         /* det at the final (Omega_p, gamma): a fresh
                              det in that precision for FP64 / FP32 /
                              BF16X3; for MIXED / MIXED_FP32 the anchored
                              model's det there (M64 at the last anchor plus
                              the low-precision Taylor corrections, which
                              enter multiplied by |s - s_a|), not a new
                              FP64 build                                   */
    int iterations;
    int converged;         /* FP64 / FP32 / BF16X3: |det| < tol (the
                              reference's test); MIXED: the last anchor's
                              Newton loop ended on its step test, the zero
                              lay within that anchor's trust radius, and
                              |det| < tol                                  */
    int n_det_evals;       /* number of M(omega) builds performed           */
    int n_fp64_builds;     /* ... of which FP64 single-omega builds         */
    int n_lp_builds;       /* ... of which FP32 builds (MIXED: Taylor passes) */
} klsp_newton_result;

/* history_out: [max_iter, 3] column-major (Omega_p, gamma, |det|); the first
 * min(iterations, max_iter) rows are filled.  May be NULL.
 *
 * precision FP64 / FP32 / BF16X3: the reference iteration, every det in
 * that precision.  precision MIXED (no counterpart in the reference; at a
 * fraction of the FP64 builds it returns an FP64 root to |d omega| <= 1e-10
 * by a calibrated estimate, the FP64 search's root where both searches
 * reach the same one; from far starts the two paths can end at different
 * roots):
 *   1. low-precision passes: at a centre w_c one BF16x3 (MIXED_FP32: FP32)
 *      contraction builds the Taylor terms T_j = G N sum W c/d^j conj W
 *      (d = w_c - Omega_res, j = 1..4 for BF16x3, 1..3 for FP32), so
 *      M(w_c + s) = T_1 - s T_2 + s^2 T_3 - ...; the zero of det(I - M) of
 *      that model within |s| <= Im(w_c)/4 (FP64 LUs only: its values at 16
 *      points on that circle, from one launch, give its Taylor coefficients;
 *      model Newton steps where that fails, moving at most Im(w_c)/2) moves
 *      the centre; passes repeat until the move's predicted model error is
 *      small enough for the finish (one pass from a start 4e-3 from the
 *      root);
 *   2. FP64 finish: one FP64 build at the low-precision root w_a anchors
 *      the model, M(w) = M64(w_a) - (s - s_a) T_2 + (s - s_a)(s + s_a) T_3
 *      (s = w - w_c, s_a = w_a - w_c), in which the low-precision terms
 *      enter only multiplied by |s - s_a| <= 1e-6, so M near the root is
 *      FP64-accurate; its zero near s_a (the same 16-point circle, radius
 *      anchor_tol below) is confirmed by Newton steps on it (FP64 LUs) until
 *      the step stops shrinking; converged if then |det| < tol.  A new anchor is built if
 *      the root moves so far from the last one that the model's slope error
 *      times that move could exceed 5e-11.  The slope error is taken along
 *      the null vectors of I - M, where a term's weight is the partial
 *      derivative of the model's det in its coefficient (central differences
 *      at the anchor, one LU launch): twice the truncation of the pass's
 *      series (its first omitted term, the ratio of the last built terms
 *      continued) plus the rounding of the low-precision terms (sixteen
 *      times the move of the pass's zero at the first anchor times
 *      |T_2| / |T_1|, at least 4e-5 for BF16x3 and 2e-6 for FP32); the new
 *      anchor is built at the anchored
 *      model's own zero (converged on first).  Passes after one that left
 *      the root within 1e-4 build three terms.
 * iterations = low-precision passes + anchors + 1; history rows record each
 * centre with its low-precision |det|, each anchor with its FP64 |det|, and
 * the final point.  Bounds (gamma >= 1e-4, Omega_p in [0, 1]) apply to
 * every move. */
int klsp_newton(klsp_ctx* ctx, const klsp_newton_params* np, klsp_newton_result* res,
                double* history_out);

/* grid_scan + find_minimum, GridScan.jl:25, :234.  det_grid_out is
 * [n_gamma, n_Op] complex column-major; Omega_p in range(Op_min, Op_max,
 * n_Op), gamma in range(g_min, g_max, n_gamma), omega = m*Omega_p + i*gamma. */
typedef struct klsp_scan_params {
    int n_Op, n_gamma;
    double Omega_p_min, Omega_p_max;
    double gamma_min, gamma_max;
    int precision;
} klsp_scan_params;

int klsp_grid_scan(klsp_ctx* ctx, const klsp_scan_params* sp, double* det_grid_out,
                   double* Omega_p_init_out, double* gamma_init_out);

/* The scan's device work as one HIP graph (no counterpart in the reference,
 * whose scan launches one det at a time, GridScan.jl:104).  By default
 * klsp_grid_scan uploads its omegas once and launches every build (weights,
 * contraction, reduction) and every LU of the scan as a single graph that
 * writes the dets to device memory; the graph is instantiated on the first
 * scan and launched again by later scans of the same size, precision and
 * problem shape on the context (a changed buffer, size or contraction
 * variant instantiates a new one).  Results are the same as launching the
 * kernels one by one, which klsp_set_scan_graph(ctx, 0) selects.  With
 * profiling on, klsp_timing.contract_ms of a graph scan is the whole graph's
 * device time.  klsp_scan_graph_stats reports how many graphs the context
 * has instantiated and launched. */
int klsp_set_scan_graph(klsp_ctx* ctx, int on);

/* Which LU computes det(I - M) (no counterpart in the reference; for tests
 * and measurement).  0 (default): batches of at most 24 matrices of order
 * N <= 512 take the step-wise LU, one launch per 16-column panel with a
 * workgroup per column block, so one det spreads over many CUs; larger
 * batches and larger N take one workgroup per matrix.  1 forces the
 * one-workgroup LU, 2 the step-wise LU (for N <= 512; larger N keep the
 * one-workgroup LU).  Both follow zgetrf's pivoting and give the same bits. */
int klsp_set_lu_form(klsp_ctx* ctx, int form);

/* Which work decomposition the BF16x3 contraction on a symmetric alpha grid
 * uses (no counterpart in the reference; for tests and measurement).
 * 0 (default): chosen per problem from how well each fills the upper
 * triangle's blocks.  1: folded column-block pairs, one block per slot.
 * 2: folded two-column strips, both blocks of a strip row per slot (half the
 * operand loads; launches of at most two weight sets).  For the same gang
 * plan both give the same bits. */
int klsp_set_bf16_form(klsp_ctx* ctx, int form);

/* Device memory the context holds between calls: W and its copies, the
 * problem arrays and the grow-only workspaces, in bytes (no counterpart in
 * the reference).  Orbit samples and quadrature temporaries of
 * klsp_compute_W / klsp_compute_orbits_W are freed before those return and
 * are not included.  A caller packing several contexts onto one device (the
 * uneven batch) budgets HBM with it. */
int klsp_device_bytes(klsp_ctx* ctx, int64_t* bytes_out);
int klsp_scan_graph_stats(klsp_ctx* ctx, int64_t* builds_out, int64_t* launches_out);

/* Real basis of a symmetric alpha grid (no counterpart in the reference;
 * the arithmetic of MatrixBuilder.jl:119-141 rearranged).  Where W[:, N-1-a]
 * = conj W[:, a] for every column a and the centre column of an odd N is
 * real -- as on the reference's grid range(-alpha_max, alpha_max, N), where
 * W_l(-alpha) = conj W_l(alpha) -- W = R T with R = [Re W(alpha >= 0) |
 * Im W(alpha > 0)] real (K x N) and T fixed, so that one-omega FP64 builds
 * take two real symmetric products R^T diag(Re c) R and R^T diag(Im c) R
 * (about a quarter of the complex product's MFMA work; same results to
 * rounding).  W from klsp_compute_W / klsp_compute_orbits_W on such a grid
 * is symmetric by construction; an uploaded W is checked value by value
 * when first needed.  on = 0 makes every build use the complex product
 * (default 1). */
int klsp_set_symmetric_basis(klsp_ctx* ctx, int on);

/* *active_out = 1 if one-omega FP64 builds of the current W use the real
 * basis (checks the symmetry of an uploaded W; needs W on the device). */
int klsp_symmetric_basis(klsp_ctx* ctx, int* active_out);

/* Autotuner (no counterpart in the reference).  Times every compiled
 * variant of the contraction that one-omega `precision` builds of the current
 * problem use -- FP64 on a symmetric alpha grid: the real-basis kernel's
 * prefetch depth and the form of its diagonal tiles; one form elsewhere --
 * with one untimed and `reps` timed builds each, keeps the fastest for the
 * context's later builds (and for later problems of the same padded width
 * on this context), and returns the variant count in *n_variants, their
 * mean contraction times in ms_out[0 .. min(count, n_max)) and the index of
 * the fastest in *chosen (ms_out and chosen may be NULL with n_max = 0). */
int klsp_autotune(klsp_ctx* ctx, int precision, int reps, int* n_variants, double* ms_out,
                  int n_max, int* chosen);

/* Name of variant `variant` of `precision` builds ("" if none). */
const char* klsp_variant_name(int precision, int variant);

/* Pins variant `variant` (as numbered by klsp_autotune) for the context's
 * later `precision` builds, e.g. a choice saved from an earlier tuning run.
 * Only FP64 (real-basis) builds have variants; anything else is
 * KLSP_ERR_INVALID_ARG. */
int klsp_set_variant(klsp_ctx* ctx, int precision, int variant);

/* ------------------------------------------------------------------ */
/* Measurement                                                         */
/* ------------------------------------------------------------------ */

/* Per-phase device timings of the most recent klsp_det / klsp_build_M call,
 * measured with HIP events on the context's stream when profiling is on. */
typedef struct klsp_timing {
    double weights_ms;     /* c = DJ*F0l/(omega - Omega_res)                */
    double contract_ms;    /* the M(omega) contraction kernel               */
    double reduce_ms;      /* K-split partial sums + G*N epilogue           */
    double lu_ms;          /* complex LU determinant                        */
    int64_t n_omega;       /* omegas processed by the call                  */
    int64_t contract_launches;
} klsp_timing;

int klsp_set_profiling(klsp_ctx* ctx, int on);
int klsp_last_timing(klsp_ctx* ctx, klsp_timing* t_out);

/* Device time of the last W quadrature (klsp_compute_W, klsp_compute_orbits_W):
 * orbit preprocessing + GEMM. */
int klsp_last_W_ms(klsp_ctx* ctx, double* ms_out);

/* Device time of the orbit integration kernel of the last
 * klsp_compute_orbits_W (the W quadrature after it is in klsp_last_W_ms). */
int klsp_last_orbit_ms(klsp_ctx* ctx, double* ms_out);

#ifdef __cplusplus
}
#endif

#endif /* KLSP_H */
