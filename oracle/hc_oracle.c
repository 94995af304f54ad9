/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C restatement of HERCULES's finite-difference time step, used as
 * the correctness checker for the MI355X kernels.  Only tests/, the smoke()
 * entry and the cpu_baseline leg of bench.py may load this library; nothing
 * in the product links or calls it.
 *
 * It mirrors the reference's algorithm AND its data layout: every field is a
 * Fortran column-major array (nx+4, ny+4, nz+4) with two ghost cells per side
 * and 1-based indices, the interior being istr3..iend3 = 3..nx+2 etc.
 * (src/module_parameters.f90:371-398, src/module_tools.f90:41-89).  One MPI
 * rank is assumed (p_row = p_col = 1), so 2DECOMP transposes are identity
 * reshapes and comm_bound is a periodic copy (src/module_boundary.f90:361-487).
 *
 * Third-party arithmetic: the reference calls FFTW 3.3.4 (fftw_plan_dft_1d,
 * FFTW_FORWARD = exp(-2 pi i jk/n), unnormalised; src/module_parameters.f90:565-572).
 * FFTW is not linked here; a mixed-radix decimation-in-time DFT of the same
 * definition is used instead (fft_rec below), pinned by the tests against
 * numpy.fft.
 *
 * Each function cites the reference lines it restates.  Pinning of every
 * function is listed in tests/test_oracle_*.py (closed forms, Fourier-mode
 * symbols, scipy banded solves, invariants).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

typedef double complex cplx;

/* Same field order and types as hc_params in include/hercules_amd.h. */
typedef struct ho_params {
    int nx, ny, nz;
    int ochannel, dp_opt, cds, dts, issk, isdamp, nzdamp, isscalar, is_ri_var, imax;
    double lx, ly, lz, zstretch, cdamp;
    double nu, alpha, got, fc, ug, vg, dt, u_mrf, t_ref, tbot, ttop, ri_str, ri_end;
    int wall_rows; /* 1: every mode gets the wall rows the reference gives mode (0,0) only */
    int precision; /* unused here: the oracle is FP64 */
    /* extension (no reference counterpart): 1 = top temperature gradient
     * dtdz_top instead of t = ttop */
    int tbc_top;
    double dtdz_top;
} ho_params;

typedef struct ho_state {
    ho_params prm;
    int nia, nja, nka, istr3, iend3, jstr3, jend3, kstr3, kend3;
    double dx, dy, dx2, dy2, dpx_coef, dpx_drive, ubar, got;
    cplx wx1, wy1;
    double rk_a, rk_b, rk_c;
    double *dz, *l_t, *l_b, *dz_t, *dz_b;
    double *u, *v, *w, *p, *phi, *t;
    double *ff, *gg, *hh, *ss, *ff0, *gg0, *hh0, *ss0, *qu, *qv, *qw;
    double *uhx, *uhy, *uhz, *vhx, *vhy, *vhz, *whx, *why, *thx, *thy, *thz;
    double *aa_u, *cc_u, *aa_w, *cc_w, *aa_p, *cc_p, *aa_t, *cc_t;
    cplx *bb_u, *bb_w, *bb_p, *bb_t;
    double *aa, *cc;
    cplx *bb;
    cplx *c_src, *c_var, *rowbuf;
    /* spectral path (cds=3): C2C spectra of the fields and right-hand sides */
    cplx *c_u, *c_v, *c_w, *c_p, *c_t, *c_phi, *c_ff, *c_gg, *c_hh, *c_ss, *c_qu, *c_qv, *c_qw;
    double div_avg, div_max, div_hist_max, time_sim;
    int stps;
    /* calculate_mean_fields running sums (interior arrays, allocated on
     * request) and the number of steps they hold */
    double *mean[11];
    int mean_steps;
    char err[256];
} ho_state;

static const double PI = 3.1415926535897932; /* src/module_parameters.f90:32 */

/* RK3 coefficients, Orlandi 2000: src/module_parameters.f90:210-212 */
static const double RKC1[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double RKC2[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};
static const double RKC3[3] = {8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0};

/* ---- index helpers (1-based, Fortran order) ---------------------------- */
#define F3(S, a, i, j, k) ((a)[((size_t)(k) - 1) * (size_t)(S)->nia * (size_t)(S)->nja + \
                               ((size_t)(j) - 1) * (size_t)(S)->nia + (size_t)(i) - 1])
/* (nx, ny, nz+4) coefficient arrays: aa, bb, cc, bb_* */
#define B3(S, a, i, j, k) ((a)[((size_t)(k) - 1) * (size_t)(S)->prm.nx * (size_t)(S)->prm.ny + \
                               ((size_t)(j) - 1) * (size_t)(S)->prm.nx + (size_t)(i) - 1])
/* (nx, ny, nz) complex scratch: c_src, c_var */
#define C3(S, a, i, j, k) ((a)[((size_t)(k) - 1) * (size_t)(S)->prm.nx * (size_t)(S)->prm.ny + \
                               ((size_t)(j) - 1) * (size_t)(S)->prm.nx + (size_t)(i) - 1])
#define V1(a, k) ((a)[(k) - 1])

static size_t nfield(const ho_state *S) { return (size_t)S->nia * S->nja * S->nka; }

/* ---- FFT: unnormalised DFT, X_k = sum_n x_n exp(sign 2 pi i nk/N) ------- */
static int smallest_factor(int n)
{
    for (int f = 2; f * f <= n; ++f)
        if (n % f == 0) return f;
    return n;
}

/* mixed-radix decimation in time; tw[e] = exp(sign 2 pi i e / N) */
static void fft_rec(const cplx *in, int n, int stride, cplx *out, const cplx *tw, int N)
{
    if (n == 1) { out[0] = in[0]; return; }
    int p = smallest_factor(n), m = n / p;
    for (int r = 0; r < p; ++r) fft_rec(in + (size_t)r * stride, m, stride * p, out + (size_t)r * m, tw, N);
    int step = N / n; /* W_n^e = W_N^(e*step) */
    cplx t[64];
    if (p > 64) { fprintf(stderr, "oracle fft: prime factor %d too large\n", p); abort(); }
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r) t[r] = out[(size_t)r * m + k] * tw[((size_t)r * k * step) % N];
        for (int q = 0; q < p; ++q) {
            cplx acc = 0;
            for (int r = 0; r < p; ++r) acc += t[r] * tw[((size_t)r * q * m * step) % N];
            out[(size_t)q * m + k] = acc;
        }
    }
}

typedef struct { int n, sign; cplx *tw; } ho_plan;

static ho_plan plan_make(int n, int sign)
{
    ho_plan pl = {n, sign, (cplx *)malloc(sizeof(cplx) * n)};
    for (int e = 0; e < n; ++e) {
        double a = 2.0 * PI * (double)e / (double)n;
        pl.tw[e] = cos(a) + (double)sign * sin(a) * I;
    }
    return pl;
}

/* x[0..n-1] with stride xs, in place, via buffers in/out of length n */
static void plan_exec(const ho_plan *pl, cplx *x, size_t xs, cplx *bin, cplx *bout)
{
    for (int e = 0; e < pl->n; ++e) bin[e] = x[(size_t)e * xs];
    fft_rec(bin, pl->n, 1, bout, pl->tw, pl->n);
    for (int e = 0; e < pl->n; ++e) x[(size_t)e * xs] = bout[e];
}

/* exported for pinning against numpy.fft */
void ho_fft1d(cplx *x, int n, int sign)
{
    ho_plan pl = plan_make(n, sign);
    cplx *a = (cplx *)malloc(sizeof(cplx) * n), *b = (cplx *)malloc(sizeof(cplx) * n);
    plan_exec(&pl, x, 1, a, b);
    free(a); free(b); free(pl.tw);
}

/* ---- state -------------------------------------------------------------- */
static double *dalloc(size_t n) { return (double *)calloc(n, sizeof(double)); }
static cplx *calloc_c(size_t n) { return (cplx *)calloc(n, sizeof(cplx)); }

void ho_destroy(ho_state *S);

/* initiate_parameters + initiate_indices + initiate_fields (allocation part)
 * src/module_parameters.f90:242-318,371-398; src/module_tools.f90:33-227 */
ho_state *ho_create(const ho_params *P)
{
    ho_state *S = (ho_state *)calloc(1, sizeof(ho_state));
    S->prm = *P;
    if (P->nx % 2 || P->ny % 2 || P->nz < 3) { ho_destroy(S); return NULL; }
    /* the damped levels kend3-nzdamp..kend3 must be interior levels */
    if (P->isdamp == 1 && (P->nzdamp < 1 || P->nzdamp > P->nz - 1)) { ho_destroy(S); return NULL; }
    if (P->tbc_top != 0 && P->tbc_top != 1) { ho_destroy(S); return NULL; }
    S->istr3 = 3; S->iend3 = P->nx + 2;
    S->jstr3 = 3; S->jend3 = P->ny + 2;
    S->kstr3 = 3; S->kend3 = P->nz + 2;
    S->nia = P->nx + 4; S->nja = P->ny + 4; S->nka = P->nz + 4;
    S->dx = P->lx / P->nx; S->dy = P->ly / P->ny;
    S->dx2 = S->dx * S->dx; S->dy2 = S->dy * S->dy;
    S->dpx_coef = P->nu * 1.0;
    S->wx1 = cexp(I * 2.0 * PI / P->nx);
    S->wy1 = cexp(I * 2.0 * PI / P->ny);
    S->got = P->got;
    if (P->dp_opt == 1) S->dpx_drive = 1.0; else { S->dpx_drive = 0.0; S->ubar = 0.0; }
    size_t nf = nfield(S), nb = (size_t)P->nx * P->ny * S->nka, nc = (size_t)P->nx * P->ny * P->nz;
    double **fl[] = {&S->u, &S->v, &S->w, &S->p, &S->phi, &S->t, &S->ff, &S->gg, &S->hh, &S->ss,
                     &S->ff0, &S->gg0, &S->hh0, &S->ss0, &S->qu, &S->qv, &S->qw, &S->uhx, &S->uhy,
                     &S->uhz, &S->vhx, &S->vhy, &S->vhz, &S->whx, &S->why, &S->thx, &S->thy, &S->thz};
    for (size_t q = 0; q < sizeof(fl) / sizeof(fl[0]); ++q) *fl[q] = dalloc(nf);
    double **v1[] = {&S->dz, &S->l_t, &S->l_b, &S->dz_t, &S->dz_b, &S->aa_u, &S->cc_u, &S->aa_w,
                     &S->cc_w, &S->aa_p, &S->cc_p, &S->aa_t, &S->cc_t};
    for (size_t q = 0; q < sizeof(v1) / sizeof(v1[0]); ++q) *v1[q] = dalloc(S->nka);
    S->bb_u = calloc_c(nb); S->bb_w = calloc_c(nb); S->bb_p = calloc_c(nb); S->bb_t = calloc_c(nb);
    S->aa = dalloc(nb); S->cc = dalloc(nb); S->bb = calloc_c(nb);
    S->c_src = calloc_c(nc); S->c_var = calloc_c(nc);
    if (P->cds == 3) {
        cplx **cs[] = {&S->c_u, &S->c_v, &S->c_w, &S->c_p, &S->c_t, &S->c_phi, &S->c_ff, &S->c_gg,
                       &S->c_hh, &S->c_ss, &S->c_qu, &S->c_qv, &S->c_qw};
        for (size_t q = 0; q < sizeof(cs) / sizeof(cs[0]); ++q) *cs[q] = calloc_c(nc);
    }
    int nmax = P->nx > P->ny ? P->nx : P->ny;
    S->rowbuf = calloc_c((size_t)4 * nmax);
    return S;
}

void ho_destroy(ho_state *S)
{
    if (!S) return;
    double *fl[] = {S->u, S->v, S->w, S->p, S->phi, S->t, S->ff, S->gg, S->hh, S->ss, S->ff0, S->gg0,
                    S->hh0, S->ss0, S->qu, S->qv, S->qw, S->uhx, S->uhy, S->uhz, S->vhx, S->vhy,
                    S->vhz, S->whx, S->why, S->thx, S->thy, S->thz, S->dz, S->l_t, S->l_b, S->dz_t,
                    S->dz_b, S->aa_u, S->cc_u, S->aa_w, S->cc_w, S->aa_p, S->cc_p, S->aa_t, S->cc_t,
                    S->aa, S->cc};
    for (size_t q = 0; q < sizeof(fl) / sizeof(fl[0]); ++q) free(fl[q]);
    free(S->bb_u); free(S->bb_w); free(S->bb_p); free(S->bb_t); free(S->bb);
    free(S->c_src); free(S->c_var); free(S->rowbuf);
    cplx *cs[] = {S->c_u, S->c_v, S->c_w, S->c_p, S->c_t, S->c_phi, S->c_ff, S->c_gg, S->c_hh, S->c_ss,
                  S->c_qu, S->c_qv, S->c_qw};
    for (size_t q = 0; q < sizeof(cs) / sizeof(cs[0]); ++q) free(cs[q]);
    for (int q = 0; q < 11; ++q) free(S->mean[q]);
    free(S);
}

/* ---- A1: get_mesh_param, src/module_parameters.f90:432-508 ------------- */
void ho_mesh(ho_state *S)
{
    const ho_params *P = &S->prm;
    int nka = S->nka;
    double zmin = 0.0, zmax = P->lz, zmean = 0.0, coeff_a = 0.0;
    double *zz = dalloc(nka);
    if (P->ochannel == 0) { zmean = 0.5 * (zmin + zmax); coeff_a = (2.0 + P->nz) * 0.5; }
    else                  { zmean = (zmin + zmax);       coeff_a = (1.0 + P->nz) * 1.0; }
    for (int k = 1; k <= P->nz + 1; ++k)
        V1(zz, k + 2) = zmean + (zmin - zmean) / tanh(2.0 * P->zstretch) *
                                tanh(2.0 * P->zstretch * (k - coeff_a) / (1.0 - coeff_a));
    for (int k = 1; k <= nka; ++k) V1(S->dz, k) = 0.0;
    for (int k = S->kstr3; k <= S->kend3; ++k) V1(S->dz, k) = V1(zz, k + 1) - V1(zz, k);
    V1(S->dz, S->kstr3 - 1) = V1(S->dz, S->kstr3);
    V1(S->dz, S->kend3 + 1) = V1(S->dz, S->kend3);
    V1(S->dz, S->kstr3 - 2) = V1(S->dz, S->kstr3);
    V1(S->dz, S->kend3 + 2) = V1(S->dz, S->kend3);
    for (int k = 1; k <= nka; ++k) { V1(S->l_t, k) = V1(S->l_b, k) = V1(S->dz_t, k) = V1(S->dz_b, k) = 0.0; }
    for (int k = S->kstr3 - 1; k <= S->kend3 + 1; ++k) {
        V1(S->l_t, k) = V1(S->dz, k) / (V1(S->dz, k) + V1(S->dz, k + 1));
        V1(S->l_b, k) = V1(S->dz, k) / (V1(S->dz, k) + V1(S->dz, k - 1));
        V1(S->dz_t, k) = 0.5 * (V1(S->dz, k) + V1(S->dz, k + 1));
        V1(S->dz_b, k) = 0.5 * (V1(S->dz, k) + V1(S->dz, k - 1));
    }
    free(zz);
}

/* ---- integer and real powers of a complex number as Fortran evaluates
 *      them: z**n (integer n) by repeated products, z**x (real x) = exp(x log z) */
static cplx cipow(cplx z, int n)
{
    cplx r = 1.0, b = z;
    int e = n < 0 ? -n : n;
    while (e) { if (e & 1) r *= b; b *= b; e >>= 1; }
    return n < 0 ? 1.0 / r : r;
}
static cplx crpow(cplx z, double x) { return x == 0.0 ? 1.0 : cpow(z, x); }

/* wavenumber index as the reference maps it: src/module_poisson_solver.f90:57-69 */
static int lwave(int i, int n) { return i <= n / 2 ? i - 1 : i - n - 1; }

/* horizontal symbols of the FD second-derivative operators */
static cplx kx_helm(const ho_state *S, cplx w1, int l, double d2, double coef)
{
    /* u/w/t equations: src/module_poisson_solver.f90:222-246 (coef = nu or alpha) */
    const ho_params *P = &S->prm;
    if (P->cds == 1) return 0.5 * P->dt * (cipow(w1, l) + cipow(w1, -l) - 2.0) * coef / d2;
    if (P->cds == 2)
        return P->dt * (-crpow(w1, 2.0 * l) + 16.0 * cipow(w1, l) - 30.0 + 16.0 * cipow(w1, -l) -
                        crpow(w1, -2.0 * l)) / 24.0 * coef / d2;
    return 0; /* cds=3 handled by caller */
}

static cplx kx_pois(const ho_state *S, cplx w1, int l, double d2)
{
    /* p equation: src/module_poisson_solver.f90:73-101 */
    const ho_params *P = &S->prm;
    if (P->cds == 1) return (cipow(w1, l) + cipow(w1, -l) - 2.0) / d2;
    if (P->cds == 2)
        return (crpow(w1, 3.0 * l) - 54.0 * crpow(w1, 2.0 * l) + 783.0 * cipow(w1, l) - 1460.0 +
                783.0 * cipow(w1, -l) - 54.0 * crpow(w1, -2.0 * l) + crpow(w1, -3.0 * l)) / 576.0 / d2;
    return 0;
}

/* kx(l), ky(m) of the u/w/t equations, evaluated once per wavenumber
 * (src/module_poisson_solver.f90:222-246 evaluate the same expressions for
 * every level) */
static void helm_symbols(const ho_state *S, double coef, cplx *kxv, cplx *kyv)
{
    const ho_params *P = &S->prm;
    for (int i = 1; i <= P->nx; ++i) {
        int l = lwave(i, P->nx);
        kxv[i - 1] = P->cds == 3 ? 0.5 * P->dt * coef * (-pow(2.0 * PI / P->lx * l, 2.0))
                                 : kx_helm(S, S->wx1, l, S->dx2, coef);
    }
    for (int j = 1; j <= P->ny; ++j) {
        int m = lwave(j, P->ny);
        kyv[j - 1] = P->cds == 3 ? 0.5 * P->dt * coef * (-pow(2.0 * PI / P->ly * m, 2.0))
                                 : kx_helm(S, S->wy1, m, S->dy2, coef);
    }
}

/* ---- A8: equation coefficients ----------------------------------------- */
/* get_p_eqn_coeff, src/module_poisson_solver.f90:32-122 */
static void coeff_p(ho_state *S)
{
    const ho_params *P = &S->prm;
    for (int k = 1; k <= S->nka; ++k) { V1(S->aa_p, k) = 0; V1(S->cc_p, k) = 0; }
    memset(S->bb_p, 0, sizeof(cplx) * (size_t)P->nx * P->ny * S->nka);
    /* kx(l), ky(m) do not depend on k; the reference re-evaluates the same
     * expressions for every level, here they are evaluated once */
    cplx *kxv = calloc_c(P->nx), *kyv = calloc_c(P->ny);
    for (int i = 1; i <= P->nx; ++i) {
        int l = lwave(i, P->nx);
        kxv[i - 1] = P->cds == 3 ? -pow(2.0 * PI / P->lx * l, 2.0) : kx_pois(S, S->wx1, l, S->dx2);
    }
    for (int j = 1; j <= P->ny; ++j) {
        int m = lwave(j, P->ny);
        kyv[j - 1] = P->cds == 3 ? -pow(2.0 * PI / P->ly * m, 2.0) : kx_pois(S, S->wy1, m, S->dy2);
    }
    for (int k = S->kstr3; k <= S->kend3; ++k) {
        V1(S->aa_p, k) = 1.0 / V1(S->dz, k) / V1(S->dz_b, k);
        V1(S->cc_p, k) = 1.0 / V1(S->dz, k) / V1(S->dz_t, k);
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i) {
                cplx kx = kxv[i - 1], ky = kyv[j - 1];
                cplx *b = &B3(S, S->bb_p, i, j, k);
                *b = kx + ky - V1(S->aa_p, k) - V1(S->cc_p, k);
                if (k == S->kstr3) { *b += V1(S->aa_p, k); V1(S->aa_p, k) = 0.0; }
                if (k == S->kend3) { *b += V1(S->cc_p, k); V1(S->cc_p, k) = 0.0; }
            }
    }
    free(kxv); free(kyv);
}

/* get_u_eqn_coeff (src/module_poisson_solver.f90:181-282), get_t_eqn_coeff
 * (:379-477).  Note the 1-D aa(k)/cc(k) are zeroed inside the (i,j) loop, so
 * only the first visited mode (i=j=1, i.e. l=m=0) receives the -aa/-cc wall
 * correction on the diagonal; the order of the loops is kept exactly. */
static void coeff_uvt(ho_state *S, double *aa, cplx *bb, double *cc, double coef, int is_t)
{
    const ho_params *P = &S->prm;
    for (int k = 1; k <= S->nka; ++k) { V1(aa, k) = 0; V1(cc, k) = 0; }
    memset(bb, 0, sizeof(cplx) * (size_t)P->nx * P->ny * S->nka);
    cplx *kxv = calloc_c(P->nx), *kyv = calloc_c(P->ny);
    helm_symbols(S, coef, kxv, kyv);
    for (int k = S->kstr3; k <= S->kend3; ++k) {
        V1(aa, k) = -0.5 * P->dt * coef / V1(S->dz, k) / V1(S->dz_b, k);
        V1(cc, k) = -0.5 * P->dt * coef / V1(S->dz, k) / V1(S->dz_t, k);
        const double a0 = V1(aa, k), c0 = V1(cc, k);
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i) {
                cplx kx = kxv[i - 1], ky = kyv[j - 1];
                cplx *b = &B3(S, bb, i, j, k);
                /* top wall row: Dirichlet (t, closed-channel u) or Neumann
                 * (open-channel u, t with tbc_top = 1) */
                const int top_dir = is_t ? P->tbc_top != 1 : P->ochannel == 0;
                if (P->wall_rows == 1) {
                    /* extension: the first mode's wall rows for every mode */
                    *b = -kx - ky - a0 - c0;
                    if (k == S->kstr3) *b -= a0;
                    if (k == S->kend3) *b += top_dir ? -c0 : c0;
                    if (k == S->kstr3) V1(aa, k) = 0.0;
                    if (k == S->kend3) V1(cc, k) = 0.0;
                    continue;
                }
                *b = -kx - ky - V1(aa, k) - V1(cc, k);
                if (k == S->kstr3) { *b -= V1(aa, k); V1(aa, k) = 0.0; }
                if (k == S->kend3) {
                    if (top_dir) { *b -= V1(cc, k); V1(cc, k) = 0.0; }
                    else                          { *b += V1(cc, k); V1(cc, k) = 0.0; }
                }
            }
    }
    free(kxv); free(kyv);
}

/* get_w_eqn_coeff, src/module_poisson_solver.f90:287-374 */
static void coeff_w(ho_state *S)
{
    const ho_params *P = &S->prm;
    double *aa = S->aa_w, *cc = S->cc_w;
    cplx *bb = S->bb_w;
    for (int k = 1; k <= S->nka; ++k) { V1(aa, k) = 0; V1(cc, k) = 0; }
    memset(bb, 0, sizeof(cplx) * (size_t)P->nx * P->ny * S->nka);
    cplx *kxv = calloc_c(P->nx), *kyv = calloc_c(P->ny);
    helm_symbols(S, P->nu, kxv, kyv);
    for (int k = S->kstr3; k <= S->kend3 - 1; ++k) {
        V1(aa, k) = -0.5 * P->dt * P->nu / V1(S->dz, k) / V1(S->dz_t, k);
        V1(cc, k) = -0.5 * P->dt * P->nu / V1(S->dz, k + 1) / V1(S->dz_t, k);
        const double a0 = V1(aa, k), c0 = V1(cc, k);
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i) {
                cplx kx = kxv[i - 1], ky = kyv[j - 1];
                B3(S, bb, i, j, k) = P->wall_rows == 1 ? -kx - ky - a0 - c0 : -kx - ky - V1(aa, k) - V1(cc, k);
                if (k == S->kstr3) V1(aa, k) = 0.0;
                if (k == S->kend3 - 1) V1(cc, k) = 0.0;
            }
    }
    free(kxv); free(kyv);
}

void ho_coefficients(ho_state *S)
{
    coeff_uvt(S, S->aa_u, S->bb_u, S->cc_u, S->prm.nu, 0);
    coeff_w(S);
    coeff_p(S);
    coeff_uvt(S, S->aa_t, S->bb_t, S->cc_t, S->prm.alpha, 1);
}

/* ---- A5: boundary conditions, src/module_boundary.f90 ------------------ */
/* comm_bound on one rank: periodic copy of two ghost layers in x then y,
 * over k = kstr3-1..kend3+1 (src/module_boundary.f90:361-487) */
static void comm_bound(ho_state *S, double *a)
{
    int is = S->istr3, ie = S->iend3, js = S->jstr3, je = S->jend3;
    for (int k = S->kstr3 - 1; k <= S->kend3 + 1; ++k) {
        for (int j = js; j <= je; ++j) {
            F3(S, a, is - 1, j, k) = F3(S, a, ie, j, k);
            F3(S, a, is - 2, j, k) = F3(S, a, ie - 1, j, k);
            F3(S, a, ie + 1, j, k) = F3(S, a, is, j, k);
            F3(S, a, ie + 2, j, k) = F3(S, a, is + 1, j, k);
        }
        for (int i = is - 2; i <= ie + 2; ++i) {
            F3(S, a, i, js - 1, k) = F3(S, a, i, je, k);
            F3(S, a, i, js - 2, k) = F3(S, a, i, je - 1, k);
            F3(S, a, i, je + 1, k) = F3(S, a, i, js, k);
            F3(S, a, i, je + 2, k) = F3(S, a, i, js + 1, k);
        }
    }
}

/* update_boundary_uvw, src/module_boundary.f90:33-128 */
static void bc_uvw(ho_state *S)
{
    const ho_params *P = &S->prm;
    int ks = S->kstr3, ke = S->kend3;
    for (int j = S->jstr3; j <= S->jend3; ++j)
        for (int i = S->istr3; i <= S->iend3; ++i) {
            if (P->ochannel == 0) {
                F3(S, S->u, i, j, ke + 1) = -F3(S, S->u, i, j, ke) - 2.0 * P->u_mrf;
                F3(S, S->v, i, j, ke + 1) = -F3(S, S->v, i, j, ke);
            } else {
                F3(S, S->u, i, j, ke + 1) = F3(S, S->u, i, j, ke);
                F3(S, S->v, i, j, ke + 1) = F3(S, S->v, i, j, ke);
            }
            F3(S, S->u, i, j, ks - 1) = -F3(S, S->u, i, j, ks) - 2.0 * P->u_mrf;
            F3(S, S->v, i, j, ks - 1) = -F3(S, S->v, i, j, ks);
            F3(S, S->w, i, j, ke) = 0.0;
            F3(S, S->w, i, j, ke + 1) = -F3(S, S->w, i, j, ke - 1);
            F3(S, S->w, i, j, ks - 1) = 0.0;
        }
    if (P->cds != 3) { comm_bound(S, S->u); comm_bound(S, S->v); comm_bound(S, S->w); }
}

/* update_boundary_p, src/module_boundary.f90:133-162 */
static void bc_p(ho_state *S, double *p)
{
    for (int j = S->jstr3; j <= S->jend3; ++j)
        for (int i = S->istr3; i <= S->iend3; ++i) {
            F3(S, p, i, j, S->kend3 + 1) = F3(S, p, i, j, S->kend3);
            F3(S, p, i, j, S->kstr3 - 1) = F3(S, p, i, j, S->kstr3);
        }
    if (S->prm.cds != 3) comm_bound(S, p);
}

/* update_boundary_t, src/module_boundary.f90:167-196.  With tbc_top = 1
 * the top ghost carries the gradient instead: t(ke+1) = t(ke) + g dz_t(ke). */
static void bc_t(ho_state *S)
{
    const double gdz = S->prm.dtdz_top * V1(S->dz_t, S->kend3);
    for (int j = S->jstr3; j <= S->jend3; ++j)
        for (int i = S->istr3; i <= S->iend3; ++i) {
            const double te = F3(S, S->t, i, j, S->kend3);
            F3(S, S->t, i, j, S->kend3 + 1) = S->prm.tbc_top == 1 ? te + gdz : 2.0 * S->prm.ttop - te;
            F3(S, S->t, i, j, S->kstr3 - 1) = 2.0 * S->prm.tbot - F3(S, S->t, i, j, S->kstr3);
        }
    if (S->prm.cds != 3) comm_bound(S, S->t);
}

/* update_boundary_uvwhxyz, src/module_boundary.f90:202-305 */
static void bc_uvwh(ho_state *S)
{
    const ho_params *P = &S->prm;
    int ks = S->kstr3, ke = S->kend3;
    double um = P->u_mrf;
    for (int j = S->jstr3; j <= S->jend3; ++j)
        for (int i = S->istr3; i <= S->iend3; ++i) {
            if (P->ochannel == 0) {
                F3(S, S->uhx, i, j, ke + 1) = -F3(S, S->uhx, i, j, ke) - 2.0 * um;
                F3(S, S->uhy, i, j, ke + 1) = -F3(S, S->uhy, i, j, ke) - 2.0 * um;
                F3(S, S->uhz, i, j, ke) = -um;
                F3(S, S->uhz, i, j, ke + 1) = -F3(S, S->uhz, i, j, ke - 1) - 2.0 * um;
                F3(S, S->vhx, i, j, ke + 1) = -F3(S, S->vhx, i, j, ke);
                F3(S, S->vhy, i, j, ke + 1) = -F3(S, S->vhy, i, j, ke);
                F3(S, S->vhz, i, j, ke) = 0.0;
                F3(S, S->vhz, i, j, ke + 1) = -F3(S, S->vhz, i, j, ke - 1);
            } else {
                F3(S, S->uhx, i, j, ke + 1) = F3(S, S->uhx, i, j, ke);
                F3(S, S->uhy, i, j, ke + 1) = F3(S, S->uhy, i, j, ke);
                F3(S, S->uhz, i, j, ke) = F3(S, S->u, i, j, ke);
                F3(S, S->uhz, i, j, ke + 1) = F3(S, S->uhz, i, j, ke);
                F3(S, S->vhx, i, j, ke + 1) = F3(S, S->vhx, i, j, ke);
                F3(S, S->vhy, i, j, ke + 1) = F3(S, S->vhy, i, j, ke);
                F3(S, S->vhz, i, j, ke) = F3(S, S->v, i, j, ke);
                F3(S, S->vhz, i, j, ke + 1) = F3(S, S->vhz, i, j, ke);
            }
            F3(S, S->whx, i, j, ke) = 0.0;
            F3(S, S->whx, i, j, ke + 1) = -F3(S, S->whx, i, j, ke - 1);
            F3(S, S->why, i, j, ke) = 0.0;
            F3(S, S->why, i, j, ke + 1) = -F3(S, S->why, i, j, ke - 1);

            F3(S, S->uhx, i, j, ks - 1) = -F3(S, S->uhx, i, j, ks) - 2.0 * um;
            F3(S, S->uhy, i, j, ks - 1) = -F3(S, S->uhy, i, j, ks) - 2.0 * um;
            F3(S, S->uhz, i, j, ks - 1) = -um;
            F3(S, S->vhx, i, j, ks - 1) = -F3(S, S->vhx, i, j, ks);
            F3(S, S->vhy, i, j, ks - 1) = -F3(S, S->vhy, i, j, ks);
            F3(S, S->vhz, i, j, ks - 1) = 0.0;
            F3(S, S->whx, i, j, ks - 1) = 0.0;
            F3(S, S->why, i, j, ks - 1) = 0.0;
        }
    if (P->cds != 3) {
        double *h[] = {S->uhx, S->uhy, S->uhz, S->vhx, S->vhy, S->vhz, S->whx, S->why};
        for (int q = 0; q < 8; ++q) comm_bound(S, h[q]);
    }
}

/* update_boundary_thxyz, src/module_boundary.f90:311-355.  With tbc_top = 1
 * the top values follow the gradient ghost: thx, thy at ke+1 shift by
 * g dz_t(ke), thz(ke) is the face value t(ke) + g dz_t(ke)/2. */
static void bc_th(ho_state *S)
{
    double tt = S->prm.ttop, tb = S->prm.tbot;
    int ks = S->kstr3, ke = S->kend3;
    const int neu = S->prm.tbc_top == 1;
    const double gdz = S->prm.dtdz_top * V1(S->dz_t, ke);
    for (int j = S->jstr3; j <= S->jend3; ++j)
        for (int i = S->istr3; i <= S->iend3; ++i) {
            if (neu) {
                F3(S, S->thx, i, j, ke + 1) = F3(S, S->thx, i, j, ke) + gdz;
                F3(S, S->thy, i, j, ke + 1) = F3(S, S->thy, i, j, ke) + gdz;
                F3(S, S->thz, i, j, ke) = F3(S, S->t, i, j, ke) + 0.5 * gdz;
                F3(S, S->thz, i, j, ke + 1) = 2.0 * F3(S, S->thz, i, j, ke) - F3(S, S->thz, i, j, ke - 1);
            } else {
                F3(S, S->thx, i, j, ke + 1) = -F3(S, S->thx, i, j, ke) + 2.0 * tt;
                F3(S, S->thy, i, j, ke + 1) = -F3(S, S->thy, i, j, ke) + 2.0 * tt;
                F3(S, S->thz, i, j, ke) = tt;
                F3(S, S->thz, i, j, ke + 1) = -F3(S, S->thz, i, j, ke - 1) + 2.0 * tt;
            }
            F3(S, S->thx, i, j, ks - 1) = -F3(S, S->thx, i, j, ks) + 2.0 * tb;
            F3(S, S->thy, i, j, ks - 1) = -F3(S, S->thy, i, j, ks) + 2.0 * tb;
            F3(S, S->thz, i, j, ks - 1) = tb;
        }
    if (S->prm.cds != 3) { comm_bound(S, S->thx); comm_bound(S, S->thy); comm_bound(S, S->thz); }
}

/* ---- A4: interpolation, src/module_tools.f90:410-512 -------------------- */
static void interp_uvw(ho_state *S)
{
    const ho_params *P = &S->prm;
    double *u = S->u, *v = S->v, *w = S->w;
#pragma omp parallel for
    for (int k = S->kstr3; k <= S->kend3; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) {
                if (P->cds == 1) {
                    F3(S, S->uhx, i, j, k) = 0.5 * (F3(S, u, i, j, k) + F3(S, u, i - 1, j, k));
                    F3(S, S->uhy, i, j, k) = 0.5 * (F3(S, u, i, j + 1, k) + F3(S, u, i, j, k));
                    F3(S, S->vhx, i, j, k) = 0.5 * (F3(S, v, i + 1, j, k) + F3(S, v, i, j, k));
                    F3(S, S->vhy, i, j, k) = 0.5 * (F3(S, v, i, j, k) + F3(S, v, i, j - 1, k));
                    F3(S, S->whx, i, j, k) = 0.5 * (F3(S, w, i + 1, j, k) + F3(S, w, i, j, k));
                    F3(S, S->why, i, j, k) = 0.5 * (F3(S, w, i, j + 1, k) + F3(S, w, i, j, k));
                } else if (P->cds == 2) {
                    F3(S, S->uhx, i, j, k) = (-F3(S, u, i + 1, j, k) + 9.0 * F3(S, u, i, j, k) +
                                              9.0 * F3(S, u, i - 1, j, k) - F3(S, u, i - 2, j, k)) / 16.0;
                    F3(S, S->uhy, i, j, k) = (-F3(S, u, i, j + 2, k) + 9.0 * F3(S, u, i, j + 1, k) +
                                              9.0 * F3(S, u, i, j, k) - F3(S, u, i, j - 1, k)) / 16.0;
                    F3(S, S->vhx, i, j, k) = (-F3(S, v, i + 2, j, k) + 9.0 * F3(S, v, i + 1, j, k) +
                                              9.0 * F3(S, v, i, j, k) - F3(S, v, i - 1, j, k)) / 16.0;
                    F3(S, S->vhy, i, j, k) = (-F3(S, v, i, j + 1, k) + 9.0 * F3(S, v, i, j, k) +
                                              9.0 * F3(S, v, i, j - 1, k) - F3(S, v, i, j - 2, k)) / 16.0;
                    F3(S, S->whx, i, j, k) = (-F3(S, w, i + 2, j, k) + 9.0 * F3(S, w, i + 1, j, k) +
                                              9.0 * F3(S, w, i, j, k) - F3(S, w, i - 1, j, k)) / 16.0;
                    F3(S, S->why, i, j, k) = (-F3(S, w, i, j + 2, k) + 9.0 * F3(S, w, i, j + 1, k) +
                                              9.0 * F3(S, w, i, j, k) - F3(S, w, i, j - 1, k)) / 16.0;
                }
                double lt = V1(S->l_t, k);
                F3(S, S->uhz, i, j, k) = F3(S, u, i, j, k + 1) * lt + F3(S, u, i, j, k) * (1.0 - lt);
                F3(S, S->vhz, i, j, k) = F3(S, v, i, j, k + 1) * lt + F3(S, v, i, j, k) * (1.0 - lt);
            }
}

static void interp_t(ho_state *S)
{
    const ho_params *P = &S->prm;
    double *t = S->t;
#pragma omp parallel for
    for (int k = S->kstr3; k <= S->kend3; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) {
                if (P->cds == 1) {
                    F3(S, S->thx, i, j, k) = 0.5 * (F3(S, t, i + 1, j, k) + F3(S, t, i, j, k));
                    F3(S, S->thy, i, j, k) = 0.5 * (F3(S, t, i, j + 1, k) + F3(S, t, i, j, k));
                } else if (P->cds == 2) {
                    F3(S, S->thx, i, j, k) = (-F3(S, t, i + 2, j, k) + 9.0 * F3(S, t, i + 1, j, k) +
                                              9.0 * F3(S, t, i, j, k) - F3(S, t, i - 1, j, k)) / 16.0;
                    F3(S, S->thy, i, j, k) = (-F3(S, t, i, j + 2, k) + 9.0 * F3(S, t, i, j + 1, k) +
                                              9.0 * F3(S, t, i, j, k) - F3(S, t, i, j - 1, k)) / 16.0;
                }
                double lt = V1(S->l_t, k);
                F3(S, S->thz, i, j, k) = F3(S, t, i, j, k + 1) * lt + F3(S, t, i, j, k) * (1.0 - lt);
            }
}

/* ---- A3: update_dpx_drive / calculate_ubar, src/module_parameters.f90:594-613,513-546 */
static void update_dpx(ho_state *S)
{
    if (S->prm.dp_opt == 1) S->dpx_drive = 1.0;
    else if (S->prm.dp_opt == 2) S->dpx_drive = S->dpx_drive - (S->ubar + S->prm.u_mrf - 1.0) * S->dpx_coef;
}

void ho_calc_ubar(ho_state *S)
{
    double s = 0.0;
    for (int k = S->kstr3; k <= S->kend3; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) s = s + F3(S, S->u, i, j, k) * V1(S->dz, k);
    s = s / S->prm.lz / S->prm.nx / S->prm.ny;
    S->ubar = s; /* one rank: ALLREDUCE sum / nprc is the identity */
}


/* ---- A6: get_momentum_cnvdiff, src/module_cnvdiff.f90:35-374 ------------ */
#define SQ(x) ((x) * (x))
static void momentum(ho_state *S)
{
    const ho_params *P = &S->prm;
    double *u = S->u, *v = S->v, *w = S->w, *pp = S->p, *t = S->t;
    double *uhx = S->uhx, *uhy = S->uhy, *uhz = S->uhz, *vhx = S->vhx, *vhy = S->vhy;
    double *vhz = S->vhz, *whx = S->whx, *why = S->why;
    double dx = S->dx, dy = S->dy, dx2 = S->dx2, dy2 = S->dy2, dt = P->dt, nu = P->nu;
    double ra = S->rk_a, rb = S->rk_b, rc = S->rk_c, um = P->u_mrf;
    update_dpx(S);                                            /* :51 */
    double dpx = S->dpx_drive;
    /* U equation :55-161 */
#pragma omp parallel for
    for (int k = S->kstr3; k <= S->kend3; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) {
                double uux = 0, uvy = 0, ux = 0, uxx = 0, uyy = 0, dpdx = 0;
                if (P->cds == 1) {
                    uux = (SQ(F3(S, uhx, i + 1, j, k)) - SQ(F3(S, uhx, i, j, k))) / dx;
                    uvy = (F3(S, uhy, i, j, k) * F3(S, vhx, i, j, k) - F3(S, uhy, i, j - 1, k) * F3(S, vhx, i, j - 1, k)) / dy;
                    ux = (F3(S, uhx, i + 1, j, k) - F3(S, uhx, i, j, k)) / dx;
                    uxx = (F3(S, u, i + 1, j, k) - 2.0 * F3(S, u, i, j, k) + F3(S, u, i - 1, j, k)) / dx2;
                    uyy = (F3(S, u, i, j + 1, k) - 2.0 * F3(S, u, i, j, k) + F3(S, u, i, j - 1, k)) / dy2;
                    dpdx = (F3(S, pp, i + 1, j, k) - F3(S, pp, i, j, k)) / dx;
                } else if (P->cds == 2) {
                    uux = (-SQ(F3(S, uhx, i + 2, j, k)) + 27.0 * SQ(F3(S, uhx, i + 1, j, k)) -
                           27.0 * SQ(F3(S, uhx, i, j, k)) + SQ(F3(S, uhx, i - 1, j, k))) / 24.0 / dx;
                    uvy = (-F3(S, uhy, i, j + 1, k) * F3(S, vhx, i, j + 1, k) + 27.0 * F3(S, uhy, i, j, k) * F3(S, vhx, i, j, k) -
                           27.0 * F3(S, uhy, i, j - 1, k) * F3(S, vhx, i, j - 1, k) + F3(S, uhy, i, j - 2, k) * F3(S, vhx, i, j - 2, k)) / 24.0 / dy;
                    ux = (-F3(S, uhx, i + 2, j, k) + 27.0 * F3(S, uhx, i + 1, j, k) - 27.0 * F3(S, uhx, i, j, k) +
                          F3(S, uhx, i - 1, j, k)) / 24.0 / dx;
                    uxx = (-F3(S, u, i + 2, j, k) + 16.0 * F3(S, u, i + 1, j, k) - 30.0 * F3(S, u, i, j, k) +
                           16.0 * F3(S, u, i - 1, j, k) - F3(S, u, i - 2, j, k)) / 12.0 / dx2;
                    uyy = (-F3(S, u, i, j + 2, k) + 16.0 * F3(S, u, i, j + 1, k) - 30.0 * F3(S, u, i, j, k) +
                           16.0 * F3(S, u, i, j - 1, k) - F3(S, u, i, j - 2, k)) / 12.0 / dy2;
                    dpdx = (-F3(S, pp, i + 2, j, k) + 27.0 * F3(S, pp, i + 1, j, k) - 27.0 * F3(S, pp, i, j, k) +
                            F3(S, pp, i - 1, j, k)) / 24.0 / dx;
                }
                double uwz = (F3(S, uhz, i, j, k) * F3(S, whx, i, j, k) - F3(S, uhz, i, j, k - 1) * F3(S, whx, i, j, k - 1)) / V1(S->dz, k);
                double uzz = ((F3(S, u, i, j, k + 1) - F3(S, u, i, j, k)) / V1(S->dz_t, k) -
                              (F3(S, u, i, j, k) - F3(S, u, i, j, k - 1)) / V1(S->dz_b, k)) / V1(S->dz, k);
                double f;
                if (P->dp_opt == 3)
                    f = -uux - uvy - uwz - um * ux + P->fc * (0.5 * (F3(S, vhx, i, j, k) + F3(S, vhx, i, j - 1, k)) - P->vg);
                else
                    f = -uux - uvy - uwz + dpx - um * ux;
                F3(S, S->ff, i, j, k) = f;
                F3(S, S->qu, i, j, k) = F3(S, u, i, j, k) + dt * (ra * f + rb * F3(S, S->ff0, i, j, k)) +
                                        rc * 0.5 * dt * nu * (uxx + uyy + uzz) - rc * dt * dpdx;
            }
    /* V equation :166-270 */
#pragma omp parallel for
    for (int k = S->kstr3; k <= S->kend3; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) {
                double vux = 0, vvy = 0, vx = 0, vxx = 0, vyy = 0, dpdy = 0;
                if (P->cds == 1) {
                    vux = (F3(S, vhx, i, j, k) * F3(S, uhy, i, j, k) - F3(S, vhx, i - 1, j, k) * F3(S, uhy, i - 1, j, k)) / dx;
                    vvy = (SQ(F3(S, vhy, i, j + 1, k)) - SQ(F3(S, vhy, i, j, k))) / dy;
                    vx = (F3(S, vhx, i, j, k) - F3(S, vhx, i - 1, j, k)) / dx;
                    vxx = (F3(S, v, i + 1, j, k) - 2.0 * F3(S, v, i, j, k) + F3(S, v, i - 1, j, k)) / dx2;
                    vyy = (F3(S, v, i, j + 1, k) - 2.0 * F3(S, v, i, j, k) + F3(S, v, i, j - 1, k)) / dy2;
                    dpdy = (F3(S, pp, i, j + 1, k) - F3(S, pp, i, j, k)) / dy;
                } else if (P->cds == 2) {
                    vux = (-F3(S, vhx, i + 1, j, k) * F3(S, uhy, i + 1, j, k) + 27.0 * F3(S, vhx, i, j, k) * F3(S, uhy, i, j, k) -
                           27.0 * F3(S, vhx, i - 1, j, k) * F3(S, uhy, i - 1, j, k) + F3(S, vhx, i - 2, j, k) * F3(S, uhy, i - 2, j, k)) / 24.0 / dx;
                    vvy = (-SQ(F3(S, vhy, i, j + 2, k)) + 27.0 * SQ(F3(S, vhy, i, j + 1, k)) -
                           27.0 * SQ(F3(S, vhy, i, j, k)) + SQ(F3(S, vhy, i, j - 1, k))) / 24.0 / dy;
                    vx = (-F3(S, vhx, i + 1, j, k) + 27.0 * F3(S, vhx, i, j, k) - 27.0 * F3(S, vhx, i - 1, j, k) +
                          F3(S, vhx, i - 2, j, k)) / 24.0 / dx;
                    vxx = (-F3(S, v, i + 2, j, k) + 16.0 * F3(S, v, i + 1, j, k) - 30.0 * F3(S, v, i, j, k) +
                           16.0 * F3(S, v, i - 1, j, k) - F3(S, v, i - 2, j, k)) / 12.0 / dx2;
                    vyy = (-F3(S, v, i, j + 2, k) + 16.0 * F3(S, v, i, j + 1, k) - 30.0 * F3(S, v, i, j, k) +
                           16.0 * F3(S, v, i, j - 1, k) - F3(S, v, i, j - 2, k)) / 12.0 / dy2;
                    dpdy = (-F3(S, pp, i, j + 2, k) + 27.0 * F3(S, pp, i, j + 1, k) - 27.0 * F3(S, pp, i, j, k) +
                            F3(S, pp, i, j - 1, k)) / 24.0 / dy;
                }
                double vwz = (F3(S, vhz, i, j, k) * F3(S, why, i, j, k) - F3(S, vhz, i, j, k - 1) * F3(S, why, i, j, k - 1)) / V1(S->dz, k);
                double vzz = ((F3(S, v, i, j, k + 1) - F3(S, v, i, j, k)) / V1(S->dz_t, k) -
                              (F3(S, v, i, j, k) - F3(S, v, i, j, k - 1)) / V1(S->dz_b, k)) / V1(S->dz, k);
                double g;
                if (P->dp_opt == 3)
                    g = -vux - vvy - vwz - um * vx + P->fc * (P->ug - 0.5 * (F3(S, uhy, i, j, k) + F3(S, uhy, i - 1, j, k)) - um);
                else
                    g = -vux - vvy - vwz - um * vx;
                F3(S, S->gg, i, j, k) = g;
                F3(S, S->qv, i, j, k) = F3(S, v, i, j, k) + dt * (ra * g + rb * F3(S, S->gg0, i, j, k)) +
                                        rc * 0.5 * dt * nu * (vxx + vyy + vzz) - rc * dt * dpdy;
            }
    /* W equation :274-370 (k < kend3) */
#pragma omp parallel for
    for (int k = S->kstr3; k <= S->kend3 - 1; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) {
                double wux = 0, wvy = 0, wx = 0, wxx = 0, wyy = 0;
                if (P->cds == 1) {
                    wux = (F3(S, whx, i, j, k) * F3(S, uhz, i, j, k) - F3(S, whx, i - 1, j, k) * F3(S, uhz, i - 1, j, k)) / dx;
                    wvy = (F3(S, why, i, j, k) * F3(S, vhz, i, j, k) - F3(S, why, i, j - 1, k) * F3(S, vhz, i, j - 1, k)) / dy;
                    wx = (F3(S, whx, i, j, k) - F3(S, whx, i - 1, j, k)) / dx;
                    wxx = (F3(S, w, i + 1, j, k) - 2.0 * F3(S, w, i, j, k) + F3(S, w, i - 1, j, k)) / dx2;
                    wyy = (F3(S, w, i, j + 1, k) - 2.0 * F3(S, w, i, j, k) + F3(S, w, i, j - 1, k)) / dy2;
                } else if (P->cds == 2) {
                    wux = (-F3(S, whx, i + 1, j, k) * F3(S, uhz, i + 1, j, k) + 27.0 * F3(S, whx, i, j, k) * F3(S, uhz, i, j, k) -
                           27.0 * F3(S, whx, i - 1, j, k) * F3(S, uhz, i - 1, j, k) + F3(S, whx, i - 2, j, k) * F3(S, uhz, i - 2, j, k)) / 24.0 / dx;
                    wvy = (-F3(S, why, i, j + 1, k) * F3(S, vhz, i, j + 1, k) + 27.0 * F3(S, why, i, j, k) * F3(S, vhz, i, j, k) -
                           27.0 * F3(S, why, i, j - 1, k) * F3(S, vhz, i, j - 1, k) + F3(S, why, i, j - 2, k) * F3(S, vhz, i, j - 2, k)) / 24.0 / dy;
                    wx = (-F3(S, whx, i + 1, j, k) + 27.0 * F3(S, whx, i, j, k) - 27.0 * F3(S, whx, i - 1, j, k) +
                          F3(S, whx, i - 2, j, k)) / 24.0 / dx;
                    wxx = (-F3(S, w, i + 2, j, k) + 16.0 * F3(S, w, i + 1, j, k) - 30.0 * F3(S, w, i, j, k) +
                           16.0 * F3(S, w, i - 1, j, k) - F3(S, w, i - 2, j, k)) / 12.0 / dx2;
                    wyy = (-F3(S, w, i, j + 2, k) + 16.0 * F3(S, w, i, j + 1, k) - 30.0 * F3(S, w, i, j, k) +
                           16.0 * F3(S, w, i, j - 1, k) - F3(S, w, i, j - 2, k)) / 12.0 / dy2;
                }
                double a1 = SQ(0.5 * (F3(S, w, i, j, k) + F3(S, w, i, j, k + 1)));
                double a2 = SQ(0.5 * (F3(S, w, i, j, k) + F3(S, w, i, j, k - 1)));
                double wwz = (a1 - a2) / V1(S->dz_t, k);
                double wzz = ((F3(S, w, i, j, k + 1) - F3(S, w, i, j, k)) / V1(S->dz, k + 1) -
                              (F3(S, w, i, j, k) - F3(S, w, i, j, k - 1)) / V1(S->dz, k)) / V1(S->dz_t, k);
                double h;
                if (P->isscalar == 1)
                    h = -wux - wvy - wwz - um * wx +
                        S->got * ((1.0 - V1(S->l_t, k)) * F3(S, t, i, j, k) + V1(S->l_t, k) * F3(S, t, i, j, k + 1) - P->t_ref);
                else
                    h = -wux - wvy - wwz - um * wx;
                double dpdz = (F3(S, pp, i, j, k + 1) - F3(S, pp, i, j, k)) / V1(S->dz_t, k);
                F3(S, S->hh, i, j, k) = h;
                F3(S, S->qw, i, j, k) = F3(S, w, i, j, k) + dt * (ra * h + rb * F3(S, S->hh0, i, j, k)) +
                                        rc * 0.5 * dt * nu * (wxx + wyy + wzz) - rc * dt * dpdz;
            }
}

/* ---- A7: get_temperature_cnvdiff, src/module_cnvdiff.f90:381-473 ------- */
static void temperature(ho_state *S, double *qt)
{
    const ho_params *P = &S->prm;
    double *u = S->u, *v = S->v, *w = S->w, *t = S->t, *thx = S->thx, *thy = S->thy, *thz = S->thz;
    double dx = S->dx, dy = S->dy, dx2 = S->dx2, dy2 = S->dy2, dt = P->dt;
#pragma omp parallel for
    for (int k = S->kstr3; k <= S->kend3; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) {
                double tux = 0, tvy = 0, tx = 0, txx = 0, tyy = 0;
                if (P->cds == 1) {
                    tux = (F3(S, thx, i, j, k) * F3(S, u, i, j, k) - F3(S, thx, i - 1, j, k) * F3(S, u, i - 1, j, k)) / dx;
                    tvy = (F3(S, thy, i, j, k) * F3(S, v, i, j, k) - F3(S, thy, i, j - 1, k) * F3(S, v, i, j - 1, k)) / dy;
                    tx = (F3(S, thx, i, j, k) - F3(S, thx, i - 1, j, k)) / dx;
                    txx = (F3(S, t, i + 1, j, k) - 2.0 * F3(S, t, i, j, k) + F3(S, t, i - 1, j, k)) / dx2;
                    tyy = (F3(S, t, i, j + 1, k) - 2.0 * F3(S, t, i, j, k) + F3(S, t, i, j - 1, k)) / dy2;
                } else if (P->cds == 2) {
                    tux = (-F3(S, thx, i + 1, j, k) * F3(S, u, i + 1, j, k) + 27.0 * F3(S, thx, i, j, k) * F3(S, u, i, j, k) -
                           27.0 * F3(S, thx, i - 1, j, k) * F3(S, u, i - 1, j, k) + F3(S, thx, i - 2, j, k) * F3(S, u, i - 2, j, k)) / 24.0 / dx;
                    tvy = (-F3(S, thy, i, j + 1, k) * F3(S, v, i, j + 1, k) + 27.0 * F3(S, thy, i, j, k) * F3(S, v, i, j, k) -
                           27.0 * F3(S, thy, i, j - 1, k) * F3(S, v, i, j - 1, k) + F3(S, thy, i, j - 2, k) * F3(S, v, i, j - 2, k)) / 24.0 / dy;
                    tx = (-F3(S, thx, i + 1, j, k) + 27.0 * F3(S, thx, i, j, k) - 27.0 * F3(S, thx, i - 1, j, k) +
                          F3(S, thx, i - 2, j, k)) / 24.0 / dx;
                    txx = (-F3(S, t, i + 2, j, k) + 16.0 * F3(S, t, i + 1, j, k) - 30.0 * F3(S, t, i, j, k) +
                           16.0 * F3(S, t, i - 1, j, k) - F3(S, t, i - 2, j, k)) / 12.0 / dx2;
                    tyy = (-F3(S, t, i, j + 2, k) + 16.0 * F3(S, t, i, j + 1, k) - 30.0 * F3(S, t, i, j, k) +
                           16.0 * F3(S, t, i, j - 1, k) - F3(S, t, i, j - 2, k)) / 12.0 / dy2;
                }
                double twz = (F3(S, thz, i, j, k) * F3(S, w, i, j, k) - F3(S, thz, i, j, k - 1) * F3(S, w, i, j, k - 1)) / V1(S->dz, k);
                double tzz = ((F3(S, t, i, j, k + 1) - F3(S, t, i, j, k)) / V1(S->dz_t, k) -
                              (F3(S, t, i, j, k) - F3(S, t, i, j, k - 1)) / V1(S->dz_b, k)) / V1(S->dz, k);
                double s = -tux - tvy - twz - P->u_mrf * tx;
                F3(S, S->ss, i, j, k) = s;
                F3(S, qt, i, j, k) = F3(S, t, i, j, k) + dt * (S->rk_a * s + S->rk_b * F3(S, S->ss0, i, j, k)) +
                                     S->rk_c * 0.5 * dt * P->alpha * (txx + tyy + tzz);
            }
}

/* ---- A11: mpi_2d_fft on one rank, src/module_tools.f90:518-700 --------- */
static void zero_nyquist(ho_state *S, cplx *c)
{
    const ho_params *P = &S->prm;
    for (int k = 1; k <= P->nz; ++k)
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i) {
                int l = lwave(i, P->nx), m = lwave(j, P->ny);
                if (l == -P->nx / 2) C3(S, c, i, j, k) = 0;
                if (m == -P->ny / 2) C3(S, c, i, j, k) = 0;
            }
}

void ho_fft2d(ho_state *S, cplx *c, int direction)
{
    const ho_params *P = &S->prm;
    if (direction == 1) zero_nyquist(S, c);
    int sign = direction == -1 ? -1 : +1;  /* FFTW_FORWARD = -1, FFTW_BACKWARD = +1 */
    ho_plan py = plan_make(P->ny, sign), px = plan_make(P->nx, sign);
    /* y pencil: FFT along j for every (i,k); forward divides by ny (:581-594) */
#pragma omp parallel
    {
        int nmax = P->nx > P->ny ? P->nx : P->ny;
        cplx *a = calloc_c(nmax), *b = calloc_c(nmax);
#pragma omp for
        for (int k = 1; k <= P->nz; ++k)
            for (int i = 1; i <= P->nx; ++i) {
                plan_exec(&py, &C3(S, c, i, 1, k), (size_t)P->nx, a, b);
                if (direction == -1)
                    for (int j = 1; j <= P->ny; ++j) C3(S, c, i, j, k) = C3(S, c, i, j, k) / (double)P->ny;
            }
        /* x pencil: FFT along i for every (j,k); forward divides by nx (:620-633) */
#pragma omp for
        for (int k = 1; k <= P->nz; ++k)
            for (int j = 1; j <= P->ny; ++j) {
                plan_exec(&px, &C3(S, c, 1, j, k), 1, a, b);
                if (direction == -1)
                    for (int i = 1; i <= P->nx; ++i) C3(S, c, i, j, k) = C3(S, c, i, j, k) / (double)P->nx;
            }
        free(a); free(b);
    }
    free(py.tw); free(px.tw);
    if (direction == -1) zero_nyquist(S, c);
}

/* ---- A9: assign_abc / assign_abc_p, src/module_navier_stokes.f90:523-619 */
static void assign_abc(ho_state *S, const double *aa_in, const cplx *bb_in, const double *cc_in, int is_p)
{
    const ho_params *P = &S->prm;
    size_t plane = (size_t)P->nx * P->ny;
    for (int k = 1; k <= S->nka; ++k)
        for (size_t q = 0; q < plane; ++q) {
            size_t o = (size_t)(k - 1) * plane + q;
            if (P->dts == 2) {
                S->aa[o] = S->rk_c * V1(aa_in, k);
                S->cc[o] = S->rk_c * V1(cc_in, k);
                S->bb[o] = is_p ? S->rk_c * bb_in[o] : 1.0 + S->rk_c * bb_in[o];
            } else {
                S->aa[o] = V1(aa_in, k);
                S->cc[o] = V1(cc_in, k);
                S->bb[o] = is_p ? bb_in[o] : 1.0 + bb_in[o];
            }
        }
}

/* ---- A10: poisson_solver_fft, src/module_poisson_solver.f90:483-658 ---- */
/* Flag handling and Thomas sweeps shared with poisson_solver_fft_spec
 * (src/module_spectral.f90:602-746, the same statements on a complex
 * source): S->c_src (z-pencil spectrum) -> S->c_var, with aa/bb/cc set by
 * assign_abc.  Levels above kmax3 stay zero. */
static int spectral_thomas(ho_state *S, int zstagger, int delnull, int tbnd, int ubnd, int *kmax3_out)
{
    const ho_params *P = &S->prm;
    int nx = P->nx, ny = P->ny, nz = P->nz, ks = S->kstr3, ke = S->kend3;
    memset(S->c_var, 0, sizeof(cplx) * (size_t)nx * ny * nz);
    int kmax3;
    if (zstagger == 1) kmax3 = ke - 1;
    else if (zstagger == 0) kmax3 = ke;
    else { snprintf(S->err, sizeof S->err, "Invalid zstagger flag!"); return 1; }
    if (delnull == 1) {
        B3(S, S->aa, 1, 1, ks) = 0.0;
        B3(S, S->cc, 1, 1, ks) = 0.0;
        C3(S, S->c_src, 1, 1, 1) = 0;
    } else if (delnull != 0) { snprintf(S->err, sizeof S->err, "delnull flag error!"); return 1; }
    if (tbnd == 1) {
        double aa_t_zstr = -0.5 * S->rk_c * P->dt * P->alpha / V1(S->dz, ks) / V1(S->dz_b, ks);
        C3(S, S->c_src, 1, 1, 1) = C3(S, S->c_src, 1, 1, 1) - 2.0 * aa_t_zstr * P->tbot;
        double cc_t_zend = -0.5 * S->rk_c * P->dt * P->alpha / V1(S->dz, ke) / V1(S->dz_t, ke);
        if (P->tbc_top == 1) /* ghost t + g dz_t: its constant part moves to the source */
            C3(S, S->c_src, 1, 1, nz) = C3(S, S->c_src, 1, 1, nz) - cc_t_zend * P->dtdz_top * V1(S->dz_t, ke);
        else
            C3(S, S->c_src, 1, 1, nz) = C3(S, S->c_src, 1, 1, nz) - 2.0 * cc_t_zend * P->ttop;
    } else if (tbnd != 0) { snprintf(S->err, sizeof S->err, "tbnd flag error!"); return 1; }
    if (ubnd == 1) {
        double aa_u_zstr = -0.5 * S->rk_c * P->dt * P->nu / V1(S->dz, ks) / V1(S->dz_b, ks);
        C3(S, S->c_src, 1, 1, 1) = C3(S, S->c_src, 1, 1, 1) + 2.0 * aa_u_zstr * P->u_mrf;
        if (P->ochannel == 0) {
            double cc_u_zend = -0.5 * S->rk_c * P->dt * P->nu / V1(S->dz, ke) / V1(S->dz_t, ke);
            C3(S, S->c_src, 1, 1, nz) = C3(S, S->c_src, 1, 1, nz) + 2.0 * cc_u_zend * P->u_mrf;
        }
    } else if (ubnd != 0) { snprintf(S->err, sizeof S->err, "ubnd flag error!"); return 1; }
    /* Thomas algorithm :611-640 */
#pragma omp parallel for
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            for (int k = ks + 1; k <= kmax3; ++k) {
                cplx tmp = B3(S, S->aa, i, j, k) / B3(S, S->bb, i, j, k - 1);
                B3(S, S->bb, i, j, k) = B3(S, S->bb, i, j, k) - B3(S, S->cc, i, j, k - 1) * tmp;
                C3(S, S->c_src, i, j, k - 2) = C3(S, S->c_src, i, j, k - 2) - C3(S, S->c_src, i, j, k - 3) * tmp;
            }
            C3(S, S->c_var, i, j, kmax3 - 2) = C3(S, S->c_src, i, j, kmax3 - 2) / B3(S, S->bb, i, j, kmax3);
            for (int k = kmax3 - 1; k >= ks; --k)
                C3(S, S->c_var, i, j, k - 2) = (C3(S, S->c_src, i, j, k - 2) - B3(S, S->cc, i, j, k) * C3(S, S->c_var, i, j, k - 1)) /
                                               B3(S, S->bb, i, j, k);
        }
    *kmax3_out = kmax3;
    return 0;
}

static int poisson_solver_fft(ho_state *S, const double *r_src, double *r_var,
                              int zstagger, int delnull, int tbnd, int ubnd)
{
    const ho_params *P = &S->prm;
    int nx = P->nx, ny = P->ny, nz = P->nz, ks = S->kstr3;
    memset(r_var, 0, sizeof(double) * nfield(S));
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) C3(S, S->c_src, i, j, k) = F3(S, r_src, i + 2, j + 2, k + 2);
    ho_fft2d(S, S->c_src, -1);
    int kmax3;
    if (spectral_thomas(S, zstagger, delnull, tbnd, ubnd, &kmax3)) return 1;
    ho_fft2d(S, S->c_var, 1);
    for (int k = ks; k <= kmax3; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) F3(S, r_var, i + 2, j + 2, k) = creal(C3(S, S->c_var, i, j, k - 2));
    return 0;
}

/* ---- A12: get_p_eqn_src, src/module_poisson_solver.f90:127-176 -------- */
static void p_src(ho_state *S, double *qp)
{
    const ho_params *P = &S->prm;
    double *u = S->u, *v = S->v, *w = S->w;
#pragma omp parallel for
    for (int k = S->kstr3; k <= S->kend3; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) {
                double dudx = 0, dvdy = 0;
                if (P->cds == 1) {
                    dudx = (F3(S, u, i, j, k) - F3(S, u, i - 1, j, k)) / S->dx;
                    dvdy = (F3(S, v, i, j, k) - F3(S, v, i, j - 1, k)) / S->dy;
                } else if (P->cds == 2) {
                    dudx = (-F3(S, u, i + 1, j, k) + 27.0 * F3(S, u, i, j, k) - 27.0 * F3(S, u, i - 1, j, k) + F3(S, u, i - 2, j, k)) / 24.0 / S->dx;
                    dvdy = (-F3(S, v, i, j + 1, k) + 27.0 * F3(S, v, i, j, k) - 27.0 * F3(S, v, i, j - 1, k) + F3(S, v, i, j - 2, k)) / 24.0 / S->dy;
                }
                double dwdz = (F3(S, w, i, j, k) - F3(S, w, i, j, k - 1)) / V1(S->dz, k);
                F3(S, qp, i, j, k) = (dudx + dvdy + dwdz) / P->dt;
            }
}

/* ---- A13: calculate_uvw, src/module_navier_stokes.f90:659-734 ---------- */
static void calc_uvw(ho_state *S)
{
    const ho_params *P = &S->prm;
    double *phi = S->phi, c = S->rk_c * P->dt;
#pragma omp parallel for
    for (int k = S->kstr3; k <= S->kend3; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) {
                if (P->cds == 1) {
                    F3(S, S->u, i, j, k) = F3(S, S->u, i, j, k) - c * (F3(S, phi, i + 1, j, k) - F3(S, phi, i, j, k)) / S->dx;
                    F3(S, S->v, i, j, k) = F3(S, S->v, i, j, k) - c * (F3(S, phi, i, j + 1, k) - F3(S, phi, i, j, k)) / S->dy;
                } else if (P->cds == 2) {
                    F3(S, S->u, i, j, k) = F3(S, S->u, i, j, k) - c * (-F3(S, phi, i + 2, j, k) + 27.0 * F3(S, phi, i + 1, j, k) -
                                                                      27.0 * F3(S, phi, i, j, k) + F3(S, phi, i - 1, j, k)) / 24.0 / S->dx;
                    F3(S, S->v, i, j, k) = F3(S, S->v, i, j, k) - c * (-F3(S, phi, i, j + 2, k) + 27.0 * F3(S, phi, i, j + 1, k) -
                                                                      27.0 * F3(S, phi, i, j, k) + F3(S, phi, i, j - 1, k)) / 24.0 / S->dy;
                }
            }
#pragma omp parallel for
    for (int k = S->kstr3; k <= S->kend3 - 1; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i)
                F3(S, S->w, i, j, k) = F3(S, S->w, i, j, k) - c * (F3(S, phi, i, j, k + 1) - F3(S, phi, i, j, k)) / V1(S->dz_t, k);
}

/* ---- A14: calculate_new_p, src/module_navier_stokes.f90:739-808 -------- */
static void calc_new_p(ho_state *S)
{
    const ho_params *P = &S->prm;
    double *phi = S->phi;
#pragma omp parallel for
    for (int k = S->kstr3; k <= S->kend3; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) {
                double pxx = 0, pyy = 0;
                if (P->cds == 1) {
                    pxx = (F3(S, phi, i + 1, j, k) - 2.0 * F3(S, phi, i, j, k) + F3(S, phi, i - 1, j, k)) / S->dx2;
                    pyy = (F3(S, phi, i, j + 1, k) - 2.0 * F3(S, phi, i, j, k) + F3(S, phi, i, j - 1, k)) / S->dy2;
                } else if (P->cds == 2) {
                    pxx = (-F3(S, phi, i + 2, j, k) + 16.0 * F3(S, phi, i + 1, j, k) - 30.0 * F3(S, phi, i, j, k) +
                           16.0 * F3(S, phi, i - 1, j, k) - F3(S, phi, i - 2, j, k)) / 12.0 / S->dx2;
                    pyy = (-F3(S, phi, i, j + 2, k) + 16.0 * F3(S, phi, i, j + 1, k) - 30.0 * F3(S, phi, i, j, k) +
                           16.0 * F3(S, phi, i, j - 1, k) - F3(S, phi, i, j - 2, k)) / 12.0 / S->dy2;
                }
                double pzz = ((F3(S, phi, i, j, k + 1) - F3(S, phi, i, j, k)) / V1(S->dz_t, k) -
                              (F3(S, phi, i, j, k) - F3(S, phi, i, j, k - 1)) / V1(S->dz_b, k)) / V1(S->dz, k);
                F3(S, S->p, i, j, k) = F3(S, S->p, i, j, k) + F3(S, phi, i, j, k) - S->rk_c * 0.5 * P->dt * P->nu * (pxx + pyy + pzz);
            }
}

/* ---- A17: screen_div_error (values only), src/module_io.f90:85-206 ----- */
static int div_error(ho_state *S, int isfinal)
{
    const ho_params *P = &S->prm;
    double *u = S->u, *v = S->v, *w = S->w, avg = 0.0, mx = 0.0;
    for (int k = S->kstr3; k <= S->kend3; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) {
                double dudx = 0, dvdy = 0;
                if (P->cds == 1) {
                    dudx = (F3(S, u, i, j, k) - F3(S, u, i - 1, j, k)) / S->dx;
                    dvdy = (F3(S, v, i, j, k) - F3(S, v, i, j - 1, k)) / S->dy;
                } else if (P->cds == 2) {
                    dudx = (-F3(S, u, i + 1, j, k) + 27.0 * F3(S, u, i, j, k) - 27.0 * F3(S, u, i - 1, j, k) + F3(S, u, i - 2, j, k)) / 24.0 / S->dx;
                    dvdy = (-F3(S, v, i, j + 1, k) + 27.0 * F3(S, v, i, j, k) - 27.0 * F3(S, v, i, j - 1, k) + F3(S, v, i, j - 2, k)) / 24.0 / S->dy;
                }
                double d = dudx + dvdy + (F3(S, w, i, j, k) - F3(S, w, i, j, k - 1)) / V1(S->dz, k);
                avg = avg + fabs(d);
                if (fabs(d) > mx) mx = d;  /* signed, as stored by the reference (:131-133) */
                if (F3(S, u, i, j, k) != F3(S, u, i, j, k) || F3(S, v, i, j, k) != F3(S, v, i, j, k) ||
                    F3(S, w, i, j, k) != F3(S, w, i, j, k)) {
                    snprintf(S->err, sizeof S->err, "NaN error! %d %d %d", i, j, k);
                    return 1;
                }
                if (isfinal == 1 && fabs(d) > S->div_hist_max) S->div_hist_max = d;
            }
    S->div_avg = avg / P->nx / P->ny / P->nz;
    S->div_max = mx;
    return 0;
}

/* ---- A18: spectral path, src/module_spectral.f90 ------------------------- */
/* Complex (nx, ny, nz) spectra in C3 layout, as the reference holds them
 * (C2C, forward normalised, Nyquist zeroed by mpi_2d_fft). */
static size_t ncplx(const ho_state *S) { return (size_t)S->prm.nx * S->prm.ny * S->prm.nz; }

/* dev_fft_linear_cmplx, src/module_spectral.f90:1274-1340 */
static void dev_linear(ho_state *S, const cplx *in, cplx *out, int ord, int dir)
{
    const ho_params *P = &S->prm;
    for (int k = 1; k <= P->nz; ++k)
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i) {
                int l = lwave(i, P->nx), m = lwave(j, P->ny);
                cplx c = C3(S, in, i, j, k);
                if (ord == 1 && dir == 1) c = c * I * 2.0 * PI / P->lx * l;
                else if (ord == 1 && dir == 2) c = c * I * 2.0 * PI / P->ly * m;
                else if (ord == 2 && dir == 1) c = -c * SQ(2.0 * PI / P->lx * l);
                else c = -c * SQ(2.0 * PI / P->ly * m);
                if (l == -P->nx / 2 || m == -P->ny / 2) c = 0;
                C3(S, out, i, j, k) = c;
            }
}

/* dev_fft_nonlinear_cmplx, src/module_spectral.f90:1061-1150: FFT of the
 * product r1*r2 (r1 interpolated to the w level when weqn=1), then the
 * first derivative in direction dir */
static void dev_nonlinear(ho_state *S, const double *r1, const double *r2, cplx *out, int dir, int weqn)
{
    const ho_params *P = &S->prm;
    for (int k = 1; k <= P->nz; ++k)
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i) {
                double a = F3(S, r1, i + 2, j + 2, k + 2);
                if (weqn == 1) {
                    double lt = V1(S->l_t, k + 2);
                    a = F3(S, r1, i + 2, j + 2, k + 3) * lt + F3(S, r1, i + 2, j + 2, k + 2) * (1.0 - lt);
                }
                C3(S, out, i, j, k) = a * F3(S, r2, i + 2, j + 2, k + 2);
            }
    ho_fft2d(S, out, -1);
    for (int k = 1; k <= P->nz; ++k)
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i) {
                int l = lwave(i, P->nx), m = lwave(j, P->ny);
                cplx c = C3(S, out, i, j, k);
                c = dir == 1 ? c * I * 2.0 * PI / P->lx * l : c * I * 2.0 * PI / P->ly * m;
                if (l == -P->nx / 2 || m == -P->ny / 2) c = 0;
                C3(S, out, i, j, k) = c;
            }
}

/* real_2_cmplx / cmplx_2_real, src/module_spectral.f90:1157-1217 */
static void real_2_cmplx(ho_state *S, const double *r, cplx *c)
{
    const ho_params *P = &S->prm;
    for (int k = 1; k <= P->nz; ++k)
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i) C3(S, c, i, j, k) = F3(S, r, i + 2, j + 2, k + 2);
    ho_fft2d(S, c, -1);
}
static void cmplx_2_real(ho_state *S, const cplx *c, double *r)
{
    const ho_params *P = &S->prm;
    memcpy(S->c_var, c, sizeof(cplx) * ncplx(S));
    ho_fft2d(S, S->c_var, 1);
    for (int k = 1; k <= P->nz; ++k)
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i) F3(S, r, i + 2, j + 2, k + 2) = creal(C3(S, S->c_var, i, j, k));
}

/* spectral_truncation (2/3 rule), src/module_spectral.f90:1222-1268 */
static void truncation(ho_state *S, cplx *c)
{
    const ho_params *P = &S->prm;
    int lc = (int)(P->nx / 3.0), mc = (int)(P->ny / 3.0);
    for (int k = 1; k <= P->nz; ++k)
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i) {
                int l = lwave(i, P->nx), m = lwave(j, P->ny);
                if (l >= lc || l <= -lc || m >= mc || m <= -mc) C3(S, c, i, j, k) = 0;
            }
}

/* out += coef * in over the whole spectrum */
static void caxpy(ho_state *S, cplx *out, cplx coef, const cplx *in)
{
    size_t n = ncplx(S);
    for (size_t q = 0; q < n; ++q) out[q] = out[q] + coef * in[q];
}

/* one physical-space term over the interior, stored into c_var and
 * transformed (the loops of get_*_cnvdiff_spec that end in mpi_2d_fft(c_var,-1)) */
#define PHYS_TERM(S, k_lo, k_hi, EXPR)                                                   \
    do {                                                                                 \
        for (int k = (k_lo); k <= (k_hi); ++k)                                           \
            for (int j = (S)->jstr3; j <= (S)->jend3; ++j)                               \
                for (int i = (S)->istr3; i <= (S)->iend3; ++i)                           \
                    C3(S, (S)->c_var, i - 2, j - 2, k - 2) = (EXPR);                     \
        ho_fft2d(S, (S)->c_var, -1);                                                     \
    } while (0)

/* the advective form r * IFFT(i k c) of the skew-symmetric terms: IFFT into
 * c_src, then the physical product into c_var, transformed */
static void skew_term(ho_state *S, const cplx *c, int dir, const double *r, int wlev)
{
    const ho_params *P = &S->prm;
    dev_linear(S, c, S->c_src, 1, dir);
    ho_fft2d(S, S->c_src, 1);
    for (int k = S->kstr3; k <= S->kend3; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) {
                double a = F3(S, r, i, j, k);
                if (wlev) a = V1(S->l_t, k) * F3(S, r, i, j, k + 1) + (1.0 - V1(S->l_t, k)) * F3(S, r, i, j, k);
                C3(S, S->c_var, i - 2, j - 2, k - 2) = a * creal(C3(S, S->c_src, i - 2, j - 2, k - 2));
            }
    ho_fft2d(S, S->c_var, -1);
    (void)P;
}

/* get_momentum_cnvdiff_spec, src/module_spectral.f90:35-502 */
static void momentum_spec(ho_state *S)
{
    const ho_params *P = &S->prm;
    double *u = S->u, *v = S->v, *w = S->w, *p = S->p, *t = S->t;
    const double coe = P->issk == 1 ? 0.5 : 1.0;
    const double dif = S->rk_c * 0.5 * P->dt * P->nu;
    size_t n = ncplx(S);
    update_dpx(S);
    /* ---- U (:61-201) */
    memcpy(S->c_qu, S->c_u, n * sizeof(cplx));
    dev_linear(S, S->c_u, S->c_var, 2, 1); caxpy(S, S->c_qu, dif, S->c_var);
    dev_linear(S, S->c_u, S->c_var, 2, 2); caxpy(S, S->c_qu, dif, S->c_var);
    PHYS_TERM(S, S->kstr3, S->kend3,
              ((F3(S, u, i, j, k + 1) - F3(S, u, i, j, k)) / V1(S->dz_t, k) -
               (F3(S, u, i, j, k) - F3(S, u, i, j, k - 1)) / V1(S->dz_b, k)) / V1(S->dz, k));
    caxpy(S, S->c_qu, dif, S->c_var);
    dev_linear(S, S->c_p, S->c_var, 1, 1); caxpy(S, S->c_qu, -S->rk_c * P->dt, S->c_var);
    caxpy(S, S->c_qu, S->rk_b * P->dt, S->c_ff);
    memset(S->c_ff, 0, n * sizeof(cplx));
    dev_linear(S, S->c_u, S->c_var, 1, 1); caxpy(S, S->c_ff, -P->u_mrf, S->c_var);
    dev_nonlinear(S, u, u, S->c_var, 1, 0); caxpy(S, S->c_ff, -coe, S->c_var);
    dev_nonlinear(S, u, v, S->c_var, 2, 0); caxpy(S, S->c_ff, -coe, S->c_var);
#define A1U(f) (F3(S, f, i, j, k + 1) * V1(S->l_t, k) + F3(S, f, i, j, k) * (1.0 - V1(S->l_t, k)))
#define A2U(f) (F3(S, f, i, j, k - 1) * V1(S->l_b, k) + F3(S, f, i, j, k) * (1.0 - V1(S->l_b, k)))
    PHYS_TERM(S, S->kstr3, S->kend3,
              1.0 / coe * (P->dp_opt == 3 ? P->fc * (F3(S, v, i, j, k) - P->vg) : S->dpx_drive) -
                  (A1U(u) * F3(S, w, i, j, k) - A2U(u) * F3(S, w, i, j, k - 1)) / V1(S->dz, k));
    caxpy(S, S->c_ff, coe, S->c_var);
    if (P->issk == 1) {
        skew_term(S, S->c_u, 1, u, 0); caxpy(S, S->c_ff, -coe, S->c_var);
        skew_term(S, S->c_u, 2, v, 0); caxpy(S, S->c_ff, -coe, S->c_var);
        PHYS_TERM(S, S->kstr3, S->kend3,
                  (A1U(u) - A2U(u)) / V1(S->dz, k) * (0.5 * (F3(S, w, i, j, k) + F3(S, w, i, j, k - 1))));
        caxpy(S, S->c_ff, -coe, S->c_var);
    }
    caxpy(S, S->c_qu, S->rk_a * P->dt, S->c_ff);
    /* ---- V (:207-339) */
    memcpy(S->c_qv, S->c_v, n * sizeof(cplx));
    dev_linear(S, S->c_v, S->c_var, 2, 1); caxpy(S, S->c_qv, dif, S->c_var);
    dev_linear(S, S->c_v, S->c_var, 2, 2); caxpy(S, S->c_qv, dif, S->c_var);
    PHYS_TERM(S, S->kstr3, S->kend3,
              ((F3(S, v, i, j, k + 1) - F3(S, v, i, j, k)) / V1(S->dz_t, k) -
               (F3(S, v, i, j, k) - F3(S, v, i, j, k - 1)) / V1(S->dz_b, k)) / V1(S->dz, k));
    caxpy(S, S->c_qv, dif, S->c_var);
    dev_linear(S, S->c_p, S->c_var, 1, 2); caxpy(S, S->c_qv, -S->rk_c * P->dt, S->c_var);
    caxpy(S, S->c_qv, S->rk_b * P->dt, S->c_gg);
    memset(S->c_gg, 0, n * sizeof(cplx));
    dev_linear(S, S->c_v, S->c_var, 1, 1); caxpy(S, S->c_gg, -P->u_mrf, S->c_var);
    dev_nonlinear(S, u, v, S->c_var, 1, 0); caxpy(S, S->c_gg, -coe, S->c_var);
    dev_nonlinear(S, v, v, S->c_var, 2, 0); caxpy(S, S->c_gg, -coe, S->c_var);
    PHYS_TERM(S, S->kstr3, S->kend3,
              1.0 / coe * (P->dp_opt == 3 ? P->fc * (P->ug - F3(S, u, i, j, k) - P->u_mrf) : 0.0) -
                  (A1U(v) * F3(S, w, i, j, k) - A2U(v) * F3(S, w, i, j, k - 1)) / V1(S->dz, k));
    caxpy(S, S->c_gg, coe, S->c_var);
    if (P->issk == 1) {
        skew_term(S, S->c_v, 1, u, 0); caxpy(S, S->c_gg, -coe, S->c_var);
        skew_term(S, S->c_v, 2, v, 0); caxpy(S, S->c_gg, -coe, S->c_var);
        PHYS_TERM(S, S->kstr3, S->kend3,
                  (A1U(v) - A2U(v)) / V1(S->dz, k) * (0.5 * (F3(S, w, i, j, k) + F3(S, w, i, j, k - 1))));
        caxpy(S, S->c_gg, -coe, S->c_var);
    }
    caxpy(S, S->c_qv, S->rk_a * P->dt, S->c_gg);
#undef A1U
#undef A2U
    /* ---- W (:344-499) */
    memcpy(S->c_qw, S->c_w, n * sizeof(cplx));
    dev_linear(S, S->c_w, S->c_var, 2, 1); caxpy(S, S->c_qw, dif, S->c_var);
    dev_linear(S, S->c_w, S->c_var, 2, 2); caxpy(S, S->c_qw, dif, S->c_var);
    PHYS_TERM(S, S->kstr3, S->kend3,
              ((F3(S, w, i, j, k + 1) - F3(S, w, i, j, k)) / V1(S->dz, k + 1) -
               (F3(S, w, i, j, k) - F3(S, w, i, j, k - 1)) / V1(S->dz, k)) / V1(S->dz_t, k));
    caxpy(S, S->c_qw, dif, S->c_var);
    PHYS_TERM(S, S->kstr3, S->kend3, (F3(S, p, i, j, k + 1) - F3(S, p, i, j, k)) / V1(S->dz_t, k));
    caxpy(S, S->c_qw, -S->rk_c * P->dt, S->c_var);
    caxpy(S, S->c_qw, S->rk_b * P->dt, S->c_hh);
    memset(S->c_hh, 0, n * sizeof(cplx));
    dev_linear(S, S->c_w, S->c_var, 1, 1); caxpy(S, S->c_hh, -P->u_mrf, S->c_var);
    dev_nonlinear(S, u, w, S->c_var, 1, 1); caxpy(S, S->c_hh, -coe, S->c_var);
    dev_nonlinear(S, v, w, S->c_var, 2, 1); caxpy(S, S->c_hh, -coe, S->c_var);
#define WP (0.5 * (F3(S, w, i, j, k) + F3(S, w, i, j, k + 1)))
#define WM (0.5 * (F3(S, w, i, j, k) + F3(S, w, i, j, k - 1)))
    PHYS_TERM(S, S->kstr3, S->kend3,
              (P->isscalar == 1 ? 1.0 / coe * (S->got * ((1.0 - V1(S->l_t, k)) * F3(S, t, i, j, k) +
                                                         V1(S->l_t, k) * F3(S, t, i, j, k + 1) - P->t_ref))
                                : 0.0) -
                  (WP * WP - WM * WM) / V1(S->dz_t, k));
    caxpy(S, S->c_hh, coe, S->c_var);
    if (P->issk == 1) {
        skew_term(S, S->c_w, 1, u, 1); caxpy(S, S->c_hh, -coe, S->c_var);
        skew_term(S, S->c_w, 2, v, 1); caxpy(S, S->c_hh, -coe, S->c_var);
        PHYS_TERM(S, S->kstr3, S->kend3, F3(S, w, i, j, k) * ((WP - WM) / V1(S->dz_t, k)));
        caxpy(S, S->c_hh, -coe, S->c_var);
    }
#undef WP
#undef WM
    caxpy(S, S->c_qw, S->rk_a * P->dt, S->c_hh);
}

/* get_temperature_cnvdiff_spec, src/module_spectral.f90:509-595 (qt into c_qu) */
static void temperature_spec(ho_state *S)
{
    const ho_params *P = &S->prm;
    double *u = S->u, *v = S->v, *w = S->w, *t = S->t;
    const double dif = S->rk_c * 0.5 * P->dt * P->alpha;
    size_t n = ncplx(S);
    memcpy(S->c_qu, S->c_t, n * sizeof(cplx));
    dev_linear(S, S->c_t, S->c_var, 2, 1); caxpy(S, S->c_qu, dif, S->c_var);
    dev_linear(S, S->c_t, S->c_var, 2, 2); caxpy(S, S->c_qu, dif, S->c_var);
    PHYS_TERM(S, S->kstr3, S->kend3,
              ((F3(S, t, i, j, k + 1) - F3(S, t, i, j, k)) / V1(S->dz_t, k) -
               (F3(S, t, i, j, k) - F3(S, t, i, j, k - 1)) / V1(S->dz_b, k)) / V1(S->dz, k));
    caxpy(S, S->c_qu, dif, S->c_var);
    caxpy(S, S->c_qu, S->rk_b * P->dt, S->c_ss);
    memset(S->c_ss, 0, n * sizeof(cplx));
    dev_linear(S, S->c_t, S->c_var, 1, 1); caxpy(S, S->c_ss, -P->u_mrf, S->c_var);
    dev_nonlinear(S, u, t, S->c_var, 1, 0); caxpy(S, S->c_ss, -1.0, S->c_var);
    dev_nonlinear(S, t, v, S->c_var, 2, 0); caxpy(S, S->c_ss, -1.0, S->c_var);
    PHYS_TERM(S, S->kstr3, S->kend3,
              ((F3(S, t, i, j, k + 1) * V1(S->l_t, k) + F3(S, t, i, j, k) * (1.0 - V1(S->l_t, k))) * F3(S, w, i, j, k) -
               (F3(S, t, i, j, k - 1) * V1(S->l_b, k) + F3(S, t, i, j, k) * (1.0 - V1(S->l_b, k))) * F3(S, w, i, j, k - 1)) /
                  V1(S->dz, k));
    caxpy(S, S->c_ss, -1.0, S->c_var);
    caxpy(S, S->c_qu, S->rk_a * P->dt, S->c_ss);
}

/* poisson_solver_fft_spec, src/module_spectral.f90:602-746 */
static int poisson_spec(ho_state *S, const cplx *src, cplx *out, int zstagger, int delnull, int tbnd, int ubnd)
{
    memcpy(S->c_src, src, sizeof(cplx) * ncplx(S));
    int kmax3;
    if (spectral_thomas(S, zstagger, delnull, tbnd, ubnd, &kmax3)) return 1;
    memcpy(out, S->c_var, sizeof(cplx) * ncplx(S));
    return 0;
}

/* get_div_spec, src/module_spectral.f90:752-795 */
static void div_spec(ho_state *S, const cplx *cu, const cplx *cv, const double *w, cplx *div)
{
    size_t n = ncplx(S);
    dev_linear(S, cu, S->c_var, 1, 1);
    memcpy(div, S->c_var, n * sizeof(cplx));
    dev_linear(S, cv, S->c_var, 1, 2);
    caxpy(S, div, 1.0, S->c_var);
    PHYS_TERM(S, S->kstr3, S->kend3, (F3(S, w, i, j, k) - F3(S, w, i, j, k - 1)) / V1(S->dz, k));
    caxpy(S, div, 1.0, S->c_var);
}

/* calculate_uvw_spec, src/module_spectral.f90:801-849 */
static void uvw_spec(ho_state *S)
{
    const ho_params *P = &S->prm;
    double *phi = S->phi;
    dev_linear(S, S->c_phi, S->c_var, 1, 1); caxpy(S, S->c_u, -S->rk_c * P->dt, S->c_var);
    dev_linear(S, S->c_phi, S->c_var, 1, 2); caxpy(S, S->c_v, -S->rk_c * P->dt, S->c_var);
    /* levels kstr3..kend3-1 only; c_var keeps its previous content at level nz */
    PHYS_TERM(S, S->kstr3, S->kend3 - 1, (F3(S, phi, i, j, k + 1) - F3(S, phi, i, j, k)) / V1(S->dz_t, k));
    for (int k = 1; k <= P->nz - 1; ++k)
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i)
                C3(S, S->c_w, i, j, k) = C3(S, S->c_w, i, j, k) - S->rk_c * P->dt * C3(S, S->c_var, i, j, k);
}

/* calculate_new_p_spec, src/module_spectral.f90:854-894 */
static void new_p_spec(ho_state *S)
{
    const ho_params *P = &S->prm;
    double *phi = S->phi;
    const double dif = -S->rk_c * 0.5 * P->dt * P->nu;
    caxpy(S, S->c_p, 1.0, S->c_phi);
    dev_linear(S, S->c_phi, S->c_var, 2, 1); caxpy(S, S->c_p, dif, S->c_var);
    dev_linear(S, S->c_phi, S->c_var, 2, 2); caxpy(S, S->c_p, dif, S->c_var);
    PHYS_TERM(S, S->kstr3, S->kend3,
              ((F3(S, phi, i, j, k + 1) - F3(S, phi, i, j, k)) / V1(S->dz_t, k) -
               (F3(S, phi, i, j, k) - F3(S, phi, i, j, k - 1)) / V1(S->dz_b, k)) / V1(S->dz, k));
    caxpy(S, S->c_p, dif, S->c_var);
}

/* screen_div_error_spec (values only), src/module_spectral.f90:899-1007 */
static int div_error_spec(ho_state *S, int isfinal)
{
    const ho_params *P = &S->prm;
    div_spec(S, S->c_u, S->c_v, S->w, S->c_src);
    ho_fft2d(S, S->c_src, 1);
    double avg = 0.0, mx = 0.0;
    for (int k = S->kstr3; k <= S->kend3; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) {
                double d = creal(C3(S, S->c_src, i - 2, j - 2, k - 2));
                avg = avg + fabs(d);
                if (fabs(d) > mx) mx = fabs(d);
                if (F3(S, S->u, i, j, k) != F3(S, S->u, i, j, k) || F3(S, S->v, i, j, k) != F3(S, S->v, i, j, k) ||
                    F3(S, S->w, i, j, k) != F3(S, S->w, i, j, k)) {
                    snprintf(S->err, sizeof S->err, "NaN error! %d %d %d", i, j, k);
                    return 1;
                }
                if (isfinal == 1 && fabs(d) > S->div_hist_max) S->div_hist_max = fabs(d);
            }
    S->div_avg = avg / P->nx / P->ny / P->nz;
    S->div_max = mx;
    return 0;
}

/* ---- assign_rk_coeff, src/module_parameters.f90:618-641 ---------------- */
/* ---- A19: rayleigh_damping, src/module_tools.f90:705-749 ---------------- */
/* Top sponge on levels kend3-nzdamp..kend3: each interior value relaxes
 * toward its plane mean by sigz = cdamp*(1-cos(pi*zd))/2.  Only the interior
 * (i,j) is written: the x/y halos and the z ghosts keep the values of the last
 * boundary update, exactly as in the reference, which calls no boundary
 * update between the damping and the stencils that read those halos
 * (src/module_navier_stokes.f90:258-275). */
static void rayleigh_damping(ho_state *S, double *a)
{
    const ho_params *P = &S->prm;
    for (int k = S->kend3 - P->nzdamp; k <= S->kend3; ++k) {
        double mean_local = 0.0;
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) mean_local += F3(S, a, i, j, k);
        mean_local = mean_local / P->nx / P->ny;
        double mean_all = mean_local / 1.0; /* MPI_ALLREDUCE over one rank, / nprc */
        double zd = (k - S->kend3 + P->nzdamp) * 1.0 / P->nzdamp;
        double sigz = P->cdamp * 0.5 * (1.0 - cos(PI * zd));
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i)
                F3(S, a, i, j, k) = F3(S, a, i, j, k) - sigz * (F3(S, a, i, j, k) - mean_all);
    }
}

static void assign_rk(ho_state *S, int t_its)
{
    if (S->prm.dts == 2) { S->rk_a = RKC1[t_its - 1]; S->rk_b = RKC2[t_its - 1]; S->rk_c = RKC3[t_its - 1]; }
    else { S->rk_a = 1.5; S->rk_b = -0.5; S->rk_c = 1.0; }
}

/* ---- A16: time_advancement_fd, src/module_navier_stokes.f90:223-365 ---- */
static void mean_fields(ho_state *S);

static int time_advancement_fd(ho_state *S)
{
    const ho_params *P = &S->prm;
    size_t nf = nfield(S);
    int t_its_max = P->dts == 2 ? 3 : 1;
    for (int t_its = 1; t_its <= t_its_max; ++t_its) {
        if (P->isdamp == 1) {                                     /* :258-266 */
            rayleigh_damping(S, S->u);
            rayleigh_damping(S, S->v);
            rayleigh_damping(S, S->w);
            rayleigh_damping(S, S->p);
            rayleigh_damping(S, S->t);
        }
        memcpy(S->ff0, S->ff, nf * sizeof(double));               /* assign_fgh_old :270 */
        memcpy(S->gg0, S->gg, nf * sizeof(double));
        memcpy(S->hh0, S->hh, nf * sizeof(double));
        interp_uvw(S);                                            /* :274 */
        bc_uvwh(S);                                               /* :275 */
        assign_rk(S, t_its);                                      /* :280 */
        momentum(S);                                              /* :284 */
        assign_abc(S, S->aa_u, S->bb_u, S->cc_u, 0);              /* :291-293 */
        if (poisson_solver_fft(S, S->qu, S->u, 0, 0, 0, 1)) return 1;
        assign_abc(S, S->aa_u, S->bb_u, S->cc_u, 0);              /* :297-299 */
        if (poisson_solver_fft(S, S->qv, S->v, 0, 0, 0, 0)) return 1;
        assign_abc(S, S->aa_w, S->bb_w, S->cc_w, 0);              /* :302-304 */
        if (poisson_solver_fft(S, S->qw, S->w, 1, 0, 0, 0)) return 1;
        bc_uvw(S);                                                /* :306 */
        p_src(S, S->qu);                                          /* :310 */
        assign_abc(S, S->aa_p, S->bb_p, S->cc_p, 1);              /* :312-314 */
        if (poisson_solver_fft(S, S->qu, S->phi, 0, 1, 0, 0)) return 1;
        bc_p(S, S->phi);                                          /* :316 */
        calc_uvw(S);                                              /* :319 */
        bc_uvw(S);                                                /* :320 */
        if (P->isscalar == 1) {                                   /* :323-341 */
            memcpy(S->ss0, S->ss, nf * sizeof(double));
            interp_t(S);
            bc_th(S);
            temperature(S, S->qu);
            assign_abc(S, S->aa_t, S->bb_t, S->cc_t, 0);
            if (poisson_solver_fft(S, S->qu, S->t, 0, 0, 1, 0)) return 1;
            bc_t(S);
        }
        if (div_error(S, 1)) return 1;                            /* :344-350 */
        calc_new_p(S);                                            /* :353 */
        bc_p(S, S->p);                                            /* :354 */
    }
    mean_fields(S);                                               /* :358-361 */
    return 0;
}

/* ---- calculate_mean_fields, src/module_tools.f90:362-404 (FD) and
 * calculate_mean_fields_spec, src/module_spectral.f90:1012-1052: pointwise
 * running sums of u + u_mrf, v, w, t, their squares, and w at the centre
 * (0.5 (w(k) + w(k-1))) times u, v, t.  The FD form takes u and v from the
 * x / y interpolants uhx, vhy, which still hold the values interpolated at
 * the start of the step's last substep (get_interp_fields_uvwhxyz, :274);
 * the spectral form takes u and v themselves.  Order of the sums:
 * u, v, w, t, uu, vv, ww, uw, vw, tt, tw. */
static void mean_fields(ho_state *S)
{
    if (!S->mean[0]) return;
    const ho_params *P = &S->prm;
    const double um = P->u_mrf;
    const int nx = P->nx, ny = P->ny;
    for (int k = S->kstr3; k <= S->kend3; ++k)
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) {
                const size_t q = ((size_t)(k - S->kstr3) * ny + (j - S->jstr3)) * nx + (i - S->istr3);
                const double u = F3(S, S->u, i, j, k), v = F3(S, S->v, i, j, k), w = F3(S, S->w, i, j, k);
                const double t = F3(S, S->t, i, j, k);
                const double wc = 0.5 * (w + F3(S, S->w, i, j, k - 1));
                const double uh = P->cds == 3 ? u : F3(S, S->uhx, i, j, k);
                const double vh = P->cds == 3 ? v : F3(S, S->vhy, i, j, k);
                S->mean[4][q] += (u + um) * (u + um);
                S->mean[5][q] += v * v;
                S->mean[6][q] += w * w;
                S->mean[7][q] += wc * (uh + um);
                S->mean[8][q] += wc * vh;
                S->mean[9][q] += t * t;
                S->mean[10][q] += t * wc;
                S->mean[0][q] += u + um;
                S->mean[1][q] += v;
                S->mean[2][q] += w;
                S->mean[3][q] += t;
            }
    S->mean_steps += 1;
}

/* ---- A18: time_advancement_sp, src/module_navier_stokes.f90:371-517 ---- */
/* rayleigh_damping_spec, src/module_spectral.f90:1345-1391: on levels
 * kend3-nzdamp..kend3 every mode relaxes toward the mean over the modes of
 * the level.  The reference passes that complex mean to a one-element real
 * MPI_ALLREDUCE (:1372), which sets only its real part; here (one rank) the
 * whole complex mean is used, mean_all = mean_local / nprc. */
static void rayleigh_damping_spec(ho_state *S, cplx *c)
{
    const ho_params *P = &S->prm;
    for (int k = S->kend3 - P->nzdamp; k <= S->kend3; ++k) {
        cplx mean_local = 0.0;
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i) mean_local += C3(S, c, i - 2, j - 2, k - 2);
        mean_local = mean_local / P->nx / P->ny;
        const cplx mean_all = mean_local / 1.0;
        const double zd = (k - S->kend3 + P->nzdamp) * 1.0 / P->nzdamp;
        const double sigz = P->cdamp * 0.5 * (1.0 - cos(PI * zd));
        for (int j = S->jstr3; j <= S->jend3; ++j)
            for (int i = S->istr3; i <= S->iend3; ++i)
                C3(S, c, i - 2, j - 2, k - 2) = C3(S, c, i - 2, j - 2, k - 2) - sigz * (C3(S, c, i - 2, j - 2, k - 2) - mean_all);
    }
}

static int time_advancement_sp(ho_state *S)
{
    const ho_params *P = &S->prm;
    int t_its_max = P->dts == 2 ? 3 : 1;
    for (int t_its = 1; t_its <= t_its_max; ++t_its) {
        if (P->isdamp == 1) {                                             /* :406-414 */
            rayleigh_damping_spec(S, S->c_u);
            rayleigh_damping_spec(S, S->c_v);
            rayleigh_damping_spec(S, S->c_w);
            rayleigh_damping_spec(S, S->c_p);
            rayleigh_damping_spec(S, S->c_t);
        }
        assign_rk(S, t_its);                                              /* :417 */
        momentum_spec(S);                                                 /* :420-421 */
        assign_abc(S, S->aa_u, S->bb_u, S->cc_u, 0);                      /* :427-429 */
        if (poisson_spec(S, S->c_qu, S->c_u, 0, 0, 0, 1)) return 1;
        assign_abc(S, S->aa_u, S->bb_u, S->cc_u, 0);                      /* :433-435 */
        if (poisson_spec(S, S->c_qv, S->c_v, 0, 0, 0, 0)) return 1;
        assign_abc(S, S->aa_w, S->bb_w, S->cc_w, 0);                      /* :438-440 */
        if (poisson_spec(S, S->c_qw, S->c_w, 1, 0, 0, 0)) return 1;
        cmplx_2_real(S, S->c_w, S->w);                                    /* :444-445 */
        bc_uvw(S);
        div_spec(S, S->c_u, S->c_v, S->w, S->c_qu);                       /* :449-450 */
        for (size_t q = 0; q < ncplx(S); ++q) S->c_qu[q] = S->c_qu[q] / P->dt;
        assign_abc(S, S->aa_p, S->bb_p, S->cc_p, 1);                      /* :452-454 */
        if (poisson_spec(S, S->c_qu, S->c_phi, 0, 1, 0, 0)) return 1;
        cmplx_2_real(S, S->c_phi, S->phi);                                /* :457-458 */
        bc_p(S, S->phi);
        new_p_spec(S);                                                    /* :461 */
        uvw_spec(S);                                                      /* :464 */
        truncation(S, S->c_u);                                            /* :467-470 */
        truncation(S, S->c_v);
        truncation(S, S->c_w);
        truncation(S, S->c_p);
        cmplx_2_real(S, S->c_u, S->u);                                    /* :473-478 */
        cmplx_2_real(S, S->c_v, S->v);
        cmplx_2_real(S, S->c_w, S->w);
        cmplx_2_real(S, S->c_p, S->p);
        bc_p(S, S->p);
        bc_uvw(S);
        if (P->isscalar == 1) {                                           /* :481-498 */
            temperature_spec(S);
            assign_abc(S, S->aa_t, S->bb_t, S->cc_t, 0);
            if (poisson_spec(S, S->c_qu, S->c_t, 0, 0, 1, 0)) return 1;
            truncation(S, S->c_t);
            cmplx_2_real(S, S->c_t, S->t);
            bc_t(S);
        }
        if (div_error_spec(S, 1)) return 1;                               /* :501-507 */
    }
    mean_fields(S);                                                       /* :510-513 */
    return 0;
}

/* ---- drive, src/module_navier_stokes.f90:31-218 (numerics only) -------- */
/* start-up: mesh, coefficients, bulk velocity and boundary values (:53-129) */
int ho_init(ho_state *S)
{
    ho_mesh(S);
    if (S->prm.is_ri_var == 1) S->got = S->prm.ri_str;  /* update_ri(0) */
    if (S->prm.dp_opt == 2) ho_calc_ubar(S);
    ho_coefficients(S);
    if (S->prm.cds == 3) {  /* :96-123: spectra of the initial fields, truncated */
        real_2_cmplx(S, S->u, S->c_u);
        real_2_cmplx(S, S->v, S->c_v);
        real_2_cmplx(S, S->w, S->c_w);
        real_2_cmplx(S, S->p, S->c_p);
        real_2_cmplx(S, S->t, S->c_t);
        real_2_cmplx(S, S->ff, S->c_ff);
        real_2_cmplx(S, S->gg, S->c_gg);
        real_2_cmplx(S, S->hh, S->c_hh);
        real_2_cmplx(S, S->ss, S->c_ss);
        truncation(S, S->c_u);
        truncation(S, S->c_v);
        truncation(S, S->c_w);
        truncation(S, S->c_p);
        truncation(S, S->c_t);
        cmplx_2_real(S, S->c_u, S->u);
        cmplx_2_real(S, S->c_v, S->v);
        cmplx_2_real(S, S->c_w, S->w);
        cmplx_2_real(S, S->c_p, S->p);
        cmplx_2_real(S, S->c_t, S->t);
    }
    bc_uvw(S);
    bc_p(S, S->p);
    bc_t(S);
    return 0;
}

/* nsteps passes of the step loop body (:139-165) */
int ho_steps(ho_state *S, int nsteps)
{
    for (int n = 0; n < nsteps; ++n) {
        S->stps += 1;
        S->time_sim += S->prm.dt;
        if (S->prm.is_ri_var == 1)
            S->got = S->prm.ri_str + (S->prm.ri_end - S->prm.ri_str) * 1.0 / S->prm.imax * S->stps;
        if (S->prm.cds == 3) {
            if (time_advancement_sp(S)) return 1;
        } else if (time_advancement_fd(S)) return 1;
        if (S->prm.dp_opt == 2) ho_calc_ubar(S);
    }
    return 0;
}

/* statistics from now on: zeroed running sums of calculate_mean_fields */
int ho_mean_enable(ho_state *S)
{
    const size_t n = (size_t)S->prm.nx * S->prm.ny * S->prm.nz;
    for (int q = 0; q < 11; ++q) {
        if (!S->mean[q]) S->mean[q] = dalloc(n);
        if (!S->mean[q]) { snprintf(S->err, sizeof S->err, "out of memory"); return 1; }
        memset(S->mean[q], 0, n * sizeof(double));
    }
    S->mean_steps = 0;
    return 0;
}

/* output_mean_fields (src/module_io.f90:546-598): the running sum divided
 * by the number of steps it holds */
int ho_get_mean(ho_state *S, int id, double *dst)
{
    if (id < 0 || id >= 11 || !S->mean[id] || S->mean_steps == 0) {
        snprintf(S->err, sizeof S->err, "no statistics %d", id);
        return 1;
    }
    const size_t n = (size_t)S->prm.nx * S->prm.ny * S->prm.nz;
    for (size_t q = 0; q < n; ++q) dst[q] = S->mean[id][q] / S->mean_steps;
    return 0;
}

/* ---- accessors for the tests -------------------------------------------- */
#ifdef _OPENMP
#include <omp.h>
int ho_num_threads(void) { return omp_get_max_threads(); }
#else
int ho_num_threads(void) { return 1; }
#endif
const char *ho_error(ho_state *S) { return S->err; }

static double *field_ptr(ho_state *S, int id)
{
    double *f[] = {S->u, S->v, S->w, S->p, S->t, S->phi, S->ff, S->gg, S->hh, S->ss,
                   S->qu, S->qv, S->qw, S->uhx, S->uhy, S->uhz, S->vhx, S->vhy, S->vhz, S->whx,
                   S->why, S->thx, S->thy, S->thz, S->ff0, S->gg0, S->hh0, S->ss0};
    if (id < 0 || id >= (int)(sizeof(f) / sizeof(f[0]))) return NULL;
    return f[id];
}

/* interior (nz, ny, nx) C-order <-> ghosted Fortran array */
int ho_set_field(ho_state *S, int id, const double *src)
{
    double *a = field_ptr(S, id);
    if (!a) return 1;
    const ho_params *P = &S->prm;
    for (int k = 1; k <= P->nz; ++k)
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i)
                F3(S, a, i + 2, j + 2, k + 2) = src[((size_t)(k - 1) * P->ny + (j - 1)) * P->nx + (i - 1)];
    return 0;
}

int ho_get_field(ho_state *S, int id, double *dst)
{
    double *a = field_ptr(S, id);
    if (!a) return 1;
    const ho_params *P = &S->prm;
    for (int k = 1; k <= P->nz; ++k)
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i)
                dst[((size_t)(k - 1) * P->ny + (j - 1)) * P->nx + (i - 1)] = F3(S, a, i + 2, j + 2, k + 2);
    return 0;
}

/* full ghosted array, Fortran order, (nx+4)*(ny+4)*(nz+4) values */
int ho_get_ghosted(ho_state *S, int id, double *dst)
{
    double *a = field_ptr(S, id);
    if (!a) return 1;
    memcpy(dst, a, nfield(S) * sizeof(double));
    return 0;
}

void ho_get_mesh(ho_state *S, double *dz, double *l_t, double *l_b, double *dz_t, double *dz_b)
{
    memcpy(dz, S->dz, sizeof(double) * S->nka);
    memcpy(l_t, S->l_t, sizeof(double) * S->nka);
    memcpy(l_b, S->l_b, sizeof(double) * S->nka);
    memcpy(dz_t, S->dz_t, sizeof(double) * S->nka);
    memcpy(dz_b, S->dz_b, sizeof(double) * S->nka);
}

/* 1-D coefficients (nz+4) and the complex bb (nx, ny, nz+4) of one equation:
 * eq 0 = u/v, 1 = w, 2 = p, 3 = t */
int ho_get_coeff(ho_state *S, int eq, double *aa, double *cc, cplx *bb)
{
    double *a[] = {S->aa_u, S->aa_w, S->aa_p, S->aa_t}, *c[] = {S->cc_u, S->cc_w, S->cc_p, S->cc_t};
    cplx *b[] = {S->bb_u, S->bb_w, S->bb_p, S->bb_t};
    if (eq < 0 || eq > 3) return 1;
    memcpy(aa, a[eq], sizeof(double) * S->nka);
    memcpy(cc, c[eq], sizeof(double) * S->nka);
    memcpy(bb, b[eq], sizeof(cplx) * (size_t)S->prm.nx * S->prm.ny * S->nka);
    return 0;
}

/* One standalone poisson_solver_fft call on interior arrays, with the
 * equation's coefficients assigned by assign_abc/assign_abc_p at rk_c.
 * eq: 0 = u (flags 0,0,0,1), 1 = v (0,0,0,0), 2 = w (1,0,0,0),
 *     3 = p (0,1,0,0), 4 = t (0,0,1,0)  (src/module_navier_stokes.f90:291-337) */
int ho_solve(ho_state *S, int eq, double rk_c, const double *src, double *out)
{
    S->rk_c = rk_c;
    double *r_src = S->qu, *r_var = S->phi;
    const ho_params *P = &S->prm;
    for (int k = 1; k <= P->nz; ++k)
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i)
                F3(S, r_src, i + 2, j + 2, k + 2) = src[((size_t)(k - 1) * P->ny + (j - 1)) * P->nx + (i - 1)];
    int rc;
    switch (eq) {
    case 0: assign_abc(S, S->aa_u, S->bb_u, S->cc_u, 0); rc = poisson_solver_fft(S, r_src, r_var, 0, 0, 0, 1); break;
    case 1: assign_abc(S, S->aa_u, S->bb_u, S->cc_u, 0); rc = poisson_solver_fft(S, r_src, r_var, 0, 0, 0, 0); break;
    case 2: assign_abc(S, S->aa_w, S->bb_w, S->cc_w, 0); rc = poisson_solver_fft(S, r_src, r_var, 1, 0, 0, 0); break;
    case 3: assign_abc(S, S->aa_p, S->bb_p, S->cc_p, 1); rc = poisson_solver_fft(S, r_src, r_var, 0, 1, 0, 0); break;
    case 4: assign_abc(S, S->aa_t, S->bb_t, S->cc_t, 0); rc = poisson_solver_fft(S, r_src, r_var, 0, 0, 1, 0); break;
    default: snprintf(S->err, sizeof S->err, "bad equation id"); return 1;
    }
    if (rc) return rc;
    for (int k = 1; k <= P->nz; ++k)
        for (int j = 1; j <= P->ny; ++j)
            for (int i = 1; i <= P->nx; ++i)
                out[((size_t)(k - 1) * P->ny + (j - 1)) * P->nx + (i - 1)] = F3(S, r_var, i + 2, j + 2, k + 2);
    return 0;
}

/* Divergence statistics of the current velocity (avg |div|, signed max) */
int ho_div_stats(ho_state *S, double *avg, double *mx)
{
    double keep = S->div_hist_max;
    int rc = S->prm.cds == 3 ? div_error_spec(S, 0) : div_error(S, 0);
    S->div_hist_max = keep;
    *avg = S->div_avg; *mx = S->div_max;
    return rc;
}

/* screen_div_error(u,v,w,isfinal,.) as the reference runs it: div_avg, the
 * signed div_max of this call and the running div_hist_max (src/module_io.f90:85-206) */
int ho_screen(ho_state *S, int isfinal, double *avg, double *mx, double *hist)
{
    if (S->prm.cds != 3) bc_uvw(S);  /* the callers refresh the ghosts first (:306, :320) */
    int rc = S->prm.cds == 3 ? div_error_spec(S, isfinal) : div_error(S, isfinal);
    *avg = S->div_avg; *mx = S->div_max; *hist = S->div_hist_max;
    return rc;
}
/* the values of the last screen the step itself ran (the final substep's) */
void ho_last_screen(ho_state *S, double *avg, double *mx, double *hist)
{
    *avg = S->div_avg; *mx = S->div_max; *hist = S->div_hist_max;
}

/* Stage-level entry points for per-function parity tests.  Each runs the
 * named reference subroutine on the current state at substep t_its. */
void ho_stage_interp(ho_state *S) { interp_uvw(S); bc_uvwh(S); }
void ho_stage_momentum(ho_state *S, int t_its) { assign_rk(S, t_its); interp_uvw(S); bc_uvwh(S); momentum(S); }
void ho_stage_p_src(ho_state *S) { p_src(S, S->qu); }
void ho_stage_bc_uvw(ho_state *S) { bc_uvw(S); }
void ho_stage_damping(ho_state *S, int id)
{
    double *f[] = {S->u, S->v, S->w, S->p, S->t};
    if (id >= 0 && id < 5) rayleigh_damping(S, f[id]);
}
void ho_stage_temperature(ho_state *S, int t_its) { assign_rk(S, t_its); interp_t(S); bc_th(S); temperature(S, S->qu); }
/* calculate_new_p at substep t_its on the current phi and p, after the
 * Neumann ghosts of phi (update_boundary_p at :316, then :353) */
void ho_stage_new_p(ho_state *S, int t_its) { assign_rk(S, t_its); bc_p(S, S->phi); calc_new_p(S); }
/* calculate_uvw at substep t_its on the current phi and u, v, w (:316-320) */
void ho_stage_project(ho_state *S, int t_its) { assign_rk(S, t_its); bc_p(S, S->phi); calc_uvw(S); bc_uvw(S); }
double ho_get_scalar(ho_state *S, int id)
{
    switch (id) {
    case 0: return S->dpx_drive;
    case 1: return S->ubar;
    case 2: return S->div_hist_max;
    case 3: return S->time_sim;
    case 4: return S->got;
    default: return NAN;
    }
}
void ho_set_scalar(ho_state *S, int id, double v)
{
    switch (id) {
    case 0: S->dpx_drive = v; break;
    case 1: S->ubar = v; break;
    case 4: S->got = v; break;
    default: break;
    }
}
/* spectral divergence of the current state as screen_div_error_spec forms it
 * (interior, C order), for the tests */
int ho_spec_divergence(ho_state *S, double *dst)
{
    if (S->prm.cds != 3) return 1;
    div_spec(S, S->c_u, S->c_v, S->w, S->c_src);
    ho_fft2d(S, S->c_src, 1);
    size_t n = ncplx(S);
    for (size_t q = 0; q < n; ++q) dst[q] = creal(S->c_src[q]);
    return 0;
}

/* spectra of the spectral path (cds=3), C order (k, j, i) complex:
 * 0 c_u 1 c_v 2 c_w 3 c_p 4 c_t 5 c_phi 6 c_ff 7 c_gg 8 c_hh 9 c_ss 10 c_qu 11 c_qv 12 c_qw */
static cplx *spec_ptr(ho_state *S, int id)
{
    cplx *t[] = {S->c_u, S->c_v, S->c_w, S->c_p, S->c_t, S->c_phi, S->c_ff, S->c_gg, S->c_hh, S->c_ss,
                 S->c_qu, S->c_qv, S->c_qw};
    return (S->prm.cds == 3 && id >= 0 && id < 13) ? t[id] : NULL;
}
int ho_get_spectrum(ho_state *S, int id, cplx *dst)
{
    cplx *a = spec_ptr(S, id);
    if (!a) return 1;
    memcpy(dst, a, sizeof(cplx) * ncplx(S));
    return 0;
}
int ho_set_spectrum(ho_state *S, int id, const cplx *src)
{
    cplx *a = spec_ptr(S, id);
    if (!a) return 1;
    memcpy(a, src, sizeof(cplx) * ncplx(S));
    return 0;
}
/* rayleigh_damping_spec of spectrum id (0 c_u .. 4 c_t) */
int ho_stage_damping_spec(ho_state *S, int id)
{
    cplx *a = spec_ptr(S, id);
    if (!a || id > 4) return 1;
    rayleigh_damping_spec(S, a);
    return 0;
}
/* get_momentum_cnvdiff_spec / get_temperature_cnvdiff_spec at substep t_its */
int ho_stage_momentum_spec(ho_state *S, int t_its)
{
    if (S->prm.cds != 3) return 1;
    assign_rk(S, t_its);
    momentum_spec(S);
    return 0;
}
int ho_stage_temperature_spec(ho_state *S, int t_its)
{
    if (S->prm.cds != 3) return 1;
    assign_rk(S, t_its);
    temperature_spec(S);
    return 0;
}
/* poisson_solver_fft_spec of one equation (flags as in ho_solve) with the
 * coefficients assigned at rk_c: c_qu -> c_phi */
int ho_stage_poisson_spec(ho_state *S, int eq, double rk_c)
{
    if (S->prm.cds != 3) return 1;
    S->rk_c = rk_c;
    cplx *src = S->c_qu, *out = S->c_phi;
    switch (eq) {
    case 0: assign_abc(S, S->aa_u, S->bb_u, S->cc_u, 0); return poisson_spec(S, src, out, 0, 0, 0, 1);
    case 1: assign_abc(S, S->aa_u, S->bb_u, S->cc_u, 0); return poisson_spec(S, src, out, 0, 0, 0, 0);
    case 2: assign_abc(S, S->aa_w, S->bb_w, S->cc_w, 0); return poisson_spec(S, src, out, 1, 0, 0, 0);
    case 3: assign_abc(S, S->aa_p, S->bb_p, S->cc_p, 1); return poisson_spec(S, src, out, 0, 1, 0, 0);
    case 4: assign_abc(S, S->aa_t, S->bb_t, S->cc_t, 0); return poisson_spec(S, src, out, 0, 0, 1, 0);
    default: snprintf(S->err, sizeof S->err, "bad equation id"); return 1;
    }
}
/* get_div_spec(c_u, c_v, w, c_qu) */
int ho_stage_div_spec(ho_state *S)
{
    if (S->prm.cds != 3) return 1;
    div_spec(S, S->c_u, S->c_v, S->w, S->c_qu);
    return 0;
}
/* calculate_uvw_spec / calculate_new_p_spec at substep t_its, with phi's
 * ghosts from update_boundary_p as time_advancement_sp sets them (:457-458) */
int ho_stage_uvw_spec(ho_state *S, int t_its)
{
    if (S->prm.cds != 3) return 1;
    assign_rk(S, t_its);
    bc_p(S, S->phi);
    uvw_spec(S);
    return 0;
}
int ho_stage_new_p_spec(ho_state *S, int t_its)
{
    if (S->prm.cds != 3) return 1;
    assign_rk(S, t_its);
    bc_p(S, S->phi);
    new_p_spec(S);
    return 0;
}
