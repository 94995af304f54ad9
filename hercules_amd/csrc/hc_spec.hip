// Spectral path (cds = 3): time_advancement_sp, src/module_navier_stokes.f90:371-517,
// and src/module_spectral.f90.  Row kernels, transforms, start-up and the
// substep sequence; the column kernels are in hc_spec_col.hip.
#include "hc_host.h"

#include "hc_spectral.h"

// ---- spectral path (cds = 3) ------------------------------------------------
template <class R>
static int launch_yio(hc_ctx *c, int sign, int nf, const cplx<R> *const *src, cplx<R> *const *dst, const int *mul)
{
    const View<R> V = view<R>(c);
    YIOArgsT<R> Y;
    std::memset(&Y, 0, sizeof Y);
    for (int q = 0; q < nf; ++q) { Y.src[q] = src[q]; Y.dst[q] = dst[q]; Y.mul[q] = mul[q]; }
    Y.nl = c->nl; Y.ny = c->ny; Y.tw_ny = V.tw_ny; Y.symx = V.symx; Y.symy = V.symy;
    dim3 g(c->nl / kYTile, c->nz, nf);
#define HC_YIO(NY_)                                                                                     \
    if (sign < 0) hipLaunchKernelGGL((k_yfft_io<NY_, -1, R>), g, dim3(256), 0, c->stream, Y);          \
    else hipLaunchKernelGGL((k_yfft_io<NY_, +1, R>), g, dim3(256), 0, c->stream, Y);
    HC_DISPATCH_N(c->ny, HC_YIO)
#undef HC_YIO
    return 0;
}

template <class C>
static int launch_truncate(hc_ctx *c, C *s)
{
    hipLaunchKernelGGL((k_truncate<C>), dim3(1024), dim3(256), 0, c->stream, s, c->nl, c->ny, c->nz,
                       (int)(c->nx / 3.0), (int)(c->ny / 3.0));
    return 0;
}

// real_2_cmplx (src/module_spectral.f90:1157-1182): x R2C then y FFT, forward normalised
template <class R>
static int spec_forward(hc_ctx *c, const R *phys, cplx<R> *spec)
{
    if (int e = launch_r2c<R>(c, phys, spec)) return e;
    return launch_yfft(c, -1, 1, spec, nullptr, nullptr, nullptr);
}

// cmplx_2_real (:1188-1217) of up to three spectra, through scratch spectra
template <class R>
int spec_inverse(hc_ctx *c, int nf, const cplx<R> *const *spec, R *const *phys)
{
    const View<R> V = view<R>(c);
    const int zero[3] = {0, 0, 0};
    cplx<R> *scr[3] = {V.sp[0], V.sp[1], V.sp[2]};
    if (int e = launch_yio<R>(c, +1, nf, spec, scr, zero)) return e;
    C2RArgsT<R> X;
    std::memset(&X, 0, sizeof X);
    for (int q = 0; q < nf; ++q) { X.src[q] = scr[q]; X.dst[q] = phys[q]; }
    return launch_c2r<R>(c, nf, X);
}

// The spectral start-up of drive (src/module_navier_stokes.f90:96-123):
// spectra of every field, 2/3-rule truncation of u, v, w, p, t, and the
// truncated fields back in physical space.
template <class R>
static int spec_init_t(hc_ctx *c, const View<R> &V)
{
    R *phys[9] = {V.u, V.v, V.w, V.p, V.t, V.fb[c->fcur][0], V.fb[c->fcur][1], V.fb[c->fcur][2], V.sb[c->scur]};
    cplx<R> *spec[9] = {V.cu, V.cv, V.cw, V.cp, V.ct, V.cff, V.cgg, V.chh, V.css};
    for (int q = 0; q < 9; ++q)
        if (int e = spec_forward<R>(c, phys[q], spec[q])) return e;
    for (int q = 0; q < 5; ++q)
        if (int e = launch_truncate(c, spec[q])) return e;
    const cplx<R> *s3[3] = {V.cu, V.cv, V.cw};
    R *p3[3] = {V.u, V.v, V.w};
    if (int e = spec_inverse<R>(c, 3, s3, p3)) return e;
    const cplx<R> *s2[2] = {V.cp, V.ct};
    R *p2[2] = {V.p, V.t};
    if (int e = spec_inverse<R>(c, 2, s2, p2)) return e;
    c->spec_dirty = false;
    HIPCHK(hipGetLastError());
    return 0;
}
int spec_init(hc_ctx *c)
{
    return c->mixed ? spec_init_t<float>(c, view<float>(c)) : spec_init_t<double>(c, view<double>(c));
}

// rayleigh_damping_spec (src/module_spectral.f90:1345-1391) of five spectra:
// on the top levels k >= kdamp0 every mode relaxes toward the level's mean
// over the modes, c <- c - sigz(k) (c - M(k)), M(k) = (1/(nx ny)) sum over
// all (l, m) of c(l, m, k).  The reference reduces that complex mean with a
// one-element real MPI_ALLREDUCE (:1372), which leaves its imaginary part
// unset; here the mean is taken whole.  For the spectrum of a real field
// (Hermitian, Nyquist modes zero) it is real: the (0, m) modes once and the
// l > 0 modes of the half spectrum twice, real parts.
struct SpecDampArgs {
    double2 *f[5];
};
__global__ __launch_bounds__(256) void k_spec_damp_means(SpecDampArgs F, double *dmp_m, int nl, int ny, int nz,
                                                         int kdamp0, double inv_nxny)
{
    __shared__ double red[256];
    const int k = kdamp0 + blockIdx.x, fid = blockIdx.y;
    const double2 *plane = F.f[fid] + (size_t)k * nl * ny;
    double s = 0.0;
    for (int q = threadIdx.x; q < nl * ny; q += 256) {
        const int l = q % nl;
        s += (l == 0 ? 1.0 : 2.0) * plane[q].x;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) dmp_m[fid * nz + k] = red[0] * inv_nxny;
}
// The undamped values go to und ([field][level - kdamp0][l, m]) first: the
// reference forms the z diffusion and the w pressure gradient from the
// physical fields, which the damping leaves alone.
__global__ __launch_bounds__(256) void k_spec_damp_apply(SpecDampArgs F, double2 *und, const double *dmp_s,
                                                         const double *dmp_m, int nl, int ny, int nz, int kdamp0)
{
    const size_t plane = (size_t)nl * ny, n = plane * (size_t)(nz - kdamp0);
    const int fid = blockIdx.y;
    double2 *f = F.f[fid] + (size_t)kdamp0 * plane;
    double2 *u = und + (size_t)fid * n;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int k = kdamp0 + (int)(q / plane);
        const double sg = dmp_s[k], m = dmp_m[fid * nz + k];
        const double2 v = f[q];
        u[q] = v;
        f[q] = make_double2(v.x - sg * (v.x - m), v.y - sg * v.y);
    }
}
static int spec_damp(hc_ctx *c)
{
    SpecDampArgs F = {{c->cu, c->cv, c->cw, c->cp, c->ct}};
    const int nlev = c->nz - c->kdamp0;
    hipLaunchKernelGGL(k_spec_damp_means, dim3(nlev, 5), dim3(256), 0, c->stream, F, c->dmp_m, c->nl, c->ny, c->nz,
                       c->kdamp0, 1.0 / ((double)c->nx * (double)c->ny));
    hipLaunchKernelGGL(k_spec_damp_apply, dim3(1024, 5), dim3(256), 0, c->stream, F, c->spund, c->dmp_s, c->dmp_m,
                       c->nl, c->ny, c->nz, c->kdamp0);
    HIPCHK(hipGetLastError());
    return 0;
}

// The physical-space products of a substep and their spectra
// (get_momentum_cnvdiff_spec, src/module_spectral.f90:35-504, which = 0:
// the y-inverted derivative spectra of the skew-symmetric form, k_spec_mom,
// the y FFT of the eight products; get_temperature_cnvdiff_spec, :509-595,
// which = 1: k_spec_t and the y FFT of its three products).
template <class R>
static int spec_products_t(hc_ctx *c, const StepArgs &A, const View<R> &V, int which)
{
    using C = cplx<R>;
    const hc_params &P = c->prm;
    StepArgsT<R> B = rows_args<R>(A, V);
    for (int q = 0; q < 8; ++q) B.sp[q] = V.sp[q];
    for (int q = 0; q < 6; ++q) B.sdv[q] = V.sp[q];
    const int rows = rows_for(c->nx);
    const dim3 grow((c->ny + rows - 1) / rows, c->nz);
    if (which == 1) {
        PROF_BEGIN(c, HC_K_RHS_T);
        // t is unchanged since the substep's start: its face interpolant of
        // the momentum pre-pass still holds (the other buffers do not)
        if (A.zorder == 6) B.zft = reinterpret_cast<const R *>(c->zbuf[3]);
#define HC_ST(NX_)                                                                                 \
    if (A.zorder == 6) hipLaunchKernelGGL((k_spec_t<NX_, R, 6>), grow, dim3(rows * NX_), 0, c->stream, B); \
    else hipLaunchKernelGGL((k_spec_t<NX_, R>), grow, dim3(rows * NX_), 0, c->stream, B);
        HC_DISPATCH_N(c->nx, HC_ST)
#undef HC_ST
        PROF_END(c, HC_K_RHS_T);
        PROF_BEGIN(c, HC_K_YFFT_T);
        if (int e = launch_yfft(c, -1, 3, V.sp[0], V.sp[1], V.sp[2], nullptr)) return e;
        PROF_END(c, HC_K_YFFT_T);
        return 0;
    }
    PROF_BEGIN(c, HC_K_RHS);
    if (A.zorder == 6) zinterp6<R>(c, B);
    if (P.issk == 1) {
        // du/dx, du/dy, dv/dx, dv/dy, dw/dx, dw/dy in y-physical space
        const C *src4[4] = {V.cu, V.cu, V.cv, V.cv};
        C *dst4[4] = {V.sp[0], V.sp[1], V.sp[2], V.sp[3]};
        const int mul4[4] = {1, 2, 1, 2};
        if (int e = launch_yio<R>(c, +1, 4, src4, dst4, mul4)) return e;
        const C *src2[2] = {V.cw, V.cw};
        C *dst2[2] = {V.sp[4], V.sp[5]};
        const int mul2[2] = {1, 2};
        if (int e = launch_yio<R>(c, +1, 2, src2, dst2, mul2)) return e;
    }
#define HC_SM(NX_)                                                                                   \
    if (A.zorder == 6) hipLaunchKernelGGL((k_spec_mom<NX_, R, 6>), grow, dim3(rows * NX_), 0, c->stream, B); \
    else hipLaunchKernelGGL((k_spec_mom<NX_, R>), grow, dim3(rows * NX_), 0, c->stream, B);
    HC_DISPATCH_N(c->nx, HC_SM)
#undef HC_SM
    PROF_END(c, HC_K_RHS);
    PROF_BEGIN(c, HC_K_YFFT_FWD);
    if (int e = launch_yfft(c, -1, 4, V.sp[0], V.sp[1], V.sp[2], V.sp[3])) return e;
    if (int e = launch_yfft(c, -1, 4, V.sp[4], V.sp[5], V.sp[6], V.sp[7])) return e;
    PROF_END(c, HC_K_YFFT_FWD);
    return 0;
}
int spec_products(hc_ctx *c, const StepArgs &A, int which)
{
    return spec_products_t<double>(c, A, view<double>(c), which);
}
int spec_forward64(hc_ctx *c, const double *phys, double2 *spec) { return spec_forward<double>(c, phys, spec); }

// rayleigh_damping_spec of one half spectrum (FP64, isdamp = 1); the
// undamped levels go to the side buffer's first field
int spec_damp_one(hc_ctx *c, double2 *f)
{
    SpecDampArgs F = {{f, nullptr, nullptr, nullptr, nullptr}};
    const int nlev = c->nz - c->kdamp0;
    hipLaunchKernelGGL(k_spec_damp_means, dim3(nlev, 1), dim3(256), 0, c->stream, F, c->dmp_m, c->nl, c->ny, c->nz,
                       c->kdamp0, 1.0 / ((double)c->nx * (double)c->ny));
    hipLaunchKernelGGL(k_spec_damp_apply, dim3(1024, 1), dim3(256), 0, c->stream, F, c->spund, c->dmp_s, c->dmp_m,
                       c->nl, c->ny, c->nz, c->kdamp0);
    HIPCHK(hipGetLastError());
    return 0;
}

// one substep of time_advancement_sp (src/module_navier_stokes.f90:393-509)
template <class R>
static int substep_sp_t(hc_ctx *c, int t_its, const View<R> &V)
{
    using C = cplx<R>;
    const hc_params &P = c->prm;
    if (c->spec_dirty)
        if (int e = spec_init_t<R>(c, V)) return e;
    // the spectra of u, v, w, p, t are damped; the physical fields keep their
    // values (src/module_navier_stokes.f90:406-414).  Damping is FP64 only
    // (the mixed variant rejects isdamp = 1 at creation).
    if constexpr (std::is_same<R, double>::value)
        if (P.isdamp == 1)
            if (int e = spec_damp(c)) return e;
    StepArgs A = base_args(c);
    if (P.dts == 2) { A.ra = kRkc1[t_its - 1]; A.rb = kRkc2[t_its - 1]; A.rc = kRkc3[t_its - 1]; }
    else { A.ra = 1.5; A.rb = -0.5; A.rc = 1.0; }
    if (P.dp_opt == 2) dpx_update_launch(c);
    if (int e = spec_products_t<R>(c, A, V, 0)) return e;
    PROF_BEGIN(c, HC_K_COLUMN);
    if (int e = launch_spec_columns<R>(c, 0, A, V)) return e;
    PROF_END(c, HC_K_COLUMN);
    PROF_BEGIN(c, HC_K_FINISH);
    {
        const C *s3[3] = {V.cu, V.cv, V.cw};
        R *p3[3] = {V.u, V.v, V.w};
        if (int e = spec_inverse<R>(c, 3, s3, p3)) return e;
    }
    PROF_END(c, HC_K_FINISH);
    if (P.isscalar == 1) {
        if (int e = spec_products_t<R>(c, A, V, 1)) return e;
        PROF_BEGIN(c, HC_K_COLUMN_T);
        if (int e = launch_spec_columns<R>(c, 1, A, V)) return e;
        PROF_END(c, HC_K_COLUMN_T);
        PROF_BEGIN(c, HC_K_FINISH_T);
        const C *s1[1] = {V.ct};
        R *p1[1] = {V.t};
        if (int e = spec_inverse<R>(c, 1, s1, p1)) return e;
        PROF_END(c, HC_K_FINISH_T);
    }
    HIPCHK(hipGetLastError());
    return 0;
}
int substep_sp(hc_ctx *c, int t_its)
{
    return c->mixed ? substep_sp_t<float>(c, t_its, view<float>(c)) : substep_sp_t<double>(c, t_its, view<double>(c));
}


int spec_divergence(hc_ctx *c)
{
    if (c->spec_dirty)
        if (int e = spec_init(c)) return e;
    const StepArgs A = base_args(c);
    if (c->mixed) {
        View<float> &M = c->m32;
        hipLaunchKernelGGL((k_spec_div<float2>), dim3(1024), dim3(256), 0, c->stream, A, (const float2 *)M.cu,
                           (const float2 *)M.cv, (const float2 *)M.cw, M.sp[3]);
        const float2 *s1[1] = {M.sp[3]};
        float *p1[1] = {M.ptmp};
        return spec_inverse<float>(c, 1, s1, p1);
    }
    hipLaunchKernelGGL((k_spec_div<double2>), dim3(1024), dim3(256), 0, c->stream, A, (const double2 *)c->cu,
                       (const double2 *)c->cv, (const double2 *)c->cw, c->sp[3]);
    const double2 *s1[1] = {c->sp[3]};
    double *p1[1] = {c->ptmp};
    return spec_inverse<double>(c, 1, s1, p1);
}

template int spec_inverse<double>(hc_ctx *, int, const double2 *const *, double *const *);
template int spec_inverse<float>(hc_ctx *, int, const float2 *const *, float *const *);
