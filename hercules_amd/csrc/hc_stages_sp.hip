// Per-stage entry points of the spectral substep, with the argument lists of
// the reference subroutines time_advancement_sp calls
// (src/module_navier_stokes.f90:417-464):
//   get_momentum_cnvdiff_spec     src/module_spectral.f90:35-504
//   get_temperature_cnvdiff_spec  src/module_spectral.f90:509-595
//   poisson_solver_fft_spec       src/module_spectral.f90:602-746
//   get_div_spec                  src/module_spectral.f90:752-795
//   calculate_uvw_spec            src/module_spectral.f90:801-849
//   calculate_new_p_spec          src/module_spectral.f90:854-894
// and the utilities the loop calls between them: real_2_cmplx, cmplx_2_real
// (:1157-1217), spectral_truncation (:1222-1268), rayleigh_damping_spec
// (:1345-1391).
// The reference's complex work arrays are full C2C spectra of one z pencil,
// c(i, j, k) with i < nx, j < ny fastest-first, forward transform divided by
// nx ny.  They cross the ABI in that layout; on the device they are the half
// spectra of the fused path (l < nx/2).  The spectra of real fields are
// Hermitian, so nothing but the x and y Nyquist modes is dropped on the way
// in; on the way out the upper half is the conjugate mirror and the Nyquist
// modes are 0 (the 2/3 rule of the substep removes them).  Wherever the
// reference forms a term from a physical-space argument (z diffusion and
// dp/dz from u, v, w, p, t; dw/dz from w; the z derivatives of phi) the
// entry point transforms that argument and uses its spectrum, so a caller
// passing damped spectra with undamped fields gets the reference's mixture.
// FP64, cds = 3, 2nd-order z.
#include "hc_host.h"

#include "hc_spectral.h"

namespace {
// full (k, j, i < nx) -> half (k, j, l < nx/2); the y Nyquist row j = ny/2
// is held at 0, as the forward y transform leaves it
__global__ __launch_bounds__(256) void k_full_to_half(const double2 *full, double2 *half, int nx, int ny, int nz)
{
    const int nl = nx / 2;
    const size_t n = (size_t)nl * ny * nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const size_t row = q / nl;  // k * ny + j
        const int l = (int)(q - row * nl), j = (int)(row % ny);
        half[q] = j == ny / 2 ? make_double2(0.0, 0.0) : full[row * nx + l];
    }
}
// half -> full: the conjugate mirror c(nx - l, -m) = conj(c(l, m)) above the
// x Nyquist column, which is 0
__global__ __launch_bounds__(256) void k_half_to_full(const double2 *half, double2 *full, int nx, int ny, int nz)
{
    const int nl = nx / 2;
    const size_t n = (size_t)nx * ny * nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const size_t row = q / nx;
        const int i = (int)(q - row * nx), j = (int)(row % ny);
        const size_t k = row / ny;
        double2 v = make_double2(0.0, 0.0);
        if (i < nl) {
            v = half[row * nl + i];
        } else if (i > nl) {
            const int jm = j == 0 ? 0 : ny - j;
            const double2 h = half[(k * ny + jm) * nl + (nx - i)];
            v = make_double2(h.x, -h.y);
        }
        full[q] = v;
    }
}

// calculate_uvw_spec: c_u -= cdt i kx c_phi, c_v -= cdt i ky c_phi, and on
// levels below the top face c_w -= cdt (phi(k+1) - phi(k)) / dz_t(k) with
// phi's own spectrum phs
__global__ __launch_bounds__(256) void k_stage_uvw_spec(StepArgs A, double2 *cu, double2 *cv, double2 *cw,
                                                        const double2 *cphi, const double2 *phs, double cdt)
{
    const size_t plane = (size_t)A.nl * A.ny, n = plane * A.nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int l = (int)(q % A.nl), j = (int)((q / A.nl) % A.ny), k = (int)(q / plane);
        const double kx = A.symx[SY_GIM * A.nl + l], ky = A.symy[SY_GIM * A.ny + j];
        const double2 ph = cphi[q];
        cu[q] = csub(cu[q], times_i(ph, cdt * kx));
        cv[q] = csub(cv[q], times_i(ph, cdt * ky));
        if (k < A.nz - 1) {
            const double g = cdt * A.lev[LV_IDZT * A.nz + k];
            const double2 a = phs[q + plane], b = phs[q], w = cw[q];
            cw[q] = make_double2(w.x - g * (a.x - b.x), w.y - g * (a.y - b.y));
        }
    }
}

// calculate_new_p_spec: c_p += c_phi + dif (-kx^2 - ky^2) c_phi + dif phi_zz,
// phi_zz from phi's spectrum with the Neumann ghosts of update_boundary_p
__global__ __launch_bounds__(256) void k_stage_new_p_spec(StepArgs A, double2 *cp, const double2 *cphi,
                                                          const double2 *phs, double dif)
{
    const size_t plane = (size_t)A.nl * A.ny, n = plane * A.nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int l = (int)(q % A.nl), j = (int)((q / A.nl) % A.ny), k = (int)(q / plane);
        const double slap = A.symx[SY_L * A.nl + l] + A.symy[SY_L * A.ny + j];
        const double2 ph = cphi[q], p0 = phs[q];
        const double2 pu = k < A.nz - 1 ? phs[q + plane] : p0, pd = k > 0 ? phs[q - plane] : p0;
        const double it = A.lev[LV_IDZT * A.nz + k], ib = A.lev[LV_IDZB * A.nz + k], iz = A.lev[LV_IDZ * A.nz + k];
        const double zx = ((pu.x - p0.x) * it - (p0.x - pd.x) * ib) * iz;
        const double zy = ((pu.y - p0.y) * it - (p0.y - pd.y) * ib) * iz;
        double2 c = cp[q];
        c = make_double2(c.x + ph.x, c.y + ph.y);
        c = make_double2(c.x + dif * slap * ph.x, c.y + dif * slap * ph.y);
        cp[q] = make_double2(c.x + dif * zx, c.y + dif * zy);
    }
}

// device scratch of one call, freed on return
struct Scratch {
    void *p = nullptr;
    ~Scratch() { if (p) (void)hipFree(p); }
    int get(size_t bytes)
    {
        HIPCHK(hipMalloc(&p, bytes));
        return 0;
    }
};

int spec_stage_supported(const hc_ctx *c)
{
    const hc_params &P = c->prm;
    if (c->mixed) return set_err(HC_ERR_UNSUPPORTED, "the per-stage entry points are FP64 (precision=0)");
    if (P.cds != 3) return set_err(HC_ERR_ARG, "cds=%d: the spectral stages need cds=3", P.cds);
    if (P.zorder == 4 || P.zorder == 6)
        return set_err(HC_ERR_UNSUPPORTED, "the per-stage entry points use the reference's 2nd-order z");
    return 0;
}

// host or device full spectrum <-> half spectrum `half` through `full` (2N doubles)
int put_spec(hc_ctx *c, double2 *full, const double *src, double2 *half)
{
    const size_t nf = (size_t)c->nx * c->ny * c->nz;
    HIPCHK(hipMemcpyAsync(full, src, nf * sizeof(double2), hipMemcpyDefault, c->stream));
    hipLaunchKernelGGL(k_full_to_half, dim3(1024), dim3(256), 0, c->stream, full, half, c->nx, c->ny, c->nz);
    HIPCHK(hipGetLastError());
    return 0;
}
int get_spec(hc_ctx *c, double2 *full, const double2 *half, double *dst)
{
    const size_t nf = (size_t)c->nx * c->ny * c->nz;
    hipLaunchKernelGGL(k_half_to_full, dim3(1024), dim3(256), 0, c->stream, half, full, c->nx, c->ny, c->nz);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dst, full, nf * sizeof(double2), hipMemcpyDefault, c->stream));
    return 0;
}
size_t half_elems(const hc_ctx *c) { return (size_t)c->nl * c->ny * c->nz; }
}  // namespace

extern "C" int hc_get_momentum_cnvdiff_spec(hc_ctx *c, double rk_a, double rk_b, double rk_c, const double *u,
                                            const double *v, const double *w, const double *p, const double *t,
                                            const double *c_u, const double *c_v, const double *c_w, const double *c_p,
                                            double *c_ff, double *c_gg, double *c_hh, double *c_qu, double *c_qv,
                                            double *c_qw, int layout)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = spec_stage_supported(c)) return e;
    if (!u || !v || !w || !p || !c_u || !c_v || !c_w || !c_p || !c_ff || !c_gg || !c_hh || !c_qu || !c_qv || !c_qw ||
        (!t && c->prm.isscalar == 1))
        return set_err(HC_ERR_ARG, "null array argument");
    const double *in[] = {u, v, w, p};
    double *dev[] = {c->u, c->v, c->w, c->p};
    for (int q = 0; q < 4; ++q)
        if (int e = put_f64(c, dev[q], in[q], layout)) return e;
    if (t)
        if (int e = put_f64(c, c->t, t, layout)) return e;
    const size_t nh = half_elems(c);
    Scratch full, und;
    if (int e = full.get(2 * nh * sizeof(double2))) return e;
    if (int e = und.get(5 * nh * sizeof(double2))) return e;
    double2 *F = static_cast<double2 *>(full.p), *U = static_cast<double2 *>(und.p);
    const double *si[] = {c_u, c_v, c_w, c_p, c_ff, c_gg, c_hh};
    double2 *sd[] = {c->cu, c->cv, c->cw, c->cp, c->cff, c->cgg, c->chh};
    for (int q = 0; q < 7; ++q)
        if (int e = put_spec(c, F, si[q], sd[q])) return e;
    // spectra of the physical u, v, w, p (and t) for the terms formed from them
    const double *ph[] = {c->u, c->v, c->w, c->p, c->t};
    for (int q = 0; q < (t ? 5 : 4); ++q)
        if (int e = spec_forward64(c, ph[q], U + q * nh)) return e;
    StepArgs A = base_args(c);
    A.ra = rk_a; A.rb = rk_b; A.rc = rk_c;
    A.spund = U; A.kdamp0 = 0; A.stage = 1;
    if (c->prm.dp_opt == 2) dpx_update_launch(c);  // update_dpx, the first statement (:51)
    if (int e = spec_products(c, A, 0)) return e;
    if (int e = launch_spec_columns<double>(c, 0, A, view<double>(c))) return e;
    double *so[] = {c_qu, c_qv, c_qw, c_ff, c_gg, c_hh};
    const double2 *sh[] = {c->cu, c->cv, c->cw, c->cff, c->cgg, c->chh};
    for (int q = 0; q < 6; ++q)
        if (int e = get_spec(c, F, sh[q], so[q])) return e;
    c->spec_dirty = true;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_get_temperature_cnvdiff_spec(hc_ctx *c, double rk_a, double rk_b, double rk_c, const double *u,
                                               const double *v, const double *w, const double *t, const double *c_t,
                                               double *c_ss, double *c_qt, int layout)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = spec_stage_supported(c)) return e;
    if (c->prm.isscalar != 1) return set_err(HC_ERR_ARG, "isscalar=0: there is no temperature");
    if (!u || !v || !w || !t || !c_t || !c_ss || !c_qt) return set_err(HC_ERR_ARG, "null array argument");
    const double *in[] = {u, v, w, t};
    double *dev[] = {c->u, c->v, c->w, c->t};
    for (int q = 0; q < 4; ++q)
        if (int e = put_f64(c, dev[q], in[q], layout)) return e;
    const size_t nh = half_elems(c);
    Scratch full, und;
    if (int e = full.get(2 * nh * sizeof(double2))) return e;
    if (int e = und.get(5 * nh * sizeof(double2))) return e;
    double2 *F = static_cast<double2 *>(full.p), *U = static_cast<double2 *>(und.p);
    if (int e = put_spec(c, F, c_t, c->ct)) return e;
    if (int e = put_spec(c, F, c_ss, c->css)) return e;
    if (int e = spec_forward64(c, c->t, U + 4 * nh)) return e;  // the physical t's z diffusion
    StepArgs A = base_args(c);
    A.ra = rk_a; A.rb = rk_b; A.rc = rk_c;
    A.spund = U; A.kdamp0 = 0; A.stage = 1;
    if (int e = spec_products(c, A, 1)) return e;
    if (int e = launch_spec_columns<double>(c, 1, A, view<double>(c))) return e;
    if (int e = get_spec(c, F, c->ct, c_qt)) return e;
    if (int e = get_spec(c, F, c->css, c_ss)) return e;
    c->spec_dirty = true;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_poisson_solver_fft_spec(hc_ctx *c, int equation, double rk_c, const double *c_src, double *c_var)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = spec_stage_supported(c)) return e;
    if (!c_src || !c_var) return set_err(HC_ERR_ARG, "null array argument");
    if (equation < HC_EQ_U || equation > HC_EQ_T) return set_err(HC_ERR_ARG, "unknown equation %d", equation);
    Scratch full;
    if (int e = full.get(2 * half_elems(c) * sizeof(double2))) return e;
    double2 *F = static_cast<double2 *>(full.p);
    if (int e = put_spec(c, F, c_src, c->s[0])) return e;
    StepArgs A = base_args(c);
    A.rc = c->prm.dts == 2 ? rk_c : 1.0;
    if (int e = launch_columns<double>(c, 3, A, view<double>(c), 1, equation)) return e;
    if (int e = get_spec(c, F, c->s[0], c_var)) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_get_div_spec(hc_ctx *c, const double *c_u, const double *c_v, const double *w, double *c_div,
                               int layout)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = spec_stage_supported(c)) return e;
    if (!c_u || !c_v || !w || !c_div) return set_err(HC_ERR_ARG, "null array argument");
    Scratch full;
    if (int e = full.get(2 * half_elems(c) * sizeof(double2))) return e;
    double2 *F = static_cast<double2 *>(full.p);
    if (int e = put_spec(c, F, c_u, c->cu)) return e;
    if (int e = put_spec(c, F, c_v, c->cv)) return e;
    if (int e = put_f64(c, c->w, w, layout)) return e;
    if (int e = spec_forward64(c, c->w, c->s[2])) return e;
    const StepArgs A = base_args(c);
    hipLaunchKernelGGL((k_spec_div<double2>), dim3(1024), dim3(256), 0, c->stream, A, (const double2 *)c->cu,
                       (const double2 *)c->cv, (const double2 *)c->s[2], c->s[3]);
    HIPCHK(hipGetLastError());
    if (int e = get_spec(c, F, c->s[3], c_div)) return e;
    c->spec_dirty = true;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_calculate_uvw_spec(hc_ctx *c, double rk_c, const double *phi, const double *c_phi, double *c_u,
                                     double *c_v, double *c_w, int layout)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = spec_stage_supported(c)) return e;
    if (!phi || !c_phi || !c_u || !c_v || !c_w) return set_err(HC_ERR_ARG, "null array argument");
    Scratch full;
    if (int e = full.get(2 * half_elems(c) * sizeof(double2))) return e;
    double2 *F = static_cast<double2 *>(full.p);
    if (int e = put_f64(c, c->ptmp, phi, layout)) return e;
    if (int e = spec_forward64(c, c->ptmp, c->s[3])) return e;
    if (int e = put_spec(c, F, c_phi, c->s[2])) return e;
    const double *si[] = {c_u, c_v, c_w};
    double2 *sd[] = {c->cu, c->cv, c->cw};
    for (int q = 0; q < 3; ++q)
        if (int e = put_spec(c, F, si[q], sd[q])) return e;
    const StepArgs A = base_args(c);
    hipLaunchKernelGGL(k_stage_uvw_spec, dim3(1024), dim3(256), 0, c->stream, A, c->cu, c->cv, c->cw,
                       (const double2 *)c->s[2], (const double2 *)c->s[3], rk_c * c->prm.dt);
    HIPCHK(hipGetLastError());
    double *so[] = {c_u, c_v, c_w};
    for (int q = 0; q < 3; ++q)
        if (int e = get_spec(c, F, sd[q], so[q])) return e;
    c->spec_dirty = true;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_calculate_new_p_spec(hc_ctx *c, double rk_c, const double *phi, const double *c_phi, double *c_p,
                                       int layout)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = spec_stage_supported(c)) return e;
    if (!phi || !c_phi || !c_p) return set_err(HC_ERR_ARG, "null array argument");
    Scratch full;
    if (int e = full.get(2 * half_elems(c) * sizeof(double2))) return e;
    double2 *F = static_cast<double2 *>(full.p);
    if (int e = put_f64(c, c->ptmp, phi, layout)) return e;
    if (int e = spec_forward64(c, c->ptmp, c->s[3])) return e;
    if (int e = put_spec(c, F, c_phi, c->s[2])) return e;
    if (int e = put_spec(c, F, c_p, c->cp)) return e;
    const StepArgs A = base_args(c);
    hipLaunchKernelGGL(k_stage_new_p_spec, dim3(1024), dim3(256), 0, c->stream, A, c->cp, (const double2 *)c->s[2],
                       (const double2 *)c->s[3], -rk_c * 0.5 * c->prm.dt * c->prm.nu);
    HIPCHK(hipGetLastError());
    if (int e = get_spec(c, F, c->cp, c_p)) return e;
    c->spec_dirty = true;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_real_2_cmplx(hc_ctx *c, const double *r_in, double *c_out, int layout)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = spec_stage_supported(c)) return e;
    if (!r_in || !c_out) return set_err(HC_ERR_ARG, "null array argument");
    Scratch full;
    if (int e = full.get(2 * half_elems(c) * sizeof(double2))) return e;
    if (int e = put_f64(c, c->ptmp, r_in, layout)) return e;
    if (int e = spec_forward64(c, c->ptmp, c->s[3])) return e;
    if (int e = get_spec(c, static_cast<double2 *>(full.p), c->s[3], c_out)) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_cmplx_2_real(hc_ctx *c, const double *c_in, double *r_out, int layout)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = spec_stage_supported(c)) return e;
    if (!c_in || !r_out) return set_err(HC_ERR_ARG, "null array argument");
    Scratch full;
    if (int e = full.get(2 * half_elems(c) * sizeof(double2))) return e;
    if (int e = put_spec(c, static_cast<double2 *>(full.p), c_in, c->s[3])) return e;
    const double2 *s1[1] = {c->s[3]};
    double *p1[1] = {c->ptmp};
    if (int e = spec_inverse<double>(c, 1, s1, p1)) return e;
    if (int e = get_f64(c, c->ptmp, r_out, layout)) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_spectral_truncation(hc_ctx *c, double *c_inout)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = spec_stage_supported(c)) return e;
    if (!c_inout) return set_err(HC_ERR_ARG, "null array argument");
    Scratch full;
    if (int e = full.get(2 * half_elems(c) * sizeof(double2))) return e;
    double2 *F = static_cast<double2 *>(full.p);
    if (int e = put_spec(c, F, c_inout, c->s[3])) return e;
    hipLaunchKernelGGL((k_truncate<double2>), dim3(1024), dim3(256), 0, c->stream, c->s[3], c->nl, c->ny, c->nz,
                       (int)(c->nx / 3.0), (int)(c->ny / 3.0));
    HIPCHK(hipGetLastError());
    if (int e = get_spec(c, F, c->s[3], c_inout)) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_rayleigh_damping_spec(hc_ctx *c, double *c_inout)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = spec_stage_supported(c)) return e;
    if (c->prm.isdamp != 1) return set_err(HC_ERR_ARG, "isdamp=0: no damped levels");
    if (!c_inout) return set_err(HC_ERR_ARG, "null array argument");
    Scratch full;
    if (int e = full.get(2 * half_elems(c) * sizeof(double2))) return e;
    double2 *F = static_cast<double2 *>(full.p);
    if (int e = put_spec(c, F, c_inout, c->s[3])) return e;
    if (int e = spec_damp_one(c, c->s[3])) return e;
    if (int e = get_spec(c, F, c->s[3], c_inout)) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}
