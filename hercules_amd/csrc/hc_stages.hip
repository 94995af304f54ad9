// Per-stage entry points of the finite-difference substep, with the argument
// lists of the reference subroutines a caller's own time_advancement_fd
// calls (src/module_navier_stokes.f90:247-355):
//   get_momentum_cnvdiff     src/module_cnvdiff.f90:35-374
//   get_temperature_cnvdiff  src/module_cnvdiff.f90:381-473
//   get_p_eqn_src            src/module_poisson_solver.f90:127-176
//   calculate_uvw            src/module_navier_stokes.f90:659-734
//   calculate_new_p          src/module_navier_stokes.f90:739-808
//   calculate_ubar(u)        src/module_parameters.f90:513-546
// Every array argument is a whole field (host or device, interior or
// ghosted layout) that crosses the ABI on each call, so these serve callers
// that keep the reference's host loop; the fused substep of
// hc_time_advancement_fd does the same work without the transfers.  The
// input fields become the context's module fields (u, v, w, p, t and the old
// right-hand sides), as the reference's arguments are its module arrays.
// Ghost cells are never read: wall rules and periodic wrap are applied on
// the fly, which is what update_boundary_* leave in the ghosts.
#include "hc_host.h"

namespace {
template <int CDS>
__global__ __launch_bounds__(256) void k_stage_momentum(StepArgs A, double *ff, double *gg, double *hh, double *qu,
                                                         double *qv, double *qw)
{
    const size_t plane = (size_t)A.nx * A.ny, n = plane * A.nz;
    const double dpx = A.dp_opt == 2 ? A.scal[0] : A.dpx;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(q / plane), r = (int)(q - (size_t)k * plane), j = r / A.nx, i = r - j * A.nx;
        Pt P{A, i, j, k};
        double f, g, h, a, b, d;
        momentum_point<CDS>(P, dpx, f, g, h, a, b, d);
        ff[q] = f; gg[q] = g; hh[q] = h;
        qu[q] = a; qv[q] = b; qw[q] = d;
    }
}

template <int CDS>
__global__ __launch_bounds__(256) void k_stage_temperature(StepArgs A, const double *ss0, double *ss, double *qt)
{
    const size_t plane = (size_t)A.nx * A.ny, n = plane * A.nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(q / plane), r = (int)(q - (size_t)k * plane), j = r / A.nx, i = r - j * A.nx;
        Pt P{A, i, j, k};
        double s, t;
        temperature_point<CDS>(P, s, t, A.use_f0 ? ss0[q] : 0.0);
        ss[q] = s;
        qt[q] = t;
    }
}

// qp = (D_x u + D_y v + D_z w) / dt
template <int CDS>
__global__ __launch_bounds__(256) void k_stage_p_src(StepArgs A, double *qp)
{
    const size_t plane = (size_t)A.nx * A.ny, n = plane * A.nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(q / plane), r = (int)(q - (size_t)k * plane), j = r / A.nx, i = r - j * A.nx;
        qp[q] = div_at<CDS, double>(A, A, i, j, k) / A.dt;
    }
}

HC_DEV double phi_at(const StepArgs &A, const double *phi, int i, int j, int k)
{
    return phi[((size_t)k * A.ny + wrapi(j, A.ny)) * A.nx + wrapi(i, A.nx)];
}

// u -= c G_x phi, v -= c G_y phi on every level; w -= c (phi(k+1) - phi(k)) /
// dz_t(k) below the top face (c = rk_c dt)
template <int CDS>
__global__ __launch_bounds__(256) void k_stage_project(StepArgs A, double *u, double *v, double *w, const double *phi,
                                                        double c)
{
    const size_t plane = (size_t)A.nx * A.ny, n = plane * A.nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(q / plane), r = (int)(q - (size_t)k * plane), j = r / A.nx, i = r - j * A.nx;
        double gx, gy;
        if (CDS == 1) {
            gx = (phi_at(A, phi, i + 1, j, k) - phi_at(A, phi, i, j, k)) / A.dx;
            gy = (phi_at(A, phi, i, j + 1, k) - phi_at(A, phi, i, j, k)) / A.dy;
        } else {
            gx = (-phi_at(A, phi, i + 2, j, k) + 27.0 * phi_at(A, phi, i + 1, j, k) - 27.0 * phi_at(A, phi, i, j, k) +
                  phi_at(A, phi, i - 1, j, k)) / 24.0 / A.dx;
            gy = (-phi_at(A, phi, i, j + 2, k) + 27.0 * phi_at(A, phi, i, j + 1, k) - 27.0 * phi_at(A, phi, i, j, k) +
                  phi_at(A, phi, i, j - 1, k)) / 24.0 / A.dy;
        }
        u[q] = u[q] - c * gx;
        v[q] = v[q] - c * gy;
        if (k < A.nz - 1)
            w[q] = w[q] - c * (phi[q + plane] - phi[q]) / A.lev[LV_DZT * A.nz + k];
    }
}

// p += phi - c/2 dt nu (phi_xx + phi_yy + phi_zz), Neumann ghosts of phi
// (update_boundary_p: phi(kstr3-1) = phi(kstr3), phi(kend3+1) = phi(kend3))
template <int CDS>
__global__ __launch_bounds__(256) void k_stage_new_p(StepArgs A, const double *phi, double *p, double c)
{
    const size_t plane = (size_t)A.nx * A.ny, n = plane * A.nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(q / plane), r = (int)(q - (size_t)k * plane), j = r / A.nx, i = r - j * A.nx;
        const double p0 = phi[q];
        double pxx, pyy;
        if (CDS == 1) {
            pxx = (phi_at(A, phi, i + 1, j, k) - 2.0 * p0 + phi_at(A, phi, i - 1, j, k)) / A.dx2;
            pyy = (phi_at(A, phi, i, j + 1, k) - 2.0 * p0 + phi_at(A, phi, i, j - 1, k)) / A.dy2;
        } else {
            pxx = (-phi_at(A, phi, i + 2, j, k) + 16.0 * phi_at(A, phi, i + 1, j, k) - 30.0 * p0 +
                   16.0 * phi_at(A, phi, i - 1, j, k) - phi_at(A, phi, i - 2, j, k)) / 12.0 / A.dx2;
            pyy = (-phi_at(A, phi, i, j + 2, k) + 16.0 * phi_at(A, phi, i, j + 1, k) - 30.0 * p0 +
                   16.0 * phi_at(A, phi, i, j - 1, k) - phi_at(A, phi, i, j - 2, k)) / 12.0 / A.dy2;
        }
        const double pu = k < A.nz - 1 ? phi[q + plane] : p0, pd = k > 0 ? phi[q - plane] : p0;
        const double pzz = ((pu - p0) / A.lev[LV_DZT * A.nz + k] - (p0 - pd) / A.lev[LV_DZB * A.nz + k]) /
                           A.lev[LV_DZ * A.nz + k];
        p[q] = p[q] + p0 - c * 0.5 * A.dt * A.nu * (pxx + pyy + pzz);
    }
}

constexpr int kStageBlocks = 2048;

// The stages restate the reference's own operators: FP64, cds = 1 or 2,
// 2nd-order z, and the fused right-hand side without damping reads.
int stage_supported(const hc_ctx *c)
{
    const hc_params &P = c->prm;
    if (c->mixed) return set_err(HC_ERR_UNSUPPORTED, "the per-stage entry points are FP64 (precision=0)");
    if (P.cds == 3) return set_err(HC_ERR_ARG, "cds=3: the finite-difference stages need cds=1 or 2");
    if (P.zorder == 4 || P.zorder == 6) return set_err(HC_ERR_UNSUPPORTED, "the per-stage entry points use the reference's 2nd-order z");
    return 0;
}

StepArgs stage_args(hc_ctx *c, double ra, double rb, double rc)
{
    StepArgs A = base_args(c);
    A.ra = ra; A.rb = rb; A.rc = rc;
    A.use_f0 = rb != 0.0;
    return A;
}
}  // namespace

extern "C" int hc_get_momentum_cnvdiff(hc_ctx *c, double rk_a, double rk_b, double rk_c, const double *u,
                                       const double *v, const double *w, const double *p, const double *t, double *ff,
                                       const double *ff0, double *gg, const double *gg0, double *hh,
                                       const double *hh0, double *qu, double *qv, double *qw, int layout)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = stage_supported(c)) return e;
    if (!u || !v || !w || !p || !ff || !ff0 || !gg || !gg0 || !hh || !hh0 || !qu || !qv || !qw ||
        (!t && c->prm.isscalar == 1))
        return set_err(HC_ERR_ARG, "null array argument");
    const int cur = c->fcur, nxt = cur ^ 1;
    const double *in[] = {u, v, w, p, ff0, gg0, hh0};
    double *dev[] = {c->u, c->v, c->w, c->p, c->fb[cur][0], c->fb[cur][1], c->fb[cur][2]};
    for (int q = 0; q < 7; ++q)
        if (int e = put_f64(c, dev[q], in[q], layout)) return e;
    if (t)
        if (int e = put_f64(c, c->t, t, layout)) return e;
    StepArgs A = stage_args(c, rk_a, rk_b, rk_c);
    A.ff0 = c->fb[cur][0]; A.gg0 = c->fb[cur][1]; A.hh0 = c->fb[cur][2];
    double *q3[3] = {(double *)c->s[0], (double *)c->s[1], (double *)c->s[2]};
    // update_dpx_drive is the first statement of get_momentum_cnvdiff (:51)
    if (c->prm.dp_opt == 2) dpx_update_launch(c);
    if (c->prm.cds == 1)
        hipLaunchKernelGGL(k_stage_momentum<1>, dim3(kStageBlocks), dim3(256), 0, c->stream, A, c->fb[nxt][0],
                           c->fb[nxt][1], c->fb[nxt][2], q3[0], q3[1], q3[2]);
    else
        hipLaunchKernelGGL(k_stage_momentum<2>, dim3(kStageBlocks), dim3(256), 0, c->stream, A, c->fb[nxt][0],
                           c->fb[nxt][1], c->fb[nxt][2], q3[0], q3[1], q3[2]);
    HIPCHK(hipGetLastError());
    c->fcur = nxt;
    double *out[] = {ff, gg, hh, qu, qv, qw};
    const double *src[] = {c->fb[nxt][0], c->fb[nxt][1], c->fb[nxt][2], q3[0], q3[1], q3[2]};
    for (int q = 0; q < 6; ++q)
        if (int e = get_f64(c, src[q], out[q], layout)) return e;
    c->spec_dirty = true;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_get_temperature_cnvdiff(hc_ctx *c, double rk_a, double rk_b, double rk_c, const double *u,
                                          const double *v, const double *w, const double *t, double *ss,
                                          const double *ss0, double *qt, int layout)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = stage_supported(c)) return e;
    if (c->prm.isscalar != 1) return set_err(HC_ERR_ARG, "get_temperature_cnvdiff needs isscalar=1");
    if (!u || !v || !w || !t || !ss || !ss0 || !qt) return set_err(HC_ERR_ARG, "null array argument");
    const int cur = c->scur, nxt = cur ^ 1;
    const double *in[] = {u, v, w, t, ss0};
    double *dev[] = {c->u, c->v, c->w, c->t, c->sb[cur]};
    for (int q = 0; q < 5; ++q)
        if (int e = put_f64(c, dev[q], in[q], layout)) return e;
    StepArgs A = stage_args(c, rk_a, rk_b, rk_c);
    double *qbuf = (double *)c->s[0];
    if (c->prm.cds == 1)
        hipLaunchKernelGGL(k_stage_temperature<1>, dim3(kStageBlocks), dim3(256), 0, c->stream, A, c->sb[cur],
                           c->sb[nxt], qbuf);
    else
        hipLaunchKernelGGL(k_stage_temperature<2>, dim3(kStageBlocks), dim3(256), 0, c->stream, A, c->sb[cur],
                           c->sb[nxt], qbuf);
    HIPCHK(hipGetLastError());
    c->scur = nxt;
    if (int e = get_f64(c, c->sb[nxt], ss, layout)) return e;
    if (int e = get_f64(c, qbuf, qt, layout)) return e;
    c->spec_dirty = true;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_get_p_eqn_src(hc_ctx *c, const double *u, const double *v, const double *w, double *qp, int layout)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = stage_supported(c)) return e;
    if (!u || !v || !w || !qp) return set_err(HC_ERR_ARG, "null array argument");
    const double *in[] = {u, v, w};
    double *dev[] = {c->u, c->v, c->w};
    for (int q = 0; q < 3; ++q)
        if (int e = put_f64(c, dev[q], in[q], layout)) return e;
    StepArgs A = base_args(c);
    double *qbuf = (double *)c->s[0];
    if (c->prm.cds == 1)
        hipLaunchKernelGGL(k_stage_p_src<1>, dim3(kStageBlocks), dim3(256), 0, c->stream, A, qbuf);
    else
        hipLaunchKernelGGL(k_stage_p_src<2>, dim3(kStageBlocks), dim3(256), 0, c->stream, A, qbuf);
    HIPCHK(hipGetLastError());
    if (int e = get_f64(c, qbuf, qp, layout)) return e;
    c->spec_dirty = true;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_calculate_uvw(hc_ctx *c, double rk_c, double *u, double *v, double *w, const double *phi, int layout)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = stage_supported(c)) return e;
    if (!u || !v || !w || !phi) return set_err(HC_ERR_ARG, "null array argument");
    const double *in[] = {u, v, w};
    double *dev[] = {c->u, c->v, c->w};
    for (int q = 0; q < 3; ++q)
        if (int e = put_f64(c, dev[q], in[q], layout)) return e;
    double *pbuf = (double *)c->s[0];
    if (int e = put_f64(c, pbuf, phi, layout)) return e;
    StepArgs A = base_args(c);
    const double cc = rk_c * c->prm.dt;
    if (c->prm.cds == 1)
        hipLaunchKernelGGL(k_stage_project<1>, dim3(kStageBlocks), dim3(256), 0, c->stream, A, c->u, c->v, c->w, pbuf,
                           cc);
    else
        hipLaunchKernelGGL(k_stage_project<2>, dim3(kStageBlocks), dim3(256), 0, c->stream, A, c->u, c->v, c->w, pbuf,
                           cc);
    HIPCHK(hipGetLastError());
    double *out[] = {u, v, w};
    const double *src[] = {c->u, c->v, c->w};
    for (int q = 0; q < 3; ++q)
        if (int e = get_f64(c, src[q], out[q], layout)) return e;
    c->spec_dirty = true;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_calculate_new_p(hc_ctx *c, double rk_c, const double *phi, double *p, int layout)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = stage_supported(c)) return e;
    if (!phi || !p) return set_err(HC_ERR_ARG, "null array argument");
    double *pbuf = (double *)c->s[0];
    if (int e = put_f64(c, c->p, p, layout)) return e;
    if (int e = put_f64(c, pbuf, phi, layout)) return e;
    StepArgs A = base_args(c);
    if (c->prm.cds == 1)
        hipLaunchKernelGGL(k_stage_new_p<1>, dim3(kStageBlocks), dim3(256), 0, c->stream, A, pbuf, c->p, rk_c);
    else
        hipLaunchKernelGGL(k_stage_new_p<2>, dim3(kStageBlocks), dim3(256), 0, c->stream, A, pbuf, c->p, rk_c);
    HIPCHK(hipGetLastError());
    if (int e = get_f64(c, c->p, p, layout)) return e;
    c->spec_dirty = true;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_calculate_ubar_of(hc_ctx *c, const double *u, int layout, double *ubar)
{
    if (!c || !u) return set_err(HC_ERR_ARG, "null argument");
    double *ubuf = (double *)c->s[0];
    if (int e = put_f64(c, ubuf, u, layout)) return e;
    if (int e = ubar_of(c, ubuf)) return e;
    if (ubar) HIPCHK(hipMemcpyAsync(ubar, c->scal + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}
