// MI355X-native HERCULES time-step hot path: the C ABI declared in
// include/hercules_amd.h, context set-up (grid metrics, equation tables,
// buffers), the FD substep sequence and the small reduction kernels.
//
// One substep of the finite-difference RK3 step (time_advancement_fd,
// src/module_navier_stokes.f90:247-355) is five launches:
//   k_rhs_mom2    right-hand sides ff,gg,hh and qu,qv,qw, fused with the
//                 x-direction real-to-complex FFT of the q's (hc_rhs.hip)
//   k_yfft<-1>    y-direction FFT of the three q spectra (hc_xfer.hip)
//   k_column_mom  per (l,m) column: Helmholtz solves for u*, v*, w*, the
//                 divergence source, the pressure Poisson solve, the
//                 projection and the pressure increment, all in spectral
//                 space (exact equivalents of get_p_eqn_src, calculate_uvw
//                 and calculate_new_p for periodic x, y) (hc_columns.hip)
//   k_yfft<+1>    inverse y FFT of u, v, w and the pressure increment
//   k_x_c2r       inverse x FFT; stores u, v, w and accumulates p
// plus, with temperature, the same chain for T.
#include "hc_host.h"

// ---------------------------------------------------------------------------
// error handling
static thread_local std::string g_err;
int set_err(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char *hc_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// small kernels
// dpx_drive update of dp_opt=2 (update_dpx_drive, src/module_parameters.f90:594-613)
__global__ void k_dpx_update(double *scal, double u_mrf, double coef)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) scal[0] = scal[0] - (scal[1] + u_mrf - 1.0) * coef;
}
void dpx_update_launch(hc_ctx *c)
{
    hipLaunchKernelGGL(k_dpx_update, dim3(1), dim3(64), 0, c->stream, c->scal, c->prm.u_mrf, c->prm.nu * 1.0);
}

// calculate_ubar (src/module_parameters.f90:513-546): per-level partial sums
// then a fixed-order final sum, so the result is deterministic.
template <class R>
__global__ __launch_bounds__(256) void k_ubar_partial(const R *__restrict__ u, const double *lev, int nx, int ny,
                                                      int nz, double *part)
{
    __shared__ double red[256];
    const int k = blockIdx.x;
    double s = 0.0;
    const size_t n = (size_t)nx * ny;
    for (size_t q = threadIdx.x; q < n; q += 256) s += u[(size_t)k * n + q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[k] = red[0] * lev[LV_DZ * nz + k];
}
// plane sums of one level per block: part[k] = sum of f over level k
template <class R>
__global__ __launch_bounds__(256) void k_plane_sum(const R *__restrict__ f, int nx, int ny, int k0, double *part)
{
    __shared__ double red[256];
    const int k = k0 + blockIdx.x;
    double s = 0.0;
    const size_t n = (size_t)nx * ny;
    for (size_t q = threadIdx.x; q < n; q += 256) s += f[(size_t)k * n + q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void k_ubar_final(const double *part, int nz, double lz, int nx, int ny, double *scal)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < nz; ++k) s += part[k];
        scal[1] = s / lz / nx / ny;
    }
}

// screen_div_error values (src/module_io.f90:85-206): per-level sums of
// |div| and max |div|, NaN flag.  cds as in the reference.  F holds the
// fields in their stored precision; the divergence is formed in FP64 with
// the metrics of D.
template <int CDS, class R>
__global__ __launch_bounds__(256) void k_div_partial(StepArgsT<R> F, StepArgs D, double *part)
{
    __shared__ double rs[256], rm[256];
    __shared__ int rn[256];
    const int k = blockIdx.x;
    double s = 0.0, mx = 0.0;
    int nan = 0;
    const size_t n = (size_t)D.nx * D.ny;
    const double dz = D.lev[LV_DZ * D.nz + k];
    for (size_t q = threadIdx.x; q < n; q += 256) {
        int j = (int)(q / D.nx), i = (int)(q - (size_t)j * D.nx);
        PtT<R> P{F, i, j, k};
        auto U = [&](int di) -> double { return (double)P.U(di, 0, 0); };
        auto V = [&](int dj) -> double { return (double)P.V(0, dj, 0); };
        double dudx = 0.0, dvdy = 0.0;
        if (CDS == 1) {
            dudx = (U(0) - U(-1)) / D.dx;
            dvdy = (V(0) - V(-1)) / D.dy;
        } else if (CDS == 2) {
            dudx = (-U(1) + 27.0 * U(0) - 27.0 * U(-1) + U(-2)) / 24.0 / D.dx;
            dvdy = (-V(1) + 27.0 * V(0) - 27.0 * V(-1) + V(-2)) / 24.0 / D.dy;
        }
        double w0 = F.w[P.at(0, 0, k)];
        double wm = k == 0 ? 0.0 : (double)F.w[P.at(0, 0, k - 1)];
        double dwdz = (w0 - wm) / dz;
        if (D.zorder == 6) {  // ZO_D over the faces, zero at the walls
            const size_t e = (size_t)ZO_D * D.nz + k;
            const int s0 = D.zst[e];
            dwdz = 0.0;
            for (int jz = 0; jz < D.zn1; ++jz) {
                const int dd = s0 + jz;
                if (dd > 0 && dd < D.nz) dwdz += D.zwt[e * kZW + jz] * (double)F.w[P.at(0, 0, dd - 1)];
            }
        }
        // spectral path: the divergence field was formed and inverted beforehand
        double d = CDS == 3 ? (double)F.aux[P.at(0, 0, k)] : dudx + dvdy + dwdz;
        s += fabs(d);
        mx = fmax(mx, fabs(d));
        double u0 = F.u[P.at(0, 0, k)], v0 = F.v[P.at(0, 0, k)];
        if (u0 != u0 || v0 != v0 || w0 != w0) nan = 1;
    }
    rs[threadIdx.x] = s;
    rm[threadIdx.x] = mx;
    rn[threadIdx.x] = nan;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            rs[threadIdx.x] += rs[threadIdx.x + w];
            rm[threadIdx.x] = fmax(rm[threadIdx.x], rm[threadIdx.x + w]);
            rn[threadIdx.x] |= rn[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * k] = rs[0];
        part[3 * k + 1] = rm[0];
        part[3 * k + 2] = rn[0];
    }
}


// Running sums of u + u_mrf, v, w, t, (u + u_mrf)^2, v^2, w^2, w_c (uh + u_mrf),
// w_c vh, t^2, t w_c with w_c = (w(k) + w(k-1))/2, w(-1) = 0; uh, vh: the
// interpolants (FD) or u, v (spectral, uhx = vhy = null).  FP64 sums.
struct MeanSums {
    double *m[HC_MEAN_COUNT];
};
template <class R>
__global__ __launch_bounds__(256) void k_mean_accum(const R *u, const R *v, const R *w, const R *t, const double *uhx,
                                                    const double *vhy, MeanSums S, int nx, int ny, int nz, double um)
{
    const size_t n = (size_t)nx * ny * nz, plane = (size_t)nx * ny;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const double uu = (double)u[q], vv = (double)v[q], ww = (double)w[q], tt = (double)t[q];
        const double wc = 0.5 * (ww + (q >= plane ? (double)w[q - plane] : 0.0));
        const double uh = uhx ? uhx[q] : uu, vh = vhy ? vhy[q] : vv;
        S.m[HC_MEAN_UU][q] += (uu + um) * (uu + um);
        S.m[HC_MEAN_VV][q] += vv * vv;
        S.m[HC_MEAN_WW][q] += ww * ww;
        S.m[HC_MEAN_UW][q] += wc * (uh + um);
        S.m[HC_MEAN_VW][q] += wc * vh;
        S.m[HC_MEAN_TT][q] += tt * tt;
        S.m[HC_MEAN_TW][q] += tt * wc;
        S.m[HC_MEAN_U][q] += uu + um;
        S.m[HC_MEAN_V][q] += vv;
        S.m[HC_MEAN_W][q] += ww;
        S.m[HC_MEAN_T][q] += tt;
    }
}
// CFL number of screen_cpu_time (src/module_io.f90:32-80): per level, the
// largest dt (|u|/dx + |v|/dy + |w|/dz(k)) over the plane (u relative to
// the moving frame, point values as the reference takes them).
template <class R>
__global__ __launch_bounds__(256) void k_cfl_partial(const R *u, const R *v, const R *w, const double *lev, int nx,
                                                     int ny, int nz, double dt, double dx, double dy, double *part)
{
    __shared__ double rm[256];
    const int k = blockIdx.x;
    const size_t n = (size_t)nx * ny, o = (size_t)k * n;
    const double idz = 1.0 / lev[LV_DZ * nz + k];
    double m = 0.0;
    for (size_t q = threadIdx.x; q < n; q += 256)
        m = fmax(m, dt * (fabs((double)u[o + q]) / dx + fabs((double)v[o + q]) / dy + fabs((double)w[o + q]) * idz));
    rm[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) rm[threadIdx.x] = fmax(rm[threadIdx.x], rm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[k] = rm[0];
}

// one z column (i, j) of a field, in FP64
template <class R>
__global__ __launch_bounds__(256) void k_get_column(const R *f, size_t plane, size_t off, int nz, double *dst)
{
    for (int k = threadIdx.x; k < nz; k += 256) dst[k] = (double)f[(size_t)k * plane + off];
}

__global__ __launch_bounds__(256) void k_scale_copy(const double *src, double *dst, size_t n, double s)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        dst[q] = src[q] * s;
}

// interior <-> ghosted (nx+4, ny+4, nz+4) Fortran layout; the interior
// array is in the stored precision R
template <class R>
__global__ void k_pack_ghosted(const double *__restrict__ g, R *dst, int nx, int ny, int nz)
{
    size_t n = (size_t)nx * ny * nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((size_t)nx * ny));
        dst[q] = R(g[((size_t)(k + 2) * (ny + 4) + (j + 2)) * (nx + 4) + (i + 2)]);
    }
}
template <class R>
__global__ void k_unpack_ghosted(const R *__restrict__ src, double *g, int nx, int ny, int nz)
{
    size_t n = (size_t)nx * ny * nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((size_t)nx * ny));
        g[((size_t)(k + 2) * (ny + 4) + (j + 2)) * (nx + 4) + (i + 2)] = (double)src[q];
    }
}
// element-wise precision conversion (rounded to nearest)
template <class S, class D>
__global__ void k_convert(const S *__restrict__ src, D *dst, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        dst[q] = D(src[q]);
}
// f[off + q] = val for q < n
__global__ void k_fill(double *f, size_t off, size_t n, double val)
{
    size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q < n) f[off + q] = val;
}
template <class R>
__global__ void k_zero_level(R *f, size_t off, size_t n)
{
    size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q < n) f[off + q] = R(0);
}

int dalloc(hc_ctx *c, void **p, size_t bytes)
{
    HIPCHK(hipMalloc(p, bytes));
    HIPCHK(hipMemsetAsync(*p, 0, bytes, c->stream));
    c->allocs.push_back(*p);
    return 0;
}

// Grid metrics, get_mesh_param (src/module_parameters.f90:432-508), in the
// reference's 1-based ghosted indexing (k = 1..nz+4).
static void mesh_1based(const hc_params &P, std::vector<double> &dz, std::vector<double> &lt,
                        std::vector<double> &lb, std::vector<double> &dzt, std::vector<double> &dzb)
{
    const int nka = P.nz + 4, ks = 3, ke = P.nz + 2;
    std::vector<double> zz(nka + 1, 0.0);
    dz.assign(nka + 2, 0.0); lt = lb = dzt = dzb = dz;
    double zmin = 0.0, zmax = P.lz, zmean, a;
    if (P.ochannel == 0) { zmean = 0.5 * (zmin + zmax); a = (2.0 + P.nz) * 0.5; }
    else { zmean = zmin + zmax; a = (1.0 + P.nz) * 1.0; }
    for (int k = 1; k <= P.nz + 1; ++k)
        zz[k + 2] = zmean + (zmin - zmean) / std::tanh(2.0 * P.zstretch) * std::tanh(2.0 * P.zstretch * (k - a) / (1.0 - a));
    for (int k = ks; k <= ke; ++k) dz[k] = zz[k + 1] - zz[k];
    dz[ks - 1] = dz[ks]; dz[ke + 1] = dz[ke]; dz[ks - 2] = dz[ks]; dz[ke + 2] = dz[ke];
    for (int k = ks - 1; k <= ke + 1; ++k) {
        lt[k] = dz[k] / (dz[k] + dz[k + 1]);
        lb[k] = dz[k] / (dz[k] + dz[k - 1]);
        dzt[k] = 0.5 * (dz[k] + dz[k + 1]);
        dzb[k] = 0.5 * (dz[k] + dz[k - 1]);
    }
}

// The faces z_f(0..nz) of get_mesh_param (src/module_parameters.f90:445-466).
static void mesh_faces(const hc_params &P, std::vector<double> &zf)
{
    const int nz = P.nz;
    zf.assign(nz + 1, 0.0);
    double zmin = 0.0, zmax = P.lz, zmean, a;
    if (P.ochannel == 0) { zmean = 0.5 * (zmin + zmax); a = (2.0 + nz) * 0.5; }
    else { zmean = zmin + zmax; a = (1.0 + nz) * 1.0; }
    for (int k = 1; k <= nz + 1; ++k)
        zf[k - 1] = zmean + (zmin - zmean) / std::tanh(2.0 * P.zstretch) * std::tanh(2.0 * P.zstretch * (k - a) / (1.0 - a));
}

// Per-level tables: metrics and the 1-D coefficients of get_{u,w,p,t}_eqn_coeff
// (src/module_poisson_solver.f90:32-477) with their wall rows, for general
// modes and for the (0,0) mode.
static void build_levels(hc_ctx *c)
{
    const hc_params &P = c->prm;
    const int nz = P.nz;
    std::vector<double> dz, lt, lb, dzt, dzb;
    mesh_1based(P, dz, lt, lb, dzt, dzb);
    std::vector<double> &T = c->h_lev;
    T.assign((size_t)LV_COUNT * nz, 0.0);
    auto at = [&](int tab, int kk) -> double & { return T[(size_t)tab * nz + kk]; };
    for (int kk = 0; kk < nz; ++kk) {
        int k = kk + 3;
        at(LV_DZ, kk) = dz[k]; at(LV_DZT, kk) = dzt[k]; at(LV_DZB, kk) = dzb[k]; at(LV_LT, kk) = lt[k];
        at(LV_IDZ, kk) = 1.0 / dz[k]; at(LV_IDZT, kk) = 1.0 / dzt[k]; at(LV_IDZB, kk) = 1.0 / dzb[k];
        at(LV_LB, kk) = lb[k];
        // u (v): aa, cc of get_u_eqn_coeff (:200-201)
        double a = -0.5 * P.dt * P.nu / dz[k] / dzb[k], cc = -0.5 * P.dt * P.nu / dz[k] / dzt[k];
        at(LV_AU, kk) = kk == 0 ? 0.0 : a;
        at(LV_CU, kk) = kk == nz - 1 ? 0.0 : cc;
        double d = -a - cc, d00 = -a - cc;
        if (kk == 0) { d = -cc; d00 = -2.0 * a - cc; }
        if (kk == nz - 1) { d = -a; d00 = P.ochannel == 0 ? -a - 2.0 * cc : -a; }
        at(LV_DU, kk) = P.wall_rows ? d00 : d; at(LV_DU00, kk) = d00;
        // t: get_t_eqn_coeff (:398-399), Dirichlet at both walls; with
        // tbc_top = 1 the top row takes the gradient ghost t + g dz_t
        a = -0.5 * P.dt * P.alpha / dz[k] / dzb[k]; cc = -0.5 * P.dt * P.alpha / dz[k] / dzt[k];
        at(LV_AT, kk) = kk == 0 ? 0.0 : a;
        at(LV_CT, kk) = kk == nz - 1 ? 0.0 : cc;
        d = d00 = -a - cc;
        if (kk == 0) { d = -cc; d00 = -2.0 * a - cc; }
        if (kk == nz - 1) { d = -a; d00 = P.tbc_top == 1 ? -a : -a - 2.0 * cc; }
        at(LV_DT, kk) = P.wall_rows ? d00 : d; at(LV_DT00, kk) = d00;
        // p: get_p_eqn_coeff (:51-52), Neumann at both walls
        a = 1.0 / dz[k] / dzb[k]; cc = 1.0 / dz[k] / dzt[k];
        at(LV_AP, kk) = kk == 0 ? 0.0 : a;
        at(LV_CP, kk) = kk == nz - 1 ? 0.0 : cc;
        d = -a - cc;
        if (kk == 0) d = -cc;
        if (kk == nz - 1) d = -a;
        at(LV_DP, kk) = d; at(LV_DP00, kk) = d;
        // w: get_w_eqn_coeff (:306-307), levels kstr3..kend3-1
        if (kk < nz - 1) {
            a = -0.5 * P.dt * P.nu / dz[k] / dzt[k]; cc = -0.5 * P.dt * P.nu / dz[k + 1] / dzt[k];
            at(LV_AW, kk) = kk == 0 ? 0.0 : a;
            at(LV_CW, kk) = kk == nz - 2 ? 0.0 : cc;
            d = -a - cc;
            if (kk == 0) d = -cc;
            if (kk == nz - 2) d = -a;
            at(LV_DW, kk) = P.wall_rows ? -a - cc : d; at(LV_DW00, kk) = -a - cc;
        }
    }
    int ks = 3, ke = nz + 2;
    c->aU0 = -0.5 * P.dt * P.nu / dz[ks] / dzb[ks];
    c->cUe = -0.5 * P.dt * P.nu / dz[ke] / dzt[ke];
    c->aT0 = -0.5 * P.dt * P.alpha / dz[ks] / dzb[ks];
    c->cTe = -0.5 * P.dt * P.alpha / dz[ke] / dzt[ke];
}

// ---- 6th-order z operators (zorder = 6) ------------------------------------
// Weights w of the stencil sum_j w_j data_j for the derivative `der` at z0,
// data_j the value (kind 0) or the first derivative (kind 1) at pts[j]: the
// n weights that make the stencil exact for every polynomial of degree < n,
// from the n x n system over the monomials ((z - z0)/h)^i, solved with
// partial pivoting in long double.
static bool fd_weights(double z0, const double *pts, const int *kind, int n, int der, double *w)
{
    long double h = 0.0L;
    for (int j = 0; j < n; ++j) h = std::max(h, std::fabs((long double)pts[j] - z0));
    if (h == 0.0L) return false;
    long double M[kZW][kZW + 1];
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            const long double sj = ((long double)pts[j] - z0) / h;
            if (kind[j] == 0) M[i][j] = std::pow(sj, (long double)i);
            else M[i][j] = i == 0 ? 0.0L : i * std::pow(sj, (long double)(i - 1)) / h;
        }
        long double f = 1.0L;
        for (int q = 2; q <= der; ++q) f *= q;
        M[i][n] = i == der ? f / std::pow(h, (long double)der) : 0.0L;
    }
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(M[r][col]) > std::fabs(M[piv][col])) piv = r;
        if (M[piv][col] == 0.0L) return false;
        if (piv != col)
            for (int q = 0; q <= n; ++q) std::swap(M[piv][q], M[col][q]);
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const long double f = M[r][col] / M[col][col];
            for (int q = col; q <= n; ++q) M[r][q] -= f * M[col][q];
        }
    }
    for (int j = 0; j < n; ++j) w[j] = (double)(M[j][n] / M[j][j]);
    return true;
}

// Stencil tables of the ZOp operators on the stretched grid: for each target
// the n consecutive data entries centred on it (shifted inwards at the walls),
// and their weights.  ZO_P is the product ZO_D o ZO_G restricted to the
// interior faces (w = 0 at the walls).
static bool build_zops(hc_ctx *c)
{
    const hc_params &P = c->prm;
    const int nz = P.nz;
    // second derivatives from zorder + 1 data, first derivatives and
    // interpolation from zorder data
    const int n2 = P.zorder + 1, n1 = P.zorder;
    std::vector<double> zf, zc;
    mesh_faces(P, zf);
    zc.resize(nz);
    for (int k = 0; k < nz; ++k) zc[k] = 0.5 * (zf[k] + zf[k + 1]);
    c->h_zst.assign((size_t)ZO_COUNT * nz, 0);
    c->h_zwt.assign((size_t)ZO_COUNT * nz * kZW, 0.0);
    // data sets: C (wall datum, centres, wall datum), F (faces), P (centres)
    auto cpos = [&](int d) { return d == 0 ? zf[0] : d == nz + 1 ? zf[nz] : zc[d - 1]; };
    auto row = [&](int op, int target, int ndata, double center, int n, double z0, int der, bool ntop,
                   const std::function<double(int)> &pos) -> bool {
        int lo = (int)std::ceil(center - 0.5 * n);
        lo = std::max(0, std::min(lo, ndata - n));
        double pts[kZW], w[kZW];
        int kind[kZW];
        for (int j = 0; j < n; ++j) {
            pts[j] = pos(lo + j);
            kind[j] = (ntop && lo + j == nz + 1) ? 1 : 0;
        }
        if (!fd_weights(z0, pts, kind, n, der, w)) return false;
        const size_t e = (size_t)op * nz + target;
        c->h_zst[e] = lo;
        for (int j = 0; j < n; ++j) c->h_zwt[e * kZW + j] = w[j];
        return true;
    };
    auto fpos = [&](int d) { return zf[d]; };
    auto ppos = [&](int d) { return zc[d]; };
    bool ok = true;
    for (int k = 0; k < nz && ok; ++k) {
        ok = ok && row(ZO_C2_DD, k, nz + 2, k + 1.0, n2, zc[k], 2, false, cpos);
        ok = ok && row(ZO_C2_DN, k, nz + 2, k + 1.0, n2, zc[k], 2, true, cpos);
        ok = ok && row(ZO_D, k, nz + 1, k + 0.5, n1, zc[k], 1, false, fpos);
        ok = ok && row(ZO_IFC, k, nz + 1, k + 0.5, n1, zc[k], 0, false, fpos);
        if (k < nz - 1) {  // interior face kw = k, at z_f(k + 1)
            ok = ok && row(ZO_F2, k, nz + 1, k + 1.0, n2, zf[k + 1], 2, false, fpos);
            ok = ok && row(ZO_G, k, nz, k + 0.5, n1, zf[k + 1], 1, false, ppos);
            ok = ok && row(ZO_I_DD, k, nz + 2, k + 1.5, n1, zf[k + 1], 0, false, cpos);
            ok = ok && row(ZO_I_DN, k, nz + 2, k + 1.5, n1, zf[k + 1], 0, true, cpos);
        } else {  // the top wall face (row nz - 1 of the interpolations)
            ok = ok && row(ZO_I_DD, k, nz + 2, k + 1.5, n1, zf[nz], 0, false, cpos);
            ok = ok && row(ZO_I_DN, k, nz + 2, k + 1.5, n1, zf[nz], 0, true, cpos);
        }
    }
    if (!ok) return false;
    // ZO_P row k = sum over the interior faces d of D(k, d) G(d - 1, .)
    std::vector<double> dense(nz);
    for (int k = 0; k < nz; ++k) {
        std::fill(dense.begin(), dense.end(), 0.0);
        const size_t ed = (size_t)ZO_D * nz + k;
        for (int j = 0; j < n1; ++j) {
            const int d = c->h_zst[ed] + j;
            if (d == 0 || d == nz) continue;
            const size_t eg = (size_t)ZO_G * nz + (d - 1);
            for (int i = 0; i < n1; ++i) dense[c->h_zst[eg] + i] += c->h_zwt[ed * kZW + j] * c->h_zwt[eg * kZW + i];
        }
        int first = nz, last = -1;
        for (int q = 0; q < nz; ++q)
            if (dense[q] != 0.0) { first = std::min(first, q); last = q; }
        if (last < first || last - first + 1 > kZW) return false;
        const size_t e = (size_t)ZO_P * nz + k;
        c->h_zst[e] = first;
        for (int q = first; q <= last; ++q) c->h_zwt[e * kZW + (q - first)] = dense[q];
    }
    // band rows of the implicit operators on their unknowns: entry (k, k + d)
    // at d + kBandMax; data outside the unknowns (wall data) are dropped.
    // C2 on the levels (data index = level + 1), F2 on the interior faces
    // (data index = w level + 1, nz - 1 unknowns), P on the levels.
    c->h_zband.assign((size_t)ZB_COUNT * nz * kBandRow, 0.0);
    const int bops[ZB_COUNT] = {ZO_C2_DD, ZO_C2_DN, ZO_F2, ZO_P, ZO_G, ZO_D, ZO_I_DD, ZO_I_DN, ZO_IFC};
    // half bands the solver is instantiated for (5 and 7); the explicit G and D
    // are applied over the full stored band, the interpolations over kIntHW
    const int bhalf[ZB_COUNT] = {5, 5, 5, 7, kBandMax, kBandMax, kIntHW, kIntHW, kIntHW};
    for (int b = 0; b < ZB_COUNT; ++b) {
        const int op = bops[b];
        const bool interp = op == ZO_I_DD || op == ZO_I_DN || op == ZO_IFC;
        // rows, the columns' range and the data index of column 0
        const int nrow = (op == ZO_F2 || op == ZO_G) ? nz - 1 : nz;
        const int ncol = (op == ZO_F2 || op == ZO_D || op == ZO_IFC) ? nz - 1 : nz;
        const int off = (op == ZO_P || op == ZO_G) ? 0 : 1;
        const int W = op == ZO_P ? kZW : (op == ZO_G || op == ZO_D || interp) ? n1 : n2;
        // the interpolations keep the wall data (columns -1 and ncol)
        const int lo = interp ? -1 : 0, hi = interp ? ncol : ncol - 1;
        for (int k = 0; k < nrow; ++k) {
            const size_t e = (size_t)op * nz + k;
            for (int j = 0; j < W; ++j) {
                const double w = c->h_zwt[e * kZW + j];
                const int lv = c->h_zst[e] + j - off;
                if (w == 0.0 || lv < lo || lv > hi) continue;
                const int d = lv - k;
                if (d < -bhalf[b] || d > bhalf[b]) return false;
                c->h_zband[((size_t)b * nz + k) * kBandRow + kBandMax + d] += w;
            }
        }
    }
    return true;
}

// Per-wavenumber symbols of the horizontal FD operators for wavenumber index
// `m` of an n-point direction with spacing h (src/module_poisson_solver.f90:73-101,
// 222-246; the stencils of src/module_cnvdiff.f90 and src/module_navier_stokes.f90:669-796).
static void build_symbols(const hc_params &P, int n, int count, double h, double len, std::vector<double> &S)
{
    S.assign((size_t)SY_COUNT * count, 0.0);
    const double h2 = h * h;
    for (int q = 0; q < count; ++q) {
        int m = q < n / 2 ? q : q - n;
        double th = 2.0 * kPi * m / n;
        double c1 = std::cos(th), c2 = std::cos(2.0 * th), c3 = std::cos(3.0 * th);
        double s1 = std::sin(th), s2 = std::sin(2.0 * th);
        double lap, pois, dre, dim, gre, gim;
        if (P.cds == 3) {
            // Fourier symbols of dev_fft_linear_cmplx and the spectral
            // coefficients (src/module_spectral.f90:1305-1320,
            // src/module_poisson_solver.f90:96-99,244-245); zero at the
            // Nyquist wavenumber, which the reference removes
            const double kw = m == -n / 2 ? 0.0 : 2.0 * kPi / len * m;
            lap = -(kw * kw);
            pois = lap;
            dre = 0.0; dim = kw;
            gre = 0.0; gim = kw;
        } else if (P.cds == 1) {
            lap = (2.0 * c1 - 2.0) / h2;
            pois = lap;
            dre = (1.0 - c1) / h; dim = s1 / h;             // (1 - e^{-i th})/h
            gre = (c1 - 1.0) / h; gim = s1 / h;             // (e^{i th} - 1)/h
        } else {
            lap = (-2.0 * c2 + 32.0 * c1 - 30.0) / 12.0 / h2;
            pois = (2.0 * c3 - 108.0 * c2 + 1566.0 * c1 - 1460.0) / 576.0 / h2;
            // (-e^{i th} + 27 - 27 e^{-i th} + e^{-2 i th}) / 24h
            dre = (-c1 + 27.0 - 27.0 * c1 + c2) / 24.0 / h;
            dim = (-s1 + 27.0 * s1 - s2) / 24.0 / h;
            // (-e^{2 i th} + 27 e^{i th} - 27 + e^{-i th}) / 24h
            gre = (-c2 + 27.0 * c1 - 27.0 + c1) / 24.0 / h;
            gim = (-s2 + 27.0 * s1 - s1) / 24.0 / h;
        }
        S[SY_HU * count + q] = 0.5 * P.dt * P.nu * lap;
        S[SY_HT * count + q] = 0.5 * P.dt * P.alpha * lap;
        S[SY_PX * count + q] = pois;
        S[SY_L * count + q] = lap;
        S[SY_DRE * count + q] = dre; S[SY_DIM * count + q] = dim;
        S[SY_GRE * count + q] = gre; S[SY_GIM * count + q] = gim;
    }
}

static std::vector<double2> twiddles(int n)
{
    std::vector<double2> t(n);
    for (int e = 0; e < n; ++e) {
        double a = 2.0 * kPi * (double)e / (double)n;
        t[e] = make_double2(std::cos(a), -std::sin(a));
    }
    return t;
}

StepArgs base_args(hc_ctx *c)
{
    StepArgs A;
    std::memset(&A, 0, sizeof A);
    const hc_params &P = c->prm;
    A.u = c->u; A.v = c->v; A.w = c->w; A.p = c->p; A.t = c->t;
    A.s0 = c->s[0]; A.s1 = c->s[1]; A.s2 = c->s[2]; A.s3 = c->s[3];
    A.lev = c->lev; A.symx = c->symx; A.symy = c->symy;
    A.tw_nl = c->tw_nl; A.tw_nx = c->tw_nx; A.tw_ny = c->tw_ny;
    A.scal = c->scal;
    A.nx = c->nx; A.ny = c->ny; A.nz = c->nz; A.nl = c->nl;
    A.ochannel = P.ochannel; A.dp_opt = P.dp_opt; A.isscalar = P.isscalar;
    A.dx = c->dx; A.dy = c->dy; A.dx2 = c->dx * c->dx; A.dy2 = c->dy * c->dy;
    A.dt = P.dt; A.nu = P.nu; A.alpha = P.alpha; A.got = c->got; A.fc = P.fc; A.ug = P.ug; A.vg = P.vg;
    A.u_mrf = P.u_mrf; A.t_ref = P.t_ref; A.tbot = P.tbot; A.ttop = P.ttop;
    A.dpx = 1.0;
    A.norm = 1.0 / ((double)c->nx * (double)c->ny);
    A.aU0 = c->aU0; A.cUe = c->cUe; A.aT0 = c->aT0; A.cTe = c->cTe;
    A.tbc_top = P.tbc_top;
    A.gdz_top = P.dtdz_top * c->h_lev[(size_t)LV_DZT * c->nz + c->nz - 1];
    A.c24x = 1.0 / (24.0 * c->dx); A.c24y = 1.0 / (24.0 * c->dy);
    A.c12x2 = 1.0 / (12.0 * c->dx * c->dx); A.c12y2 = 1.0 / (12.0 * c->dy * c->dy);
    A.idx = 1.0 / c->dx; A.idy = 1.0 / c->dy; A.idx2 = 1.0 / (c->dx * c->dx); A.idy2 = 1.0 / (c->dy * c->dy);
    A.idt = 1.0 / P.dt;
    A.dmp_s = c->dmp_s; A.dmp_m = c->dmp_m; A.kdamp0 = c->kdamp0;
    A.klo = 0; A.nlev = c->nz;
    A.cu = c->cu; A.cv = c->cv; A.cw = c->cw; A.cp = c->cp; A.ct = c->ct;
    A.cff = c->cff; A.cgg = c->cgg; A.chh = c->chh; A.css = c->css;
    for (int q = 0; q < 8; ++q) A.sp[q] = c->sp[q];
    for (int q = 0; q < 6; ++q) A.sdv[q] = c->sp[q];
    A.coe = P.issk == 1 ? 0.5 : 1.0;  // coe_sk, src/module_spectral.f90:55-59
    A.issk = P.issk == 1;
    A.lc = (int)(c->nx / 3.0); A.mc = (int)(c->ny / 3.0);
    A.rgroup = c->tune.rgroup;
    A.zorder = (P.zorder == 4 || P.zorder == 6) ? 6 : 2;  // 6: the table-driven z operators
    A.zst = c->zst; A.zwt = c->zwt;
    A.zn1 = P.zorder == 4 ? 4 : 6;
    A.zn2 = A.zn1 + 1;
    A.zband = c->zband; A.zscr = c->zscr;
    A.zspc = c->zbuf[0] ? reinterpret_cast<double2 *>(c->zbuf[0]) : nullptr;
    A.spund = c->spund;
    return A;
}

// one substep t_its (1..3 for RK3, 1 for AB2) over the buffers of V (stored
// precision R); Rayleigh damping only with FP64 storage
template <class R>
static int substep_t(hc_ctx *c, int t_its, const View<R> &V)
{
    using C = cplx<R>;
    const hc_params &P = c->prm;
    StepArgs A = base_args(c);
    if (P.dts == 2) { A.ra = kRkc1[t_its - 1]; A.rb = kRkc2[t_its - 1]; A.rc = kRkc3[t_its - 1]; }
    else { A.ra = 1.5; A.rb = -0.5; A.rc = 1.0; }
    A.use_f0 = A.rb != 0.0;
    int nxt = c->fcur ^ 1;
    StepArgsT<R> B = rows_args<R>(A, V);
    B.ff0 = V.fb[c->fcur][0]; B.gg0 = V.fb[c->fcur][1]; B.hh0 = V.fb[c->fcur][2];
    B.ff = V.fb[nxt][0]; B.gg = V.fb[nxt][1]; B.hh = V.fb[nxt][2];
    if (P.dp_opt == 2) dpx_update_launch(c);
    const bool damp = P.isdamp == 1;
    if (damp && !std::is_same<R, double>::value) return set_err(HC_ERR_UNSUPPORTED, "damping needs FP64 storage");
    PROF_BEGIN(c, HC_K_RHS);
    if (int e = rhs_momentum<R>(c, A, B)) return e;
    PROF_END(c, HC_K_RHS);
    if (c->mean_on && t_its == (P.dts == 2 ? 3 : 1))
        if (int e = mean_interp<R>(c, B)) return e;
    C *const *S = V.s;
    PROF_BEGIN(c, HC_K_YFFT_FWD);
    if (int e = launch_yfft(c, -1, 3, S[0], S[1], S[2], S[3])) return e;
    PROF_END(c, HC_K_YFFT_FWD);
    PROF_BEGIN(c, HC_K_COLUMN);
    if (int e = launch_columns<R>(c, 0, A, V, 4)) return e;
    PROF_END(c, HC_K_COLUMN);
    PROF_BEGIN(c, HC_K_YFFT_INV);
    if (int e = launch_yfft(c, +1, 4, S[0], S[1], S[2], S[3])) return e;
    PROF_END(c, HC_K_YFFT_INV);
    C2RArgsT<R> X;
    std::memset(&X, 0, sizeof X);
    X.src[0] = S[0]; X.src[1] = S[1]; X.src[2] = S[2]; X.src[3] = S[3];
    X.dst[0] = V.u; X.dst[1] = V.v; X.dst[2] = V.w; X.dst[3] = V.p;
    X.mode[3] = damp ? 2 : 1;  // calculate_new_p adds to the damped p
    if constexpr (std::is_same<R, double>::value) { X.dmp_s = c->dmp_s; X.dmp_mean = c->dmp_m + 3 * c->nz; }
    X.kdamp0 = c->kdamp0;
    PROF_BEGIN(c, HC_K_FINISH);
    if (int e = launch_c2r<R>(c, 4, X)) return e;
    PROF_END(c, HC_K_FINISH);
    c->fcur = nxt;
    if (P.isscalar == 1) {
        int snx = c->scur ^ 1;
        B.ss0 = V.sb[c->scur]; B.ss = V.sb[snx];
        PROF_BEGIN(c, HC_K_RHS_T);
        if (int e = rhs_temperature<R>(c, B)) return e;
        PROF_END(c, HC_K_RHS_T);
        PROF_BEGIN(c, HC_K_YFFT_T);
        if (int e = launch_yfft(c, -1, 1, S[0], nullptr, nullptr, nullptr)) return e;
        PROF_END(c, HC_K_YFFT_T);
        PROF_BEGIN(c, HC_K_COLUMN_T);
        if (int e = launch_columns<R>(c, 1, A, V, 1)) return e;
        PROF_END(c, HC_K_COLUMN_T);
        PROF_BEGIN(c, HC_K_YFFT_T);
        if (int e = launch_yfft(c, +1, 1, S[0], nullptr, nullptr, nullptr)) return e;
        PROF_END(c, HC_K_YFFT_T);
        C2RArgsT<R> XT;
        std::memset(&XT, 0, sizeof XT);
        XT.src[0] = S[0]; XT.dst[0] = V.t;
        PROF_BEGIN(c, HC_K_FINISH_T);
        if (int e = launch_c2r<R>(c, 1, XT)) return e;
        PROF_END(c, HC_K_FINISH_T);
        c->scur = snx;
    }
    HIPCHK(hipGetLastError());
    return 0;
}

static int substep(hc_ctx *c, int t_its)
{
    return c->mixed ? substep_t<float>(c, t_its, view<float>(c)) : substep_t<double>(c, t_its, view<double>(c));
}

int ubar_of(hc_ctx *c, const double *u)
{
    hipLaunchKernelGGL(k_ubar_partial<double>, dim3(c->nz), dim3(256), 0, c->stream, u, c->lev, c->nx, c->ny, c->nz,
                       c->part);
    hipLaunchKernelGGL(k_ubar_final, dim3(1), dim3(64), 0, c->stream, c->part, c->nz, c->prm.lz, c->nx, c->ny, c->scal);
    HIPCHK(hipGetLastError());
    return 0;
}

static int ubar_launch(hc_ctx *c)
{
    if (c->mixed)
        hipLaunchKernelGGL(k_ubar_partial<float>, dim3(c->nz), dim3(256), 0, c->stream, c->m32.u, c->lev, c->nx, c->ny,
                           c->nz, c->part);
    else
        hipLaunchKernelGGL(k_ubar_partial<double>, dim3(c->nz), dim3(256), 0, c->stream, c->u, c->lev, c->nx, c->ny,
                           c->nz, c->part);
    hipLaunchKernelGGL(k_ubar_final, dim3(1), dim3(64), 0, c->stream, c->part, c->nz, c->prm.lz, c->nx, c->ny, c->scal);
    HIPCHK(hipGetLastError());
    return 0;
}

// ---- C ABI ----------------------------------------------------------------
extern "C" int hc_create(const hc_params *prm, hc_ctx **out)
{
    if (!prm || !out) return set_err(HC_ERR_ARG, "null argument");
    *out = nullptr;
    const hc_params &P = *prm;
    // initiate_parameters checks (src/module_parameters.f90:256-270)
    if (P.nx % 2 != 0 || P.ny % 2 != 0) return set_err(HC_ERR_ARG, "Please use even numbers for nx and ny! Quit!");
    if (P.nz < 3) return set_err(HC_ERR_ARG, "nz must be at least 3");
    if (P.cds < 1 || P.cds > 3) return set_err(HC_ERR_ARG, "cds must be 1, 2 or 3");
    if (P.dts != 1 && P.dts != 2) return set_err(HC_ERR_ARG, "dts must be 1 (AB2) or 2 (RK3)");
    if (P.dp_opt < 1 || P.dp_opt > 3) return set_err(HC_ERR_ARG, "dp_opt must be 1, 2 or 3");
    if (P.ochannel != 0 && P.ochannel != 1) return set_err(HC_ERR_ARG, "ochannel must be 0 or 1");
    if (P.isscalar != 0 && P.isscalar != 1) return set_err(HC_ERR_ARG, "isscalar error!");
    if (P.isdamp != 0 && P.isdamp != 1) return set_err(HC_ERR_ARG, "isdamp must be 0 or 1");
    // the damped levels kend3-nzdamp..kend3 must be interior levels
    if (P.isdamp == 1 && (P.nzdamp < 1 || P.nzdamp > P.nz - 1))
        return set_err(HC_ERR_ARG, "nzdamp=%d must be in 1..nz-1 when isdamp=1", P.nzdamp);
    if (P.wall_rows != 0 && P.wall_rows != 1) return set_err(HC_ERR_ARG, "wall_rows must be 0 or 1");
    if (P.precision != HC_FP64 && P.precision != HC_MIXED) return set_err(HC_ERR_ARG, "precision must be 0 or 1");
    if (P.tbc_top != 0 && P.tbc_top != 1) return set_err(HC_ERR_ARG, "tbc_top must be 0 or 1");
    if (P.precision == HC_MIXED && P.isdamp == 1)
        return set_err(HC_ERR_UNSUPPORTED, "precision=1 (mixed) does not cover Rayleigh damping");
    if (P.zorder != 0 && P.zorder != 2 && P.zorder != 4 && P.zorder != 6)
        return set_err(HC_ERR_ARG, "zorder must be 2, 4 or 6");
    const bool hiz = P.zorder == 4 || P.zorder == 6;
    if (hiz) {
        if (P.isdamp == 1) return set_err(HC_ERR_UNSUPPORTED, "zorder=%d does not cover Rayleigh damping", P.zorder);
        if (P.nz < 16) return set_err(HC_ERR_UNSUPPORTED, "zorder=%d needs nz >= 16", P.zorder);
    }
    if (!fft_len_ok(P.nx) || !fft_len_ok(P.ny))
        return set_err(HC_ERR_UNSUPPORTED, "nx, ny must be one of 16,32,48,64,96,128,192,256,384,512");
    {
        const int ty = rhs_tile_rows(P.nx);
        if (P.ny % ty != 0) return set_err(HC_ERR_UNSUPPORTED, "ny=%d must be a multiple of %d for nx=%d", P.ny, ty, P.nx);
    }
    int L = col_levels(P.nz);
    if (L < 0) return set_err(HC_ERR_UNSUPPORTED, "nz=%d too large", P.nz);
    hc_ctx *c = new hc_ctx();
    c->prm = P;
    c->nx = P.nx; c->ny = P.ny; c->nz = P.nz; c->nl = P.nx / 2; c->Lcol = L;
    c->dx = P.lx / P.nx; c->dy = P.ly / P.ny;
    c->got = P.got;
    c->time_sim = 0.0; c->stps = 0; c->fcur = 0; c->scur = 0;
    hipError_t he = hipGetDevice(&c->device);
    if (he == hipSuccess) he = hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device);
    if (he == hipSuccess) he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (he != hipSuccess) { delete c; return set_err(HC_ERR_HIP, "HIP device/stream: %s", hipGetErrorString(he)); }
    {
        // allow the multi-field column tiles (up to kColLds of dynamic LDS)
        hipError_t ae = hipSuccess;
        if (ae == hipSuccess) ae = columns_set_lds(L, (int)kColLds);
        if (ae == hipSuccess) ae = spec_set_lds(L, (int)kColLds);
        if (ae == hipSuccess) ae = z6_set_lds(L, (int)kColLds);
        c->big_lds = ae == hipSuccess;
        (void)hipGetLastError();
    }
    const size_t N = (size_t)c->nx * c->ny * c->nz;
    int e = 0;
    double **phys[] = {&c->u, &c->v, &c->w, &c->p, &c->t, &c->fb[0][0], &c->fb[0][1], &c->fb[0][2],
                       &c->fb[1][0], &c->fb[1][1], &c->fb[1][2], &c->sb[0], &c->sb[1]};
    for (auto pp : phys) if (!e) e = dalloc(c, (void **)pp, N * sizeof(double));
    for (int q = 0; q < 4 && !e; ++q) e = dalloc(c, (void **)&c->s[q], (size_t)c->nl * c->ny * c->nz * sizeof(double2));
    build_levels(c);
    std::vector<double> sx, sy;
    build_symbols(P, c->nx, c->nl, c->dx, P.lx, sx);
    build_symbols(P, c->ny, c->ny, c->dy, P.ly, sy);
    auto tnl = twiddles(c->nl), tnx = twiddles(c->nx), tny = twiddles(c->ny);
    if (!e) e = dalloc(c, (void **)&c->lev, c->h_lev.size() * sizeof(double));
    if (!e) e = dalloc(c, (void **)&c->symx, sx.size() * sizeof(double));
    if (!e) e = dalloc(c, (void **)&c->symy, sy.size() * sizeof(double));
    if (!e) e = dalloc(c, (void **)&c->tw_nl, tnl.size() * sizeof(double2));
    if (!e) e = dalloc(c, (void **)&c->tw_nx, tnx.size() * sizeof(double2));
    if (!e) e = dalloc(c, (void **)&c->tw_ny, tny.size() * sizeof(double2));
    if (!e) e = dalloc(c, (void **)&c->scal, 8 * sizeof(double));
    if (!e) e = dalloc(c, (void **)&c->part, 3 * (size_t)c->nz * sizeof(double));
    if (!e) e = dalloc(c, (void **)&c->dmp_s, (size_t)c->nz * sizeof(double));
    if (!e) e = dalloc(c, (void **)&c->dmp_m, 5 * (size_t)c->nz * sizeof(double));
    if (!e && P.cds == 3) {
        const size_t hs = (size_t)c->nl * c->ny * c->nz * sizeof(double2);
        double2 **sp[] = {&c->cu, &c->cv, &c->cw, &c->cp, &c->ct, &c->cff, &c->cgg, &c->chh, &c->css,
                          &c->sp[4], &c->sp[5], &c->sp[6], &c->sp[7]};
        for (auto pp : sp) if (!e) e = dalloc(c, (void **)pp, hs);
        for (int q = 0; q < 4; ++q) c->sp[q] = c->s[q];
        // undamped spectra of u, v, w, p, t on the nzdamp + 1 damped levels
        if (!e && P.isdamp == 1)
            e = dalloc(c, (void **)&c->spund, 5 * (size_t)(P.nzdamp + 1) * c->nl * c->ny * sizeof(double2));
        if (!e) e = dalloc(c, (void **)&c->ptmp, N * sizeof(double));
    }
    if (!e && hiz) {
        if (!build_zops(c)) e = set_err(HC_ERR_ARG, "order-%d z stencils: singular system", P.zorder);
        if (!e) e = dalloc(c, (void **)&c->zst, c->h_zst.size() * sizeof(int));
        if (!e) e = dalloc(c, (void **)&c->zwt, c->h_zwt.size() * sizeof(double));
        if (!e) e = dalloc(c, (void **)&c->zband, c->h_zband.size() * sizeof(double));
        if (!e) e = dalloc(c, (void **)&c->zscr, (size_t)12 * c->nz * c->nl * c->ny * sizeof(double));
        for (int q = 0; q < 4 && !e; ++q)
            if (q < 3 || P.isscalar == 1) e = dalloc(c, (void **)&c->zbuf[q], N * sizeof(double));
    }
    if (!e && P.precision == HC_MIXED) {
        c->mixed = true;
        c->lo_l = std::min(HC_LOWK_L, c->nl / kColWaves);
        c->lo_m = std::min(HC_LOWK_M, c->ny / 2 - 1);
        View<float> &M = c->m32;
        float **phys32[] = {&M.u, &M.v, &M.w, &M.p, &M.t, &M.fb[0][0], &M.fb[0][1], &M.fb[0][2],
                            &M.fb[1][0], &M.fb[1][1], &M.fb[1][2], &M.sb[0], &M.sb[1]};
        for (auto pp : phys32) if (!e) e = dalloc(c, (void **)pp, N * sizeof(float));
        for (int q = 0; q < 4 && !e; ++q)
            e = dalloc(c, (void **)&M.s[q], (size_t)c->nl * c->ny * c->nz * sizeof(float2));
        if (!e) e = dalloc(c, (void **)&M.lev, c->h_lev.size() * sizeof(float));
        if (!e) e = dalloc(c, (void **)&M.symx, sx.size() * sizeof(float));
        if (!e) e = dalloc(c, (void **)&M.symy, sy.size() * sizeof(float));
        if (!e) e = dalloc(c, (void **)&M.tw_nl, tnl.size() * sizeof(float2));
        if (!e) e = dalloc(c, (void **)&M.tw_nx, tnx.size() * sizeof(float2));
        if (!e) e = dalloc(c, (void **)&M.tw_ny, tny.size() * sizeof(float2));
        if (!e && P.cds == 3) {
            const size_t hs = (size_t)c->nl * c->ny * c->nz * sizeof(float2);
            float2 **sp[] = {&M.cu, &M.cv, &M.cw, &M.cp, &M.ct, &M.cff, &M.cgg, &M.chh, &M.css,
                             &M.sp[4], &M.sp[5], &M.sp[6], &M.sp[7]};
            for (auto pp : sp) if (!e) e = dalloc(c, (void **)pp, hs);
            for (int q = 0; q < 4; ++q) M.sp[q] = M.s[q];
            if (!e) e = dalloc(c, (void **)&M.ptmp, N * sizeof(float));
        }
    }
    if (e) { hc_destroy(c); return e; }
    // sponge strength sigz of rayleigh_damping (src/module_tools.f90:735-736)
    // on the top nzdamp+1 levels, zero below
    c->kdamp0 = P.isdamp == 1 ? c->nz - 1 - P.nzdamp : c->nz;
    std::vector<double> hs(c->nz, 0.0);
    for (int kk = c->kdamp0; kk < c->nz; ++kk) {
        double zd = (kk - c->kdamp0) * 1.0 / P.nzdamp;
        hs[kk] = P.cdamp * 0.5 * (1.0 - std::cos(kPi * zd));
    }
    double scal0[8] = {P.dp_opt == 1 ? 1.0 : 0.0, 0, 0, 0, 0, 0, 0, 0};
    hipError_t r = hipMemcpyAsync(c->lev, c->h_lev.data(), c->h_lev.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (r == hipSuccess) r = hipMemcpyAsync(c->symx, sx.data(), sx.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (r == hipSuccess) r = hipMemcpyAsync(c->symy, sy.data(), sy.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (r == hipSuccess) r = hipMemcpyAsync(c->tw_nl, tnl.data(), tnl.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream);
    if (r == hipSuccess) r = hipMemcpyAsync(c->tw_nx, tnx.data(), tnx.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream);
    if (r == hipSuccess) r = hipMemcpyAsync(c->tw_ny, tny.data(), tny.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream);
    if (r == hipSuccess) r = hipMemcpyAsync(c->scal, scal0, sizeof scal0, hipMemcpyHostToDevice, c->stream);
    if (r == hipSuccess) r = hipMemcpyAsync(c->dmp_s, hs.data(), hs.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (r == hipSuccess && c->zst)
        r = hipMemcpyAsync(c->zst, c->h_zst.data(), c->h_zst.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (r == hipSuccess && c->zwt)
        r = hipMemcpyAsync(c->zwt, c->h_zwt.data(), c->h_zwt.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (r == hipSuccess && c->zband)
        r = hipMemcpyAsync(c->zband, c->h_zband.data(), c->h_zband.size() * sizeof(double), hipMemcpyHostToDevice,
                           c->stream);
    // FP32 copies of the tables for the mixed variant's row kernels
    std::vector<float> flev(c->h_lev.begin(), c->h_lev.end()), fsx(sx.begin(), sx.end()), fsy(sy.begin(), sy.end());
    auto f2 = [](const std::vector<double2> &t) {
        std::vector<float2> o(t.size());
        for (size_t q = 0; q < t.size(); ++q) o[q] = make_float2((float)t[q].x, (float)t[q].y);
        return o;
    };
    std::vector<float2> fnl = f2(tnl), fnx = f2(tnx), fny = f2(tny);
    if (c->mixed) {
        View<float> &M = c->m32;
        if (r == hipSuccess) r = hipMemcpyAsync(M.lev, flev.data(), flev.size() * sizeof(float), hipMemcpyHostToDevice, c->stream);
        if (r == hipSuccess) r = hipMemcpyAsync(M.symx, fsx.data(), fsx.size() * sizeof(float), hipMemcpyHostToDevice, c->stream);
        if (r == hipSuccess) r = hipMemcpyAsync(M.symy, fsy.data(), fsy.size() * sizeof(float), hipMemcpyHostToDevice, c->stream);
        if (r == hipSuccess) r = hipMemcpyAsync(M.tw_nl, fnl.data(), fnl.size() * sizeof(float2), hipMemcpyHostToDevice, c->stream);
        if (r == hipSuccess) r = hipMemcpyAsync(M.tw_nx, fnx.data(), fnx.size() * sizeof(float2), hipMemcpyHostToDevice, c->stream);
        if (r == hipSuccess) r = hipMemcpyAsync(M.tw_ny, fny.data(), fny.size() * sizeof(float2), hipMemcpyHostToDevice, c->stream);
    }
    if (r == hipSuccess) r = hipStreamSynchronize(c->stream);
    if (r != hipSuccess) { hc_destroy(c); return set_err(HC_ERR_HIP, "table upload: %s", hipGetErrorString(r)); }
    *out = c;
    return 0;
}

extern "C" int hc_destroy(hc_ctx *c)
{
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->stage) (void)hipFree(c->stage);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

template <class R>
static R *field_ptr(hc_ctx *c, int f)
{
    const View<R> V = view<R>(c);
    switch (f) {
    case HC_U: return V.u;
    case HC_V: return V.v;
    case HC_W: return V.w;
    case HC_P: return V.p;
    case HC_T: return V.t;
    case HC_FF: return V.fb[c->fcur][0];
    case HC_GG: return V.fb[c->fcur][1];
    case HC_HH: return V.fb[c->fcur][2];
    case HC_SS: return V.sb[c->scur];
    default: return nullptr;
    }
}

static int ensure_stage(hc_ctx *c)
{
    if (c->stage) return 0;
    size_t n = (size_t)(c->nx + 4) * (c->ny + 4) * (c->nz + 4);
    HIPCHK(hipMalloc((void **)&c->stage, n * sizeof(double)));
    return 0;
}

// Host or device FP64 data -> a field in the stored precision R.
template <class R>
static int set_field_t(hc_ctx *c, R *d, const double *src, int layout)
{
    const size_t N = (size_t)c->nx * c->ny * c->nz;
    if (layout == HC_LAYOUT_INTERIOR) {
        if constexpr (std::is_same<R, double>::value) {
            HIPCHK(hipMemcpyAsync(d, src, N * sizeof(double), hipMemcpyDefault, c->stream));
        } else {
            if (int e = ensure_stage(c)) return e;
            HIPCHK(hipMemcpyAsync(c->stage, src, N * sizeof(double), hipMemcpyDefault, c->stream));
            hipLaunchKernelGGL((k_convert<double, R>), dim3(1024), dim3(256), 0, c->stream, c->stage, d, N);
        }
    } else if (layout == HC_LAYOUT_GHOSTED) {
        if (int e = ensure_stage(c)) return e;
        size_t n = (size_t)(c->nx + 4) * (c->ny + 4) * (c->nz + 4);
        HIPCHK(hipMemcpyAsync(c->stage, src, n * sizeof(double), hipMemcpyDefault, c->stream));
        hipLaunchKernelGGL((k_pack_ghosted<R>), dim3(1024), dim3(256), 0, c->stream, c->stage, d, c->nx, c->ny, c->nz);
    } else {
        return set_err(HC_ERR_ARG, "unknown layout %d", layout);
    }
    return 0;
}

// A field in the stored precision R -> host or device FP64 data.
template <class R>
static int get_field_t(hc_ctx *c, const R *d, double *dst, int layout)
{
    const size_t N = (size_t)c->nx * c->ny * c->nz;
    if (layout == HC_LAYOUT_INTERIOR) {
        if constexpr (std::is_same<R, double>::value) {
            HIPCHK(hipMemcpyAsync(dst, d, N * sizeof(double), hipMemcpyDefault, c->stream));
        } else {
            if (int e = ensure_stage(c)) return e;
            hipLaunchKernelGGL((k_convert<R, double>), dim3(1024), dim3(256), 0, c->stream, d, c->stage, N);
            HIPCHK(hipMemcpyAsync(dst, c->stage, N * sizeof(double), hipMemcpyDefault, c->stream));
        }
    } else if (layout == HC_LAYOUT_GHOSTED) {
        if (int e = ensure_stage(c)) return e;
        size_t n = (size_t)(c->nx + 4) * (c->ny + 4) * (c->nz + 4);
        // keep the caller's ghost cells: bring the whole array, overwrite the interior
        HIPCHK(hipMemcpyAsync(c->stage, dst, n * sizeof(double), hipMemcpyDefault, c->stream));
        hipLaunchKernelGGL((k_unpack_ghosted<R>), dim3(1024), dim3(256), 0, c->stream, d, c->stage, c->nx, c->ny, c->nz);
        HIPCHK(hipMemcpyAsync(dst, c->stage, n * sizeof(double), hipMemcpyDefault, c->stream));
    } else {
        return set_err(HC_ERR_ARG, "unknown layout %d", layout);
    }
    return 0;
}

int put_f64(hc_ctx *c, double *dev, const double *src, int layout) { return set_field_t<double>(c, dev, src, layout); }
int get_f64(hc_ctx *c, const double *dev, double *dst, int layout) { return get_field_t<double>(c, dev, dst, layout); }

extern "C" int hc_set_field(hc_ctx *c, int field, const double *src, int layout)
{
    if (!c || !src) return set_err(HC_ERR_ARG, "null argument");
    if (field < HC_U || field > HC_SS) return set_err(HC_ERR_ARG, "unknown field %d", field);
    int e = c->mixed ? set_field_t<float>(c, field_ptr<float>(c, field), src, layout)
                     : set_field_t<double>(c, field_ptr<double>(c, field), src, layout);
    if (e) return e;
    // spectral path: the spectra are formed again from the physical fields
    // before the next substep, as at start-up
    c->spec_dirty = true;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_get_field(hc_ctx *c, int field, double *dst, int layout)
{
    if (!c || !dst) return set_err(HC_ERR_ARG, "null argument");
    if (field < HC_U || field > HC_SS) return set_err(HC_ERR_ARG, "unknown field %d", field);
    if (c->prm.cds == 3 && field == HC_P && !c->spec_dirty) {
        // the spectral path keeps p as a spectrum: cmplx_2_real(c_p, p) (:476)
        int e = 0;
        if (c->mixed) {
            const float2 *s1[1] = {c->m32.cp};
            float *p1[1] = {c->m32.p};
            e = spec_inverse<float>(c, 1, s1, p1);
        } else {
            const double2 *s1[1] = {c->cp};
            double *p1[1] = {c->p};
            e = spec_inverse<double>(c, 1, s1, p1);
        }
        if (e) return e;
    }
    int e = c->mixed ? get_field_t<float>(c, field_ptr<float>(c, field), dst, layout)
                     : get_field_t<double>(c, field_ptr<double>(c, field), dst, layout);
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_init(hc_ctx *c)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    const hc_params &P = c->prm;
    if (P.is_ri_var == 1) c->got = P.ri_str;  // update_ri(0)
    // update_boundary_uvw forces w = 0 on the top face (src/module_boundary.f90:107)
    size_t plane = (size_t)c->nx * c->ny;
    if (c->mixed)
        hipLaunchKernelGGL(k_zero_level<float>, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, c->stream,
                           c->m32.w, (size_t)(c->nz - 1) * plane, plane);
    else
        hipLaunchKernelGGL(k_zero_level<double>, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, c->stream, c->w,
                           (size_t)(c->nz - 1) * plane, plane);
    if (P.dp_opt == 2)
        if (int e = ubar_launch(c)) return e;
    if (P.cds == 3)
        if (int e = spec_init(c)) return e;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// calculate_mean_fields at the end of a step (src/module_navier_stokes.f90:358-361, 510-513)
static int mean_accumulate(hc_ctx *c)
{
    if (!c->mean_on) return 0;
    MeanSums S;
    for (int q = 0; q < HC_MEAN_COUNT; ++q) S.m[q] = c->mean[q];
    const double *uh = c->prm.cds == 3 ? nullptr : c->mean_uhx, *vh = c->prm.cds == 3 ? nullptr : c->mean_vhy;
    if (c->mixed) {
        const View<float> &M = c->m32;
        hipLaunchKernelGGL(k_mean_accum<float>, dim3(2048), dim3(256), 0, c->stream, M.u, M.v, M.w, M.t, uh, vh, S,
                           c->nx, c->ny, c->nz, c->prm.u_mrf);
    } else {
        hipLaunchKernelGGL(k_mean_accum<double>, dim3(2048), dim3(256), 0, c->stream, c->u, c->v, c->w, c->t, uh, vh,
                           S, c->nx, c->ny, c->nz, c->prm.u_mrf);
    }
    c->mean_steps += 1;
    HIPCHK(hipGetLastError());
    return 0;
}

// The substeps of one step (time_advancement_fd / _sp) and
// calculate_mean_fields; with the screen on, screen_div_error(u,v,w,1,.)
// after each substep (src/module_navier_stokes.f90:344-350, 505-509) into
// record slot t_its.
static int step_body(hc_ctx *c)
{
    const hc_params &P = c->prm;
    const int nsub = P.dts == 2 ? 3 : 1;
    for (int s = 1; s <= nsub; ++s) {
        if (int e = P.cds == 3 ? substep_sp(c, s) : substep(c, s)) return e;
        if (c->screen)
            if (int e = screen_queue(c, s, 1)) return e;
    }
    return mean_accumulate(c);
}

// After a screened step: the reference stops at the end of the first
// substep whose fields hold a NaN ('NaN error!', src/module_io.f90:135-140).
static int screen_check(hc_ctx *c)
{
    if (!c->screen) return 0;
    double rec[16];
    HIPCHK(hipMemcpyAsync(rec, c->scr + 4, sizeof rec, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int nsub = c->prm.dts == 2 ? 3 : 1;
    for (int s = 1; s <= nsub; ++s)
        if (rec[4 * s + 3] != 0.0)
            return set_err(HC_ERR_NAN, "NaN error! (substep %d of step %d)", s, c->stps);
    return 0;
}

static int one_step(hc_ctx *c)
{
    const hc_params &P = c->prm;
    c->stps += 1;
    c->time_sim += P.dt;
    if (P.is_ri_var == 1) c->got = P.ri_str + (P.ri_end - P.ri_str) * 1.0 / P.imax * c->stps;
    return step_body(c);
}

extern "C" int hc_time_advancement_fd(hc_ctx *c)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (c->prm.cds == 3) return set_err(HC_ERR_ARG, "cds=3: the spectral path is hc_time_advancement_sp");
    if (int e = step_body(c)) return e;
    return screen_check(c);
}

extern "C" int hc_time_advancement_sp(hc_ctx *c)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (c->prm.cds != 3) return set_err(HC_ERR_ARG, "cds=%d: time_advancement_sp needs cds=3", c->prm.cds);
    if (int e = step_body(c)) return e;
    return screen_check(c);
}

extern "C" int hc_update_ri(hc_ctx *c, int stps)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    const hc_params &P = c->prm;
    if (P.imax < 1) return set_err(HC_ERR_ARG, "update_ri needs imax >= 1");
    c->got = P.ri_str + (P.ri_end - P.ri_str) * 1.0 / P.imax * stps;
    return 0;
}

extern "C" int hc_drive_steps(hc_ctx *c, int nsteps)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    for (int n = 0; n < nsteps; ++n) {
        if (int e = one_step(c)) return e;
        if (c->prm.dp_opt == 2)
            if (int e = ubar_launch(c)) return e;
        if (int e = screen_check(c)) return e;
    }
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int hc_screen_substeps(hc_ctx *c, int on)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (on != 0 && on != 1) return set_err(HC_ERR_ARG, "on must be 0 or 1");
    if (on)
        if (int e = screen_alloc(c)) return e;
    c->screen = on;
    return 0;
}

extern "C" int hc_get_screen(hc_ctx *c, int substep, double *div_avg, double *div_max, double *div_hist_max)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (substep < 0 || substep > 3) return set_err(HC_ERR_ARG, "substep must be 0..3");
    if (!c->scr) return set_err(HC_ERR_ARG, "no screen has run on this context");
    double rec[4];
    HIPCHK(hipMemcpyAsync(rec, c->scr + 4 + 4 * substep, sizeof rec, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (div_avg) *div_avg = rec[0];
    if (div_max) *div_max = rec[1];
    if (div_hist_max) *div_hist_max = rec[2];
    if (rec[3] != 0.0) return set_err(HC_ERR_NAN, "NaN error!");
    return 0;
}

extern "C" int hc_screen_div_error_ref(hc_ctx *c, int isfinal, double *div_avg, double *div_max, double *div_hist_max)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (isfinal != 0 && isfinal != 1) return set_err(HC_ERR_ARG, "isfinal must be 0 or 1");
    if (int e = screen_queue(c, 0, isfinal)) return e;
    return hc_get_screen(c, 0, div_avg, div_max, div_hist_max);
}

extern "C" int hc_get_column(hc_ctx *c, int field, int i, int j, double *dst)
{
    if (!c || !dst) return set_err(HC_ERR_ARG, "null argument");
    if (i < 0 || i >= c->nx || j < 0 || j >= c->ny) return set_err(HC_ERR_ARG, "column (%d, %d) outside the grid", i, j);
    const size_t plane = (size_t)c->nx * c->ny, off = (size_t)j * c->nx + i;
    if (c->prm.cds == 3 && field == HC_P && !c->spec_dirty) {  // p lives as a spectrum: form it first
        std::vector<double> full(plane * c->nz);
        if (int e = hc_get_field(c, HC_P, full.data(), HC_LAYOUT_INTERIOR)) return e;
        for (int k = 0; k < c->nz; ++k) dst[k] = full[(size_t)k * plane + off];
        return 0;
    }
    if (c->mixed) {
        const float *f = field_ptr<float>(c, field);
        if (!f) return set_err(HC_ERR_ARG, "no field %d", field);
        hipLaunchKernelGGL(k_get_column<float>, dim3(1), dim3(256), 0, c->stream, f, plane, off, c->nz, c->part);
    } else {
        const double *f = field_ptr<double>(c, field);
        if (!f) return set_err(HC_ERR_ARG, "no field %d", field);
        hipLaunchKernelGGL(k_get_column<double>, dim3(1), dim3(256), 0, c->stream, f, plane, off, c->nz, c->part);
    }
    HIPCHK(hipMemcpyAsync(dst, c->part, (size_t)c->nz * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_screen_cfl(hc_ctx *c, double *cfl_max)
{
    if (!c || !cfl_max) return set_err(HC_ERR_ARG, "null argument");
    const hc_params &P = c->prm;
    if (c->mixed) {
        const View<float> &M = c->m32;
        hipLaunchKernelGGL(k_cfl_partial<float>, dim3(c->nz), dim3(256), 0, c->stream, M.u, M.v, M.w, c->lev, c->nx,
                           c->ny, c->nz, P.dt, c->dx, c->dy, c->part);
    } else {
        hipLaunchKernelGGL(k_cfl_partial<double>, dim3(c->nz), dim3(256), 0, c->stream, c->u, c->v, c->w, c->lev,
                           c->nx, c->ny, c->nz, P.dt, c->dx, c->dy, c->part);
    }
    std::vector<double> h(c->nz);
    HIPCHK(hipMemcpyAsync(h.data(), c->part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double m = 0.0;
    for (double x : h) m = std::max(m, x);
    *cfl_max = m;
    return 0;
}

extern "C" int hc_get_mesh(hc_ctx *c, double *zc, double *dz, double *l_t, double *l_b)
{
    if (!c || !zc || !dz || !l_t || !l_b) return set_err(HC_ERR_ARG, "null argument");
    std::vector<double> zf;
    mesh_faces(c->prm, zf);
    for (int k = 0; k < c->nz; ++k) {
        zc[k] = 0.5 * (zf[k + 1] + zf[k]);
        dz[k] = c->h_lev[(size_t)LV_DZ * c->nz + k];
        l_t[k] = c->h_lev[(size_t)LV_LT * c->nz + k];
        l_b[k] = c->h_lev[(size_t)LV_LB * c->nz + k];
    }
    return 0;
}

extern "C" int hc_mean_fields(hc_ctx *c, int on)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (on != 0 && on != 1) return set_err(HC_ERR_ARG, "on must be 0 or 1");
    if (on) {
        const size_t N = (size_t)c->nx * c->ny * c->nz;
        if (!c->mean[0]) {
            for (int q = 0; q < HC_MEAN_COUNT; ++q)
                if (int e = dalloc(c, (void **)&c->mean[q], N * sizeof(double))) return e;
            if (c->prm.cds != 3) {
                if (int e = dalloc(c, (void **)&c->mean_uhx, N * sizeof(double))) return e;
                if (int e = dalloc(c, (void **)&c->mean_vhy, N * sizeof(double))) return e;
            }
        } else {
            for (int q = 0; q < HC_MEAN_COUNT; ++q) HIPCHK(hipMemsetAsync(c->mean[q], 0, N * sizeof(double), c->stream));
        }
        c->mean_steps = 0;
    }
    c->mean_on = on == 1;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_get_mean_field(hc_ctx *c, int stat, double *dst, int layout)
{
    if (!c || !dst) return set_err(HC_ERR_ARG, "null argument");
    if (stat < 0 || stat >= HC_MEAN_COUNT) return set_err(HC_ERR_ARG, "statistic %d", stat);
    if (!c->mean[0] || c->mean_steps == 0) return set_err(HC_ERR_ARG, "no statistics accumulated (hc_mean_fields)");
    const size_t N = (size_t)c->nx * c->ny * c->nz;
    double *tmp = c->sb[c->scur ^ 1];  // the free ss slot
    hipLaunchKernelGGL(k_scale_copy, dim3(1024), dim3(256), 0, c->stream, c->mean[stat], tmp, N,
                       1.0 / (double)c->mean_steps);
    if (int e = get_field_t<double>(c, tmp, dst, layout)) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// poisson_solver_fft over the stored precision R.  src/dst: the caller's
// FP64 arrays (device interior arrays are used in place when R is double).
template <class R>
static int poisson_t(hc_ctx *c, int equation, double rk_c, const double *r_src, double *r_var, int layout)
{
    const size_t N = (size_t)c->nx * c->ny * c->nz;
    const View<R> V = view<R>(c);
    // the physical work buffer: the ss ping-pong slot not holding ss
    R *work = V.sb[c->scur ^ 1];
    bool dev_io = false;
    hipPointerAttribute_t at;
    if (layout == HC_LAYOUT_INTERIOR && hipPointerGetAttributes(&at, r_src) == hipSuccess &&
        at.type == hipMemoryTypeDevice)
        dev_io = true;
    (void)hipGetLastError();
    const R *src = work;
    if (dev_io && std::is_same<R, double>::value) {
        src = reinterpret_cast<const R *>(r_src);
    } else if (dev_io) {
        hipLaunchKernelGGL((k_convert<double, R>), dim3(1024), dim3(256), 0, c->stream, r_src, work, N);
    } else {
        if (int e = set_field_t<R>(c, work, r_src, layout)) return e;
    }
    PROF_BEGIN(c, HC_K_SOLVE_XR2C);
    if (int e = launch_r2c<R>(c, src, V.s[0])) return e;
    PROF_END(c, HC_K_SOLVE_XR2C);
    PROF_BEGIN(c, HC_K_SOLVE_YFFT);
    if (int e = launch_yfft(c, -1, 1, V.s[0], nullptr, nullptr, nullptr)) return e;
    PROF_END(c, HC_K_SOLVE_YFFT);
    StepArgs A = base_args(c);
    A.rc = c->prm.dts == 2 ? rk_c : 1.0;
    PROF_BEGIN(c, HC_K_SOLVE_COLUMN);
    if (int e = launch_columns<R>(c, 3, A, V, 1, equation)) return e;
    PROF_END(c, HC_K_SOLVE_COLUMN);
    PROF_BEGIN(c, HC_K_SOLVE_YFFT);
    if (int e = launch_yfft(c, +1, 1, V.s[0], nullptr, nullptr, nullptr)) return e;
    PROF_END(c, HC_K_SOLVE_YFFT);
    C2RArgsT<R> X;
    std::memset(&X, 0, sizeof X);
    X.src[0] = V.s[0];
    const bool direct = dev_io && std::is_same<R, double>::value;
    X.dst[0] = direct ? reinterpret_cast<R *>(r_var) : work;
    PROF_BEGIN(c, HC_K_SOLVE_XC2R);
    if (int e = launch_c2r<R>(c, 1, X)) return e;
    PROF_END(c, HC_K_SOLVE_XC2R);
    if (dev_io && !direct) {
        hipLaunchKernelGGL((k_convert<R, double>), dim3(1024), dim3(256), 0, c->stream, work, r_var, N);
    } else if (!dev_io) {
        if (int e = get_field_t<R>(c, work, r_var, layout)) return e;
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int hc_poisson_solver_fft(hc_ctx *c, int equation, double rk_c, const double *r_src, double *r_var, int layout)
{
    if (!c || !r_src || !r_var) return set_err(HC_ERR_ARG, "null argument");
    if (equation < HC_EQ_U || equation > HC_EQ_T) return set_err(HC_ERR_ARG, "unknown equation %d", equation);
    if (layout != HC_LAYOUT_INTERIOR && layout != HC_LAYOUT_GHOSTED) return set_err(HC_ERR_ARG, "unknown layout %d", layout);
    return c->mixed ? poisson_t<float>(c, equation, rk_c, r_src, r_var, layout)
                    : poisson_t<double>(c, equation, rk_c, r_src, r_var, layout);
}

extern "C" int hc_calculate_ubar(hc_ctx *c, double *ubar)
{
    if (!c || !ubar) return set_err(HC_ERR_ARG, "null argument");
    if (int e = ubar_launch(c)) return e;
    HIPCHK(hipMemcpyAsync(ubar, c->scal + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_screen_div_error(hc_ctx *c, double *div_avg, double *div_max)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    StepArgs A = base_args(c);
    if (c->prm.cds == 3) {
        // screen_div_error_spec (src/module_spectral.f90:899-1007): the
        // spectral divergence of (c_u, c_v, w) in physical space
        if (int e = spec_divergence(c)) return e;
        if (c->mixed) {
            StepArgsT<float> F = rows_args<float>(A, c->m32);
            F.aux = c->m32.ptmp;
            hipLaunchKernelGGL((k_div_partial<3, float>), dim3(c->nz), dim3(256), 0, c->stream, F, A, c->part);
        } else {
            A.aux = c->ptmp;
            hipLaunchKernelGGL((k_div_partial<3, double>), dim3(c->nz), dim3(256), 0, c->stream, A, A, c->part);
        }
    } else if (c->mixed) {
        StepArgsT<float> F = rows_args<float>(A, c->m32);
        if (c->prm.cds == 1)
            hipLaunchKernelGGL((k_div_partial<1, float>), dim3(c->nz), dim3(256), 0, c->stream, F, A, c->part);
        else
            hipLaunchKernelGGL((k_div_partial<2, float>), dim3(c->nz), dim3(256), 0, c->stream, F, A, c->part);
    } else if (c->prm.cds == 1) {
        hipLaunchKernelGGL((k_div_partial<1, double>), dim3(c->nz), dim3(256), 0, c->stream, A, A, c->part);
    } else {
        hipLaunchKernelGGL((k_div_partial<2, double>), dim3(c->nz), dim3(256), 0, c->stream, A, A, c->part);
    }
    std::vector<double> h(3 * (size_t)c->nz);
    HIPCHK(hipMemcpyAsync(h.data(), c->part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double s = 0.0, m = 0.0;
    bool nan = false;
    for (int k = 0; k < c->nz; ++k) {
        s += h[3 * k];
        m = std::fmax(m, h[3 * k + 1]);
        nan = nan || h[3 * k + 2] != 0.0;
    }
    if (div_avg) *div_avg = s / c->nx / c->ny / c->nz;
    if (div_max) *div_max = m;
    if (nan) return set_err(HC_ERR_NAN, "NaN error!");
    return 0;
}

extern "C" int hc_wall_shear(hc_ctx *c, double *tau_bottom, double *tau_top)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    const hc_params &P = c->prm;
    if (c->mixed) {
        hipLaunchKernelGGL(k_plane_sum<float>, dim3(1), dim3(256), 0, c->stream, c->m32.u, c->nx, c->ny, 0, c->part);
        hipLaunchKernelGGL(k_plane_sum<float>, dim3(1), dim3(256), 0, c->stream, c->m32.u, c->nx, c->ny, c->nz - 1,
                           c->part + 1);
    } else {
        hipLaunchKernelGGL(k_plane_sum<double>, dim3(1), dim3(256), 0, c->stream, c->u, c->nx, c->ny, 0, c->part);
        hipLaunchKernelGGL(k_plane_sum<double>, dim3(1), dim3(256), 0, c->stream, c->u, c->nx, c->ny, c->nz - 1,
                           c->part + 1);
    }
    double h[2];
    HIPCHK(hipMemcpyAsync(h, c->part, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double plane = (double)c->nx * (double)c->ny;
    const double dz0 = c->h_lev[(size_t)LV_DZ * c->nz], dze = c->h_lev[(size_t)LV_DZ * c->nz + c->nz - 1];
    if (tau_bottom) *tau_bottom = 2.0 * P.nu * (h[0] / plane + P.u_mrf) / dz0;
    if (tau_top) *tau_top = P.ochannel == 0 ? 2.0 * P.nu * (h[1] / plane + P.u_mrf) / dze : 0.0;
    return 0;
}

extern "C" int hc_get_scalar(hc_ctx *c, int id, double *value)
{
    if (!c || !value) return set_err(HC_ERR_ARG, "null argument");
    double sc[2];
    switch (id) {
    case 0:
    case 1:
        HIPCHK(hipMemcpyAsync(sc, c->scal, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        *value = sc[id];
        return 0;
    case 2: *value = c->time_sim; return 0;
    case 3: *value = c->got; return 0;
    case 4: *value = c->stps; return 0;
    case 5: *value = col_tiles(c, 4, c->mixed ? sizeof(float2) : sizeof(double2)); return 0;
    case 6: *value = c->tune.rgroup; return 0;
    case 7: *value = c->tune.ytile; return 0;
    case 8: *value = c->tune.c2r_pts; return 0;
    case 9: *value = c->tune.col_remap; return 0;
    case 10: *value = (double)c->mean_steps; return 0;
    case 11: *value = c->tune.rhs_xt; return 0;
    default: return set_err(HC_ERR_ARG, "unknown scalar %d", id);
    }
}

extern "C" void *hc_stream(hc_ctx *c) { return c ? (void *)c->stream : nullptr; }
extern "C" void *hc_field_device_ptr(hc_ctx *c, int field)
{
    if (!c) return nullptr;
    return c->mixed ? (void *)field_ptr<float>(c, field) : (void *)field_ptr<double>(c, field);
}

extern "C" int hc_synchronize(hc_ctx *c)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_profile_enable(hc_ctx *c, int on)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    c->prof = on != 0;
    c->ev_used = 0;
    c->ev_kid.clear();
    for (int k = 0; k < HC_KERNEL_COUNT; ++k) { c->prof_ms[k] = 0.0; c->prof_n[k] = 0; }
    return 0;
}

extern "C" int hc_profile_read(hc_ctx *c, int kernel, double *total_ms, long *launches)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (kernel < 0 || kernel >= HC_KERNEL_COUNT) return set_err(HC_ERR_ARG, "unknown kernel id %d", kernel);
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t q = 0; q < c->ev_kid.size(); ++q) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev_pool[2 * q], c->ev_pool[2 * q + 1]));
        c->prof_ms[c->ev_kid[q]] += ms;
        c->prof_n[c->ev_kid[q]] += 1;
    }
    c->ev_kid.clear();
    c->ev_used = 0;
    if (total_ms) *total_ms = c->prof_ms[kernel];
    if (launches) *launches = c->prof_n[kernel];
    return 0;
}

template <class R>
static int project_t(hc_ctx *c)
{
    const View<R> V = view<R>(c);
    size_t plane = (size_t)c->nx * c->ny;
    hipLaunchKernelGGL(k_zero_level<R>, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, c->stream, V.w,
                       (size_t)(c->nz - 1) * plane, plane);
    {
        const R *src3[3] = {V.u, V.v, V.w};
        cplx<R> *dst3[3] = {V.s[0], V.s[1], V.s[2]};
        if (int e = launch_r2c_n<R>(c, 3, src3, dst3)) return e;
    }
    if (int e = launch_yfft(c, -1, 3, V.s[0], V.s[1], V.s[2], nullptr)) return e;
    StepArgs A = base_args(c);
    A.rc = 1.0;
    if (int e = launch_columns<R>(c, 2, A, V, 1)) return e;
    if (int e = launch_yfft(c, +1, 3, V.s[0], V.s[1], V.s[2], nullptr)) return e;
    C2RArgsT<R> X;
    std::memset(&X, 0, sizeof X);
    X.src[0] = V.s[0]; X.src[1] = V.s[1]; X.src[2] = V.s[2];
    X.dst[0] = V.u; X.dst[1] = V.v; X.dst[2] = V.w;
    return launch_c2r<R>(c, 3, X);
}

extern "C" int hc_project(hc_ctx *c)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = c->mixed ? project_t<float>(c) : project_t<double>(c)) return e;
    c->spec_dirty = true;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// ---- ensembles --------------------------------------------------------------
struct hc_ensemble {
    std::vector<hc_ctx *> m;
    bool use_graph = false, graph_used = false;
    hipStream_t master = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<char> cap_mean;  // each member's statistics switch when the graph was captured
};

// one pass of the drive loop body for one member (hc_drive_steps' body)
static int member_step(hc_ctx *c)
{
    const int scr = c->screen;
    c->screen = 0;
    const int e0 = one_step(c);
    c->screen = scr;
    if (e0) return e0;
    if (c->prm.dp_opt == 2)
        if (int e = ubar_launch(c)) return e;
    return 0;
}

extern "C" int hc_ensemble_create(hc_ctx *const *members, int n, int use_graph, hc_ensemble **out)
{
    if (!members || n < 1 || !out) return set_err(HC_ERR_ARG, "null argument or empty ensemble");
    *out = nullptr;
    hc_ensemble *e = new hc_ensemble();
    for (int q = 0; q < n; ++q) {
        if (!members[q]) { delete e; return set_err(HC_ERR_ARG, "null member %d", q); }
        e->m.push_back(members[q]);
    }
    e->use_graph = use_graph != 0;
    hipError_t he = hipStreamCreateWithFlags(&e->master, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
    e->join.assign(n, nullptr);
    for (int q = 0; q < n && he == hipSuccess; ++q) he = hipEventCreateWithFlags(&e->join[q], hipEventDisableTiming);
    if (he != hipSuccess) {
        hc_ensemble_destroy(e);
        return set_err(HC_ERR_HIP, "ensemble streams/events: %s", hipGetErrorString(he));
    }
    *out = e;
    return 0;
}

static bool ensemble_graphable(const hc_ensemble *e)
{
    for (const hc_ctx *c : e->m)
        if (c->prm.is_ri_var == 1 || c->prof) return false;
    return true;
}

// Capture two drive steps of every member into one graph: the master stream
// forks to the member streams and joins them back.  The host-side counters
// that one_step advances are restored afterwards; replays advance them.
static int ensemble_capture(hc_ensemble *e)
{
    struct Saved { int stps, fcur, scur; double time_sim, got; long mean_steps; };
    std::vector<Saved> sv;
    e->cap_mean.clear();
    for (hc_ctx *c : e->m) {
        if (c->prm.cds == 3 && c->spec_dirty)
            if (int er = spec_init(c)) return er;
        HIPCHK(hipStreamSynchronize(c->stream));
        sv.push_back({c->stps, c->fcur, c->scur, c->time_sim, c->got, c->mean_steps});
        e->cap_mean.push_back(c->mean_on ? 1 : 0);
    }
    HIPCHK(hipStreamBeginCapture(e->master, hipStreamCaptureModeThreadLocal));
    int err = 0;
    hipError_t he = hipEventRecord(e->fork, e->master);
    for (size_t q = 0; q < e->m.size() && !err && he == hipSuccess; ++q) {
        hc_ctx *c = e->m[q];
        he = hipStreamWaitEvent(c->stream, e->fork, 0);
        for (int s = 0; s < 2 && !err && he == hipSuccess; ++s) err = member_step(c);
        if (!err && he == hipSuccess) he = hipEventRecord(e->join[q], c->stream);
        if (!err && he == hipSuccess) he = hipStreamWaitEvent(e->master, e->join[q], 0);
    }
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(e->master, &g);
    for (size_t q = 0; q < e->m.size(); ++q) {
        hc_ctx *c = e->m[q];
        c->stps = sv[q].stps; c->fcur = sv[q].fcur; c->scur = sv[q].scur;
        c->time_sim = sv[q].time_sim; c->got = sv[q].got; c->mean_steps = sv[q].mean_steps;
    }
    if (err) { if (g) (void)hipGraphDestroy(g); return err; }
    if (he != hipSuccess || ce != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return set_err(HC_ERR_HIP, "ensemble capture: %s", hipGetErrorString(he != hipSuccess ? he : ce));
    }
    hipGraphExec_t x = nullptr;
    hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(g);
        return set_err(HC_ERR_HIP, "ensemble graph instantiate: %s", hipGetErrorString(ie));
    }
    e->graph = g;
    e->exec = x;
    return 0;
}

extern "C" int hc_ensemble_steps(hc_ensemble *e, int nsteps)
{
    if (!e) return set_err(HC_ERR_ARG, "null ensemble");
    if (nsteps < 0) return set_err(HC_ERR_ARG, "nsteps < 0");
    int done = 0;
    e->graph_used = false;
    if (e->use_graph && ensemble_graphable(e) && nsteps >= 2) {
        // a member's statistics switch changed since the capture: capture again
        bool stale = false;
        for (size_t q = 0; q < e->m.size() && e->exec; ++q)
            stale = stale || (e->cap_mean[q] != (e->m[q]->mean_on ? 1 : 0));
        if (stale) {
            (void)hipGraphExecDestroy(e->exec);
            (void)hipGraphDestroy(e->graph);
            e->exec = nullptr;
            e->graph = nullptr;
        }
        if (!e->exec)
            if (int er = ensemble_capture(e)) return er;
        for (hc_ctx *c : e->m)
            if (c->prm.cds == 3 && c->spec_dirty) return set_err(HC_ERR_ARG, "member changed after capture");
        for (; done + 2 <= nsteps; done += 2) {
            HIPCHK(hipGraphLaunch(e->exec, e->master));
            for (hc_ctx *c : e->m) {
                c->stps += 2;
                c->time_sim += 2.0 * c->prm.dt;
                if (c->mean_on) c->mean_steps += 2;
            }
        }
        HIPCHK(hipStreamSynchronize(e->master));
        e->graph_used = true;
    }
    // remaining steps (or all of them): direct launches, member after member,
    // each on its own stream so that the members still overlap
    for (; done < nsteps; ++done)
        for (hc_ctx *c : e->m)
            if (int er = member_step(c)) return er;
    for (hc_ctx *c : e->m) HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int hc_ensemble_graph_used(const hc_ensemble *e) { return e && e->graph_used ? 1 : 0; }

extern "C" int hc_ensemble_destroy(hc_ensemble *e)
{
    if (!e) return 0;
    if (e->master) (void)hipStreamSynchronize(e->master);
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    for (hipEvent_t v : e->join)
        if (v) (void)hipEventDestroy(v);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->master) (void)hipStreamDestroy(e->master);
    delete e;
    return 0;
}

// ---- discrete adjoint ---------------------------------------------------------
static int adjoint_supported(const hc_ctx *c)
{
    const hc_params &P = c->prm;
    if (c->mixed || P.cds == 3 || P.isscalar == 1 || P.isdamp == 1 || P.dp_opt == 2 || P.zorder == 4 || P.zorder == 6)
        return set_err(HC_ERR_UNSUPPORTED,
                       "the adjoint covers FP64, cds=1,2, isscalar=0, isdamp=0, dp_opt=1,3, 2nd-order z");
    return 0;
}

static int ensure_adjoint(hc_ctx *c)
{
    if (c->adj[0]) return 0;
    const size_t bytes = (size_t)c->nx * c->ny * c->nz * sizeof(double);
    int e = 0;
    for (int q = 0; q < 4 && !e; ++q) e = dalloc(c, (void **)&c->adj[q], bytes);
    for (int a = 0; a < 2; ++a)
        for (int q = 0; q < 3 && !e; ++q) e = dalloc(c, (void **)&c->adf[a][q], bytes);
    for (int q = 0; q < 3 && !e; ++q) e = dalloc(c, (void **)&c->adq[q], bytes);
    for (int q = 0; q < 3 && !e; ++q) e = dalloc(c, (void **)&c->adfb[q], bytes);
    for (int q = 0; q < 8 && !e; ++q) e = dalloc(c, (void **)&c->adface[q], bytes);
    return e;
}

static double *adjoint_ptr(hc_ctx *c, int f)
{
    switch (f) {
    case HC_U: return c->adj[0];
    case HC_V: return c->adj[1];
    case HC_W: return c->adj[2];
    case HC_P: return c->adj[3];
    case HC_FF: return c->adf[c->adcur][0];
    case HC_GG: return c->adf[c->adcur][1];
    case HC_HH: return c->adf[c->adcur][2];
    default: return nullptr;
    }
}



extern "C" int hc_substep_fd(hc_ctx *c, int t_its)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (c->prm.cds == 3) return set_err(HC_ERR_ARG, "cds=3: no finite-difference substep");
    const int nsub = c->prm.dts == 2 ? 3 : 1;
    if (t_its < 1 || t_its > nsub) return set_err(HC_ERR_ARG, "t_its=%d outside 1..%d", t_its, nsub);
    return substep(c, t_its);
}

extern "C" int hc_set_adjoint(hc_ctx *c, int field, const double *src, int layout)
{
    if (!c || !src) return set_err(HC_ERR_ARG, "null argument");
    if (int e = adjoint_supported(c)) return e;
    if (int e = ensure_adjoint(c)) return e;
    double *d = adjoint_ptr(c, field);
    if (!d) return set_err(HC_ERR_ARG, "no adjoint of field %d", field);
    if (int e = set_field_t<double>(c, d, src, layout)) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_get_adjoint(hc_ctx *c, int field, double *dst, int layout)
{
    if (!c || !dst) return set_err(HC_ERR_ARG, "null argument");
    if (int e = adjoint_supported(c)) return e;
    if (int e = ensure_adjoint(c)) return e;
    double *d = adjoint_ptr(c, field);
    if (!d) return set_err(HC_ERR_ARG, "no adjoint of field %d", field);
    if (int e = get_field_t<double>(c, d, dst, layout)) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hc_adjoint_substep_fd(hc_ctx *c, int t_its)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = adjoint_supported(c)) return e;
    const int nsub = c->prm.dts == 2 ? 3 : 1;
    if (t_its < 1 || t_its > nsub) return set_err(HC_ERR_ARG, "t_its=%d outside 1..%d", t_its, nsub);
    if (int e = ensure_adjoint(c)) return e;
    if (int e = adjoint_substep(c, t_its)) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// Gradient of the bottom wall shear after nsteps steps with respect to the
// initial u, v, w, p: forward run with a checkpoint of (u, v, w, p) every
// ckpt_every steps, then segment by segment (last first) the segment's
// substep inputs are recomputed from its checkpoint and the substep adjoints
// applied in reverse order.  RK3 only (its first substep does not read
// f_old, so a step's input is (u, v, w, p)).
extern "C" int hc_wall_shear_gradient(hc_ctx *c, int nsteps, int ckpt_every, double *loss)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (int e = adjoint_supported(c)) return e;
    const hc_params &P = c->prm;
    if (P.dts != 2) return set_err(HC_ERR_UNSUPPORTED, "the gradient driver needs RK3 (dts=2)");
    if (nsteps < 1 || ckpt_every < 1) return set_err(HC_ERR_ARG, "nsteps and ckpt_every must be >= 1");
    if (int e = ensure_adjoint(c)) return e;
    QuietSteps quiet(c);
    const size_t N = (size_t)c->nx * c->ny * c->nz, fb = N * sizeof(double);
    const int nseg = (nsteps + ckpt_every - 1) / ckpt_every, nsub = 3;
    const int seglen = std::min(ckpt_every, nsteps);
    std::vector<void *> mem;
    auto alloc = [&](size_t bytes) -> double * {
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        mem.push_back(p);
        return (double *)p;
    };
    auto release = [&]() {
        (void)hipStreamSynchronize(c->stream);
        for (void *p : mem) (void)hipFree(p);
        mem.clear();
    };
    double *ck = alloc((size_t)nseg * 4 * fb);
    double *seg = alloc((size_t)seglen * nsub * 4 * fb);
    if (!ck || !seg) { release(); return set_err(HC_ERR_HIP, "checkpoint storage (%zu MB) not available",
                                                 ((size_t)nseg * 4 + (size_t)seglen * nsub * 4) * fb >> 20); }
    double *fld[4] = {c->u, c->v, c->w, c->p};
    auto save = [&](double *dst) -> hipError_t {
        hipError_t r = hipSuccess;
        for (int q = 0; q < 4 && r == hipSuccess; ++q)
            r = hipMemcpyAsync(dst + q * N, fld[q], fb, hipMemcpyDeviceToDevice, c->stream);
        return r;
    };
    auto load = [&](const double *src) -> hipError_t {
        hipError_t r = hipSuccess;
        for (int q = 0; q < 4 && r == hipSuccess; ++q)
            r = hipMemcpyAsync(fld[q], src + q * N, fb, hipMemcpyDeviceToDevice, c->stream);
        return r;
    };
    const int stps0 = c->stps, fcur0 = c->fcur;
    const double time0 = c->time_sim, got0 = c->got;
    int err = 0;
    hipError_t he = hipSuccess;
    // forward with checkpoints; the last segment's substep inputs are kept
    // as they are computed, so the backward sweep starts without recomputing
    const int last0 = (nseg - 1) * ckpt_every;
    int idx_last = 0;
    for (int n = 0; n < nsteps && !err && he == hipSuccess; ++n) {
        if (n % ckpt_every == 0) he = save(ck + (size_t)(n / ckpt_every) * 4 * N);
        if (n < last0) {
            if (he == hipSuccess) err = one_step(c);
            continue;
        }
        c->stps += 1;  // the bookkeeping of one_step
        c->time_sim += P.dt;
        if (P.is_ri_var == 1) c->got = P.ri_str + (P.ri_end - P.ri_str) * 1.0 / P.imax * c->stps;
        for (int s = 1; s <= nsub && !err && he == hipSuccess; ++s) {
            he = save(seg + (size_t)(idx_last++) * 4 * N);
            if (he == hipSuccess) err = substep(c, s);
        }
    }
    double tau[2] = {0.0, 0.0};
    if (!err && he == hipSuccess) err = hc_wall_shear(c, &tau[0], &tau[1]);
    // seed: d tau_bottom / d u(level 0) = 2 nu / (dz(kstr3) nx ny)
    if (!err && he == hipSuccess) {
        for (int q = 0; q < 4 && he == hipSuccess; ++q) he = hipMemsetAsync(c->adj[q], 0, fb, c->stream);
        for (int a = 0; a < 2; ++a)
            for (int q = 0; q < 3 && he == hipSuccess; ++q) he = hipMemsetAsync(c->adf[a][q], 0, fb, c->stream);
        const size_t plane = (size_t)c->nx * c->ny;
        const double seed = 2.0 * P.nu / (c->h_lev[(size_t)LV_DZ * c->nz] * (double)plane);
        hipLaunchKernelGGL(k_fill, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, c->stream, c->adj[0],
                           (size_t)0, plane, seed);
    }
    // backward over the segments; the transposed substeps read the stored
    // inputs in place (the context's field pointers are pointed at them)
    for (int sg = nseg - 1; sg >= 0 && !err && he == hipSuccess; --sg) {
        const int n0 = sg * ckpt_every, n1 = std::min(nsteps, n0 + ckpt_every);
        int idx = idx_last;
        if (sg != nseg - 1) {
            he = load(ck + (size_t)sg * 4 * N);
            idx = 0;
            for (int n = n0; n < n1 && !err && he == hipSuccess; ++n)
                for (int s = 1; s <= nsub && !err && he == hipSuccess; ++s) {
                    he = save(seg + (size_t)(idx++) * 4 * N);
                    if (he == hipSuccess) err = substep(c, s);
                }
        }
        for (int q = idx - 1; q >= 0 && !err && he == hipSuccess; --q) {
            double *st = seg + (size_t)q * 4 * N;
            c->u = st; c->v = st + N; c->w = st + 2 * N; c->p = st + 3 * N;
            err = adjoint_substep(c, q % nsub + 1);
            c->u = fld[0]; c->v = fld[1]; c->w = fld[2]; c->p = fld[3];
        }
    }
    // leave the initial state in the context
    if (!err && he == hipSuccess) he = load(ck);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    release();
    c->stps = stps0; c->fcur = fcur0; c->time_sim = time0; c->got = got0;
    if (err) return err;
    if (he != hipSuccess) return set_err(HC_ERR_HIP, "gradient: %s", hipGetErrorString(he));
    if (loss) *loss = tau[0];
    return 0;
}

// ---- autotuner ------------------------------------------------------------------
// Launch shapes that leave every result bit-identical (each tile, transform
// row or column is computed the same way whatever the block shape): the RHS
// tile-order group, the y-transform columns per block, the inverse-x points
// per block and the column-kernel block order.  Coordinate descent over the
// knobs, each trial timing whole steps from the same saved state with HIP
// events; a candidate replaces the current choice when it is more than 1%
// faster.  The state and step counters are restored afterwards.
extern "C" int hc_autotune(hc_ctx *c, int steps_per_trial)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    if (steps_per_trial < 1) steps_per_trial = 2;
    const hc_params &P = c->prm;
    const size_t N = (size_t)c->nx * c->ny * c->nz;
    const size_t eb = c->mixed ? sizeof(float) : sizeof(double);
    std::vector<void *> fields;
    if (c->mixed) {
        View<float> &M = c->m32;
        void *f[] = {M.u, M.v, M.w, M.p, M.t, M.fb[0][0], M.fb[0][1], M.fb[0][2], M.fb[1][0], M.fb[1][1], M.fb[1][2],
                     M.sb[0], M.sb[1]};
        fields.assign(f, f + 13);
    } else {
        void *f[] = {c->u, c->v, c->w, c->p, c->t, c->fb[0][0], c->fb[0][1], c->fb[0][2], c->fb[1][0], c->fb[1][1],
                     c->fb[1][2], c->sb[0], c->sb[1]};
        fields.assign(f, f + 13);
    }
    if (P.cds == 3) return 0;  // the spectral path's launch shapes are fixed
    QuietSteps quiet(c);
    void *save = nullptr;
    HIPCHK(hipMalloc(&save, fields.size() * N * eb + 8 * sizeof(double)));
    char *sv = (char *)save;
    hipError_t he = hipSuccess;
    for (size_t q = 0; q < fields.size() && he == hipSuccess; ++q)
        he = hipMemcpyAsync(sv + q * N * eb, fields[q], N * eb, hipMemcpyDeviceToDevice, c->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(sv + fields.size() * N * eb, c->scal, 8 * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
    const int stps0 = c->stps, fcur0 = c->fcur, scur0 = c->scur;
    const double time0 = c->time_sim, got0 = c->got;
    auto restore = [&]() -> hipError_t {
        hipError_t r = hipSuccess;
        for (size_t q = 0; q < fields.size() && r == hipSuccess; ++q)
            r = hipMemcpyAsync(fields[q], sv + q * N * eb, N * eb, hipMemcpyDeviceToDevice, c->stream);
        if (r == hipSuccess)
            r = hipMemcpyAsync(c->scal, sv + fields.size() * N * eb, 8 * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream);
        c->stps = stps0; c->fcur = fcur0; c->scur = scur0; c->time_sim = time0; c->got = got0;
        return r;
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (he == hipSuccess) he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    int err = 0;
    auto step = [&]() -> int {
        if (int e = one_step(c)) return e;
        if (P.dp_opt == 2) return ubar_launch(c);
        return 0;
    };
    auto trial = [&](float &ms) -> int {
        if ((he = restore()) != hipSuccess) return HC_ERR_HIP;
        if (int e = step()) return e;
        if ((he = hipEventRecord(e0, c->stream)) != hipSuccess) return HC_ERR_HIP;
        for (int n = 0; n < steps_per_trial; ++n)
            if (int e = step()) return e;
        if ((he = hipEventRecord(e1, c->stream)) != hipSuccess) return HC_ERR_HIP;
        if ((he = hipEventSynchronize(e1)) != hipSuccess) return HC_ERR_HIP;
        if ((he = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return HC_ERR_HIP;
        return 0;
    };
    // candidates per knob
    const int ty = P.nx == 16 ? 16 : P.nx == 32 ? 32 : P.nx == 64 ? 16 : P.nx == 128 ? 8 : P.nx == 192 ? 4 :
                   P.nx == 256 ? 4 : 1;
    const int nty = c->ny / ty;
    std::vector<int> groups = {0};
    if (P.nx <= 256)
        for (int g : {1, 2, 4, 8, 16, 32})
            if (nty % g == 0 && g <= nty) groups.push_back(g);
    std::vector<int> ytiles = {0, kYTile};
    if (c->nl % (2 * kYTile) == 0) ytiles.push_back(2 * kYTile);
    std::vector<int> xts = {0};
    if (P.zorder == 2)
        for (int sh = 1; sh <= 4; ++sh)
            if (rhs_xt_fits(sh, c->nx, c->ny)) xts.push_back(sh);
    int *knob[5] = {&c->tune.rgroup, &c->tune.ytile, &c->tune.c2r_pts, &c->tune.col_remap, &c->tune.rhs_xt};
    std::vector<int> cand[5] = {groups, ytiles, {1024, 512, 2048}, {0, 1}, xts};
    float best = 0.0f;
    if (he == hipSuccess) err = trial(best);
    for (int kb = 0; kb < 5 && !err && he == hipSuccess; ++kb) {
        const int keep = *knob[kb];
        int chosen = keep;
        for (int v : cand[kb]) {
            if (v == keep) continue;
            *knob[kb] = v;
            float ms = 0.0f;
            if ((err = trial(ms)) || he != hipSuccess) break;
            if (ms < 0.99f * best) { best = ms; chosen = v; }
        }
        *knob[kb] = chosen;
    }
    if (he == hipSuccess) he = restore();
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(save);
    if (err) return err == HC_ERR_HIP && he != hipSuccess ? set_err(HC_ERR_HIP, "autotune: %s", hipGetErrorString(he)) : err;
    if (he != hipSuccess) return set_err(HC_ERR_HIP, "autotune: %s", hipGetErrorString(he));
    return 0;
}

// Set one launch shape directly (knob 0 RHS tile-order group, 1 y-transform
// columns per block, 2 inverse-x points per block, 3 column block order,
// 4 momentum tile shape),
// e.g. to reuse a previous hc_autotune choice.
extern "C" int hc_set_tuning(hc_ctx *c, int knob, int value)
{
    if (!c) return set_err(HC_ERR_ARG, "null context");
    switch (knob) {
    case 0: {
        const hc_params &P = c->prm;
        const int ty = P.nx == 16 ? 16 : P.nx == 32 ? 32 : P.nx == 64 ? 16 : P.nx == 128 ? 8 : P.nx == 192 ? 4 :
                       P.nx == 256 ? 4 : 1;
        const int nty = c->ny / ty;
        if (value < 0 || (value > 0 && (nty % value != 0)))
            return set_err(HC_ERR_ARG, "RHS tile group %d must divide %d", value, nty);
        c->tune.rgroup = value;
        return 0;
    }
    case 1:
        if (value != 0 && value != kYTile && value != 2 * kYTile)
            return set_err(HC_ERR_ARG, "y-transform columns per block must be 0, %d or %d", kYTile, 2 * kYTile);
        c->tune.ytile = value;
        return 0;
    case 2:
        if (value != 512 && value != 1024 && value != 2048)
            return set_err(HC_ERR_ARG, "inverse-x points per block must be 512, 1024 or 2048");
        c->tune.c2r_pts = value;
        return 0;
    case 3:
        if (value != 0 && value != 1) return set_err(HC_ERR_ARG, "column block order must be 0 or 1");
        c->tune.col_remap = value;
        return 0;
    case 4:
        if (value != 0 && !rhs_xt_fits(value, c->nx, c->ny))
            return set_err(HC_ERR_ARG, "momentum tile shape %d must be 0, or 1-4 with tiles dividing %dx%d", value,
                           c->nx, c->ny);
        c->tune.rhs_xt = value;
        return 0;
    default:
        return set_err(HC_ERR_ARG, "unknown launch-shape knob %d", knob);
    }
}
