// Kernels of the spectral time step (cds = 3): time_advancement_sp,
// src/module_navier_stokes.f90:371-517, and src/module_spectral.f90.
//
// Spectra are half spectra s[k][j][l] (l < nx/2) of the reference's C2C
// spectra: every field is real, so the l < 0 half is the conjugate mirror,
// and every operator of the path (Fourier derivatives, the 2/3 rule, the
// Helmholtz and Poisson solves with real symmetric coefficients) keeps that
// symmetry.  z-direction finite differences commute with the horizontal
// transforms, so the reference's physical-space z terms (uzz, dpdz, dwdz,
// dphidz, dphidz2) are applied to the spectra column by column, with the
// wall ghosts of update_boundary_uvw / _p / _t carried by the (0,0) mode.
//
// Included by hc_core.hip after the column helpers.
#pragma once

// ---------------------------------------------------------------------------
// y transforms from a source spectrum into a destination spectrum, with an
// optional Fourier first derivative applied first (dev_fft_linear_cmplx with
// ord = 1, src/module_spectral.f90:1305-1311).
struct YIOArgs {
    const double2 *src[4];
    double2 *dst[4];
    int mul[4];  // 0: none, 1: i kx(l), 2: i ky(m)
    int nl, ny;
    const double2 *tw_ny;
    const double *symx, *symy;
};
HC_DEV double2 times_i(double2 v, double k) { return make_double2(-k * v.y, k * v.x); }

template <int NY, int SIGN>
__global__ __launch_bounds__(256) void k_yfft_io(YIOArgs A)
{
    __shared__ double2 tile[NY * kYPad];
    const int f = blockIdx.z;
    const double2 *src = A.src[f];
    double2 *dst = A.dst[f];
    const int mul = A.mul[f];
    const int l0 = blockIdx.x * kYTile, k = blockIdx.y, tid = threadIdx.x;
    __shared__ double2 twy[NY];
    for (int q = tid; q < NY; q += 256) twy[q] = A.tw_ny[q];
    constexpr int NIT = (NY * kYTile + 255) / 256;
    double2 vv[NIT];
#pragma unroll
    for (int m = 0; m < NIT; ++m) {
        const int p = tid + m * 256, j = p / kYTile, ll = p - j * kYTile;
        vv[m] = p < NY * kYTile ? src[((size_t)k * NY + j) * A.nl + l0 + ll] : czero();
    }
#pragma unroll
    for (int m = 0; m < NIT; ++m) {
        const int p = tid + m * 256, j = p / kYTile, ll = p - j * kYTile;
        if (p >= NY * kYTile) continue;
        double2 v = vv[m];
        if (mul == 1) v = times_i(v, A.symx[SY_GIM * A.nl + l0 + ll]);
        else if (mul == 2) v = times_i(v, A.symy[SY_GIM * A.ny + j]);
        if (SIGN > 0 && j == NY / 2) v = czero();
        tile[j * kYPad + ll] = v;
    }
    __syncthreads();
    lds_fft<NY, kYTile, 1, kYPad, 256, SIGN>(tile, twy, tid);
    for (int p = tid; p < NY * kYTile; p += 256) {
        int j = p / kYTile, ll = p - j * kYTile;
        double2 v = tile[j * kYPad + ll];
        if (SIGN < 0 && j == NY / 2) v = czero();
        dst[((size_t)k * NY + j) * A.nl + l0 + ll] = v;
    }
}

// ---------------------------------------------------------------------------
// Physical-space part of get_momentum_cnvdiff_spec
// (src/module_spectral.f90:35-502).  Block: rows_per_block full x rows of
// level k, one point per thread.  Inputs: the truncated physical u, v, w, t
// (wall ghosts by the reference's rules) and, for the skew-symmetric form,
// the y-inverted spectra of du/dx, du/dy, dv/dx, dv/dy, dw/dx, dw/dy, whose
// x inverse is done here in LDS.  Outputs, x-transformed (forward
// normalised) into sp[0..7]:
//   0 u*u   (d/dx in the u equation)     1 u*v  (d/dy in u, d/dx in v)
//   2 v*v   (d/dy in v)                  3 R_u  4 R_v
//   5 u_w*w (d/dx in w)                  6 v_w*w (d/dy in w)   7 R_w
// where R collects every term the reference transforms without a
// derivative, already divided by coe_sk:
//   R_u = tmp2/coe - duwdz - [issk](u dudx + v dudy + w dudz)   (:98-197)
//   R_v = tmp2/coe - dvwdz - [issk](u dvdx + v dvdy + w dvdz)   (:279-339)
//   R_w = buoy/coe - dwwdz - [issk](u dwdx + v dwdy + w dwdz)   (:426-496)
template <int NX>
__global__ __launch_bounds__(1024) void k_spec_mom(StepArgs A)
{
    constexpr int ROWS = rows_per_block<NX>();
    constexpr int T = ROWS * NX;
    static_assert(T <= 1024, "one point per thread");
    constexpr int NL = NX / 2;
    __shared__ double2 sq_[8 * ROWS * NL];
    double *sd = reinterpret_cast<double *>(sq_);
    const int k = blockIdx.y, j0 = blockIdx.x * ROWS, tid = threadIdx.x;
    const int i = tid % NX, jj = tid / NX, j = j0 + jj;
    const bool in = j < A.ny;
    if (A.issk) {
        for (int p = tid; p < 6 * ROWS * NL; p += T) {
            int row = p / NL, l = p - row * NL, f = row / ROWS, r = row - f * ROWS, jr = j0 + r;
            double2 z = czero();
            if (jr < A.ny) {
                const double2 *s = A.sdv[f] + ((size_t)k * A.ny + jr) * NL;
                double2 xl = s[l];
                double2 xt = l == 0 ? czero() : cconj(s[NL - l]);
                z = c2r_pre(xl, xt, A.tw_nx[l]);
            }
            sq_[row * NL + l] = z;
        }
        __syncthreads();
        lds_fft<NL, 6 * ROWS, NL, 1, T, +1>(sq_, A.tw_nl, tid);
    }
    double out[8];
#pragma unroll
    for (int f = 0; f < 8; ++f) out[f] = 0.0;
    if (in) {
        Pt g(A, i, j, k);
        const double u = g.U(0, 0, 0), up = g.U(0, 0, 1), um_ = g.U(0, 0, -1);
        const double v = g.V(0, 0, 0), vp = g.V(0, 0, 1), vm = g.V(0, 0, -1);
        const double w = g.W(0, 0, 0), wp = g.W(0, 0, 1), wm = g.W(0, 0, -1);
        const double lt = A.lev[LV_LT * A.nz + k], lb = A.lev[LV_LB * A.nz + k];
        const double idz = A.lev[LV_IDZ * A.nz + k], idzt = A.lev[LV_IDZT * A.nz + k];
        const double coe = A.coe;
        const double dpx = A.dp_opt == 2 ? A.scal[0] : A.dpx;
        // face values of u and v (dev_fft_nonlinear_cmplx weqn=1 uses a1 too)
        const double a1 = up * lt + u * (1.0 - lt), a2 = um_ * lb + u * (1.0 - lb);
        const double a1v = vp * lt + v * (1.0 - lt), a2v = vm * lb + v * (1.0 - lb);
        out[0] = u * u;
        out[1] = u * v;
        out[2] = v * v;
        out[5] = a1 * w;
        out[6] = a1v * w;
        const double tmp2u = A.dp_opt == 3 ? A.fc * (v - A.vg) : dpx;
        const double tmp2v = A.dp_opt == 3 ? A.fc * (A.ug - u - A.u_mrf) : 0.0;
        double ru = 1.0 / coe * tmp2u - (a1 * w - a2 * wm) / A.lev[LV_DZ * A.nz + k];
        double rv = 1.0 / coe * tmp2v - (a1v * w - a2v * wm) / A.lev[LV_DZ * A.nz + k];
        const double hp = 0.5 * (w + wp), hm = 0.5 * (w + wm);
        const double wwz = (hp * hp - hm * hm) / A.lev[LV_DZT * A.nz + k];
        double rw;
        if (A.isscalar) {
            const double t = g.T(0, 0, 0), tp = g.T(0, 0, 1);
            rw = 1.0 / coe * (A.got * ((1.0 - lt) * t + lt * tp - A.t_ref)) - wwz;
        } else {
            rw = -wwz;
        }
        if (A.issk) {
            auto D = [&](int f) { return sd[(f * ROWS + jj) * NX + i]; };
            ru -= u * D(0);
            ru -= v * D(1);
            ru -= (a1 - a2) / A.lev[LV_DZ * A.nz + k] * (0.5 * (w + wm));
            rv -= u * D(2);
            rv -= v * D(3);
            rv -= (a1v - a2v) / A.lev[LV_DZ * A.nz + k] * (0.5 * (w + wm));
            rw -= a1 * D(4);
            rw -= a1v * D(5);
            rw -= w * ((hp - hm) / A.lev[LV_DZT * A.nz + k]);
        }
        (void)idz;
        (void)idzt;
        out[3] = ru;
        out[4] = rv;
        out[7] = rw;
    }
    __syncthreads();  // derivative rows consumed: reuse the LDS for the outputs
#pragma unroll
    for (int f = 0; f < 8; ++f) sd[(f * ROWS + jj) * NX + i] = out[f];
    __syncthreads();
    lds_fft<NL, 8 * ROWS, NL, 1, T, -1>(sq_, A.tw_nl, tid);
    for (int p = tid; p < 8 * ROWS * NL; p += T) {
        int row = p / NL, l = p - row * NL, f = row / ROWS, r = row - f * ROWS, jr = j0 + r;
        if (jr < A.ny) A.sp[f][((size_t)k * A.ny + jr) * NL + l] = r2c_post<NL>(sq_ + row * NL, l, A.tw_nx, A.norm);
    }
}

// Physical-space part of get_temperature_cnvdiff_spec
// (src/module_spectral.f90:509-595) with the updated u, v, w: sp[0] = u*t
// (d/dx), sp[1] = t*v (d/dy), sp[2] = -dTwdz.
template <int NX>
__global__ __launch_bounds__(1024) void k_spec_t(StepArgs A)
{
    constexpr int ROWS = rows_per_block<NX>();
    constexpr int T = ROWS * NX;
    static_assert(T <= 1024, "one point per thread");
    constexpr int NL = NX / 2;
    __shared__ double2 sq_[3 * ROWS * NL];
    double *sd = reinterpret_cast<double *>(sq_);
    const int k = blockIdx.y, j0 = blockIdx.x * ROWS, tid = threadIdx.x;
    const int i = tid % NX, jj = tid / NX, j = j0 + jj;
    double out[3] = {0.0, 0.0, 0.0};
    if (j < A.ny) {
        Pt g(A, i, j, k);
        const double t = g.T(0, 0, 0), tp = g.T(0, 0, 1), tm = g.T(0, 0, -1);
        const double u = g.U(0, 0, 0), v = g.V(0, 0, 0), w = g.W(0, 0, 0), wm = g.W(0, 0, -1);
        const double lt = A.lev[LV_LT * A.nz + k], lb = A.lev[LV_LB * A.nz + k];
        const double a1 = tp * lt + t * (1.0 - lt), a2 = tm * lb + t * (1.0 - lb);
        out[0] = u * t;
        out[1] = t * v;
        out[2] = -((a1 * w - a2 * wm) / A.lev[LV_DZ * A.nz + k]);
    }
#pragma unroll
    for (int f = 0; f < 3; ++f) sd[(f * ROWS + jj) * NX + i] = out[f];
    __syncthreads();
    lds_fft<NL, 3 * ROWS, NL, 1, T, -1>(sq_, A.tw_nl, tid);
    for (int p = tid; p < 3 * ROWS * NL; p += T) {
        int row = p / NL, l = p - row * NL, f = row / ROWS, r = row - f * ROWS, jr = j0 + r;
        if (jr < A.ny) A.sp[f][((size_t)k * A.ny + jr) * NL + l] = r2c_post<NL>(sq_ + row * NL, l, A.tw_nx, A.norm);
    }
}

// ---------------------------------------------------------------------------
// Spectral column kernels.  The z second difference of a lane-distributed
// column, with the wall ghosts of update_boundary_uvw / _p / _t
// (src/module_boundary.f90:33-196) as they appear in spectral space: a ghost
// -f(k0) - 2 u_mrf becomes -c(k0) - 2 u_mrf in the (0,0) mode and -c(k0)
// elsewhere.  kind: 0 u, 1 v, 2 p (Neumann), 3 t (Dirichlet, or the
// gradient ghost at the top with tbc_top = 1).
HC_DEV double2 zghost(const StepArgs &A, int kind, bool bottom, double2 c, bool is00)
{
    switch (kind) {
    case 0: {
        const double s = is00 ? 2.0 * A.u_mrf : 0.0;
        if (bottom || A.ochannel == 0) return make_double2(-c.x - s, -c.y);
        return c;
    }
    case 1:
        if (bottom || A.ochannel == 0) return make_double2(-c.x, -c.y);
        return c;
    case 2:
        return c;
    default: {
        if (!bottom && A.tbc_top)  // gradient ghost t + g dz_t
            return make_double2(c.x + (is00 ? A.gdz_top : 0.0), c.y);
        const double s = is00 ? 2.0 * (bottom ? A.tbot : A.ttop) : 0.0;
        return make_double2(s - c.x, -c.y);
    }
    }
}

// value of level r+1 of this lane (runtime r), or the next lane's first
template <int L>
HC_DEV double2 col_up(const ColMap &m, const double2 (&c)[L], int r, double2 next)
{
    double2 cand = czero();
#pragma unroll
    for (int q = 0; q < L; ++q)
        if (q == r + 1) cand = c[q];
    return r + 1 < m.cnt ? cand : next;
}

// q += coef * ((c(k+1)-c(k))/dz_t - (c(k)-c(k-1))/dz_b)/dz
template <int L>
HC_DEV void add_dzz(const StepArgs &A, const ColMap &m, int kind, bool is00, const double2 (&c)[L], double coef,
                    double2 (&q)[L])
{
    const int nz = A.nz;
    const double2 prev = shfl_up_c(last_of<L>(c, m.cnt), 1);
    const double2 next = shfl_dn_c(c[0], 1);
#pragma unroll
    for (int r = 0; r < L; ++r) {
        if (r >= m.cnt) continue;
        const int k = m.k0 + r;
        const double2 cm = k == 0 ? zghost(A, kind, true, c[r], is00) : (r == 0 ? prev : c[r > 0 ? r - 1 : 0]);
        const double2 cp = k == nz - 1 ? zghost(A, kind, false, c[r], is00) : col_up<L>(m, c, r, next);
        const double idzt = A.lev[LV_IDZT * nz + k], idzb = A.lev[LV_IDZB * nz + k], idz = A.lev[LV_IDZ * nz + k];
        const double zx = ((cp.x - c[r].x) * idzt - (c[r].x - cm.x) * idzb) * idz;
        const double zy = ((cp.y - c[r].y) * idzt - (c[r].y - cm.y) * idzb) * idz;
        q[r].x += coef * zx;
        q[r].y += coef * zy;
    }
}

// w levels k < kend3: q += coef * ((w(k+1)-w(k))/dz(k+1) - (w(k)-w(k-1))/dz(k))/dz_t(k),
// w(k0-1) = 0 and w(kend3) = 0 (src/module_spectral.f90:357-370)
template <int L>
HC_DEV void add_dzz_w(const StepArgs &A, const ColMap &m, const double2 (&c)[L], double coef, double2 (&q)[L])
{
    const int nz = A.nz;
    const double2 prev = shfl_up_c(last_of<L>(c, m.cnt), 1);
    const double2 next = shfl_dn_c(c[0], 1);
#pragma unroll
    for (int r = 0; r < L; ++r) {
        if (r >= m.cnt) continue;
        const int k = m.k0 + r;
        if (k >= nz - 1) continue;
        const double2 cm = k == 0 ? czero() : (r == 0 ? prev : c[r > 0 ? r - 1 : 0]);
        const double2 cp = k + 1 == nz - 1 ? czero() : col_up<L>(m, c, r, next);
        const double idzp = A.lev[LV_IDZ * nz + k + 1], idz = A.lev[LV_IDZ * nz + k], idzt = A.lev[LV_IDZT * nz + k];
        const double zx = ((cp.x - c[r].x) * idzp - (c[r].x - cm.x) * idz) * idzt;
        const double zy = ((cp.y - c[r].y) * idzp - (c[r].y - cm.y) * idz) * idzt;
        q[r].x += coef * zx;
        q[r].y += coef * zy;
    }
}

// One velocity component of get_momentum_cnvdiff_spec assembled in spectral
// space (e: 0 u, 1 v, 2 w):
//   q  = c + dif (-kx^2-ky^2) c + dif Dzz c - rc dt grad_e p + rb dt f_old + ra dt f_new
//   f_new = -u_mrf i kx c - coe (i kx X + i ky Y) + coe R
// with X, Y, R the transformed products of k_spec_mom.  f_new replaces f_old.
// With four LDS sub-tiles (A.ntile >= 4) the component moves in three
// global phases: {c, X, Y, R} in, {f_old} in, {f_new} out.
template <int L>
HC_DEV void spec_assemble(const StepArgs &A, double2 *tile, const ColMap &m, size_t plane, size_t col0, int e,
                          bool is00, double kx, double ky, double hxy, const double2 (&cp)[L], double2 (&q)[L])
{
    const int nz = A.nz;
    const double dt = A.dt, coe = A.coe;
    const double dif = A.rc * 0.5 * dt * A.nu;
    const double2 *cs = e == 0 ? A.cu : e == 1 ? A.cv : A.cw;
    const int ix = e == 0 ? 0 : e == 1 ? 1 : 5, iy = e == 0 ? 1 : e == 1 ? 2 : 6, ir = e == 0 ? 3 : e == 1 ? 4 : 7;
    double2 *fo = e == 0 ? A.cff : e == 1 ? A.cgg : A.chh;
    const bool batch = A.ntile >= 4;
    const size_t sub = (size_t)nz * kColWaves;
    double2 x[L], f[L];
    if (batch) {
        const double2 *srcs[4] = {cs, A.sp[ix], A.sp[iy], A.sp[ir]};
        col_fill_n<L, 4>(srcs, tile, nz, plane, col0);
        __syncthreads();
        col_take<L>(tile, m, x);
    } else {
        col_load<L>(cs, tile, nz, plane, col0, m, x);
    }
#pragma unroll
    for (int r = 0; r < L; ++r) {
        q[r] = make_double2(x[r].x + hxy * x[r].x, x[r].y + hxy * x[r].y);
        f[r] = times_i(x[r], -A.u_mrf * kx);
    }
    if (e == 2) add_dzz_w<L>(A, m, x, dif, q);
    else add_dzz<L>(A, m, e, is00, x, dif, q);
    if (batch) col_take<L>(tile + sub, m, x);
    else col_load<L>(A.sp[ix], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = csub(f[r], times_i(x[r], coe * kx));
    if (batch) col_take<L>(tile + 2 * sub, m, x);
    else col_load<L>(A.sp[iy], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = csub(f[r], times_i(x[r], coe * ky));
    if (batch) col_take<L>(tile + 3 * sub, m, x);
    else col_load<L>(A.sp[ir], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = make_double2(f[r].x + coe * x[r].x, f[r].y + coe * x[r].y);
    // pressure gradient
    const double cdt = A.rc * dt;
    if (e < 2) {
        const double kk = e == 0 ? kx : ky;
#pragma unroll
        for (int r = 0; r < L; ++r) q[r] = csub(q[r], times_i(cp[r], cdt * kk));
    } else {
        const double2 next = shfl_dn_c(cp[0], 1);
#pragma unroll
        for (int r = 0; r < L; ++r) {
            const int k = m.k0 + r;
            if (r >= m.cnt || k >= nz - 1) continue;
            const double2 pn = col_up<L>(m, cp, r, next);
            const double g = cdt * A.lev[LV_IDZT * nz + k];
            q[r] = make_double2(q[r].x - g * (pn.x - cp[r].x), q[r].y - g * (pn.y - cp[r].y));
        }
    }
    if (batch) {
        __syncthreads();
        col_fill(fo, tile, nz, plane, col0);
        __syncthreads();
        col_take<L>(tile, m, x);
    } else {
        col_load<L>(fo, tile, nz, plane, col0, m, x);
    }
    const double rb = A.rb * dt, ra = A.ra * dt;
#pragma unroll
    for (int r = 0; r < L; ++r) {
        q[r] = make_double2(q[r].x + rb * x[r].x, q[r].y + rb * x[r].y);
        q[r] = make_double2(q[r].x + ra * f[r].x, q[r].y + ra * f[r].y);
    }
    if (batch) {
        col_put<L>(tile + sub, m, f);
        __syncthreads();
        col_drain(fo, tile + sub, nz, plane, col0);
        __syncthreads();
    } else {
        col_store<L>(fo, tile, nz, plane, col0, m, f);
    }
}

// Spectral substep column: assembly of qu, qv, qw, the u, v, w Helmholtz
// solves, divergence, pressure Poisson, projection and new pressure of
// time_advancement_sp (src/module_navier_stokes.f90:417-470) with the
// spectral forms of get_div_spec, calculate_new_p_spec, calculate_uvw_spec
// (src/module_spectral.f90:752-894), then the 2/3-rule truncation (:1222-1268).
template <int L>
__global__ __launch_bounds__(64 * kColWaves) void k_column_spec(StepArgs A)
{
    extern __shared__ double2 tile[];
    const int l = blockIdx.x * kColWaves + (threadIdx.x >> 6), jm = blockIdx.y;
    const ColMap m = col_map(A.nz);
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    const bool is00 = (l == 0 && jm == 0);
    const int mw = jm < ny / 2 ? jm : jm - ny;
    const bool kept = l < A.lc && mw < A.mc && mw > -A.mc;
    const size_t plane = (size_t)ny * nl, col0 = (size_t)jm * nl + (size_t)blockIdx.x * kColWaves;
    const double kx = A.symx[SY_GIM * nl + l], ky = A.symy[SY_GIM * ny + jm];
    const double hxy = A.rc * (A.symx[SY_HU * nl + l] + A.symy[SY_HU * ny + jm]);
    double2 cpr[L], duv[2][L], dw[1][L], dp[1][L];
    col_load<L>(A.cp, tile, nz, plane, col0, m, cpr);
    spec_assemble<L>(A, tile, m, plane, col0, 0, is00, kx, ky, hxy, cpr, duv[0]);
    spec_assemble<L>(A, tile, m, plane, col0, 1, is00, kx, ky, hxy, cpr, duv[1]);
    spec_assemble<L>(A, tile, m, plane, col0, 2, is00, kx, ky, hxy, cpr, dw[0]);
    if (kept) {
        double a[L], b[L], c[L];
        assemble<L>(A, 0, l, jm, is00, m, a, b, c);
        if (is00) fixups<L>(A, 0, m, duv[0]);
        wave_tridiag<L, 2>(m.cnt, a, b, c, duv);
#pragma unroll
        for (int r = 0; r < L; ++r)
            if (r < m.cnt && m.k0 + r == nz - 1) dw[0][r] = czero();
        assemble<L>(A, 1, l, jm, is00, m, a, b, c);
        wave_tridiag<L, 1>(m.cnt, a, b, c, dw);
        col_divergence<L>(A, m, l, jm, duv[0], duv[1], dw[0], A.idt, dp[0]);
        assemble<L>(A, 2, l, jm, is00, m, a, b, c);
        if (is00) fixups<L>(A, 2, m, dp[0]);
        wave_tridiag_refined<L, 1>(m.cnt, a, b, c, dp);
        const double cdt = A.rc * A.dt;
        // new pressure first (:461), then the projection (:464)
        const double slap = A.symx[SY_L * nl + l] + A.symy[SY_L * ny + jm];
        const double cnu = A.rc * 0.5 * A.dt * A.nu;
        double2 lz[L];
#pragma unroll
        for (int r = 0; r < L; ++r) lz[r] = czero();
        add_dzz<L>(A, m, 2, is00, dp[0], 1.0, lz);
#pragma unroll
        for (int r = 0; r < L; ++r) {
            if (r >= m.cnt) continue;
            const double2 ph = dp[0][r];
            const double lx = slap * ph.x + lz[r].x, ly = slap * ph.y + lz[r].y;
            cpr[r] = make_double2(cpr[r].x + ph.x - cnu * lx, cpr[r].y + ph.y - cnu * ly);
        }
        col_project<L>(A, m, l, jm, cdt, dp[0], duv[0], duv[1], dw[0]);
    } else {
#pragma unroll
        for (int r = 0; r < L; ++r) {
            duv[0][r] = duv[1][r] = dw[0][r] = cpr[r] = czero();
        }
    }
    if (A.ntile >= 4) {
        const size_t sub = (size_t)nz * kColWaves;
        col_put<L>(tile, m, duv[0]);
        col_put<L>(tile + sub, m, duv[1]);
        col_put<L>(tile + 2 * sub, m, dw[0]);
        col_put<L>(tile + 3 * sub, m, cpr);
        __syncthreads();
        col_drain(A.cu, tile, nz, plane, col0);
        col_drain(A.cv, tile + sub, nz, plane, col0);
        col_drain(A.cw, tile + 2 * sub, nz, plane, col0);
        col_drain(A.cp, tile + 3 * sub, nz, plane, col0);
    } else {
        col_store<L>(A.cu, tile, nz, plane, col0, m, duv[0]);
        col_store<L>(A.cv, tile, nz, plane, col0, m, duv[1]);
        col_store<L>(A.cw, tile, nz, plane, col0, m, dw[0]);
        col_store<L>(A.cp, tile, nz, plane, col0, m, cpr);
    }
}

// Spectral temperature column (src/module_navier_stokes.f90:481-495):
// qt assembled from c_t and the products of k_spec_t, Dirichlet solve with
// tbnd, truncation.  ss_new replaces ss_old.
template <int L>
__global__ __launch_bounds__(64 * kColWaves) void k_column_spec_t(StepArgs A)
{
    extern __shared__ double2 tile[];
    const int l = blockIdx.x * kColWaves + (threadIdx.x >> 6), jm = blockIdx.y;
    const ColMap m = col_map(A.nz);
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    const bool is00 = (l == 0 && jm == 0);
    const int mw = jm < ny / 2 ? jm : jm - ny;
    const bool kept = l < A.lc && mw < A.mc && mw > -A.mc;
    const size_t plane = (size_t)ny * nl, col0 = (size_t)jm * nl + (size_t)blockIdx.x * kColWaves;
    const double kx = A.symx[SY_GIM * nl + l], ky = A.symy[SY_GIM * ny + jm];
    const double hxy = A.rc * (A.symx[SY_HT * nl + l] + A.symy[SY_HT * ny + jm]);
    const double dif = A.rc * 0.5 * A.dt * A.alpha;
    double2 x[L], f[L], d[1][L];
    col_load<L>(A.ct, tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) {
        d[0][r] = make_double2(x[r].x + hxy * x[r].x, x[r].y + hxy * x[r].y);
        f[r] = times_i(x[r], -A.u_mrf * kx);
    }
    add_dzz<L>(A, m, 3, is00, x, dif, d[0]);
    col_load<L>(A.sp[0], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = csub(f[r], times_i(x[r], kx));
    col_load<L>(A.sp[1], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = csub(f[r], times_i(x[r], ky));
    col_load<L>(A.sp[2], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = cadd(f[r], x[r]);
    col_load<L>(A.css, tile, nz, plane, col0, m, x);
    const double rb = A.rb * A.dt, ra = A.ra * A.dt;
#pragma unroll
    for (int r = 0; r < L; ++r) {
        d[0][r] = make_double2(d[0][r].x + rb * x[r].x, d[0][r].y + rb * x[r].y);
        d[0][r] = make_double2(d[0][r].x + ra * f[r].x, d[0][r].y + ra * f[r].y);
    }
    col_store<L>(A.css, tile, nz, plane, col0, m, f);
    if (kept) {
        double a[L], b[L], c[L];
        assemble<L>(A, 3, l, jm, is00, m, a, b, c);
        if (is00) fixups<L>(A, 3, m, d[0]);
        wave_tridiag<L, 1>(m.cnt, a, b, c, d);
    } else {
#pragma unroll
        for (int r = 0; r < L; ++r) d[0][r] = czero();
    }
    col_store<L>(A.ct, tile, nz, plane, col0, m, d[0]);
}

// 2/3-rule truncation of a half spectrum in place (spectral_truncation,
// src/module_spectral.f90:1222-1268), grid over (l, j, k)
__global__ __launch_bounds__(256) void k_truncate(double2 *s, int nl, int ny, int nz, int lc, int mc)
{
    const size_t n = (size_t)nl * ny * nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int l = (int)(q % nl), j = (int)((q / nl) % ny);
        const int mw = j < ny / 2 ? j : j - ny;
        if (l >= lc || mw >= mc || mw <= -mc) s[q] = czero();
    }
}

// Spectral divergence i kx c_u + i ky c_v + (c_w(k) - c_w(k-1))/dz(k), c_w(-1) = 0
// (get_div_spec, src/module_spectral.f90:752-795) into dst
__global__ __launch_bounds__(256) void k_spec_div(StepArgs A, double2 *dst)
{
    const size_t n = (size_t)A.nl * A.ny * A.nz;
    const size_t plane = (size_t)A.nl * A.ny;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int l = (int)(q % A.nl), j = (int)((q / A.nl) % A.ny), k = (int)(q / plane);
        const double kx = A.symx[SY_GIM * A.nl + l], ky = A.symy[SY_GIM * A.ny + j];
        double2 d = cadd(times_i(A.cu[q], kx), times_i(A.cv[q], ky));
        const double2 wm = k == 0 ? czero() : A.cw[q - plane];
        const double idz = A.lev[LV_IDZ * A.nz + k];
        d.x += (A.cw[q].x - wm.x) * idz;
        d.y += (A.cw[q].y - wm.y) * idz;
        dst[q] = d;
    }
}
