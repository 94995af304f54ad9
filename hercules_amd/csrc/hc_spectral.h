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
// Included by hc_spec.hip and hc_spec_col.hip after the column helpers.
#pragma once

// ---------------------------------------------------------------------------
// y transforms from a source spectrum into a destination spectrum, with an
// optional Fourier first derivative applied first (dev_fft_linear_cmplx with
// ord = 1, src/module_spectral.f90:1305-1311).
template <class R>
struct YIOArgsT {
    const cplx<R> *src[4];
    cplx<R> *dst[4];
    int mul[4];  // 0: none, 1: i kx(l), 2: i ky(m)
    int nl, ny;
    const cplx<R> *tw_ny;
    const R *symx, *symy;
};
template <class C> HC_DEV C times_i(C v, real_of<C> k) { return mk2<C>(-k * v.y, k * v.x); }

template <int NY, int SIGN, class R>
__global__ __launch_bounds__(256) void k_yfft_io(YIOArgsT<R> A)
{
    using C = cplx<R>;
    __shared__ C tile[NY * kYPad];
    const int f = blockIdx.z;
    const C *src = A.src[f];
    C *dst = A.dst[f];
    const int mul = A.mul[f];
    const int l0 = blockIdx.x * kYTile, k = blockIdx.y, tid = threadIdx.x;
    __shared__ C twy[NY];
    for (int q = tid; q < NY; q += 256) twy[q] = A.tw_ny[q];
    constexpr int NIT = (NY * kYTile + 255) / 256;
    C vv[NIT];
#pragma unroll
    for (int m = 0; m < NIT; ++m) {
        const int p = tid + m * 256, j = p / kYTile, ll = p - j * kYTile;
        vv[m] = p < NY * kYTile ? src[((size_t)k * NY + j) * A.nl + l0 + ll] : czero_t<C>();
    }
#pragma unroll
    for (int m = 0; m < NIT; ++m) {
        const int p = tid + m * 256, j = p / kYTile, ll = p - j * kYTile;
        if (p >= NY * kYTile) continue;
        C v = vv[m];
        if (mul == 1) v = times_i(v, A.symx[SY_GIM * A.nl + l0 + ll]);
        else if (mul == 2) v = times_i(v, A.symy[SY_GIM * A.ny + j]);
        if (SIGN > 0 && j == NY / 2) v = czero_t<C>();
        tile[j * kYPad + ll] = v;
    }
    __syncthreads();
    lds_fft<NY, kYTile, 1, kYPad, 256, SIGN>(tile, twy, tid);
    for (int p = tid; p < NY * kYTile; p += 256) {
        int j = p / kYTile, ll = p - j * kYTile;
        C v = tile[j * kYPad + ll];
        if (SIGN < 0 && j == NY / 2) v = czero_t<C>();
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
template <int NX, class R, int ZO = 2>
__global__ __launch_bounds__(1024) void k_spec_mom(StepArgsT<R> A)
{
    using C = cplx<R>;
    constexpr int ROWS = rows_per_block<NX>();
    constexpr int T = ROWS * NX;
    static_assert(T <= 1024, "one point per thread");
    constexpr int NL = NX / 2;
    __shared__ C sq_[8 * ROWS * NL];
    R *sd = reinterpret_cast<R *>(sq_);
    const int k = blockIdx.y, j0 = blockIdx.x * ROWS, tid = threadIdx.x;
    const int i = tid % NX, jj = tid / NX, j = j0 + jj;
    const bool in = j < A.ny;
    if (A.issk) {
        for (int p = tid; p < 6 * ROWS * NL; p += T) {
            int row = p / NL, l = p - row * NL, f = row / ROWS, r = row - f * ROWS, jr = j0 + r;
            C z = czero_t<C>();
            if (jr < A.ny) {
                const C *s = A.sdv[f] + ((size_t)k * A.ny + jr) * NL;
                C xl = s[l];
                C xt = l == 0 ? czero_t<C>() : cconj(s[NL - l]);
                z = c2r_pre(xl, xt, A.tw_nx[l]);
            }
            sq_[row * NL + l] = z;
        }
        __syncthreads();
        lds_fft<NL, 6 * ROWS, NL, 1, T, +1>(sq_, A.tw_nl, tid);
    }
    R out[8];
#pragma unroll
    for (int f = 0; f < 8; ++f) out[f] = R(0.0);
    if (in && ZO == 6) {
        // 6th-order z (zorder = 6): the same products and terms with the
        // face values, z derivatives and z interpolants of the ZOp tables
        PtT<R> g(A, i, j, k);
        const R u = g.U(0, 0, 0), v = g.V(0, 0, 0), w = g.W(0, 0, 0);
        const R coe = A.coe;
        const R dpx = A.dp_opt == 2 ? R(A.scal[0]) : A.dpx;
        const R a1 = g.uf6(k), a1v = g.vf6(k);  // u, v at this level's w face
        out[0] = u * u;
        out[1] = u * v;
        out[2] = v * v;
        out[5] = a1 * w;
        out[6] = a1v * w;
        const R tmp2u = A.dp_opt == 3 ? A.fc * (v - A.vg) : dpx;
        const R tmp2v = A.dp_opt == 3 ? A.fc * (A.ug - u - A.u_mrf) : R(0.0);
        // d(u w)/dz, d(v w)/dz and du/dz, dv/dz at this level from the face values
        int s0;
        const double *wd = g.zw(ZO_D, k, s0);
        R uwz = R(0.0), vwz = R(0.0), duz = R(0.0), dvz = R(0.0);
#pragma unroll 1
        for (int q = 0; q < A.zn1; ++q) {  // faces s0-1 .. s0+zn1-2, all within -1..nz-1
            const int kw = s0 + q - 1;
            const R uf = g.uf6(kw), vf = g.vf6(kw);
            const R wf = (kw < 0 || kw >= A.nz - 1) ? R(0.0) : g.W(0, 0, kw - k);
            uwz += R(wd[q]) * uf * wf;
            vwz += R(wd[q]) * vf * wf;
            duz += R(wd[q]) * uf;
            dvz += R(wd[q]) * vf;
        }
        R ru = R(1.0) / coe * tmp2u - uwz;
        R rv = R(1.0) / coe * tmp2v - vwz;
        R rw = k < A.nz - 1 ? -g.wwz6() : R(0.0);
        if (A.isscalar && k < A.nz - 1) rw += R(1.0) / coe * (A.got * (g.thz6(0) - A.t_ref));
        if (A.issk) {
            auto D = [&](int f) { return sd[(f * ROWS + jj) * NX + i]; };
            const R wc = g.whz6(k);  // w at this level's centre
            ru -= u * D(0) + v * D(1) + duz * wc;
            rv -= u * D(2) + v * D(3) + dvz * wc;
            if (k < A.nz - 1) {
                int sg;
                const double *wg = g.zw(ZO_G, k, sg);
                R dwz = R(0.0);
#pragma unroll 1
                for (int q = 0; q < A.zn1; ++q) dwz += R(wg[q]) * g.whz6(sg + q);
                rw -= a1 * D(4) + a1v * D(5) + w * dwz;
            }
        }
        out[3] = ru;
        out[4] = rv;
        out[7] = rw;
    }
    if (in && ZO != 6) {
        PtT<R> g(A, i, j, k);
        const R u = g.U(0, 0, 0), up = g.U(0, 0, 1), um_ = g.U(0, 0, -1);
        const R v = g.V(0, 0, 0), vp = g.V(0, 0, 1), vm = g.V(0, 0, -1);
        const R w = g.W(0, 0, 0), wp = g.W(0, 0, 1), wm = g.W(0, 0, -1);
        const R lt = A.lev[LV_LT * A.nz + k], lb = A.lev[LV_LB * A.nz + k];
        const R idz = A.lev[LV_IDZ * A.nz + k], idzt = A.lev[LV_IDZT * A.nz + k];
        const R coe = A.coe;
        const R dpx = A.dp_opt == 2 ? R(A.scal[0]) : A.dpx;
        // face values of u and v (dev_fft_nonlinear_cmplx weqn=1 uses a1 too)
        const R a1 = up * lt + u * (R(1.0) - lt), a2 = um_ * lb + u * (R(1.0) - lb);
        const R a1v = vp * lt + v * (R(1.0) - lt), a2v = vm * lb + v * (R(1.0) - lb);
        out[0] = u * u;
        out[1] = u * v;
        out[2] = v * v;
        out[5] = a1 * w;
        out[6] = a1v * w;
        const R tmp2u = A.dp_opt == 3 ? A.fc * (v - A.vg) : dpx;
        const R tmp2v = A.dp_opt == 3 ? A.fc * (A.ug - u - A.u_mrf) : R(0.0);
        R ru = R(1.0) / coe * tmp2u - (a1 * w - a2 * wm) / A.lev[LV_DZ * A.nz + k];
        R rv = R(1.0) / coe * tmp2v - (a1v * w - a2v * wm) / A.lev[LV_DZ * A.nz + k];
        const R hp = R(0.5) * (w + wp), hm = R(0.5) * (w + wm);
        const R wwz = (hp * hp - hm * hm) / A.lev[LV_DZT * A.nz + k];
        R rw;
        if (A.isscalar) {
            const R t = g.T(0, 0, 0), tp = g.T(0, 0, 1);
            rw = R(1.0) / coe * (A.got * ((R(1.0) - lt) * t + lt * tp - A.t_ref)) - wwz;
        } else {
            rw = -wwz;
        }
        if (A.issk) {
            auto D = [&](int f) { return sd[(f * ROWS + jj) * NX + i]; };
            ru -= u * D(0);
            ru -= v * D(1);
            ru -= (a1 - a2) / A.lev[LV_DZ * A.nz + k] * (R(0.5) * (w + wm));
            rv -= u * D(2);
            rv -= v * D(3);
            rv -= (a1v - a2v) / A.lev[LV_DZ * A.nz + k] * (R(0.5) * (w + wm));
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
template <int NX, class R, int ZO = 2>
__global__ __launch_bounds__(1024) void k_spec_t(StepArgsT<R> A)
{
    using C = cplx<R>;
    constexpr int ROWS = rows_per_block<NX>();
    constexpr int T = ROWS * NX;
    static_assert(T <= 1024, "one point per thread");
    constexpr int NL = NX / 2;
    __shared__ C sq_[3 * ROWS * NL];
    R *sd = reinterpret_cast<R *>(sq_);
    const int k = blockIdx.y, j0 = blockIdx.x * ROWS, tid = threadIdx.x;
    const int i = tid % NX, jj = tid / NX, j = j0 + jj;
    R out[3] = {R(0.0), R(0.0), R(0.0)};
    if (j < A.ny) {
        PtT<R> g(A, i, j, k);
        const R t = g.T(0, 0, 0), tp = g.T(0, 0, 1), tm = g.T(0, 0, -1);
        const R u = g.U(0, 0, 0), v = g.V(0, 0, 0), w = g.W(0, 0, 0), wm = g.W(0, 0, -1);
        const R lt = A.lev[LV_LT * A.nz + k], lb = A.lev[LV_LB * A.nz + k];
        const R a1 = tp * lt + t * (R(1.0) - lt), a2 = tm * lb + t * (R(1.0) - lb);
        out[0] = u * t;
        out[1] = t * v;
        if constexpr (ZO == 6) out[2] = -g.twz6();
        else out[2] = -((a1 * w - a2 * wm) / A.lev[LV_DZ * A.nz + k]);
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
// gradient ghost at the top with tbc_top = 1).  C is the complex type of the
// solve (double2, or float2 in the FP32 blocks of the mixed variant).
template <class C>
HC_DEV C zghost(const StepArgs &A, int kind, bool bottom, C c, bool is00)
{
    using Re = real_of<C>;
    switch (kind) {
    case 0: {
        const Re s = is00 ? Re(2.0 * A.u_mrf) : Re(0);
        if (bottom || A.ochannel == 0) return mk2<C>(-c.x - s, -c.y);
        return c;
    }
    case 1:
        if (bottom || A.ochannel == 0) return mk2<C>(-c.x, -c.y);
        return c;
    case 2:
        return c;
    default: {
        if (!bottom && A.tbc_top)  // gradient ghost t + g dz_t
            return mk2<C>(c.x + (is00 ? Re(A.gdz_top) : Re(0)), c.y);
        const Re s = is00 ? Re(2.0 * (bottom ? A.tbot : A.ttop)) : Re(0);
        return mk2<C>(s - c.x, -c.y);
    }
    }
}

// value of level r+1 of this lane (runtime r), or the next lane's first
template <int L, class C>
HC_DEV C col_up(const ColMap &m, const C (&c)[L], int r, C next)
{
    C cand = czero_t<C>();
#pragma unroll
    for (int q = 0; q < L; ++q)
        if (q == r + 1) cand = c[q];
    return r + 1 < m.cnt ? cand : next;
}

// q += coef * ((c(k+1)-c(k))/dz_t - (c(k)-c(k-1))/dz_b)/dz
template <int L, class C>
HC_DEV void add_dzz(const StepArgs &A, const ColMap &m, int kind, bool is00, const C (&c)[L], real_of<C> coef,
                    C (&q)[L])
{
    using Re = real_of<C>;
    const int nz = A.nz;
    const C prev = shfl_up_c(last_of<L>(c, m.cnt), 1);
    const C next = shfl_dn_c(c[0], 1);
#pragma unroll
    for (int r = 0; r < L; ++r) {
        if (r >= m.cnt) continue;
        const int k = m.k0 + r;
        const C cm = k == 0 ? zghost(A, kind, true, c[r], is00) : (r == 0 ? prev : c[r > 0 ? r - 1 : 0]);
        const C cp = k == nz - 1 ? zghost(A, kind, false, c[r], is00) : col_up<L>(m, c, r, next);
        const Re idzt = Re(A.lev[LV_IDZT * nz + k]), idzb = Re(A.lev[LV_IDZB * nz + k]), idz = Re(A.lev[LV_IDZ * nz + k]);
        const Re zx = ((cp.x - c[r].x) * idzt - (c[r].x - cm.x) * idzb) * idz;
        const Re zy = ((cp.y - c[r].y) * idzt - (c[r].y - cm.y) * idzb) * idz;
        q[r].x += coef * zx;
        q[r].y += coef * zy;
    }
}

// w levels k < kend3: q += coef * ((w(k+1)-w(k))/dz(k+1) - (w(k)-w(k-1))/dz(k))/dz_t(k),
// w(k0-1) = 0 and w(kend3) = 0 (src/module_spectral.f90:357-370)
template <int L, class C>
HC_DEV void add_dzz_w(const StepArgs &A, const ColMap &m, const C (&c)[L], real_of<C> coef, C (&q)[L])
{
    using Re = real_of<C>;
    const int nz = A.nz;
    const C prev = shfl_up_c(last_of<L>(c, m.cnt), 1);
    const C next = shfl_dn_c(c[0], 1);
#pragma unroll
    for (int r = 0; r < L; ++r) {
        if (r >= m.cnt) continue;
        const int k = m.k0 + r;
        if (k >= nz - 1) continue;
        const C cm = k == 0 ? czero_t<C>() : (r == 0 ? prev : c[r > 0 ? r - 1 : 0]);
        const C cp = k + 1 == nz - 1 ? czero_t<C>() : col_up<L>(m, c, r, next);
        const Re idzp = Re(A.lev[LV_IDZ * nz + k + 1]), idz = Re(A.lev[LV_IDZ * nz + k]);
        const Re idzt = Re(A.lev[LV_IDZT * nz + k]);
        const Re zx = ((cp.x - c[r].x) * idzp - (c[r].x - cm.x) * idz) * idzt;
        const Re zy = ((cp.y - c[r].y) * idzp - (c[r].y - cm.y) * idz) * idzt;
        q[r].x += coef * zx;
        q[r].y += coef * zy;
    }
}

// Spectra of the spectral column kernels in the stored precision, and the
// block selection of the mixed variant (as ColIO).
template <class RIO>
struct SpecIO {
    cplx<RIO> *cu, *cv, *cw, *cp, *ct, *cff, *cgg, *chh, *css;
    cplx<RIO> *sp[8];
    int sel, lo_l, lo_m, remap;
};
template <class RIO>
HC_DEV ColPos spec_pos(const SpecIO<RIO> &io, int ny)
{
    ColIO<RIO> c;
    c.sel = io.sel; c.lo_l = io.lo_l; c.lo_m = io.lo_m; c.remap = io.remap;
    return col_pos(c, ny);
}

// Levels k >= kdamp0 of field fid (0 u, 1 v, 2 w, 3 p, 4 t) as they were
// before the spectral Rayleigh damping of the substep.  The reference damps
// only the spectra; its z diffusion (src/module_spectral.f90:83, 225, 367,
// 547) and w pressure gradient (:382) are formed from the physical fields,
// which keep the undamped values.  No-op without damping.
template <int L, class CC>
HC_DEV void spec_undamped(const StepArgs &A, const ColMap &m, size_t plane, size_t col, int fid, CC (&x)[L])
{
    if (!A.spund) return;
    const cplx<double> *u = A.spund + (size_t)fid * (A.nz - A.kdamp0) * plane;
#pragma unroll
    for (int r = 0; r < L; ++r) {
        const int k = m.k0 + r;
        if (r < m.cnt && k >= A.kdamp0) x[r] = ccast<CC>(u[(size_t)(k - A.kdamp0) * plane + col]);
    }
}

// One velocity component of get_momentum_cnvdiff_spec assembled in spectral
// space (e: 0 u, 1 v, 2 w):
//   q  = c + dif (-kx^2-ky^2) c + dif Dzz c - rc dt grad_e p + rb dt f_old + ra dt f_new
//   f_new = -u_mrf i kx c - coe (i kx X + i ky Y) + coe R
// with X, Y, R the transformed products of k_spec_mom.  f_new replaces f_old.
// With four LDS sub-tiles (A.ntile >= 4) the component moves in three
// global phases: {c, X, Y, R} in, {f_old} in, {f_new} out.
template <int L, class CI, class CC>
HC_DEV void spec_assemble(const StepArgs &A, const SpecIO<real_of<CI>> &io, CI *tile, const ColMap &m, size_t plane,
                          size_t col0, int e, bool is00, real_of<CC> kx, real_of<CC> ky, real_of<CC> hxy,
                          const CC (&cp)[L], CC (&q)[L])
{
    using Re = real_of<CC>;
    const int nz = A.nz;
    const Re dt = Re(A.dt), coe = Re(A.coe);
    const Re dif = Re(A.rc * 0.5 * A.dt * A.nu);
    const CI *cs = e == 0 ? io.cu : e == 1 ? io.cv : io.cw;
    const int ix = e == 0 ? 0 : e == 1 ? 1 : 5, iy = e == 0 ? 1 : e == 1 ? 2 : 6, ir = e == 0 ? 3 : e == 1 ? 4 : 7;
    CI *fo = e == 0 ? io.cff : e == 1 ? io.cgg : io.chh;
    const bool batch = A.ntile >= 4;
    const size_t sub = (size_t)nz * kColWaves;
    CC x[L], f[L];
    if (batch) {
        const CI *srcs[4] = {cs, io.sp[ix], io.sp[iy], io.sp[ir]};
        col_fill_n<L, 4>(srcs, tile, nz, plane, col0);
        __syncthreads();
        col_take<L>(tile, m, x);
    } else {
        col_load<L>(cs, tile, nz, plane, col0, m, x);
    }
    const Re um = Re(A.u_mrf);
#pragma unroll
    for (int r = 0; r < L; ++r) {
        q[r] = mk2<CC>(x[r].x + hxy * x[r].x, x[r].y + hxy * x[r].y);
        f[r] = times_i(x[r], -um * kx);
    }
    spec_undamped<L>(A, m, plane, col0 + (threadIdx.x >> 6), e, x);
    if (e == 2) add_dzz_w<L>(A, m, x, dif, q);
    else add_dzz<L>(A, m, e, is00, x, dif, q);
    if (batch) col_take<L>(tile + sub, m, x);
    else col_load<L>(io.sp[ix], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = csub(f[r], times_i(x[r], coe * kx));
    if (batch) col_take<L>(tile + 2 * sub, m, x);
    else col_load<L>(io.sp[iy], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = csub(f[r], times_i(x[r], coe * ky));
    if (batch) col_take<L>(tile + 3 * sub, m, x);
    else col_load<L>(io.sp[ir], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = mk2<CC>(f[r].x + coe * x[r].x, f[r].y + coe * x[r].y);
    // pressure gradient
    const Re cdt = Re(A.rc * A.dt);
    if (e < 2) {
        const Re kk = e == 0 ? kx : ky;
#pragma unroll
        for (int r = 0; r < L; ++r) q[r] = csub(q[r], times_i(cp[r], cdt * kk));
    } else {
        const CC next = shfl_dn_c(cp[0], 1);
#pragma unroll
        for (int r = 0; r < L; ++r) {
            const int k = m.k0 + r;
            if (r >= m.cnt || k >= nz - 1) continue;
            const CC pn = col_up<L>(m, cp, r, next);
            const Re g = cdt * Re(A.lev[LV_IDZT * nz + k]);
            q[r] = mk2<CC>(q[r].x - g * (pn.x - cp[r].x), q[r].y - g * (pn.y - cp[r].y));
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
    const Re rb = Re(A.rb * A.dt), ra = Re(A.ra * A.dt);
#pragma unroll
    for (int r = 0; r < L; ++r) {
        q[r] = mk2<CC>(q[r].x + rb * x[r].x, q[r].y + rb * x[r].y);
        q[r] = mk2<CC>(q[r].x + ra * f[r].x, q[r].y + ra * f[r].y);
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
// RIO: stored precision of the spectra; RC: the arithmetic of the column.
template <int L, class RIO, class RC>
__global__ __launch_bounds__(64 * kColWaves) void k_column_spec(StepArgs A, SpecIO<RIO> io)
{
    using CI = cplx<RIO>;
    using CC = cplx<RC>;
    extern __shared__ double2 tile_raw[];
    CI *tile = reinterpret_cast<CI *>(tile_raw);
    const ColPos pos = spec_pos(io, A.ny);
    if (pos.skip) return;  // block-uniform, before any barrier
    const int l = pos.bx * kColWaves + (threadIdx.x >> 6), jm = pos.jm;
    const ColMap m = col_map(A.nz);
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    const bool is00 = (l == 0 && jm == 0);
    const int mw = jm < ny / 2 ? jm : jm - ny;
    const bool kept = l < A.lc && mw < A.mc && mw > -A.mc;
    const size_t plane = (size_t)ny * nl, col0 = (size_t)jm * nl + (size_t)pos.bx * kColWaves;
    const RC kx = RC(A.symx[SY_GIM * nl + l]), ky = RC(A.symy[SY_GIM * ny + jm]);
    const RC hxy = RC(A.rc * (A.symx[SY_HU * nl + l] + A.symy[SY_HU * ny + jm]));
    CC cpr[L], duv[2][L], dw[1][L], dp[1][L];
    col_load<L>(io.cp, tile, nz, plane, col0, m, cpr);
    spec_assemble<L>(A, io, tile, m, plane, col0, 0, is00, kx, ky, hxy, cpr, duv[0]);
    spec_assemble<L>(A, io, tile, m, plane, col0, 1, is00, kx, ky, hxy, cpr, duv[1]);
    if (A.spund) {
        // w pressure gradient from the undamped p; the damped one returns for
        // the new pressure
        spec_undamped<L>(A, m, plane, col0 + (threadIdx.x >> 6), 3, cpr);
        spec_assemble<L>(A, io, tile, m, plane, col0, 2, is00, kx, ky, hxy, cpr, dw[0]);
        col_load<L>(io.cp, tile, nz, plane, col0, m, cpr);
    } else {
        spec_assemble<L>(A, io, tile, m, plane, col0, 2, is00, kx, ky, hxy, cpr, dw[0]);
    }
    if (A.stage) {  // get_momentum_cnvdiff_spec alone: c_qu, c_qv, c_qw in place of c_u, c_v, c_w
        col_store<L>(io.cu, tile, nz, plane, col0, m, duv[0]);
        col_store<L>(io.cv, tile, nz, plane, col0, m, duv[1]);
        col_store<L>(io.cw, tile, nz, plane, col0, m, dw[0]);
        return;
    }
    if (kept) {
        RC a[L], b[L], c[L];
        assemble<L>(A, 0, l, jm, is00, m, a, b, c);
        if (is00) fixups<L>(A, 0, m, duv[0]);
        wave_tridiag<L, 2>(m.cnt, a, b, c, duv);
#pragma unroll
        for (int r = 0; r < L; ++r)
            if (r < m.cnt && m.k0 + r == nz - 1) dw[0][r] = czero_t<CC>();
        assemble<L>(A, 1, l, jm, is00, m, a, b, c);
        wave_tridiag<L, 1>(m.cnt, a, b, c, dw);
        col_divergence<L>(A, m, l, jm, duv[0], duv[1], dw[0], RC(A.idt), dp[0]);
        assemble<L>(A, 2, l, jm, is00, m, a, b, c);
        if (is00) fixups<L>(A, 2, m, dp[0]);
        wave_tridiag_refined<L, 1>(m.cnt, a, b, c, dp);
        const RC cdt = RC(A.rc * A.dt);
        // new pressure first (:461), then the projection (:464)
        const RC slap = RC(A.symx[SY_L * nl + l] + A.symy[SY_L * ny + jm]);
        const RC cnu = RC(A.rc * 0.5 * A.dt * A.nu);
        CC lz[L];
#pragma unroll
        for (int r = 0; r < L; ++r) lz[r] = czero_t<CC>();
        add_dzz<L>(A, m, 2, is00, dp[0], RC(1), lz);
#pragma unroll
        for (int r = 0; r < L; ++r) {
            if (r >= m.cnt) continue;
            const CC ph = dp[0][r];
            const RC lx = slap * ph.x + lz[r].x, ly = slap * ph.y + lz[r].y;
            cpr[r] = mk2<CC>(cpr[r].x + ph.x - cnu * lx, cpr[r].y + ph.y - cnu * ly);
        }
        col_project<L>(A, m, l, jm, cdt, dp[0], duv[0], duv[1], dw[0]);
    } else {
#pragma unroll
        for (int r = 0; r < L; ++r) {
            duv[0][r] = duv[1][r] = dw[0][r] = cpr[r] = czero_t<CC>();
        }
    }
    if (A.ntile >= 4) {
        const size_t sub = (size_t)nz * kColWaves;
        col_put<L>(tile, m, duv[0]);
        col_put<L>(tile + sub, m, duv[1]);
        col_put<L>(tile + 2 * sub, m, dw[0]);
        col_put<L>(tile + 3 * sub, m, cpr);
        __syncthreads();
        col_drain(io.cu, tile, nz, plane, col0);
        col_drain(io.cv, tile + sub, nz, plane, col0);
        col_drain(io.cw, tile + 2 * sub, nz, plane, col0);
        col_drain(io.cp, tile + 3 * sub, nz, plane, col0);
    } else {
        col_store<L>(io.cu, tile, nz, plane, col0, m, duv[0]);
        col_store<L>(io.cv, tile, nz, plane, col0, m, duv[1]);
        col_store<L>(io.cw, tile, nz, plane, col0, m, dw[0]);
        col_store<L>(io.cp, tile, nz, plane, col0, m, cpr);
    }
}

// Spectral temperature column (src/module_navier_stokes.f90:481-495):
// qt assembled from c_t and the products of k_spec_t, Dirichlet solve with
// tbnd, truncation.  ss_new replaces ss_old.
template <int L, class RIO, class RC>
__global__ __launch_bounds__(64 * kColWaves) void k_column_spec_t(StepArgs A, SpecIO<RIO> io)
{
    using CI = cplx<RIO>;
    using CC = cplx<RC>;
    extern __shared__ double2 tile_raw[];
    CI *tile = reinterpret_cast<CI *>(tile_raw);
    const ColPos pos = spec_pos(io, A.ny);
    if (pos.skip) return;
    const int l = pos.bx * kColWaves + (threadIdx.x >> 6), jm = pos.jm;
    const ColMap m = col_map(A.nz);
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    const bool is00 = (l == 0 && jm == 0);
    const int mw = jm < ny / 2 ? jm : jm - ny;
    const bool kept = l < A.lc && mw < A.mc && mw > -A.mc;
    const size_t plane = (size_t)ny * nl, col0 = (size_t)jm * nl + (size_t)pos.bx * kColWaves;
    const RC kx = RC(A.symx[SY_GIM * nl + l]), ky = RC(A.symy[SY_GIM * ny + jm]);
    const RC hxy = RC(A.rc * (A.symx[SY_HT * nl + l] + A.symy[SY_HT * ny + jm]));
    const RC dif = RC(A.rc * 0.5 * A.dt * A.alpha);
    const RC um = RC(A.u_mrf);
    CC x[L], f[L], d[1][L];
    col_load<L>(io.ct, tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) {
        d[0][r] = mk2<CC>(x[r].x + hxy * x[r].x, x[r].y + hxy * x[r].y);
        f[r] = times_i(x[r], -um * kx);
    }
    spec_undamped<L>(A, m, plane, col0 + (threadIdx.x >> 6), 4, x);
    add_dzz<L>(A, m, 3, is00, x, dif, d[0]);
    col_load<L>(io.sp[0], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = csub(f[r], times_i(x[r], kx));
    col_load<L>(io.sp[1], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = csub(f[r], times_i(x[r], ky));
    col_load<L>(io.sp[2], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = cadd(f[r], x[r]);
    col_load<L>(io.css, tile, nz, plane, col0, m, x);
    const RC rb = RC(A.rb * A.dt), ra = RC(A.ra * A.dt);
#pragma unroll
    for (int r = 0; r < L; ++r) {
        d[0][r] = mk2<CC>(d[0][r].x + rb * x[r].x, d[0][r].y + rb * x[r].y);
        d[0][r] = mk2<CC>(d[0][r].x + ra * f[r].x, d[0][r].y + ra * f[r].y);
    }
    col_store<L>(io.css, tile, nz, plane, col0, m, f);
    if (A.stage) {  // get_temperature_cnvdiff_spec alone: c_qt in place of c_t
        col_store<L>(io.ct, tile, nz, plane, col0, m, d[0]);
        return;
    }
    if (kept) {
        RC a[L], b[L], c[L];
        assemble<L>(A, 3, l, jm, is00, m, a, b, c);
        if (is00) fixups<L>(A, 3, m, d[0]);
        wave_tridiag<L, 1>(m.cnt, a, b, c, d);
    } else {
#pragma unroll
        for (int r = 0; r < L; ++r) d[0][r] = czero_t<CC>();
    }
    col_store<L>(io.ct, tile, nz, plane, col0, m, d[0]);
}

// 2/3-rule truncation of a half spectrum in place (spectral_truncation,
// src/module_spectral.f90:1222-1268), grid over (l, j, k)
template <class C>
__global__ __launch_bounds__(256) void k_truncate(C *s, int nl, int ny, int nz, int lc, int mc)
{
    const size_t n = (size_t)nl * ny * nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int l = (int)(q % nl), j = (int)((q / nl) % ny);
        const int mw = j < ny / 2 ? j : j - ny;
        if (l >= lc || mw >= mc || mw <= -mc) s[q] = czero_t<C>();
    }
}

// Spectral divergence i kx c_u + i ky c_v + (c_w(k) - c_w(k-1))/dz(k), c_w(-1) = 0
// (get_div_spec, src/module_spectral.f90:752-795) into dst, formed in FP64
template <class C>
__global__ __launch_bounds__(256) void k_spec_div(StepArgs A, const C *cu, const C *cv, const C *cw, C *dst)
{
    const size_t n = (size_t)A.nl * A.ny * A.nz;
    const size_t plane = (size_t)A.nl * A.ny;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int l = (int)(q % A.nl), j = (int)((q / A.nl) % A.ny), k = (int)(q / plane);
        const double kx = A.symx[SY_GIM * A.nl + l], ky = A.symy[SY_GIM * A.ny + j];
        double2 d = cadd(times_i(ccast<double2>(cu[q]), kx), times_i(ccast<double2>(cv[q]), ky));
        if (A.zorder == 6) {  // ZO_D over the faces, zero at the walls
            const size_t e = (size_t)ZO_D * A.nz + k;
            const int s0 = A.zst[e];
            for (int jz = 0; jz < A.zn1; ++jz) {
                const int lv = s0 + jz - 1;
                if (lv < 0 || lv >= A.nz - 1) continue;
                const double2 wv = ccast<double2>(cw[q + (size_t)(lv - k) * plane]);
                d.x += A.zwt[e * kZW + jz] * wv.x;
                d.y += A.zwt[e * kZW + jz] * wv.y;
            }
        } else {
            const double2 w0 = ccast<double2>(cw[q]);
            const double2 wm = k == 0 ? czero() : ccast<double2>(cw[q - plane]);
            const double idz = A.lev[LV_IDZ * A.nz + k];
            d.x += (w0.x - wm.x) * idz;
            d.y += (w0.y - wm.y) * idz;
        }
        dst[q] = ccast<C>(d);
    }
}
