// 6th-order wall-normal operators (hc_params.zorder = 6), column side.
//
// The reference is 2nd order in z (src/module_cnvdiff.f90:89,121,
// src/module_poisson_solver.f90:32-122).  With zorder = 6 the implicit
// operators of a column are
//   u, v, t:  I + rc (-hxy - kappa Z)      Z = ZO_C2_* (7-point d2/dz2, centres)
//   w:        I + rc (-hxy - kappa Z)      Z = ZO_F2   (7-point d2/dz2, faces)
//   p:        rc (hxy + ZO_P)              ZO_P = ZO_D o ZO_G
// with hxy the horizontal symbol the 2nd-order rows use.  They are banded
// (half band 5, and 7 for ZO_P) and solved directly, one thread per column,
// by k_z6_band (hc_band6.h).  A column pass of the step is therefore a
// sequence of launches: right-hand sides (spectral path), the Helmholtz
// solves, the divergence, the pressure solve, then the projection and the
// pressure increment.  The explicit operator products of the divergence,
// projection and increment run one column per wavefront through the
// wavefront's own slice of an LDS sub-tile.  FP64 only.
#pragma once

#include "hc_band6.h"

// The 6th-order interpolants of the substep's fields, once per point, for
// the right-hand-side kernels: u, v at the w face of level k (the top wall
// face included), w at the centre of level k, t at the interior w face k.
// A carries no interpolant buffers, so the accessor evaluates them.
__global__ __launch_bounds__(256) void k_zinterp6(StepArgs A, double *fu, double *fv, double *cw, double *ft)
{
    const size_t n = (size_t)A.nx * A.ny * A.nz, plane = (size_t)A.nx * A.ny;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(q / plane), j = (int)((q - (size_t)k * plane) / A.nx), i = (int)(q % A.nx);
        const Pt g(A, i, j, k);
        fu[q] = g.uf6(k);
        fv[q] = g.vf6(k);
        cw[q] = g.whz6(k);
        if (ft) ft[q] = k < A.nz - 1 ? g.thz6(0) : 0.0;
    }
}

// y = Z x for the ZOp `op` on this wavefront's column.  Data index d of
// the table maps to column level d - off; levels outside 0..ndata-1 are
// zero (the homogeneous wall data).  Rows k >= nrows give y = 0.
template <int L, int W>
HC_DEV void z6_apply(const StepArgs &A, int op, int off, int ndata, int nrows, const ColMap &m, double2 *buf,
                     const double2 (&x)[L], double2 (&y)[L])
{
    // buf is this wavefront's own slice of the tile: ordering the LDS
    // accesses within the wavefront is enough (no block barrier)
    const int nz = A.nz;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int r = 0; r < L; ++r)
        if (r < m.cnt) buf[m.k0 + r] = x[r];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int r = 0; r < L; ++r) {
        double2 acc = make_double2(0.0, 0.0);
        const int k = m.k0 + r;
        if (r < m.cnt && k < nrows) {
            const size_t e = (size_t)op * nz + k;
            const int s = A.zst[e];
            const double *w = A.zwt + e * kZW;
#pragma unroll
            for (int j = 0; j < W; ++j) {
                const int lv = s + j - off;
                if (lv >= 0 && lv < ndata) {
                    const double2 v = buf[lv];
                    acc.x += w[j] * v.x;
                    acc.y += w[j] * v.y;
                }
            }
        }
        y[r] = acc;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Operator of equation eq (0 u/v, 1 w, 2 p, 3 t) of column (l, jm): its
// ZOp table, horizontal symbol and diffusivity.
struct Z6Op {
    int eq, op;
    double hxy, kappa, rc, b0;
    bool delnull;
};
HC_DEV Z6Op z6_op(const StepArgs &A, int eq, int l, int jm)
{
    Z6Op o;
    o.eq = eq;
    o.rc = A.rc;
    o.b0 = 0.0;
    o.delnull = false;
    const int nl = A.nl, ny = A.ny;
    switch (eq) {
    case 0:
        o.op = A.ochannel ? ZO_C2_DN : ZO_C2_DD;
        o.hxy = A.symx[SY_HU * nl + l] + A.symy[SY_HU * ny + jm];
        o.kappa = 0.5 * A.dt * A.nu;
        break;
    case 1:
        o.op = ZO_F2;
        o.hxy = A.symx[SY_HU * nl + l] + A.symy[SY_HU * ny + jm];
        o.kappa = 0.5 * A.dt * A.nu;
        break;
    case 2:
        o.op = ZO_P;
        o.hxy = A.symx[SY_PX * nl + l] + A.symy[SY_PX * ny + jm];
        o.kappa = 0.0;
        break;
    default:
        o.op = A.tbc_top ? ZO_C2_DN : ZO_C2_DD;
        o.hxy = A.symx[SY_HT * nl + l] + A.symy[SY_HT * ny + jm];
        o.kappa = 0.5 * A.dt * A.alpha;
        break;
    }
    return o;
}
// (0,0)-mode wall data of the u and t equations: the Crank-Nicolson
// implicit half of the wall terms, rc kappa w_wall (wall value or gradient),
// for the rows whose stencils reach a wall datum.
template <int L>
HC_DEV void z6_fixups(const StepArgs &A, const Z6Op &o, const ColMap &m, double2 (&d)[L])
{
    const int nz = A.nz;
    double bot, top;
    bool has_top;
    if (o.eq == 0) {
        bot = -A.u_mrf;
        top = -A.u_mrf;
        has_top = A.ochannel == 0;  // open top: du/dz = 0
    } else if (o.eq == 3) {
        bot = A.tbot;
        top = A.tbc_top ? A.gdz_top * A.lev[LV_IDZT * nz + nz - 1] : A.ttop;
        has_top = true;
    } else {
        return;
    }
    const double ck = o.rc * o.kappa;
#pragma unroll
    for (int r = 0; r < L; ++r) {
        if (r >= m.cnt) continue;
        const int k = m.k0 + r;
        const size_t e = (size_t)o.op * nz + k;
        const int s = A.zst[e];
        const double *w = A.zwt + e * kZW;
        if (s == 0) d[r].x += ck * w[0] * bot;
        const int jt = nz + 1 - s;
        if (has_top && jt < 7) d[r].x += ck * w[jt] * top;
    }
}

// Divergence of (u, v, w) times `scale`: horizontal symbols, ZO_D in z.
template <int L>
HC_DEV void z6_divergence(const StepArgs &A, const ColMap &m, double2 *buf, int l, int jm, const double2 (&u)[L],
                          const double2 (&v)[L], const double2 (&w)[L], double scale, double2 (&out)[L])
{
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    z6_apply<L, 6>(A, ZO_D, 1, nz - 1, nz, m, buf, w, out);
    const double2 sdx = make_double2(A.symx[SY_DRE * nl + l], A.symx[SY_DIM * nl + l]);
    const double2 sdy = make_double2(A.symy[SY_DRE * ny + jm], A.symy[SY_DIM * ny + jm]);
#pragma unroll
    for (int r = 0; r < L; ++r) {
        const double2 dv = cadd(cadd(cmul(sdx, u[r]), cmul(sdy, v[r])), out[r]);
        out[r] = r < m.cnt ? make_double2(dv.x * scale, dv.y * scale) : make_double2(0.0, 0.0);
    }
}

// u, v, w -= cdt grad phi: horizontal symbols, ZO_G at the interior faces.
template <int L>
HC_DEV void z6_project(const StepArgs &A, const ColMap &m, double2 *buf, int l, int jm, double cdt,
                       const double2 (&phi)[L], double2 (&u)[L], double2 (&v)[L], double2 (&w)[L])
{
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    double2 gz[L];
    z6_apply<L, 6>(A, ZO_G, 0, nz, nz - 1, m, buf, phi, gz);
    const double2 sgx = make_double2(A.symx[SY_GRE * nl + l], A.symx[SY_GIM * nl + l]);
    const double2 sgy = make_double2(A.symy[SY_GRE * ny + jm], A.symy[SY_GIM * ny + jm]);
#pragma unroll
    for (int r = 0; r < L; ++r) {
        if (r >= m.cnt) continue;
        u[r] = csub(u[r], cscale(cmul(sgx, phi[r]), cdt));
        v[r] = csub(v[r], cscale(cmul(sgy, phi[r]), cdt));
        if (m.k0 + r < nz - 1) { w[r].x -= cdt * gz[r].x; w[r].y -= cdt * gz[r].y; }
    }
}

// Divergence of (s0, s1, s2) = (u, v, w) times `scale` into dst (ZO_D over
// the interior faces, w(nz-1) taken as the top wall's zero): get_p_eqn_src
// with the 6th-order z operator.
template <int L>
__global__ __launch_bounds__(64 * kColWaves) void k_z6_div(StepArgs A, ColIO<double> io, double2 *dst, double scale)
{
    extern __shared__ double2 tile_raw[];
    double2 *tile = tile_raw;
    const ColPos pos = col_pos(io, A.ny);
    if (pos.skip) return;
    const int wave = threadIdx.x >> 6;
    const int l = pos.bx * kColWaves + wave, jm = pos.jm;
    const ColMap m = col_map(A.nz);
    const size_t plane = (size_t)A.ny * A.nl, col0 = (size_t)jm * A.nl + (size_t)pos.bx * kColWaves;
    double2 *buf = tile + (size_t)(kColWaves + wave) * A.nz;
    double2 u[L], v[L], w[L], d[L];
    col_load<L>(io.s0, tile, A.nz, plane, col0, m, u);
    col_load<L>(io.s1, tile, A.nz, plane, col0, m, v);
    col_load<L>(io.s2, tile, A.nz, plane, col0, m, w);
#pragma unroll
    for (int r = 0; r < L; ++r)
        if (r < m.cnt && m.k0 + r == A.nz - 1) w[r] = make_double2(0.0, 0.0);
    z6_divergence<L>(A, m, buf, l, jm, u, v, w, scale, d);
    col_store<L>(dst, tile, A.nz, plane, col0, m, d);
}

// Projection u, v, w -= cdt grad phi (calculate_uvw) with phi in `phi`;
// mode 0 (FD substep): cdt = rc dt, then the pressure increment
// phi - rc dt nu / 2 (hxy_L + ZO_P) phi (calculate_new_p) replaces phi;
// mode 1 (hc_project): cdt = 1; mode 2 (spectral substep): cdt = rc dt, the
// increment is added to c_p (s3), and the columns outside the 2/3 rule are
// zeroed.
template <int L>
__global__ __launch_bounds__(64 * kColWaves) void k_z6_proj(StepArgs A, ColIO<double> io, double2 *phi, int mode)
{
    extern __shared__ double2 tile_raw[];
    double2 *tile = tile_raw;
    const ColPos pos = col_pos(io, A.ny);
    if (pos.skip) return;
    const int wave = threadIdx.x >> 6;
    const int l = pos.bx * kColWaves + wave, jm = pos.jm;
    const ColMap m = col_map(A.nz);
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    const size_t plane = (size_t)ny * nl, col0 = (size_t)jm * nl + (size_t)pos.bx * kColWaves;
    double2 *buf = tile + (size_t)(kColWaves + wave) * nz;
    const int mw = jm < ny / 2 ? jm : jm - ny;
    const bool kept = mode != 2 || (l < A.lc && mw < A.mc && mw > -A.mc);
    const double2 zero = make_double2(0.0, 0.0);
    double2 ph[L], u[L], v[L], w[L];
    col_load<L>(phi, tile, nz, plane, col0, m, ph);
    col_load<L>(io.s0, tile, nz, plane, col0, m, u);
    col_load<L>(io.s1, tile, nz, plane, col0, m, v);
    col_load<L>(io.s2, tile, nz, plane, col0, m, w);
    const double cdt = mode == 1 ? 1.0 : A.rc * A.dt;
    z6_project<L>(A, m, buf, l, jm, cdt, ph, u, v, w);
    if (!kept) {
#pragma unroll
        for (int r = 0; r < L; ++r) u[r] = v[r] = w[r] = zero;
    }
    col_store<L>(io.s0, tile, nz, plane, col0, m, u);
    col_store<L>(io.s1, tile, nz, plane, col0, m, v);
    col_store<L>(io.s2, tile, nz, plane, col0, m, w);
    if (mode == 1) return;
    double2 x[L];
    z6_apply<L, kZW>(A, ZO_P, 0, nz, nz, m, buf, ph, x);
    const double slap = A.symx[SY_L * nl + l] + A.symy[SY_L * ny + jm];
    const double cnu = A.rc * 0.5 * A.dt * A.nu;
#pragma unroll
    for (int r = 0; r < L; ++r)
        x[r] = make_double2(ph[r].x - cnu * (slap * ph[r].x + x[r].x), ph[r].y - cnu * (slap * ph[r].y + x[r].y));
    if (mode == 0) {
#pragma unroll
        for (int r = 0; r < L; ++r)
            if (r >= m.cnt) x[r] = zero;
        col_store<L>(io.s3, tile, nz, plane, col0, m, x);
        return;
    }
    double2 p[L];
    col_load<L>(io.s3, tile, nz, plane, col0, m, p);
#pragma unroll
    for (int r = 0; r < L; ++r) p[r] = (kept && r < m.cnt) ? cadd(p[r], x[r]) : zero;
    col_store<L>(io.s3, tile, nz, plane, col0, m, p);
}

// ---- spectral path (cds = 3) with the 6th-order z operators ---------------
// One velocity component of the spectral column assembly (spec_assemble)
// with ZO_C2_* / ZO_F2 for the explicit half of the z diffusion (plus the
// (0,0)-mode wall data of u) and ZO_G for the w pressure gradient.
template <int L>
HC_DEV void spec_assemble6(const StepArgs &A, const SpecIO<double> &io, double2 *tile, double2 *buf, const ColMap &m,
                           size_t plane, size_t col0, int e, int l, int jm, bool is00, double kx, double ky,
                           double hxy, double2 (&q)[L])
{
    const int nz = A.nz;
    const double coe = A.coe;
    const double dif = A.rc * 0.5 * A.dt * A.nu;
    const double2 *cs = e == 0 ? io.cu : e == 1 ? io.cv : io.cw;
    const int ix = e == 0 ? 0 : e == 1 ? 1 : 5, iy = e == 0 ? 1 : e == 1 ? 2 : 6, ir = e == 0 ? 3 : e == 1 ? 4 : 7;
    double2 *fo = e == 0 ? io.cff : e == 1 ? io.cgg : io.chh;
    double2 x[L], f[L], t[L];
    col_load<L>(cs, tile, nz, plane, col0, m, x);
    const double um = A.u_mrf;
#pragma unroll
    for (int r = 0; r < L; ++r) {
        q[r] = make_double2(x[r].x + hxy * x[r].x, x[r].y + hxy * x[r].y);
        f[r] = times_i(x[r], -um * kx);
    }
    const Z6Op o = z6_op(A, e == 2 ? 1 : 0, l, jm);
    if (e == 2) z6_apply<L, 7>(A, ZO_F2, 1, nz - 1, nz - 1, m, buf, x, t);
    else z6_apply<L, 7>(A, o.op, 1, nz, nz, m, buf, x, t);
#pragma unroll
    for (int r = 0; r < L; ++r) { q[r].x += dif * t[r].x; q[r].y += dif * t[r].y; }
    if (is00 && e == 0) z6_fixups<L>(A, o, m, q);
    col_load<L>(io.sp[ix], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = csub(f[r], times_i(x[r], coe * kx));
    col_load<L>(io.sp[iy], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = csub(f[r], times_i(x[r], coe * ky));
    col_load<L>(io.sp[ir], tile, nz, plane, col0, m, x);
#pragma unroll
    for (int r = 0; r < L; ++r) f[r] = make_double2(f[r].x + coe * x[r].x, f[r].y + coe * x[r].y);
    const double cdt = A.rc * A.dt;
    col_load<L>(io.cp, tile, nz, plane, col0, m, x);  // the pressure of the substep's start
    if (e < 2) {
        const double kk = e == 0 ? kx : ky;
#pragma unroll
        for (int r = 0; r < L; ++r) q[r] = csub(q[r], times_i(x[r], cdt * kk));
    } else {
        z6_apply<L, 6>(A, ZO_G, 0, nz, nz - 1, m, buf, x, t);
#pragma unroll
        for (int r = 0; r < L; ++r) { q[r].x -= cdt * t[r].x; q[r].y -= cdt * t[r].y; }
    }
    col_load<L>(fo, tile, nz, plane, col0, m, x);
    const double rb = A.rb * A.dt, ra = A.ra * A.dt;
#pragma unroll
    for (int r = 0; r < L; ++r) {
        q[r] = make_double2(q[r].x + rb * x[r].x + ra * f[r].x, q[r].y + rb * x[r].y + ra * f[r].y);
    }
    col_store<L>(fo, tile, nz, plane, col0, m, f);
}

// Right-hand sides of the spectral substep's u, v, w Helmholtz systems
// (get_momentum_cnvdiff_spec with the 6th-order z operators), in place of
// c_u, c_v, c_w; w's top wall level is zero.  The new right-hand sides f
// replace the old ones.
template <int L>
__global__ __launch_bounds__(64 * kColWaves) void k_spec6_rhs(StepArgs A, SpecIO<double> io)
{
    extern __shared__ double2 tile_raw[];
    double2 *tile = tile_raw;
    const ColPos pos = spec_pos(io, A.ny);
    if (pos.skip) return;
    const int wave = threadIdx.x >> 6;
    const int l = pos.bx * kColWaves + wave, jm = pos.jm;
    const ColMap m = col_map(A.nz);
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    const bool is00 = (l == 0 && jm == 0);
    const size_t plane = (size_t)ny * nl, col0 = (size_t)jm * nl + (size_t)pos.bx * kColWaves;
    const double kx = A.symx[SY_GIM * nl + l], ky = A.symy[SY_GIM * ny + jm];
    const double hxy = A.rc * (A.symx[SY_HU * nl + l] + A.symy[SY_HU * ny + jm]);
    double2 *buf = tile + (size_t)(kColWaves + wave) * nz;
    double2 q[L];
    spec_assemble6<L>(A, io, tile, buf, m, plane, col0, 0, l, jm, is00, kx, ky, hxy, q);
    col_store<L>(io.cu, tile, nz, plane, col0, m, q);
    spec_assemble6<L>(A, io, tile, buf, m, plane, col0, 1, l, jm, is00, kx, ky, hxy, q);
    col_store<L>(io.cv, tile, nz, plane, col0, m, q);
    spec_assemble6<L>(A, io, tile, buf, m, plane, col0, 2, l, jm, is00, kx, ky, hxy, q);
#pragma unroll
    for (int r = 0; r < L; ++r)
        if (r < m.cnt && m.k0 + r == nz - 1) q[r] = make_double2(0.0, 0.0);
    col_store<L>(io.cw, tile, nz, plane, col0, m, q);
}

// Right-hand side of the spectral temperature system
// (get_temperature_cnvdiff_spec with the 6th-order z operator) in place of
// c_t; the new right-hand side s replaces the old one in c_ss.
template <int L>
__global__ __launch_bounds__(64 * kColWaves) void k_spec6_t_rhs(StepArgs A, SpecIO<double> io)
{
    extern __shared__ double2 tile_raw[];
    double2 *tile = tile_raw;
    const ColPos pos = spec_pos(io, A.ny);
    if (pos.skip) return;
    const int wave = threadIdx.x >> 6;
    const int l = pos.bx * kColWaves + wave, jm = pos.jm;
    const ColMap m = col_map(A.nz);
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    const bool is00 = (l == 0 && jm == 0);
    const size_t plane = (size_t)ny * nl, col0 = (size_t)jm * nl + (size_t)pos.bx * kColWaves;
    const double kx = A.symx[SY_GIM * nl + l], ky = A.symy[SY_GIM * ny + jm];
    const double hxy = A.rc * (A.symx[SY_HT * nl + l] + A.symy[SY_HT * ny + jm]);
    const double dif = A.rc * 0.5 * A.dt * A.alpha;
    const double um = A.u_mrf;
    double2 x[L], f[L], d[L], t[L];
    col_load<L>(io.ct, tile, nz, plane, col0, m, x);
    double2 *buf = tile + (size_t)(kColWaves + wave) * nz;
#pragma unroll
    for (int r = 0; r < L; ++r) {
        d[r] = make_double2(x[r].x + hxy * x[r].x, x[r].y + hxy * x[r].y);
        f[r] = times_i(x[r], -um * kx);
    }
    const Z6Op o = z6_op(A, 3, l, jm);
    z6_apply<L, 7>(A, o.op, 1, nz, nz, m, buf, x, t);
#pragma unroll
    for (int r = 0; r < L; ++r) { d[r].x += dif * t[r].x; d[r].y += dif * t[r].y; }
    if (is00) z6_fixups<L>(A, o, m, d);  // explicit half of the wall data
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
    const double rb = A.rb * A.dt, ra = A.ra * A.dt;
#pragma unroll
    for (int r = 0; r < L; ++r) d[r] = make_double2(d[r].x + rb * x[r].x + ra * f[r].x, d[r].y + rb * x[r].y + ra * f[r].y);
    col_store<L>(io.css, tile, nz, plane, col0, m, f);
    col_store<L>(io.ct, tile, nz, plane, col0, m, d);
}
