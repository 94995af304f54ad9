// Column-kernel building blocks: one (l, m) column of a half spectrum per
// wavefront, levels spread over the lanes, global loads and stores through
// LDS tiles; row assembly of the implicit z systems and the spectral-space
// divergence, projection and pressure increment.
#pragma once

#include "hc_xfer.h"

// ---------------------------------------------------------------------------
// Column kernels: one (l, m) column per wavefront, eight consecutive l per
// block, levels spread over the lanes.
constexpr int kColWaves = 8;

struct ColMap {
    int lane, cnt, k0;
};
HC_DEV ColMap col_map(int nz)
{
    ColMap m;
    m.lane = threadIdx.x & 63;
    int base = nz / 64, rem = nz % 64;
    m.cnt = base + (m.lane < rem ? 1 : 0);
    m.k0 = m.lane * base + (m.lane < rem ? m.lane : rem);
    return m;
}

// The eight columns of a block pass through an LDS tile of nz rows x 8
// complex (one 128-byte row per level) so that global loads and stores are
// whole cache lines.  Element (k, w) sits in slot (w + k/4) mod 8 of row k:
// the lanes of a wavefront, which own consecutive groups of ~4 levels,
// then hit distinct bank groups.
HC_DEV int col_tidx(int k, int w) { return k * kColWaves + ((w + (k >> 2)) & (kColWaves - 1)); }

// The tile holds the stored precision (CI: double2, or float2 in the mixed
// variant); the registers of a column hold the solve's precision (CC), the
// conversion happens between the two.
template <int L, class CI, class CC>
HC_DEV void col_load(const CI *__restrict__ src, CI *tile, int nz, size_t plane, size_t col0, const ColMap &m,
                     CC (&d)[L])
{
    const int w = threadIdx.x >> 6;
    for (int q = threadIdx.x; q < nz * kColWaves; q += 64 * kColWaves) {
        int k = q / kColWaves, c = q - k * kColWaves;
        tile[col_tidx(k, c)] = src[(size_t)k * plane + col0 + c];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < L; ++r) d[r] = r < m.cnt ? ccast<CC>(tile[col_tidx(m.k0 + r, w)]) : czero_t<CC>();
    __syncthreads();
}

template <int L, class CI, class CC>
HC_DEV void col_store(CI *__restrict__ dst, CI *tile, int nz, size_t plane, size_t col0, const ColMap &m,
                      const CC (&d)[L])
{
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < L; ++r)
        if (r < m.cnt) tile[col_tidx(m.k0 + r, w)] = ccast<CI>(d[r]);
    __syncthreads();
    for (int q = threadIdx.x; q < nz * kColWaves; q += 64 * kColWaves) {
        int k = q / kColWaves, c = q - k * kColWaves;
        dst[(size_t)k * plane + col0 + c] = tile[col_tidx(k, c)];
    }
    __syncthreads();
}

// Multi-field transfers: fields f = 0..n-1 use sub-tiles f * nz * kColWaves
// of the tile, so several fields share one global phase and one barrier.
template <class CI>
HC_DEV void col_fill(const CI *__restrict__ src, CI *sub, int nz, size_t plane, size_t col0)
{
    for (int q = threadIdx.x; q < nz * kColWaves; q += 64 * kColWaves) {
        int k = q / kColWaves, c = q - k * kColWaves;
        sub[col_tidx(k, c)] = src[(size_t)k * plane + col0 + c];
    }
}
// NF fields at once: every load is issued before the first LDS store
// (nz <= 64 L, so each thread moves at most L elements per field)
template <int L, int NF, class CI>
HC_DEV void col_fill_n(const CI *const (&src)[NF], CI *tile, int nz, size_t plane, size_t col0)
{
    const size_t sub = (size_t)nz * kColWaves;
    CI v[NF][L];
#pragma unroll
    for (int it = 0; it < L; ++it) {
        const int q = threadIdx.x + it * 64 * kColWaves;
        const int k = q / kColWaves, c = q - k * kColWaves;
#pragma unroll
        for (int f = 0; f < NF; ++f)
            v[f][it] = q < nz * kColWaves ? src[f][(size_t)k * plane + col0 + c] : czero_t<CI>();
    }
#pragma unroll
    for (int it = 0; it < L; ++it) {
        const int q = threadIdx.x + it * 64 * kColWaves;
        const int k = q / kColWaves, c = q - k * kColWaves;
        if (q < nz * kColWaves) {
#pragma unroll
            for (int f = 0; f < NF; ++f) tile[f * sub + col_tidx(k, c)] = v[f][it];
        }
    }
}
template <int L, class CI, class CC>
HC_DEV void col_take(const CI *sub, const ColMap &m, CC (&d)[L])
{
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < L; ++r) d[r] = r < m.cnt ? ccast<CC>(sub[col_tidx(m.k0 + r, w)]) : czero_t<CC>();
}
template <int L, class CI, class CC>
HC_DEV void col_put(CI *sub, const ColMap &m, const CC (&d)[L])
{
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < L; ++r)
        if (r < m.cnt) sub[col_tidx(m.k0 + r, w)] = ccast<CI>(d[r]);
}
template <class CI>
HC_DEV void col_drain(CI *__restrict__ dst, const CI *sub, int nz, size_t plane, size_t col0)
{
    for (int q = threadIdx.x; q < nz * kColWaves; q += 64 * kColWaves) {
        int k = q / kColWaves, c = q - k * kColWaves;
        dst[(size_t)k * plane + col0 + c] = sub[col_tidx(k, c)];
    }
}

// value of level (k0+cnt-1) from register array v (runtime index)
template <int L, class C>
HC_DEV C last_of(const C (&v)[L], int cnt)
{
    C r = czero_t<C>();
#pragma unroll
    for (int q = 0; q < L; ++q)
        if (q == cnt - 1) r = v[q];
    return r;
}

// Assemble the rows of one equation for this lane's levels.  eq: 0 u/v,
// 1 w, 2 p, 3 t.  The diagonal base of the (0,0) mode differs at the wall
// rows because get_{u,w,t}_eqn_coeff zero the 1-D aa(k)/cc(k) while visiting
// the first mode (src/module_poisson_solver.f90:257-272,361-366,455-469).
// The rows are formed in FP64 and rounded once to the solve's type Re.
template <int L, class Re>
HC_DEV void assemble(const StepArgs &A, int eq, int l, int jm, bool is00, const ColMap &m, Re (&a)[L],
                     Re (&b)[L], Re (&c)[L])
{
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    const double rc = A.rc;
    int ta, tc, td, td00;
    double hxy;
    switch (eq) {
    case 0: ta = LV_AU; tc = LV_CU; td = LV_DU; td00 = LV_DU00;
        hxy = A.symx[SY_HU * nl + l] + A.symy[SY_HU * ny + jm]; break;
    case 1: ta = LV_AW; tc = LV_CW; td = LV_DW; td00 = LV_DW00;
        hxy = A.symx[SY_HU * nl + l] + A.symy[SY_HU * ny + jm]; break;
    case 2: ta = LV_AP; tc = LV_CP; td = LV_DP; td00 = LV_DP00;
        hxy = A.symx[SY_PX * nl + l] + A.symy[SY_PX * ny + jm]; break;
    default: ta = LV_AT; tc = LV_CT; td = LV_DT; td00 = LV_DT00;
        hxy = A.symx[SY_HT * nl + l] + A.symy[SY_HT * ny + jm]; break;
    }
#pragma unroll
    for (int r = 0; r < L; ++r) {
        a[r] = Re(0); b[r] = Re(1); c[r] = Re(0);
        if (r < m.cnt) {
            int k = m.k0 + r;
            if (eq == 1 && k == nz - 1) continue;  // w at kend3: trivial row, w = 0
            a[r] = Re(rc * A.lev[ta * nz + k]);
            c[r] = Re(rc * A.lev[tc * nz + k]);
            double dg = A.lev[(is00 ? td00 : td) * nz + k];
            if (eq == 2) {
                b[r] = Re(rc * (hxy + dg));
                if (is00 && k == 0) { a[r] = Re(0); c[r] = Re(0); }  // delnull (:558-564)
            } else {
                b[r] = Re(1.0 + rc * (-hxy + dg));
            }
        }
    }
}

// (0,0)-mode right-hand-side fix-ups of poisson_solver_fft (:572-607)
template <int L, class C>
HC_DEV void fixups(const StepArgs &A, int eq, const ColMap &m, C (&d)[L])
{
    using Re = real_of<C>;
#pragma unroll
    for (int r = 0; r < L; ++r) {
        if (r >= m.cnt) continue;
        int k = m.k0 + r;
        if (eq == 0) {  // ubnd
            if (k == 0) d[r].x += Re(2.0 * (A.rc * A.aU0) * A.u_mrf);
            if (k == A.nz - 1 && A.ochannel == 0) d[r].x += Re(2.0 * (A.rc * A.cUe) * A.u_mrf);
        } else if (eq == 3) {  // tbnd
            if (k == 0) d[r].x -= Re(2.0 * (A.rc * A.aT0) * A.tbot);
            if (k == A.nz - 1)  // Dirichlet ghost 2 ttop - t, or gradient ghost t + g dz_t
                d[r].x -= A.tbc_top ? Re((A.rc * A.cTe) * A.gdz_top) : Re(2.0 * (A.rc * A.cTe) * A.ttop);
        } else if (eq == 2) {  // delnull
            if (k == 0) d[r] = czero_t<C>();
        }
    }
}

// Spectral divergence of (u, v, w) for this lane's levels, times `scale`:
// D_x u + D_y v + (w(k) - w(k-1))/dz(k), w(-1) = 0 (get_p_eqn_src,
// src/module_poisson_solver.f90:141-172, for periodic x, y).
template <int L, class C>
HC_DEV void col_divergence(const StepArgs &A, const ColMap &m, int l, int jm, const C (&u)[L], const C (&v)[L],
                           const C (&w)[L], real_of<C> scale, C (&out)[L])
{
    using Re = real_of<C>;
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    const C sdx = mk2<C>(Re(A.symx[SY_DRE * nl + l]), Re(A.symx[SY_DIM * nl + l]));
    const C sdy = mk2<C>(Re(A.symy[SY_DRE * ny + jm]), Re(A.symy[SY_DIM * ny + jm]));
    C wprev = shfl_up_c(last_of<L>(w, m.cnt), 1);
    if (m.lane == 0) wprev = czero_t<C>();
#pragma unroll
    for (int r = 0; r < L; ++r) {
        out[r] = czero_t<C>();
        if (r < m.cnt) {
            int k = m.k0 + r;
            C wm = r == 0 ? wprev : w[r > 0 ? r - 1 : 0];
            Re idz = Re(A.lev[LV_IDZ * nz + k]);
            C dv = cadd(cmul(sdx, u[r]), cmul(sdy, v[r]));
            C dwz = csub(w[r], wm);
            dv.x += dwz.x * idz;
            dv.y += dwz.y * idz;
            out[r] = cscale(dv, scale);
        }
    }
}

// phi(k+1) seen from level r of this lane (Neumann copy at the top wall)
template <int L, class C>
HC_DEV C col_next(const ColMap &m, int nz, const C (&ph)[L], int r, C pnext)
{
    int k = m.k0 + r;
    if (k >= nz - 1) return ph[r];
    C cand = czero_t<C>();
#pragma unroll
    for (int q = 0; q < L; ++q)
        if (q == r + 1) cand = ph[q];
    return r + 1 < m.cnt ? cand : pnext;
}

// Projection u -= cdt G_x phi, v -= cdt G_y phi, w(k) -= cdt (phi(k+1)-phi(k))/dz_t(k)
// for k < kend3 (calculate_uvw, src/module_navier_stokes.f90:659-734).
template <int L, class C>
HC_DEV void col_project(const StepArgs &A, const ColMap &m, int l, int jm, real_of<C> cdt, const C (&ph)[L],
                        C (&u)[L], C (&v)[L], C (&w)[L])
{
    using Re = real_of<C>;
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    const C sgx = mk2<C>(Re(A.symx[SY_GRE * nl + l]), Re(A.symx[SY_GIM * nl + l]));
    const C sgy = mk2<C>(Re(A.symy[SY_GRE * ny + jm]), Re(A.symy[SY_GIM * ny + jm]));
    C pnext = shfl_dn_c(ph[0], 1);
#pragma unroll
    for (int r = 0; r < L; ++r) {
        if (r < m.cnt) {
            int k = m.k0 + r;
            u[r] = csub(u[r], cscale(cmul(sgx, ph[r]), cdt));
            v[r] = csub(v[r], cscale(cmul(sgy, ph[r]), cdt));
            if (k < nz - 1) {
                C gz = csub(col_next<L>(m, nz, ph, r, pnext), ph[r]);
                Re f = cdt * Re(A.lev[LV_IDZT * nz + k]);
                w[r] = mk2<C>(w[r].x - f * gz.x, w[r].y - f * gz.y);
            }
        }
    }
}

// Spectra and column selection of a column-kernel launch.  The columns of
// one block are kColWaves consecutive l at one y mode jm.  sel = 0: every
// block of the grid; 1: only the low-wavenumber blocks (l < 8 * lo_l, |m|
// <= lo_m), grid (lo_l, 2 lo_m + 1); 2: every block except those.  The mixed
// precision variant solves the low-wavenumber blocks in FP64 and the rest
// in FP32.
template <class RIO>
struct ColIO {
    cplx<RIO> *s0, *s1, *s2, *s3;
    int sel, lo_l, lo_m;
    int remap;  // 1: XCD-contiguous block order (neighbouring l blocks share an L2)
};
struct ColPos {
    int bx, jm;
    bool skip;
};
template <class RIO>
HC_DEV ColPos col_pos(const ColIO<RIO> &io, int ny)
{
    ColPos q;
    q.bx = blockIdx.x;
    q.jm = blockIdx.y;
    q.skip = false;
    if (io.remap && io.sel != 1) {
        const int nb = gridDim.x * gridDim.y;
        const int logical = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, nb);
        q.bx = logical % gridDim.x;
        q.jm = logical / gridDim.x;
    }
    if (io.sel == 1) {
        const int by = blockIdx.y;
        q.jm = by <= io.lo_m ? by : ny - (2 * io.lo_m + 1) + by;
    } else if (io.sel == 2) {
        q.skip = q.bx < io.lo_l && (q.jm <= io.lo_m || q.jm >= ny - io.lo_m);
    }
    return q;
}
