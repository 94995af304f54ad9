// HIP kernels of the HERCULES finite-difference time step (MI355X, gfx950).
//
// Data layout in HBM (all FP64):
//   physical fields   f[k][j][i]        i < nx fastest, k < nz, interior only;
//                                       z ghost values are derived on the fly
//                                       from the reference's wall rules and
//                                       x/y halos by periodic index wrap.
//   spectra           s[k][j][l]        l < nx/2 complex, j < ny (m = j or
//                                       j-ny, the Nyquist row j = ny/2 is 0);
//                                       the x Nyquist bin l = nx/2 is never
//                                       stored because the reference zeroes it.
//
// Transform normalisation follows mpi_2d_fft (src/module_tools.f90:518-700):
// forward divides by nx*ny, inverse is unnormalised.
#pragma once
#include "hc_device.h"

// per-level tables, each nz long, stacked
enum {
    LV_DZ = 0, LV_DZT, LV_DZB, LV_LT,
    LV_AU, LV_CU, LV_DU, LV_DU00,
    LV_AW, LV_CW, LV_DW, LV_DW00,
    LV_AP, LV_CP, LV_DP, LV_DP00,
    LV_AT, LV_CT, LV_DT, LV_DT00,
    LV_IDZ, LV_IDZT, LV_IDZB,  // 1/dz, 1/dz_t, 1/dz_b
    LV_LB,                     // l_b, bottom-face interpolation weight
    LV_COUNT
};
// per-wavenumber tables (x: nx/2 entries, y: ny entries), stacked
enum {
    SY_HU = 0,  // Helmholtz symbol with nu:    dt*nu/2 * (second-derivative symbol)
    SY_HT,      // Helmholtz symbol with alpha
    SY_PX,      // Poisson symbol: divergence o gradient
    SY_L,       // second-derivative symbol used by calculate_new_p
    SY_DRE, SY_DIM,  // divergence symbol (p point from u points)
    SY_GRE, SY_GIM,  // gradient symbol (u point from p points)
    SY_COUNT
};

// 6th-order z operators (zorder = 6).  Each row holds up to kZW weights of
// consecutive data entries starting at zst.  Data indices:
//   centre data with wall values ("C"): 0 bottom datum (at the bottom wall
//     face), 1..nz levels 0..nz-1, nz+1 top datum (the top wall value, or the
//     wall gradient for a Neumann top);
//   face data ("F"): 0..nz the faces z_f(0..nz), i.e. w levels -1..nz-1,
//     whose two wall entries are zero;
//   plain centres ("P"): 0..nz-1 the levels.
// Rows: levels k for centre targets, w levels kw = 0..nz-2 for face targets.
enum ZOp {
    ZO_C2_DD = 0,  // d2/dz2 at centres from C data, value at both walls (7 points)
    ZO_C2_DN,      // the same with a Neumann top (top datum = wall gradient)
    ZO_F2,         // d2/dz2 at interior faces from F data (7 points)
    ZO_G,          // d/dz at interior faces from P data (6 points)
    ZO_D,          // d/dz at centres from F data (6 points)
    ZO_I_DD,       // centre-to-face interpolation from C data (6 points); row nz-1: the top wall face
    ZO_I_DN,       // the same with a Neumann top (row nz-1 extrapolates to the wall)
    ZO_IFC,        // face-to-centre interpolation from F data (6 points)
    ZO_P,          // ZO_D o ZO_G: the pressure operator on P data (up to kZW points)
    ZO_COUNT
};
constexpr int kZW = 16;
// Band rows of the implicit 6th-order operators on their unknowns, for the
// direct solves: nz rows of kBandRow entries each, entry (k, k + d) at
// d + kBandMax.  ZB_F2 has nz - 1 rows (the interior faces).
constexpr int kBandMax = 7;
constexpr int kBandRow = 2 * kBandMax + 1;
// ZB_G: faces kw = 0..nz-2 from the levels; ZB_D: levels from the interior
// faces (w levels 0..nz-2, the wall faces being zero).  ZB_IU_DD, ZB_IU_DN
// (ZO_I_DD, ZO_I_DN: faces kw = 0..nz-1 from the levels, the wall data at
// levels -1 and nz kept) and ZB_IFC (ZO_IFC: levels from the faces) are the
// explicit interpolations, within kIntHW of the diagonal.
enum { ZB_C2_DD = 0, ZB_C2_DN, ZB_F2, ZB_P, ZB_G, ZB_D, ZB_IU_DD, ZB_IU_DN, ZB_IFC, ZB_COUNT };
constexpr int kIntHW = 5;


// Arguments of the step kernels.  R is the storage and arithmetic type of
// the fields, spectra and coefficients (double, or float in the mixed
// precision variant); the device scalars of the mean pressure gradient and
// bulk velocity stay FP64.
template <class R>
struct StepArgsT {
    using Real = R;
    using Cx = cplx<R>;
    // physical fields
    const R *u, *v, *w, *p, *t, *ff0, *gg0, *hh0, *ss0;
    R *ff, *gg, *hh, *ss;
    R *uo, *vo, *wo, *po, *to;  // outputs of the finish kernels
    // spectra
    Cx *s0, *s1, *s2, *s3;
    // tables
    const R *lev, *symx, *symy;
    const Cx *tw_nl, *tw_nx, *tw_ny;
    const double *scal;  // device scalars: [0] dpx_drive, [1] ubar
    int nx, ny, nz, nl;
    int ochannel, dp_opt, isscalar, use_f0;
    R dx, dy, dx2, dy2, dt, nu, alpha, got, fc, ug, vg, u_mrf, t_ref, tbot, ttop;
    R ra, rb, rc, dpx, norm;
    // reciprocal stencil factors: 1/(24 dx), 1/(12 dx^2), 1/dx, 1/dx^2 (same in y), 1/dt
    R c24x, c12x2, c24y, c12y2, idx, idx2, idy, idy2, idt;
    R aU0, cUe, aT0, cTe;  // unmodified wall coefficients for the (0,0)-mode fix-ups
    // top temperature condition: tbc_top 0 t = ttop, 1 dT/dz given; gdz_top =
    // dtdz_top * dz_t(kend3), the ghost's offset t(kend3+1) - t(kend3)
    int tbc_top;
    R gdz_top;
    // Rayleigh damping of the top levels k >= kdamp0: sponge strength per
    // level and plane means per (field u,v,w,p,t; level); dmask selects the
    // fields an accessor reads damped (bit f for field f).
    const R *dmp_s, *dmp_m;
    int kdamp0, dmask;
    // levels klo .. klo+nlev-1 of a right-hand-side launch: all levels, or in
    // a damped substep the levels below the sponge (the tiled kernels) and
    // the levels that read a damped value (the accessor kernels)
    int klo, nlev;
    // spectral path (cds = 3): state spectra, product spectra, y-inverted
    // derivative spectra of the skew-symmetric terms, and the 2/3-rule limits
    Cx *cu, *cv, *cw, *cp, *ct, *cff, *cgg, *chh, *css;
    Cx *sp[8];
    const Cx *sdv[6];
    R coe;
    int issk, lc, mc;
    const R *aux;  // a physical field read by k_div_partial<3> (the spectral divergence)
    int ntile;     // column kernels: number of nz x kColWaves sub-tiles in the LDS
    int stage;     // spectral column kernels: 1 = assemble the right-hand sides only (per-stage entries)
    int rgroup;    // RHS tile order: y tiles per group (0: the largest of 16, 8, 4, 1 dividing ny/TY)
    // 6th-order z operators (zorder = 6; also 4): per operator and target row,
    // the first data index of the stencil (zst) and kZW weights (zwt)
    int zorder;
    int zn1, zn2;  // stencil widths: first derivatives / interpolation, second derivatives
    const int *zst;
    const double *zwt;
    // 6th-order interpolants of the substep's fields, formed once by
    // k_zinterp6 (null: the accessors evaluate them): u, v at the w face of
    // each level (the top wall face included), w at the centres, t at the
    // interior w faces
    const R *zfu, *zfv, *zcw, *zft;
    // direct banded solves of the 6th-order systems: band tables of the
    // operators (ZB_*), scratch for the U rows, and a half spectrum of scratch
    const double *zband;
    double *zscr;
    cplx<double> *zspc;
    // spectral Rayleigh damping (cds = 3, FP64): the spectra of u, v, w, p, t
    // on levels kdamp0..nz-1 as they were before the damping, [field][level][l,m]
    // (null without damping)
    const cplx<double> *spund;
};
using StepArgs = StepArgsT<double>;

HC_DEV int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
template <class Re> HC_DEV Re sq(Re x) { return x * x; }

// ---------------------------------------------------------------------------
// Point accessors with the reference's ghost rules
// (src/module_boundary.f90:33-196) and periodic x/y wrap (comm_bound).
// Mid-point interpolants, src/module_tools.f90:410-512, with the ghost
// overrides of update_boundary_uvwhxyz / _thxyz (src/module_boundary.f90:202-355),
// written once over the raw field accessors U, V, W, T of the derived class.
template <class D>
struct Interp {
    HC_DEV const D &d() const { return static_cast<const D &>(*this); }
    template <int CDS> HC_DEV auto uhx(int di, int dj, int dk) const
    {
        const D &q = d();
        using Re = typename D::Real;
        if (CDS == 1) return Re(0.5) * (q.U(di, dj, dk) + q.U(di - 1, dj, dk));
        return (-q.U(di + 1, dj, dk) + Re(9.0) * q.U(di, dj, dk) + Re(9.0) * q.U(di - 1, dj, dk) - q.U(di - 2, dj, dk)) / Re(16.0);
    }
    template <int CDS> HC_DEV auto uhy(int di, int dj, int dk) const
    {
        const D &q = d();
        using Re = typename D::Real;
        if (CDS == 1) return Re(0.5) * (q.U(di, dj + 1, dk) + q.U(di, dj, dk));
        return (-q.U(di, dj + 2, dk) + Re(9.0) * q.U(di, dj + 1, dk) + Re(9.0) * q.U(di, dj, dk) - q.U(di, dj - 1, dk)) / Re(16.0);
    }
    template <int CDS> HC_DEV auto vhx(int di, int dj, int dk) const
    {
        const D &q = d();
        using Re = typename D::Real;
        if (CDS == 1) return Re(0.5) * (q.V(di + 1, dj, dk) + q.V(di, dj, dk));
        return (-q.V(di + 2, dj, dk) + Re(9.0) * q.V(di + 1, dj, dk) + Re(9.0) * q.V(di, dj, dk) - q.V(di - 1, dj, dk)) / Re(16.0);
    }
    template <int CDS> HC_DEV auto vhy(int di, int dj, int dk) const
    {
        const D &q = d();
        using Re = typename D::Real;
        if (CDS == 1) return Re(0.5) * (q.V(di, dj, dk) + q.V(di, dj - 1, dk));
        return (-q.V(di, dj + 1, dk) + Re(9.0) * q.V(di, dj, dk) + Re(9.0) * q.V(di, dj - 1, dk) - q.V(di, dj - 2, dk)) / Re(16.0);
    }
    template <int CDS> HC_DEV auto whx(int di, int dj, int dk) const
    {
        const D &q = d();
        using Re = typename D::Real;
        int kk = q.k + dk;
        if (kk < 0 || kk == q.A.nz - 1) return Re(0.0);
        if (CDS == 1) return Re(0.5) * (q.W(di + 1, dj, dk) + q.W(di, dj, dk));
        return (-q.W(di + 2, dj, dk) + Re(9.0) * q.W(di + 1, dj, dk) + Re(9.0) * q.W(di, dj, dk) - q.W(di - 1, dj, dk)) / Re(16.0);
    }
    template <int CDS> HC_DEV auto why(int di, int dj, int dk) const
    {
        const D &q = d();
        using Re = typename D::Real;
        int kk = q.k + dk;
        if (kk < 0 || kk == q.A.nz - 1) return Re(0.0);
        if (CDS == 1) return Re(0.5) * (q.W(di, dj + 1, dk) + q.W(di, dj, dk));
        return (-q.W(di, dj + 2, dk) + Re(9.0) * q.W(di, dj + 1, dk) + Re(9.0) * q.W(di, dj, dk) - q.W(di, dj - 1, dk)) / Re(16.0);
    }
    HC_DEV auto uhz(int di, int dj, int dk) const
    {
        const D &q = d();
        using Re = typename D::Real;
        int kk = q.k + dk;
        if (kk < 0) return -q.A.u_mrf;
        if (kk == q.A.nz - 1) return q.A.ochannel == 0 ? -q.A.u_mrf : q.U(di, dj, dk);
        Re lt = q.lev(LV_LT, kk);
        return q.U(di, dj, dk + 1) * lt + q.U(di, dj, dk) * (Re(1.0) - lt);
    }
    HC_DEV auto vhz(int di, int dj, int dk) const
    {
        const D &q = d();
        using Re = typename D::Real;
        int kk = q.k + dk;
        if (kk < 0) return Re(0.0);
        if (kk == q.A.nz - 1) return q.A.ochannel == 0 ? Re(0.0) : q.V(di, dj, dk);
        Re lt = q.lev(LV_LT, kk);
        return q.V(di, dj, dk + 1) * lt + q.V(di, dj, dk) * (Re(1.0) - lt);
    }
    template <int CDS> HC_DEV auto thx(int di, int dj, int dk) const
    {
        const D &q = d();
        using Re = typename D::Real;
        if (CDS == 1) return Re(0.5) * (q.T(di + 1, dj, dk) + q.T(di, dj, dk));
        return (-q.T(di + 2, dj, dk) + Re(9.0) * q.T(di + 1, dj, dk) + Re(9.0) * q.T(di, dj, dk) - q.T(di - 1, dj, dk)) / Re(16.0);
    }
    template <int CDS> HC_DEV auto thy(int di, int dj, int dk) const
    {
        const D &q = d();
        using Re = typename D::Real;
        if (CDS == 1) return Re(0.5) * (q.T(di, dj + 1, dk) + q.T(di, dj, dk));
        return (-q.T(di, dj + 2, dk) + Re(9.0) * q.T(di, dj + 1, dk) + Re(9.0) * q.T(di, dj, dk) - q.T(di, dj - 1, dk)) / Re(16.0);
    }
    HC_DEV auto thz(int dk) const
    {
        const D &q = d();
        using Re = typename D::Real;
        int kk = q.k + dk;
        if (kk < 0) return q.A.tbot;
        if (kk == q.A.nz - 1) return q.A.tbc_top ? q.T(0, 0, dk) + Re(0.5) * q.A.gdz_top : q.A.ttop;
        Re lt = q.lev(LV_LT, kk);
        return q.T(0, 0, dk + 1) * lt + q.T(0, 0, dk) * (Re(1.0) - lt);
    }

    // ---- 6th-order z (zorder = 6): stencils of the ZOp tables over the
    // data of u, v, t (with their wall values), w (with its zero wall
    // values) and p.  Only levels inside 0..nz-1 are read.
    // width of a stencil (wide: second derivatives); table entries beyond it are zero
    HC_DEV int zn(bool wide) const { return wide ? d().A.zn2 : d().A.zn1; }
    HC_DEV const double *zw(int op, int row, int &s) const
    {
        const D &q = d();
        const size_t e = (size_t)op * q.A.nz + row;
        s = q.A.zst[e];
        return q.A.zwt + e * kZW;
    }
    // centre data (C convention) of u, v, t at column offset (di, dj)
    HC_DEV auto udat(int di, int dj, int dd) const
    {
        const D &q = d();
        using Re = typename D::Real;
        if (dd == 0) return Re(-q.A.u_mrf);
        if (dd == q.A.nz + 1) return q.A.ochannel == 0 ? Re(-q.A.u_mrf) : Re(0.0);  // value, or du/dz = 0
        return q.U(di, dj, dd - 1 - q.k);
    }
    HC_DEV auto vdat(int di, int dj, int dd) const
    {
        const D &q = d();
        using Re = typename D::Real;
        if (dd == 0 || dd == q.A.nz + 1) return Re(0.0);
        return q.V(di, dj, dd - 1 - q.k);
    }
    HC_DEV auto tdat(int dd) const
    {
        const D &q = d();
        using Re = typename D::Real;
        if (dd == 0) return Re(q.A.tbot);
        if (dd == q.A.nz + 1)  // ttop, or dT/dz = gdz_top / dz_t(kend3)
            return q.A.tbc_top ? Re(q.A.gdz_top * q.lev(LV_IDZT, q.A.nz - 1)) : Re(q.A.ttop);
        return q.T(0, 0, dd - 1 - q.k);
    }
    // face data (F convention) of w
    HC_DEV auto wdat(int di, int dj, int dd) const
    {
        const D &q = d();
        using Re = typename D::Real;
        if (dd == 0 || dd == q.A.nz) return Re(0.0);
        return q.W(di, dj, dd - 1 - q.k);
    }
    // u, v, t at the face of w level k + dk (interior faces), 6 points
    HC_DEV auto uhz6(int di, int dj, int dk) const
    {
        const D &q = d();
        using Re = typename D::Real;
        const int kw = q.k + dk;
        if (kw < 0 || kw >= q.A.nz - 1) return q.uhz(di, dj, dk);  // wall faces: the wall rule
        if (q.A.zfu) return Re(q.A.zfu[q.at(di, dj, kw)]);
        int s;
        const double *w = zw(q.A.ochannel ? ZO_I_DN : ZO_I_DD, kw, s);
        Re acc = Re(0.0);
#pragma unroll
        for (int j = 0; j < 6; ++j)
            if (j < zn(false)) acc += Re(w[j]) * udat(di, dj, s + j);
        return acc;
    }
    HC_DEV auto vhz6(int di, int dj, int dk) const
    {
        const D &q = d();
        using Re = typename D::Real;
        const int kw = q.k + dk;
        if (kw < 0 || kw >= q.A.nz - 1) return q.vhz(di, dj, dk);
        if (q.A.zfv) return Re(q.A.zfv[q.at(di, dj, kw)]);
        int s;
        const double *w = zw(q.A.ochannel ? ZO_I_DN : ZO_I_DD, kw, s);
        Re acc = Re(0.0);
#pragma unroll
        for (int j = 0; j < 6; ++j)
            if (j < zn(false)) acc += Re(w[j]) * vdat(di, dj, s + j);
        return acc;
    }
    HC_DEV auto thz6(int dk) const
    {
        const D &q = d();
        using Re = typename D::Real;
        const int kw = q.k + dk;
        if (kw < 0 || kw >= q.A.nz - 1) return q.thz(dk);
        if (q.A.zft) return Re(q.A.zft[q.at(0, 0, kw)]);
        int s;
        const double *w = zw(q.A.tbc_top ? ZO_I_DN : ZO_I_DD, kw, s);
        Re acc = Re(0.0);
#pragma unroll
        for (int j = 0; j < 6; ++j)
            if (j < zn(false)) acc += Re(w[j]) * tdat(s + j);
        return acc;
    }
    // u, v at the face of w level kw = -1..nz-1 (both walls included), for
    // the advective z terms of the spectral path
    HC_DEV auto uf6(int kw) const
    {
        const D &q = d();
        using Re = typename D::Real;
        if (kw < 0) return Re(-q.A.u_mrf);
        if (q.A.zfu) return Re(q.A.zfu[q.at(0, 0, kw)]);
        int s;
        const double *w = zw(q.A.ochannel ? ZO_I_DN : ZO_I_DD, kw, s);
        Re acc = Re(0.0);
#pragma unroll
        for (int j = 0; j < 6; ++j)
            if (j < zn(false)) acc += Re(w[j]) * udat(0, 0, s + j);
        return acc;
    }
    HC_DEV auto vf6(int kw) const
    {
        const D &q = d();
        using Re = typename D::Real;
        if (kw < 0) return Re(0.0);
        if (q.A.zfv) return Re(q.A.zfv[q.at(0, 0, kw)]);
        int s;
        const double *w = zw(q.A.ochannel ? ZO_I_DN : ZO_I_DD, kw, s);
        Re acc = Re(0.0);
#pragma unroll
        for (int j = 0; j < 6; ++j)
            if (j < zn(false)) acc += Re(w[j]) * vdat(0, 0, s + j);
        return acc;
    }
    template <int ZO> HC_DEV auto uhzO(int di, int dj, int dk) const
    {
        // through the accessor, which may override the interpolant (PtD)
        if constexpr (ZO == 6) return d().uhz6(di, dj, dk);
        else return d().uhz(di, dj, dk);
    }
    template <int ZO> HC_DEV auto vhzO(int di, int dj, int dk) const
    {
        if constexpr (ZO == 6) return d().vhz6(di, dj, dk);
        else return d().vhz(di, dj, dk);
    }
    // w at centre level c (absolute), 6 points of face data
    HC_DEV auto whz6(int c) const
    {
        using Re = typename D::Real;
        if (d().A.zcw) return Re(d().A.zcw[d().at(0, 0, c)]);
        int s;
        const double *w = zw(ZO_IFC, c, s);
        Re acc = Re(0.0);
#pragma unroll
        for (int j = 0; j < 6; ++j)
            if (j < zn(false)) acc += Re(w[j]) * wdat(0, 0, s + j);
        return acc;
    }
    // d(u w)/dz, d(v w)/dz, d(t w)/dz at this centre level: the face fluxes
    // (6th-order interpolant times the x / y interpolated w) differentiated
    // by ZO_D; the wall fluxes vanish with w
    template <int CDS> HC_DEV auto uwz6() const
    {
        const D &q = d();
        using Re = typename D::Real;
        int s;
        const double *w = zw(ZO_D, q.k, s);
        Re acc = Re(0.0);
#pragma unroll 1
        for (int j = 0; j < 6; ++j) {
            if (j >= zn(false)) break;
            const int dd = s + j;
            if (dd == 0 || dd == q.A.nz) continue;
            const int dk = dd - 1 - q.k;
            acc += Re(w[j]) * uhz6(0, 0, dk) * q.template whx<CDS>(0, 0, dk);
        }
        return acc;
    }
    template <int CDS> HC_DEV auto vwz6() const
    {
        const D &q = d();
        using Re = typename D::Real;
        int s;
        const double *w = zw(ZO_D, q.k, s);
        Re acc = Re(0.0);
#pragma unroll 1
        for (int j = 0; j < 6; ++j) {
            if (j >= zn(false)) break;
            const int dd = s + j;
            if (dd == 0 || dd == q.A.nz) continue;
            const int dk = dd - 1 - q.k;
            acc += Re(w[j]) * vhz6(0, 0, dk) * q.template why<CDS>(0, 0, dk);
        }
        return acc;
    }
    HC_DEV auto twz6() const
    {
        const D &q = d();
        using Re = typename D::Real;
        int s;
        const double *w = zw(ZO_D, q.k, s);
        Re acc = Re(0.0);
#pragma unroll 1
        for (int j = 0; j < 6; ++j) {
            if (j >= zn(false)) break;
            const int dd = s + j;
            if (dd == 0 || dd == q.A.nz) continue;
            const int dk = dd - 1 - q.k;
            acc += Re(w[j]) * thz6(dk) * q.W(0, 0, dk);
        }
        return acc;
    }
    // second derivatives at this level: u, v, t at centres, w at its face
    HC_DEV auto uzz6() const
    {
        const D &q = d();
        using Re = typename D::Real;
        int s;
        const double *w = zw(q.A.ochannel ? ZO_C2_DN : ZO_C2_DD, q.k, s);
        Re acc = Re(0.0);
#pragma unroll
        for (int j = 0; j < 7; ++j)
            if (j < zn(true)) acc += Re(w[j]) * udat(0, 0, s + j);
        return acc;
    }
    HC_DEV auto vzz6() const
    {
        const D &q = d();
        using Re = typename D::Real;
        int s;
        const double *w = zw(q.A.ochannel ? ZO_C2_DN : ZO_C2_DD, q.k, s);
        Re acc = Re(0.0);
#pragma unroll
        for (int j = 0; j < 7; ++j)
            if (j < zn(true)) acc += Re(w[j]) * vdat(0, 0, s + j);
        return acc;
    }
    HC_DEV auto tzz6() const
    {
        const D &q = d();
        using Re = typename D::Real;
        int s;
        const double *w = zw(q.A.tbc_top ? ZO_C2_DN : ZO_C2_DD, q.k, s);
        Re acc = Re(0.0);
#pragma unroll
        for (int j = 0; j < 7; ++j)
            if (j < zn(true)) acc += Re(w[j]) * tdat(s + j);
        return acc;
    }
    HC_DEV auto wzz6() const
    {
        const D &q = d();
        using Re = typename D::Real;
        int s;
        const double *w = zw(ZO_F2, q.k, s);
        Re acc = Re(0.0);
#pragma unroll
        for (int j = 0; j < 7; ++j)
            if (j < zn(true)) acc += Re(w[j]) * wdat(0, 0, s + j);
        return acc;
    }
    // d(w^2)/dz at this face from w interpolated to the centres, and dp/dz
    HC_DEV auto wwz6() const
    {
        const D &q = d();
        using Re = typename D::Real;
        int s;
        const double *w = zw(ZO_G, q.k, s);
        Re acc = Re(0.0);
#pragma unroll 1
        for (int j = 0; j < 6; ++j)
            if (j < zn(false)) acc += Re(w[j]) * sq(whz6(s + j));
        return acc;
    }
    HC_DEV auto pz6() const
    {
        const D &q = d();
        using Re = typename D::Real;
        int s;
        const double *w = zw(ZO_G, q.k, s);
        Re acc = Re(0.0);
#pragma unroll
        for (int j = 0; j < 6; ++j)
            if (j < zn(false)) acc += Re(w[j]) * q.P(0, 0, s + j - q.k);
        return acc;
    }
};

// Global-memory accessor: any offset, ghost values by the wall rules.
template <class R>
struct PtT : Interp<PtT<R>> {
    using Real = R;
    const StepArgsT<R> &A;
    int i, j, k;
    HC_DEV PtT(const StepArgsT<R> &a, int i_, int j_, int k_) : A(a), i(i_), j(j_), k(k_) {}
    HC_DEV size_t at(int di, int dj, int kk) const
    {
        return ((size_t)kk * A.ny + wrapi(j + dj, A.ny)) * A.nx + wrapi(i + di, A.nx);
    }
    HC_DEV R U(int di, int dj, int dk) const
    {
        int kk = k + dk;
        if (kk < 0) return -A.u[at(di, dj, 0)] - R(2.0) * A.u_mrf;
        if (kk >= A.nz) {
            R top = A.u[at(di, dj, A.nz - 1)];
            return A.ochannel == 0 ? -top - R(2.0) * A.u_mrf : top;
        }
        return A.u[at(di, dj, kk)];
    }
    HC_DEV R V(int di, int dj, int dk) const
    {
        int kk = k + dk;
        if (kk < 0) return -A.v[at(di, dj, 0)];
        if (kk >= A.nz) {
            R top = A.v[at(di, dj, A.nz - 1)];
            return A.ochannel == 0 ? -top : top;
        }
        return A.v[at(di, dj, kk)];
    }
    HC_DEV R W(int di, int dj, int dk) const
    {
        int kk = k + dk;
        if (kk < 0 || kk == A.nz - 1) return R(0.0);
        if (kk >= A.nz) return -A.w[at(di, dj, A.nz - 2)];
        return A.w[at(di, dj, kk)];
    }
    HC_DEV R P(int di, int dj, int dk) const
    {
        int kk = k + dk;
        kk = kk < 0 ? 0 : (kk >= A.nz ? A.nz - 1 : kk);
        return A.p[at(di, dj, kk)];
    }
    HC_DEV R T(int di, int dj, int dk) const
    {
        int kk = k + dk;
        if (kk < 0) return R(2.0) * A.tbot - A.t[at(di, dj, 0)];
        if (kk >= A.nz) {
            const R te = A.t[at(di, dj, A.nz - 1)];
            return A.tbc_top ? te + A.gdz_top : R(2.0) * A.ttop - te;
        }
        return A.t[at(di, dj, kk)];
    }
    HC_DEV R lev(int tab, int kk) const { return A.lev[tab * A.nz + kk]; }
    HC_DEV size_t center() const { return at(0, 0, k); }
    HC_DEV R f0() const { return A.use_f0 ? A.ff0[center()] : R(0.0); }
    HC_DEV R g0() const { return A.use_f0 ? A.gg0[center()] : R(0.0); }
    HC_DEV R h0() const { return A.use_f0 ? A.hh0[center()] : R(0.0); }
};
using Pt = PtT<double>;

// Global-memory accessor of a substep that starts with rayleigh_damping
// (src/module_tools.f90:705-749, called at src/module_navier_stokes.f90:258-266).
// The reference damps the interior of the top levels in place and refreshes
// no ghost before the stencils run, so a read sees
//   interior (i,j) of a damped level:  v - sigz*(v - plane mean)
//   x/y halo (wrapped index):          the undamped value
//   z ghost:                           the wall rule applied to undamped values.
// Interpolants are computed on the interior and copied to their halos after
// (comm_bound in update_boundary_uvwhxyz, src/module_boundary.f90:202-355), so
// an interpolant at a wrapped offset is evaluated at the wrapped centre.
struct PtD : Interp<PtD> {
    using Real = double;
    const StepArgs &A;
    int i, j, k;
    HC_DEV PtD(const StepArgs &a, int i_, int j_, int k_) : A(a), i(i_), j(j_), k(k_) {}
    HC_DEV double raw(const double *f, int x, int y, int kk) const
    {
        return f[((size_t)kk * A.ny + wrapi(y, A.ny)) * A.nx + wrapi(x, A.nx)];
    }
    HC_DEV double val(int fid, const double *f, int di, int dj, int kk) const
    {
        const int x = i + di, y = j + dj;
        double v = raw(f, x, y, kk);
        if (kk >= A.kdamp0 && ((A.dmask >> fid) & 1) && x >= 0 && x < A.nx && y >= 0 && y < A.ny) {
            const double s = A.dmp_s[kk], m = A.dmp_m[fid * A.nz + kk];
            v = v - s * (v - m);
        }
        return v;
    }
    HC_DEV double U(int di, int dj, int dk) const
    {
        int kk = k + dk;
        if (kk < 0) return -raw(A.u, i + di, j + dj, 0) - 2.0 * A.u_mrf;
        if (kk >= A.nz) {
            double top = raw(A.u, i + di, j + dj, A.nz - 1);
            return A.ochannel == 0 ? -top - 2.0 * A.u_mrf : top;
        }
        return val(0, A.u, di, dj, kk);
    }
    HC_DEV double V(int di, int dj, int dk) const
    {
        int kk = k + dk;
        if (kk < 0) return -raw(A.v, i + di, j + dj, 0);
        if (kk >= A.nz) {
            double top = raw(A.v, i + di, j + dj, A.nz - 1);
            return A.ochannel == 0 ? -top : top;
        }
        return val(1, A.v, di, dj, kk);
    }
    HC_DEV double W(int di, int dj, int dk) const
    {
        int kk = k + dk;
        if (kk < 0 || kk == A.nz - 1) return 0.0;
        if (kk >= A.nz) return -raw(A.w, i + di, j + dj, A.nz - 2);
        return val(2, A.w, di, dj, kk);
    }
    HC_DEV double P(int di, int dj, int dk) const
    {
        int kk = k + dk;
        if (kk < 0) return raw(A.p, i + di, j + dj, 0);
        if (kk >= A.nz) return raw(A.p, i + di, j + dj, A.nz - 1);
        return val(3, A.p, di, dj, kk);
    }
    HC_DEV double T(int di, int dj, int dk) const
    {
        int kk = k + dk;
        if (kk < 0) return 2.0 * A.tbot - raw(A.t, i + di, j + dj, 0);
        if (kk >= A.nz) {
            const double te = raw(A.t, i + di, j + dj, A.nz - 1);
            return A.tbc_top ? te + A.gdz_top : 2.0 * A.ttop - te;
        }
        return val(4, A.t, di, dj, kk);
    }
    HC_DEV PtD centre(int di, int dj) const { return PtD(A, wrapi(i + di, A.nx), wrapi(j + dj, A.ny), k); }
#define HC_PTD_INTERP(NAME)                                                                       \
    template <int CDS> HC_DEV double NAME(int di, int dj, int dk) const                           \
    {                                                                                             \
        const PtD c = centre(di, dj);                                                             \
        return c.Interp<PtD>::template NAME<CDS>(0, 0, dk);                                       \
    }
    HC_PTD_INTERP(uhx)
    HC_PTD_INTERP(uhy)
    HC_PTD_INTERP(vhx)
    HC_PTD_INTERP(vhy)
    HC_PTD_INTERP(whx)
    HC_PTD_INTERP(why)
    HC_PTD_INTERP(thx)
    HC_PTD_INTERP(thy)
#undef HC_PTD_INTERP
    HC_DEV double uhz(int di, int dj, int dk) const { return centre(di, dj).Interp<PtD>::uhz(0, 0, dk); }
    HC_DEV double vhz(int di, int dj, int dk) const { return centre(di, dj).Interp<PtD>::vhz(0, 0, dk); }
    HC_DEV double lev(int tab, int kk) const { return A.lev[tab * A.nz + kk]; }
    HC_DEV size_t center() const { return ((size_t)k * A.ny + j) * A.nx + i; }
    HC_DEV double f0() const { return A.use_f0 ? A.ff0[center()] : 0.0; }
    HC_DEV double g0() const { return A.use_f0 ? A.gg0[center()] : 0.0; }
    HC_DEV double h0() const { return A.use_f0 ? A.hh0[center()] : 0.0; }
};

// ---------------------------------------------------------------------------
// Momentum right-hand side at one point, get_momentum_cnvdiff
// (src/module_cnvdiff.f90:35-374).  Returns ff,gg,hh and qu,qv,qw.
template <int CDS, int ZO = 2, class Q>
HC_DEV void momentum_point(const Q &q, typename Q::Real dpx, typename Q::Real &f, typename Q::Real &g,
                           typename Q::Real &h, typename Q::Real &qu, typename Q::Real &qv, typename Q::Real &qw)
{
    using Re = typename Q::Real;
    const auto &A = q.A;
    const Re um = A.u_mrf;
    const int k = q.k;
    const Re idz = q.lev(LV_IDZ, k), idzt = q.lev(LV_IDZT, k), idzb = q.lev(LV_IDZB, k);
    // ---- U equation (:55-161)
    {
        Re uux, uvy, ux, uxx, uyy, dpdx;
        if (CDS == 1) {
            uux = (sq(q.template uhx<1>(1, 0, 0)) - sq(q.template uhx<1>(0, 0, 0))) * A.idx;
            uvy = (q.template uhy<1>(0, 0, 0) * q.template vhx<1>(0, 0, 0) - q.template uhy<1>(0, -1, 0) * q.template vhx<1>(0, -1, 0)) * A.idy;
            ux = (q.template uhx<1>(1, 0, 0) - q.template uhx<1>(0, 0, 0)) * A.idx;
            uxx = (q.U(1, 0, 0) - Re(2.0) * q.U(0, 0, 0) + q.U(-1, 0, 0)) * A.idx2;
            uyy = (q.U(0, 1, 0) - Re(2.0) * q.U(0, 0, 0) + q.U(0, -1, 0)) * A.idy2;
            dpdx = (q.P(1, 0, 0) - q.P(0, 0, 0)) * A.idx;
        } else {
            Re h2 = q.template uhx<2>(2, 0, 0), h1 = q.template uhx<2>(1, 0, 0), h0 = q.template uhx<2>(0, 0, 0), hm = q.template uhx<2>(-1, 0, 0);
            uux = (-sq(h2) + Re(27.0) * sq(h1) - Re(27.0) * sq(h0) + sq(hm)) * A.c24x;
            uvy = (-q.template uhy<2>(0, 1, 0) * q.template vhx<2>(0, 1, 0) + Re(27.0) * q.template uhy<2>(0, 0, 0) * q.template vhx<2>(0, 0, 0) -
                   Re(27.0) * q.template uhy<2>(0, -1, 0) * q.template vhx<2>(0, -1, 0) + q.template uhy<2>(0, -2, 0) * q.template vhx<2>(0, -2, 0)) * A.c24y;
            ux = (-h2 + Re(27.0) * h1 - Re(27.0) * h0 + hm) * A.c24x;
            uxx = (-q.U(2, 0, 0) + Re(16.0) * q.U(1, 0, 0) - Re(30.0) * q.U(0, 0, 0) + Re(16.0) * q.U(-1, 0, 0) - q.U(-2, 0, 0)) * A.c12x2;
            uyy = (-q.U(0, 2, 0) + Re(16.0) * q.U(0, 1, 0) - Re(30.0) * q.U(0, 0, 0) + Re(16.0) * q.U(0, -1, 0) - q.U(0, -2, 0)) * A.c12y2;
            dpdx = (-q.P(2, 0, 0) + Re(27.0) * q.P(1, 0, 0) - Re(27.0) * q.P(0, 0, 0) + q.P(-1, 0, 0)) * A.c24x;
        }
        Re u0 = q.U(0, 0, 0), uwz, uzz;
        if constexpr (ZO == 6) {
            uwz = q.template uwz6<CDS>();
            uzz = q.uzz6();
        } else {
            uwz = (q.uhz(0, 0, 0) * q.template whx<CDS>(0, 0, 0) - q.uhz(0, 0, -1) * q.template whx<CDS>(0, 0, -1)) * idz;
            uzz = ((q.U(0, 0, 1) - u0) * idzt - (u0 - q.U(0, 0, -1)) * idzb) * idz;
        }
        if (A.dp_opt == 3)
            f = -uux - uvy - uwz - um * ux + A.fc * (Re(0.5) * (q.template vhx<CDS>(0, 0, 0) + q.template vhx<CDS>(0, -1, 0)) - A.vg);
        else
            f = -uux - uvy - uwz + dpx - um * ux;
        Re f0 = q.f0();
        qu = u0 + A.dt * (A.ra * f + A.rb * f0) + A.rc * Re(0.5) * A.dt * A.nu * (uxx + uyy + uzz) - A.rc * A.dt * dpdx;
    }
    // ---- V equation (:166-270)
    {
        Re vux, vvy, vx, vxx, vyy, dpdy;
        if (CDS == 1) {
            vux = (q.template vhx<1>(0, 0, 0) * q.template uhy<1>(0, 0, 0) - q.template vhx<1>(-1, 0, 0) * q.template uhy<1>(-1, 0, 0)) * A.idx;
            vvy = (sq(q.template vhy<1>(0, 1, 0)) - sq(q.template vhy<1>(0, 0, 0))) * A.idy;
            vx = (q.template vhx<1>(0, 0, 0) - q.template vhx<1>(-1, 0, 0)) * A.idx;
            vxx = (q.V(1, 0, 0) - Re(2.0) * q.V(0, 0, 0) + q.V(-1, 0, 0)) * A.idx2;
            vyy = (q.V(0, 1, 0) - Re(2.0) * q.V(0, 0, 0) + q.V(0, -1, 0)) * A.idy2;
            dpdy = (q.P(0, 1, 0) - q.P(0, 0, 0)) * A.idy;
        } else {
            Re x1 = q.template vhx<2>(1, 0, 0), x0 = q.template vhx<2>(0, 0, 0), xm = q.template vhx<2>(-1, 0, 0), xm2 = q.template vhx<2>(-2, 0, 0);
            vux = (-x1 * q.template uhy<2>(1, 0, 0) + Re(27.0) * x0 * q.template uhy<2>(0, 0, 0) - Re(27.0) * xm * q.template uhy<2>(-1, 0, 0) +
                   xm2 * q.template uhy<2>(-2, 0, 0)) * A.c24x;
            vvy = (-sq(q.template vhy<2>(0, 2, 0)) + Re(27.0) * sq(q.template vhy<2>(0, 1, 0)) - Re(27.0) * sq(q.template vhy<2>(0, 0, 0)) +
                   sq(q.template vhy<2>(0, -1, 0))) * A.c24y;
            vx = (-x1 + Re(27.0) * x0 - Re(27.0) * xm + xm2) * A.c24x;
            vxx = (-q.V(2, 0, 0) + Re(16.0) * q.V(1, 0, 0) - Re(30.0) * q.V(0, 0, 0) + Re(16.0) * q.V(-1, 0, 0) - q.V(-2, 0, 0)) * A.c12x2;
            vyy = (-q.V(0, 2, 0) + Re(16.0) * q.V(0, 1, 0) - Re(30.0) * q.V(0, 0, 0) + Re(16.0) * q.V(0, -1, 0) - q.V(0, -2, 0)) * A.c12y2;
            dpdy = (-q.P(0, 2, 0) + Re(27.0) * q.P(0, 1, 0) - Re(27.0) * q.P(0, 0, 0) + q.P(0, -1, 0)) * A.c24y;
        }
        Re v0 = q.V(0, 0, 0), vwz, vzz;
        if constexpr (ZO == 6) {
            vwz = q.template vwz6<CDS>();
            vzz = q.vzz6();
        } else {
            vwz = (q.vhz(0, 0, 0) * q.template why<CDS>(0, 0, 0) - q.vhz(0, 0, -1) * q.template why<CDS>(0, 0, -1)) * idz;
            vzz = ((q.V(0, 0, 1) - v0) * idzt - (v0 - q.V(0, 0, -1)) * idzb) * idz;
        }
        if (A.dp_opt == 3)
            g = -vux - vvy - vwz - um * vx + A.fc * (A.ug - Re(0.5) * (q.template uhy<CDS>(0, 0, 0) + q.template uhy<CDS>(-1, 0, 0)) - um);
        else
            g = -vux - vvy - vwz - um * vx;
        Re g0 = q.g0();
        qv = v0 + A.dt * (A.ra * g + A.rb * g0) + A.rc * Re(0.5) * A.dt * A.nu * (vxx + vyy + vzz) - A.rc * A.dt * dpdy;
    }
    // ---- W equation (:274-370), k < kend3
    if (k < A.nz - 1) {
        Re wux, wvy, wx, wxx, wyy;
        if (CDS == 1) {
            wux = (q.template whx<1>(0, 0, 0) * q.template uhzO<ZO>(0, 0, 0) - q.template whx<1>(-1, 0, 0) * q.template uhzO<ZO>(-1, 0, 0)) * A.idx;
            wvy = (q.template why<1>(0, 0, 0) * q.template vhzO<ZO>(0, 0, 0) - q.template why<1>(0, -1, 0) * q.template vhzO<ZO>(0, -1, 0)) * A.idy;
            wx = (q.template whx<1>(0, 0, 0) - q.template whx<1>(-1, 0, 0)) * A.idx;
            wxx = (q.W(1, 0, 0) - Re(2.0) * q.W(0, 0, 0) + q.W(-1, 0, 0)) * A.idx2;
            wyy = (q.W(0, 1, 0) - Re(2.0) * q.W(0, 0, 0) + q.W(0, -1, 0)) * A.idy2;
        } else {
            Re x1 = q.template whx<2>(1, 0, 0), x0 = q.template whx<2>(0, 0, 0), xm = q.template whx<2>(-1, 0, 0), xm2 = q.template whx<2>(-2, 0, 0);
            wux = (-x1 * q.template uhzO<ZO>(1, 0, 0) + Re(27.0) * x0 * q.template uhzO<ZO>(0, 0, 0) - Re(27.0) * xm * q.template uhzO<ZO>(-1, 0, 0) + xm2 * q.template uhzO<ZO>(-2, 0, 0)) * A.c24x;
            wvy = (-q.template why<2>(0, 1, 0) * q.template vhzO<ZO>(0, 1, 0) + Re(27.0) * q.template why<2>(0, 0, 0) * q.template vhzO<ZO>(0, 0, 0) -
                   Re(27.0) * q.template why<2>(0, -1, 0) * q.template vhzO<ZO>(0, -1, 0) + q.template why<2>(0, -2, 0) * q.template vhzO<ZO>(0, -2, 0)) * A.c24y;
            wx = (-x1 + Re(27.0) * x0 - Re(27.0) * xm + xm2) * A.c24x;
            wxx = (-q.W(2, 0, 0) + Re(16.0) * q.W(1, 0, 0) - Re(30.0) * q.W(0, 0, 0) + Re(16.0) * q.W(-1, 0, 0) - q.W(-2, 0, 0)) * A.c12x2;
            wyy = (-q.W(0, 2, 0) + Re(16.0) * q.W(0, 1, 0) - Re(30.0) * q.W(0, 0, 0) + Re(16.0) * q.W(0, -1, 0) - q.W(0, -2, 0)) * A.c12y2;
        }
        Re w0 = q.W(0, 0, 0), wwz, wzz, dpdz, th;
        if constexpr (ZO == 6) {
            wwz = q.wwz6();
            wzz = q.wzz6();
            dpdz = q.pz6();
            th = A.isscalar ? q.thz6(0) : Re(0.0);
        } else {
            Re wp = q.W(0, 0, 1), wm = q.W(0, 0, -1);
            Re a1 = sq(Re(0.5) * (w0 + wp)), a2 = sq(Re(0.5) * (w0 + wm));
            wwz = (a1 - a2) * idzt;
            wzz = ((wp - w0) * q.lev(LV_IDZ, k + 1) - (w0 - wm) * idz) * idzt;
            dpdz = (q.P(0, 0, 1) - q.P(0, 0, 0)) * idzt;
            Re lt = q.lev(LV_LT, k);
            th = A.isscalar ? (Re(1.0) - lt) * q.T(0, 0, 0) + lt * q.T(0, 0, 1) : Re(0.0);
        }
        if (A.isscalar) {
            h = -wux - wvy - wwz - um * wx + A.got * (th - A.t_ref);
        } else {
            h = -wux - wvy - wwz - um * wx;
        }
        Re h0 = q.h0();
        qw = w0 + A.dt * (A.ra * h + A.rb * h0) + A.rc * Re(0.5) * A.dt * A.nu * (wxx + wyy + wzz) - A.rc * A.dt * dpdz;
    } else {
        h = Re(0.0);
        qw = Re(0.0);
    }
}

// Temperature right-hand side at one point, get_temperature_cnvdiff
// (src/module_cnvdiff.f90:381-473), evaluated with the updated u, v, w.
template <int CDS, int ZO = 2, class Q>
HC_DEV void temperature_point(const Q &q, typename Q::Real &s, typename Q::Real &qt, typename Q::Real s0)
{
    using Re = typename Q::Real;
    const auto &A = q.A;
    const int k = q.k;
    Re tux, tvy, tx, txx, tyy;
    if (CDS == 1) {
        tux = (q.template thx<1>(0, 0, 0) * q.U(0, 0, 0) - q.template thx<1>(-1, 0, 0) * q.U(-1, 0, 0)) * A.idx;
        tvy = (q.template thy<1>(0, 0, 0) * q.V(0, 0, 0) - q.template thy<1>(0, -1, 0) * q.V(0, -1, 0)) * A.idy;
        tx = (q.template thx<1>(0, 0, 0) - q.template thx<1>(-1, 0, 0)) * A.idx;
        txx = (q.T(1, 0, 0) - Re(2.0) * q.T(0, 0, 0) + q.T(-1, 0, 0)) * A.idx2;
        tyy = (q.T(0, 1, 0) - Re(2.0) * q.T(0, 0, 0) + q.T(0, -1, 0)) * A.idy2;
    } else {
        Re x1 = q.template thx<2>(1, 0, 0), x0 = q.template thx<2>(0, 0, 0), xm = q.template thx<2>(-1, 0, 0), xm2 = q.template thx<2>(-2, 0, 0);
        tux = (-x1 * q.U(1, 0, 0) + Re(27.0) * x0 * q.U(0, 0, 0) - Re(27.0) * xm * q.U(-1, 0, 0) + xm2 * q.U(-2, 0, 0)) * A.c24x;
        tvy = (-q.template thy<2>(0, 1, 0) * q.V(0, 1, 0) + Re(27.0) * q.template thy<2>(0, 0, 0) * q.V(0, 0, 0) -
               Re(27.0) * q.template thy<2>(0, -1, 0) * q.V(0, -1, 0) + q.template thy<2>(0, -2, 0) * q.V(0, -2, 0)) * A.c24y;
        tx = (-x1 + Re(27.0) * x0 - Re(27.0) * xm + xm2) * A.c24x;
        txx = (-q.T(2, 0, 0) + Re(16.0) * q.T(1, 0, 0) - Re(30.0) * q.T(0, 0, 0) + Re(16.0) * q.T(-1, 0, 0) - q.T(-2, 0, 0)) * A.c12x2;
        tyy = (-q.T(0, 2, 0) + Re(16.0) * q.T(0, 1, 0) - Re(30.0) * q.T(0, 0, 0) + Re(16.0) * q.T(0, -1, 0) - q.T(0, -2, 0)) * A.c12y2;
    }
    Re t0 = q.T(0, 0, 0), twz, tzz;
    if constexpr (ZO == 6) {
        twz = q.twz6();
        tzz = q.tzz6();
    } else {
        twz = (q.thz(0) * q.W(0, 0, 0) - q.thz(-1) * q.W(0, 0, -1)) * q.lev(LV_IDZ, k);
        tzz = ((q.T(0, 0, 1) - t0) * q.lev(LV_IDZT, k) - (t0 - q.T(0, 0, -1)) * q.lev(LV_IDZB, k)) * q.lev(LV_IDZ, k);
    }
    s = -tux - tvy - twz - A.u_mrf * tx;
    qt = t0 + A.dt * (A.ra * s + A.rb * s0) + A.rc * Re(0.5) * A.dt * A.alpha * (txx + tyy + tzz);
}

// ---------------------------------------------------------------------------
// Row transforms in LDS.  Rows of NX reals are packed as NX/2 complex.
// R2C post-processing for rows already transformed by an NX/2-point FFT:
// X(l) = (Z(l)+conj Z(NL-l))/2 + e^{-2 pi i l/NX} (Z(l)-conj Z(NL-l))/(2i)
template <int NL, class C>
HC_DEV C r2c_post(const C *row, int l, const C *__restrict__ tw_nx, real_of<C> norm)
{
    using Re = real_of<C>;
    const Re half = Re(0.5);
    C z = row[l], zr = cconj(row[(NL - l) % NL]);
    C e = mk2<C>(half * (z.x + zr.x), half * (z.y + zr.y));
    C dd = csub(z, zr);
    C o = mk2<C>(half * dd.y, -half * dd.x);  // (Z - conj Z')/(2i)
    C x = cadd(e, cmul(tw_nx[l], o));
    return cscale(x, norm);
}
// C2R pre-processing: Z(l) = A + i e^{+2 pi i l/NX} B, A = X(l) + X~, B = X(l) - X~,
// X~ = conj X(NL-l) (zero for l = 0, whose partner is the zeroed Nyquist bin).
template <class C>
HC_DEV C c2r_pre(C xl, C xt, C twl)
{
    C a = cadd(xl, xt), b = csub(xl, xt);
    C e = cmul(cconj(twl), b);          // e^{+i theta} B
    return mk2<C>(a.x - e.y, a.y + e.x);  // A + i e
}

// ---------------------------------------------------------------------------
// LDS tile of the momentum right-hand side: full x rows j0..j0+TY-1 of
// level k and the halo rows each stencil reaches, staged once per block.
template <int NX, int TY>
struct RhsTile {
    static constexpr int RU = TY + 6;   // u, v, w at k:   rows j0-3 .. j0+TY+2
    static constexpr int RP = TY + 3;   // p at k:         rows j0-1 .. j0+TY+1
    static constexpr int RWM = TY + 3;  // w at k-1:       rows j0-1 .. j0+TY+1
    static constexpr int RVP = TY + 3;  // v at k+1:       rows j0-2 .. j0+TY
    static constexpr int RUP = TY;      // u at k+1:       rows j0   .. j0+TY-1
    static constexpr int OU = 0, OV = RU, OW = 2 * RU, OP = 3 * RU, OWM = OP + RP, OVP = OWM + RWM,
                         OUP = OVP + RVP, NROWS = OUP + RUP;
};

template <int NX>
HC_DEV int wrapx(int i)
{
    if constexpr ((NX & (NX - 1)) == 0) return i & (NX - 1);
    else return i < 0 ? i + NX : (i >= NX ? i - NX : i);
}

// Value of staged row r (of RhsTile) at column i, with the wall rules of
// update_boundary_uvw / _p (src/module_boundary.f90:33-162) for the ghost
// levels k-1 = -1 and k+1 = nz.
template <int NX, int TY, class R>
HC_DEV R rhs_stage_value(const StepArgsT<R> &A, int r, int i, int j0, int k)
{
    using Tl = RhsTile<NX, TY>;
    const int ny = A.ny, nz = A.nz;
    auto at = [&](int jj, int kk) -> size_t { return ((size_t)kk * ny + wrapi(jj, ny)) * NX + i; };
    if (r < Tl::OV) return A.u[at(j0 - 3 + r, k)];
    if (r < Tl::OW) return A.v[at(j0 - 3 + (r - Tl::OV), k)];
    if (r < Tl::OP) return A.w[at(j0 - 3 + (r - Tl::OW), k)];
    if (r < Tl::OWM) return A.p[at(j0 - 1 + (r - Tl::OP), k)];
    if (r < Tl::OVP) return k == 0 ? R(0.0) : A.w[at(j0 - 1 + (r - Tl::OWM), k - 1)];
    if (r < Tl::OUP) {
        int jj = j0 - 2 + (r - Tl::OVP);
        if (k + 1 < nz) return A.v[at(jj, k + 1)];
        R top = A.v[at(jj, nz - 1)];
        return A.ochannel == 0 ? -top : top;
    }
    int jj = j0 + (r - Tl::OUP);
    if (k + 1 < nz) return A.u[at(jj, k + 1)];
    R top = A.u[at(jj, nz - 1)];
    return A.ochannel == 0 ? -top - R(2.0) * A.u_mrf : top;
}

// LDS accessor: the offsets momentum_point uses, read from the staged rows;
// the single-point values at k-1 / k+1 come from registers.
// Single-point values of one RHS point at levels k-1 / k+1 and the previous
// substep's right-hand sides, loaded ahead of the LDS staging barrier.
template <class R>
struct PointVals {
    R u_km, v_km, w_kp, p_kp, t_k, t_kp, f0, g0, h0;
    HC_DEV void load(const StepArgsT<R> &a, int i, int j, int k)
    {
        PtT<R> g(a, i, j, k);
        u_km = g.U(0, 0, -1);
        v_km = g.V(0, 0, -1);
        const bool in = k + 1 < a.nz;
        w_kp = in ? g.W(0, 0, 1) : R(0.0);
        p_kp = g.P(0, 0, 1);
        t_k = a.isscalar ? g.T(0, 0, 0) : R(0.0);
        t_kp = (a.isscalar && in) ? g.T(0, 0, 1) : R(0.0);
        f0 = g.f0();
        g0 = g.g0();
        h0 = g.h0();
    }
};

template <int NX, int TY, class R>
struct Ld : Interp<Ld<NX, TY, R>> {
    using Real = R;
    using Tl = RhsTile<NX, TY>;
    const StepArgsT<R> &A;
    const R *sm;
    int i, jj, j, k;
    R u_km, v_km, w_kp, p_kp, t_k, t_kp, f0_, g0_, h0_;
    HC_DEV Ld(const StepArgsT<R> &a, const R *s, int i_, int jj_, int j_, int k_, const PointVals<R> &pv)
        : A(a), sm(s), i(i_), jj(jj_), j(j_), k(k_), u_km(pv.u_km), v_km(pv.v_km), w_kp(pv.w_kp),
          p_kp(pv.p_kp), t_k(pv.t_k), t_kp(pv.t_kp), f0_(pv.f0), g0_(pv.g0), h0_(pv.h0) {}
    HC_DEV R f0() const { return f0_; }
    HC_DEV R g0() const { return g0_; }
    HC_DEV R h0() const { return h0_; }
    HC_DEV R row(int base, int r, int di) const { return sm[(base + r) * NX + wrapx<NX>(i + di)]; }
    HC_DEV R U(int di, int dj, int dk) const
    {
        if (dk == 0) return row(Tl::OU, jj + 3 + dj, di);
        if (dk > 0) return row(Tl::OUP, jj + dj, di);
        return u_km;
    }
    HC_DEV R V(int di, int dj, int dk) const
    {
        if (dk == 0) return row(Tl::OV, jj + 3 + dj, di);
        if (dk > 0) return row(Tl::OVP, jj + 2 + dj, di);
        return v_km;
    }
    HC_DEV R W(int di, int dj, int dk) const
    {
        if (dk == 0) return row(Tl::OW, jj + 3 + dj, di);
        if (dk < 0) return row(Tl::OWM, jj + 1 + dj, di);
        return w_kp;
    }
    HC_DEV R P(int di, int dj, int dk) const
    {
        if (dk == 0) return row(Tl::OP, jj + 1 + dj, di);
        return p_kp;
    }
    HC_DEV R T(int, int, int dk) const { return dk == 0 ? t_k : t_kp; }
    HC_DEV R lev(int tab, int kk) const { return A.lev[tab * A.nz + kk]; }
    HC_DEV size_t center() const { return ((size_t)k * A.ny + j) * NX + i; }
};

// Block id remap so that each of the 8 XCDs (which receive hardware block
// ids round-robin) works through one contiguous range of logical tiles and
// its L2 sees neighbouring rows and levels.  A permutation of [0, n).
HC_DEV int xcd_remap(int b, int n)
{
    constexpr int X = 8;
    int q = n / X, r = n % X, x = b % X, idx = b / X;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + idx;
}

// ---------------------------------------------------------------------------
// Two-phase momentum tile.  Level k, full x rows j0..j0+TY-1.  Phase 1
// evaluates every mid-point interpolant and convective flux product of the
// tile once into LDS rows; phase 2 forms the flux divergences, Laplacians and
// pressure gradients of get_momentum_cnvdiff (src/module_cnvdiff.f90:35-374)
// from those rows.  The interpolants and their wall overrides are the ones of
// get_interp_fields_uvwhxyz + update_boundary_uvwhxyz
// (src/module_tools.f90:410-470, src/module_boundary.f90:202-290).
template <int NX, int TY>
struct Tile2 {
    static constexpr int RU = TY + 6;  // raw u, v, w: rows j0-3 .. j0+TY+2
    static constexpr int RP = TY + 3;  // raw p:       rows j0-1 .. j0+TY+1
    static constexpr int OU = 0, OV = RU, OW = 2 * RU, OP = 3 * RU, NRAW = 3 * RU + RP;
    static constexpr int OUHX = NRAW;           // uhx          rows j0   .. j0+TY-1
    static constexpr int OWHX = OUHX + TY;      // whx          rows j0   .. j0+TY-1
    static constexpr int OFUW = OWHX + TY;      // uhz*whx      rows j0   .. j0+TY-1
    static constexpr int OVHX = OFUW + TY;      // vhx          rows j0-1 .. j0+TY-1
    static constexpr int OVHY = OVHX + TY + 1;  // vhy          rows j0-1 .. j0+TY+1
    static constexpr int OFUV = OVHY + TY + 3;  // uhy*vhx      rows j0-2 .. j0+TY
    static constexpr int OFWV = OFUV + TY + 3;  // why*vhz      rows j0-2 .. j0+TY
    static constexpr int NROWS = OFWV + TY + 3;
    static constexpr int NH1 = 2 * TY + 4;      // rows of the VHX + VHY block
};

// Raw row r of a Tile2 at column i (interior values of level k).
template <int NX, int TY, class R>
HC_DEV R tile2_raw(const StepArgsT<R> &A, int r, int i, int j0, int k)
{
    using Tl = Tile2<NX, TY>;
    auto at = [&](int jj) -> size_t { return ((size_t)k * A.ny + wrapi(jj, A.ny)) * NX + i; };
    if (r < Tl::OV) return A.u[at(j0 - 3 + r)];
    if (r < Tl::OW) return A.v[at(j0 - 3 + (r - Tl::OV))];
    if (r < Tl::OP) return A.w[at(j0 - 3 + (r - Tl::OW))];
    return A.p[at(j0 - 1 + (r - Tl::OP))];
}

// x interpolants on one LDS row: to the face between i-1 and i (uhx form)
// and between i and i+1 (vhx / whx form); y interpolants on a column of rows.
template <int CDS, int NX, class R>
HC_DEV R ixm(const R *row, int i)
{
    if (CDS == 1) return R(0.5) * (row[wrapx<NX>(i)] + row[wrapx<NX>(i - 1)]);
    return (-row[wrapx<NX>(i + 1)] + R(9.0) * row[wrapx<NX>(i)] + R(9.0) * row[wrapx<NX>(i - 1)] - row[wrapx<NX>(i - 2)]) / R(16.0);
}
template <int CDS, int NX, class R>
HC_DEV R ixp(const R *row, int i)
{
    if (CDS == 1) return R(0.5) * (row[wrapx<NX>(i + 1)] + row[wrapx<NX>(i)]);
    return (-row[wrapx<NX>(i + 2)] + R(9.0) * row[wrapx<NX>(i + 1)] + R(9.0) * row[wrapx<NX>(i)] - row[wrapx<NX>(i - 1)]) / R(16.0);
}
// The same on a row indexed directly (a tile that carries its x halos).
template <int CDS, class R>
HC_DEV R ixm_d(const R *row, int i)
{
    if (CDS == 1) return R(0.5) * (row[i] + row[i - 1]);
    return (-row[i + 1] + R(9.0) * row[i] + R(9.0) * row[i - 1] - row[i - 2]) / R(16.0);
}
template <int CDS, class R>
HC_DEV R ixp_d(const R *row, int i)
{
    if (CDS == 1) return R(0.5) * (row[i + 1] + row[i]);
    return (-row[i + 2] + R(9.0) * row[i + 1] + R(9.0) * row[i] - row[i - 1]) / R(16.0);
}
// col points at row r of the staged block, i at column i: value at row r+d is col[d*NX]
template <int CDS, int NX, class R>
HC_DEV R iym(const R *col)  // vhy form: between r-1 and r
{
    if (CDS == 1) return R(0.5) * (col[0] + col[-NX]);
    return (-col[NX] + R(9.0) * col[0] + R(9.0) * col[-NX] - col[-2 * NX]) / R(16.0);
}
template <int CDS, int NX, class R>
HC_DEV R iyp(const R *col)  // uhy / why form: between r and r+1
{
    if (CDS == 1) return R(0.5) * (col[NX] + col[0]);
    return (-col[2 * NX] + R(9.0) * col[NX] + R(9.0) * col[0] - col[-NX]) / R(16.0);
}

// Single-point values of one tile point, loaded ahead of the staging barrier:
// the k-1 / k+1 neighbours, the previous right-hand sides and w at level k-1
// where the k-1 interpolants whx, why need it.
template <class R>
struct PointVals2 {
    R u_km, v_km, u_kp, v_kp, w_kp, p_kp, t_k, t_kp, f0, g0, h0;
    R wkx[4];  // w(k-1) at x-1, x, x+1, x+2 (row j)
    R wky[3];  // w(k-1) at rows j-1, j+1, j+2 (column x)
    // Every load is issued from a clamped level before any value is used;
    // the wall rules of update_boundary_uvw / _p / _t
    // (src/module_boundary.f90:33-196) are then applied by selects on the
    // block-uniform level k.
    HC_DEV void load(const StepArgsT<R> &a, int i, int j, int k, bool with_old = true)
    {
        const int nx = a.nx, ny = a.ny, nz = a.nz;
        const bool bot = k == 0, top = k == nz - 1;
        const int km = bot ? 0 : k - 1, kp = top ? nz - 1 : k + 1;
        auto at = [&](int kk, int jj, int ii) -> unsigned {
            return ((unsigned)kk * (unsigned)ny + (unsigned)wrapi(jj, ny)) * (unsigned)nx + (unsigned)wrapi(ii, nx);
        };
        const unsigned ck = at(k, j, i), ckm = at(km, j, i), ckp = at(kp, j, i);
        const R ukm = a.u[ckm], vkm = a.v[ckm], ukp = a.u[ckp], vkp = a.v[ckp];
        const R wkp = a.w[ckp], pkp = a.p[ckp];
        const R tk = a.isscalar ? a.t[ck] : R(0.0), tkp = a.isscalar ? a.t[ckp] : R(0.0);
        R o0 = R(0.0), o1 = R(0.0), o2 = R(0.0);
        if (with_old && a.use_f0) { o0 = a.ff0[ck]; o1 = a.gg0[ck]; o2 = a.hh0[ck]; }
        R wx[4], wy[3];
#pragma unroll
        for (int d = 0; d < 4; ++d) wx[d] = a.w[at(km, j, i + d - 1)];
        wy[0] = a.w[at(km, j - 1, i)];
        wy[1] = a.w[at(km, j + 1, i)];
        wy[2] = a.w[at(km, j + 2, i)];
        const R um = a.u_mrf;
        u_km = bot ? -ukm - R(2.0) * um : ukm;
        v_km = bot ? -vkm : vkm;
        u_kp = top ? (a.ochannel == 0 ? -ukp - R(2.0) * um : ukp) : ukp;
        v_kp = top ? (a.ochannel == 0 ? -vkp : vkp) : vkp;
        w_kp = (top || k + 1 == nz - 1) ? R(0.0) : wkp;
        p_kp = pkp;
        t_k = tk;
        t_kp = top ? R(0.0) : tkp;
        f0 = o0;
        g0 = o1;
        h0 = o2;
#pragma unroll
        for (int d = 0; d < 4; ++d) wkx[d] = bot ? R(0.0) : wx[d];
#pragma unroll
        for (int d = 0; d < 3; ++d) wky[d] = bot ? R(0.0) : wy[d];
    }
};

// Phase 2 of the two-phase tile at point (i, j0+jj): ff, gg, hh and
// qu, qv, qw.  Same terms and coefficients as momentum_point.  The LDS rows
// have NX entries; with WRAP they are whole periodic x rows and i is the
// global x index, otherwise they hold a tile with its x halos and i is the
// point's index within the row.
template <int CDS, int NX, int TY, class R, bool WRAP = true>
HC_DEV void momentum_tile(const StepArgsT<R> &A, const R *sm, int i, int jj, int k, const PointVals2<R> &pv, R dpx,
                          R &f, R &g, R &h, R &qu, R &qv, R &qw)
{
    using Tl = Tile2<NX, TY>;
    auto X = [](int x) { if constexpr (WRAP) return wrapx<NX>(x); else return x; };
    auto RW = [&](int base, int r, int di) -> R { return sm[(base + r) * NX + X(i + di)]; };
    auto U = [&](int di, int dj) { return RW(Tl::OU, jj + 3 + dj, di); };
    auto V = [&](int di, int dj) { return RW(Tl::OV, jj + 3 + dj, di); };
    auto W = [&](int di, int dj) { return RW(Tl::OW, jj + 3 + dj, di); };
    auto P = [&](int di, int dj) { return RW(Tl::OP, jj + 1 + dj, di); };
    auto UHX = [&](int di) { return RW(Tl::OUHX, jj, di); };
    auto WHX = [&](int di) { return RW(Tl::OWHX, jj, di); };
    auto FUW = [&](int di) { return RW(Tl::OFUW, jj, di); };
    auto VHX = [&](int di, int dj) { return RW(Tl::OVHX, jj + 1 + dj, di); };
    auto VHY = [&](int dj) { return RW(Tl::OVHY, jj + 1 + dj, 0); };
    auto FUV = [&](int di, int dj) { return RW(Tl::OFUV, jj + 2 + dj, di); };
    auto FWV = [&](int dj) { return RW(Tl::OFWV, jj + 2 + dj, 0); };
    const R um = A.u_mrf;
    const R idz = A.lev[LV_IDZ * A.nz + k], idzt = A.lev[LV_IDZT * A.nz + k], idzb = A.lev[LV_IDZB * A.nz + k];
    const R u0 = U(0, 0), v0 = V(0, 0), w0 = W(0, 0);
    // k-1 flux products at the point (uhz*whx, vhz*why), wall rules at k-1 = -1
    R fuw_km, fwv_km;
    if (k == 0) {
        fuw_km = -um * R(0.0);
        fwv_km = R(0.0) * R(0.0);
    } else {
        const R lt = A.lev[LV_LT * A.nz + k - 1];
        const R uhz = u0 * lt + pv.u_km * (R(1.0) - lt);
        const R vhz = v0 * lt + pv.v_km * (R(1.0) - lt);
        R whx, why;
        if (CDS == 1) {
            whx = R(0.5) * (pv.wkx[2] + pv.wkx[1]);
            why = R(0.5) * (pv.wky[1] + pv.wkx[1]);
        } else {
            whx = (-pv.wkx[3] + R(9.0) * pv.wkx[2] + R(9.0) * pv.wkx[1] - pv.wkx[0]) / R(16.0);
            why = (-pv.wky[2] + R(9.0) * pv.wky[1] + R(9.0) * pv.wkx[1] - pv.wky[0]) / R(16.0);
        }
        fuw_km = uhz * whx;
        fwv_km = vhz * why;
    }
    // ---- U equation (:55-161)
    {
        R uux, uvy, ux, uxx, uyy, dpdx;
        if (CDS == 1) {
            const R h1 = UHX(1), h0 = UHX(0);
            uux = (sq(h1) - sq(h0)) * A.idx;
            uvy = (FUV(0, 0) - FUV(0, -1)) * A.idy;
            ux = (h1 - h0) * A.idx;
            uxx = (U(1, 0) - R(2.0) * u0 + U(-1, 0)) * A.idx2;
            uyy = (U(0, 1) - R(2.0) * u0 + U(0, -1)) * A.idy2;
            dpdx = (P(1, 0) - P(0, 0)) * A.idx;
        } else {
            const R h2 = UHX(2), h1 = UHX(1), h0 = UHX(0), hm = UHX(-1);
            uux = (-sq(h2) + R(27.0) * sq(h1) - R(27.0) * sq(h0) + sq(hm)) * A.c24x;
            uvy = (-FUV(0, 1) + R(27.0) * FUV(0, 0) - R(27.0) * FUV(0, -1) + FUV(0, -2)) * A.c24y;
            ux = (-h2 + R(27.0) * h1 - R(27.0) * h0 + hm) * A.c24x;
            uxx = (-U(2, 0) + R(16.0) * U(1, 0) - R(30.0) * u0 + R(16.0) * U(-1, 0) - U(-2, 0)) * A.c12x2;
            uyy = (-U(0, 2) + R(16.0) * U(0, 1) - R(30.0) * u0 + R(16.0) * U(0, -1) - U(0, -2)) * A.c12y2;
            dpdx = (-P(2, 0) + R(27.0) * P(1, 0) - R(27.0) * P(0, 0) + P(-1, 0)) * A.c24x;
        }
        const R uwz = (FUW(0) - fuw_km) * idz;
        const R uzz = ((pv.u_kp - u0) * idzt - (u0 - pv.u_km) * idzb) * idz;
        if (A.dp_opt == 3)
            f = -uux - uvy - uwz - um * ux + A.fc * (R(0.5) * (VHX(0, 0) + VHX(0, -1)) - A.vg);
        else
            f = -uux - uvy - uwz + dpx - um * ux;
        qu = u0 + A.dt * (A.ra * f + A.rb * pv.f0) + A.rc * R(0.5) * A.dt * A.nu * (uxx + uyy + uzz) - A.rc * A.dt * dpdx;
    }
    // ---- V equation (:166-270)
    {
        R vux, vvy, vx, vxx, vyy, dpdy;
        if (CDS == 1) {
            vux = (FUV(0, 0) - FUV(-1, 0)) * A.idx;
            vvy = (sq(VHY(1)) - sq(VHY(0))) * A.idy;
            vx = (VHX(0, 0) - VHX(-1, 0)) * A.idx;
            vxx = (V(1, 0) - R(2.0) * v0 + V(-1, 0)) * A.idx2;
            vyy = (V(0, 1) - R(2.0) * v0 + V(0, -1)) * A.idy2;
            dpdy = (P(0, 1) - P(0, 0)) * A.idy;
        } else {
            vux = (-FUV(1, 0) + R(27.0) * FUV(0, 0) - R(27.0) * FUV(-1, 0) + FUV(-2, 0)) * A.c24x;
            vvy = (-sq(VHY(2)) + R(27.0) * sq(VHY(1)) - R(27.0) * sq(VHY(0)) + sq(VHY(-1))) * A.c24y;
            vx = (-VHX(1, 0) + R(27.0) * VHX(0, 0) - R(27.0) * VHX(-1, 0) + VHX(-2, 0)) * A.c24x;
            vxx = (-V(2, 0) + R(16.0) * V(1, 0) - R(30.0) * v0 + R(16.0) * V(-1, 0) - V(-2, 0)) * A.c12x2;
            vyy = (-V(0, 2) + R(16.0) * V(0, 1) - R(30.0) * v0 + R(16.0) * V(0, -1) - V(0, -2)) * A.c12y2;
            dpdy = (-P(0, 2) + R(27.0) * P(0, 1) - R(27.0) * P(0, 0) + P(0, -1)) * A.c24y;
        }
        const R vwz = (FWV(0) - fwv_km) * idz;
        const R vzz = ((pv.v_kp - v0) * idzt - (v0 - pv.v_km) * idzb) * idz;
        if (A.dp_opt == 3) {
            // uhy at (i, j) and (i-1, j), from the staged u rows
            const R *cu = sm + (Tl::OU + jj + 3) * NX;
            const R a0 = iyp<CDS, NX>(cu + X(i)), a1 = iyp<CDS, NX>(cu + X(i - 1));
            g = -vux - vvy - vwz - um * vx + A.fc * (A.ug - R(0.5) * (a0 + a1) - um);
        } else {
            g = -vux - vvy - vwz - um * vx;
        }
        qv = v0 + A.dt * (A.ra * g + A.rb * pv.g0) + A.rc * R(0.5) * A.dt * A.nu * (vxx + vyy + vzz) - A.rc * A.dt * dpdy;
    }
    // ---- W equation (:274-370), k < kend3
    if (k < A.nz - 1) {
        R wux, wvy, wx, wxx, wyy;
        if (CDS == 1) {
            wux = (FUW(0) - FUW(-1)) * A.idx;
            wvy = (FWV(0) - FWV(-1)) * A.idy;
            wx = (WHX(0) - WHX(-1)) * A.idx;
            wxx = (W(1, 0) - R(2.0) * w0 + W(-1, 0)) * A.idx2;
            wyy = (W(0, 1) - R(2.0) * w0 + W(0, -1)) * A.idy2;
        } else {
            wux = (-FUW(1) + R(27.0) * FUW(0) - R(27.0) * FUW(-1) + FUW(-2)) * A.c24x;
            wvy = (-FWV(1) + R(27.0) * FWV(0) - R(27.0) * FWV(-1) + FWV(-2)) * A.c24y;
            wx = (-WHX(1) + R(27.0) * WHX(0) - R(27.0) * WHX(-1) + WHX(-2)) * A.c24x;
            wxx = (-W(2, 0) + R(16.0) * W(1, 0) - R(30.0) * w0 + R(16.0) * W(-1, 0) - W(-2, 0)) * A.c12x2;
            wyy = (-W(0, 2) + R(16.0) * W(0, 1) - R(30.0) * w0 + R(16.0) * W(0, -1) - W(0, -2)) * A.c12y2;
        }
        const R wp = pv.w_kp, wm = pv.wkx[1];
        const R a1 = sq(R(0.5) * (w0 + wp)), a2 = sq(R(0.5) * (w0 + wm));
        const R wwz = (a1 - a2) * idzt;
        const R wzz = ((wp - w0) * A.lev[LV_IDZ * A.nz + k + 1] - (w0 - wm) * idz) * idzt;
        if (A.isscalar) {
            const R lt = A.lev[LV_LT * A.nz + k];
            h = -wux - wvy - wwz - um * wx + A.got * ((R(1.0) - lt) * pv.t_k + lt * pv.t_kp - A.t_ref);
        } else {
            h = -wux - wvy - wwz - um * wx;
        }
        const R dpdz = (pv.p_kp - P(0, 0)) * idzt;
        qw = w0 + A.dt * (A.ra * h + A.rb * pv.h0) + A.rc * R(0.5) * A.dt * A.nu * (wxx + wyy + wzz) - A.rc * A.dt * dpdz;
    } else {
        h = R(0.0);
        qw = R(0.0);
    }
}

// LDS accessor of the tiled temperature right-hand side: t rows j0-3 ..
// j0+TY+2, u rows j0 .. j0+TY-1 and v rows j0-2 .. j0+TY of level k, plus
// the point's t and w at k-1 / k+1 with the wall rules of update_boundary_t
// and update_boundary_uvw (src/module_boundary.f90:33-196) already applied.
// Offers the offsets temperature_point reads.
template <int NX, int TY, class R>
struct LdT : Interp<LdT<NX, TY, R>> {
    using Real = R;
    static constexpr int OT = 0, OU = TY + 6, OV = OU + TY, NROWS = OV + TY + 3;
    const StepArgsT<R> &A;
    const R *sm;
    int i, jj, j, k;
    R t_km, t_kp, w_k, w_km;
    HC_DEV LdT(const StepArgsT<R> &a, const R *s, int i_, int jj_, int j_, int k_)
        : A(a), sm(s), i(i_), jj(jj_), j(j_), k(k_) {}
    HC_DEV R row(int base, int r, int di) const { return sm[(base + r) * NX + wrapx<NX>(i + di)]; }
    HC_DEV R U(int di, int, int) const { return row(OU, jj, di); }
    HC_DEV R V(int di, int dj, int) const { return row(OV, jj + 2 + dj, di); }
    HC_DEV R W(int, int, int dk) const { return dk == 0 ? w_k : w_km; }
    HC_DEV R T(int di, int dj, int dk) const
    {
        if (dk == 0) return row(OT, jj + 3 + dj, di);
        return dk > 0 ? t_kp : t_km;
    }
    HC_DEV R lev(int tab, int kk) const { return A.lev[tab * A.nz + kk]; }
    HC_DEV size_t center() const { return ((size_t)k * A.ny + j) * NX + i; }
};

// ---------------------------------------------------------------------------
// Discrete divergence at interior point (i, j, k) as screen_div_error forms it
// (src/module_io.f90:103-128): x and y differences of order cds, the z
// difference (w(k) - w(k-1))/dz(k) with w = 0 below the first level; with
// zorder = 6 the z derivative is the face-to-centre operator ZO_D.  F holds
// the fields in their stored precision R, D the FP64 metrics; the spectral
// path (CDS = 3) reads the divergence formed beforehand from F.aux.
template <int CDS, class R>
HC_DEV double div_at(const StepArgsT<R> &F, const StepArgs &D, int i, int j, int k)
{
    PtT<R> P{F, i, j, k};
    if (CDS == 3) return (double)F.aux[P.at(0, 0, k)];
    auto U = [&](int di) -> double { return (double)P.U(di, 0, 0); };
    auto V = [&](int dj) -> double { return (double)P.V(0, dj, 0); };
    double dudx, dvdy;
    if (CDS == 1) {
        dudx = (U(0) - U(-1)) / D.dx;
        dvdy = (V(0) - V(-1)) / D.dy;
    } else {
        dudx = (-U(1) + 27.0 * U(0) - 27.0 * U(-1) + U(-2)) / 24.0 / D.dx;
        dvdy = (-V(1) + 27.0 * V(0) - 27.0 * V(-1) + V(-2)) / 24.0 / D.dy;
    }
    double dwdz;
    if (D.zorder == 6) {
        const size_t e = (size_t)ZO_D * D.nz + k;
        const int s0 = D.zst[e];
        dwdz = 0.0;
        for (int jz = 0; jz < D.zn1; ++jz) {
            const int dd = s0 + jz;
            if (dd > 0 && dd < D.nz) dwdz += D.zwt[e * kZW + jz] * (double)F.w[P.at(0, 0, dd - 1)];
        }
    } else {
        const double w0 = F.w[P.at(0, 0, k)];
        const double wm = k == 0 ? 0.0 : (double)F.w[P.at(0, 0, k - 1)];
        dwdz = (w0 - wm) / D.lev[LV_DZ * D.nz + k];
    }
    return dudx + dvdy + dwdz;
}
