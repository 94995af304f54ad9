// 6th-order wall-normal operators (hc_params.zorder = 6), column side.
//
// The reference is 2nd order in z (src/module_cnvdiff.f90:89,121,
// src/module_poisson_solver.f90:32-122).  With zorder = 6 the implicit
// operators of a column are
//   u, v, t:  I + rc (-hxy - kappa Z)      Z = ZO_C2_* (7-point d2/dz2, centres)
//   w:        I + rc (-hxy - kappa Z)      Z = ZO_F2   (7-point d2/dz2, faces)
//   p:        rc (hxy + ZO_P)              ZO_P = ZO_D o ZO_G
// with hxy the horizontal symbol the 2nd-order rows use.  They are banded
// (half band 5, and 7 for ZO_P) and solved directly, one thread per column
// (k_z6_band, hc_band6.h).  A column pass of the step is a sequence of
// launches, each one thread per (l, m) column streaming along z:
// right-hand sides (spectral path), the Helmholtz solves, the divergence,
// the pressure solve, then the projection and the pressure increment.
// Arithmetic is FP64; the spectra are stored in FP64 or (mixed variant)
// FP32, the pressure correction phi always in FP64.
#pragma once

#include "hc_band6.h"

// The 6th-order interpolants of the substep's fields, once per point, for
// the right-hand-side kernels: u, v at the w face of level k (the top wall
// face included), w at the centre of level k, t at the interior w face k.
// One thread per (i, j) column, field (blockIdx.y: u, v, w, t) and z chunk
// (blockIdx.z of gridDim.z) streams along z with a window of kIntWin levels,
// k-kIntHW .. k+kIntHW around the row k being formed, and applies the band
// rows ZB_IU_* and ZB_IFC (block-uniform values).  The window holds the wall
// data of the C convention at levels -1 and nz (u: -u_mrf, and -u_mrf or
// du/dz = 0 at the top; v: 0; t: tbot, and ttop or the wall gradient) and
// zero beyond; w is zero outside its interior faces 0..nz-2.  The band rows
// list the stencil weights in the stencils' own order with zeros between
// them, so each value equals the accessor's sum term for term, in the stored
// type R.
constexpr int kIntWin = 2 * kIntHW + 1;
template <class R>
__global__ __launch_bounds__(256) void k_zinterp6(StepArgsT<R> A, R *fu, R *fv, R *cw, R *ft)
{
    constexpr int PF = 8;
    const int nz = A.nz, f = (int)blockIdx.y;
    const size_t plane = (size_t)A.nx * A.ny;
    const size_t col = blockIdx.x * (size_t)256 + threadIdx.x;
    if (col >= plane || (f == 3 && !ft)) return;
    const int csz = (nz + (int)gridDim.z - 1) / (int)gridDim.z;
    const int kb = (int)blockIdx.z * csz, ke = kb + csz < nz ? kb + csz : nz;
    const int ib = kb - kIntHW;
    const R zero = R(0.0);
    // the field's data, its stored levels lo .. nlev-1, wall data at levels
    // -1 and nz, band and output
    const R *src;
    R *dst, wb = zero, wt = zero;
    int nlev = nz, band = A.ochannel ? ZB_IU_DN : ZB_IU_DD;
    if (f == 0) {
        src = A.u; dst = fu;
        wb = R(-A.u_mrf);
        wt = A.ochannel == 0 ? R(-A.u_mrf) : zero;
    } else if (f == 1) {
        src = A.v; dst = fv;
    } else if (f == 2) {
        src = A.w; dst = cw; nlev = nz - 1; band = ZB_IFC;
    } else {
        src = A.t; dst = ft;
        wb = A.tbot;
        wt = A.tbc_top ? R(A.gdz_top * A.lev[LV_IDZT * nz + nz - 1]) : A.ttop;
        band = A.tbc_top ? ZB_IU_DN : ZB_IU_DD;
    }
    if (f == 2) wb = wt = zero;  // w: zero at both wall faces
    // level li (block-uniform: the selection branches on scalars)
    auto load = [&](int li) __attribute__((always_inline)) {
        return (li >= 0 && li < nlev) ? src[(size_t)li * plane + col] : (li == -1 ? wb : (li == nz ? wt : zero));
    };
    R win[kIntWin];
    R ring[PF];
#pragma unroll
    for (int t = 0; t < kIntWin; ++t) win[t] = zero;
#pragma unroll
    for (int u = 0; u < PF; ++u) ring[u] = load(ib + u);
    // iteration i: level i enters the window; row k = i - kIntHW is formed
    for (int i0 = ib; i0 < ke + kIntHW; i0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int i = i0 + u;
#pragma unroll
            for (int t = 0; t < kIntWin - 1; ++t) win[t] = win[t + 1];
            win[kIntWin - 1] = ring[u];
            ring[u] = load(i + PF);
            const int k = i - kIntHW;
            if (k >= kb && k < ke) {
                R acc = zero;
                if (f != 3 || k < nz - 1) {
                    const double *b = A.zband + ((size_t)band * nz + k) * kBandRow + kBandMax - kIntHW;
#pragma unroll
                    for (int t = 0; t < kIntWin; ++t) acc += R(b[t]) * win[t];
                }
                dst[(size_t)k * plane + col] = acc;
            }
        }
    }
}
