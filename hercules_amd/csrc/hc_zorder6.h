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
// A carries no interpolant buffers, so the accessor evaluates them (in the
// stored type R).
template <class R>
__global__ __launch_bounds__(256) void k_zinterp6(StepArgsT<R> A, R *fu, R *fv, R *cw, R *ft)
{
    const size_t n = (size_t)A.nx * A.ny * A.nz, plane = (size_t)A.nx * A.ny;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(q / plane), j = (int)((q - (size_t)k * plane) / A.nx), i = (int)(q % A.nx);
        const PtT<R> g(A, i, j, k);
        fu[q] = g.uf6(k);
        fv[q] = g.vf6(k);
        cw[q] = g.whz6(k);
        if (ft) ft[q] = k < A.nz - 1 ? g.thz6(0) : R(0.0);
    }
}
