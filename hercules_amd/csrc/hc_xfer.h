// Launch constants and argument blocks of the row kernels and of the x / y
// transforms, shared by the kernels and the host code that launches them.
#pragma once

#include "hc_kernels.h"

// ---------------------------------------------------------------------------
// kernels
constexpr int kRhsThreads = 256;
constexpr int kRowPoints = 1024;  // points of one block of the row kernels

template <int NX>
constexpr int rows_per_block() { return NX >= kRowPoints ? 1 : kRowPoints / NX; }

// Two-phase momentum RHS + x R2C of qu,qv,qw, one point per thread.
// Staging: raw u, v, w, p rows of level k (Tile2).  Phase 1: every
// interpolant and flux product of the tile once into LDS.  Phase 2: the
// point's right-hand side from those rows (momentum_tile).  The staging area
// then holds the q rows for the x FFT.
template <int NX>
constexpr int rhs2_ty() { return NX == 16 ? 16 : NX == 32 ? 32 : NX == 48 ? 16 : NX == 64 ? 16 : NX == 96 ? 8 : NX == 128 ? 8 : 4; }
template <int NX>
constexpr int rhs2_threads() { return NX * rhs2_ty<NX>(); }
// waves per SIMD the register allocation must allow: with FP32 tiles two
// 1024-thread blocks fit in the LDS of one CU when each thread keeps to 64
// VGPRs
#ifndef HC_RHS2_F32_WAVES
#define HC_RHS2_F32_WAVES 8
#endif
template <int NX, class R>
constexpr int rhs2_waves() { return sizeof(R) == 4 ? HC_RHS2_F32_WAVES : 1; }

// Inverse x transform (C2R) of F spectra rows into physical rows.
// mode[f]: 0 store, 1 accumulate (dst += value), 2 accumulate onto the
// damped old value on levels k >= kdamp0 (dst = dst - s (dst - mean) + value).
template <class R>
struct C2RArgsT {
    const cplx<R> *src[4];
    R *dst[4];
    int mode[4];
    int ny, nz;
    const cplx<R> *tw_nl, *tw_nx;
    const R *dmp_s, *dmp_mean;
    int kdamp0;
};
using C2RArgs = C2RArgsT<double>;
// fields of one batched x R2C launch (blockIdx.z selects the field)
template <class R>
struct R2CFields {
    const R *src[4];
    cplx<R> *dst[4];
};
// points per field of one inverse-x block
constexpr int kC2RPoints = 1024;
template <int NX, int F, int PTS>
constexpr int c2r_rows() { return PTS / NX / F > 0 ? PTS / NX / F : 1; }


// y-direction FFT of up to 4 spectra, in place.  Block: TL consecutive l of
// level k, all j.  Forward zeroes the Nyquist row after the transform,
// inverse before it (src/module_tools.f90:538-572,659-693).
constexpr int kYTile = 8;
template <class R>
struct YArgsT {
    cplx<R> *f[4];
    int nl, nz;
    const cplx<R> *tw_ny;
};
using YArgs = YArgsT<double>;
// LDS rows of the tile are padded by one complex so that the transform's
// strided accesses spread over the banks.
constexpr int kYPad = kYTile + 1;
