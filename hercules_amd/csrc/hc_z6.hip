// 6th-order z operators (zorder = 6): interpolants of the right-hand side,
// FD and spectral column kernels.
#include "hc_host.h"

#include "hc_spectral.h"
#include "hc_zorder6.h"

// 6th-order interpolants of the current fields into zbuf, and their
// addresses in the row-kernel arguments B
void zinterp6(hc_ctx *c, const StepArgs &A, StepArgs &B)
{
    hipLaunchKernelGGL(k_zinterp6, dim3(2048), dim3(256), 0, c->stream, A, c->zbuf[0], c->zbuf[1], c->zbuf[2],
                       c->zbuf[3]);
    B.zfu = c->zbuf[0]; B.zfv = c->zbuf[1]; B.zcw = c->zbuf[2]; B.zft = c->zbuf[3];
}

namespace {
BandSys band_sys(double2 *f0, double2 *f1, double *scr, int eq, int fix, int trunc)
{
    BandSys b;
    b.f[0] = f0;
    b.f[1] = f1;
    b.scr = scr;
    b.nr = f1 ? 2 : 1;
    b.eq = eq;
    b.fix = fix;
    b.trunc = trunc;
    return b;
}

// Direct solves of nsys independent system groups, one thread per column;
// the pressure (half band 7) runs alone.
void band_launch(hipStream_t st, const StepArgs &A, const BandArgs &B, int nsys)
{
    const int ncol = A.nl * A.ny;
    const dim3 g((unsigned)((ncol + 255) / 256), (unsigned)nsys);
    if (B.s[0].eq == 2) hipLaunchKernelGGL(k_z6_band<7>, g, dim3(256), 0, st, A, B);
    else hipLaunchKernelGGL(k_z6_band<5>, g, dim3(256), 0, st, A, B);
}

// U-row scratch of the second system group (the first starts at A.zscr):
// 6 doubles per level and column for half band 5
size_t band_off(const StepArgs &A) { return (size_t)6 * A.nz * A.nl * A.ny; }

template <int L>
void col6_seq(int kind, dim3 g, size_t lds, hipStream_t st, const StepArgs &A, const ColIO<double> &io, int eq)
{
    const dim3 b(64 * kColWaves);
    BandArgs B;
    switch (kind) {
    case 0:  // FD substep: u, v and w Helmholtz solves, divergence, pressure, projection and increment
        B.s[0] = band_sys(io.s0, io.s1, A.zscr, 0, 1, 0);
        B.s[1] = band_sys(io.s2, nullptr, A.zscr + band_off(A), 1, 0, 0);
        band_launch(st, A, B, 2);
        hipLaunchKernelGGL(k_z6_div<L>, g, b, lds, st, A, io, io.s3, A.idt);
        B.s[0] = band_sys(io.s3, nullptr, A.zscr, 2, 1, 0);
        band_launch(st, A, B, 1);
        hipLaunchKernelGGL(k_z6_proj<L>, g, b, lds, st, A, io, io.s3, 0);
        break;
    case 1:  // temperature
        B.s[0] = band_sys(io.s0, nullptr, A.zscr, 3, 1, 0);
        band_launch(st, A, B, 1);
        break;
    case 2:  // projection onto divergence-free fields (rk_c = 1)
        hipLaunchKernelGGL(k_z6_div<L>, g, b, lds, st, A, io, io.s3, 1.0);
        B.s[0] = band_sys(io.s3, nullptr, A.zscr, 2, 1, 0);
        band_launch(st, A, B, 1);
        hipLaunchKernelGGL(k_z6_proj<L>, g, b, lds, st, A, io, io.s3, 1);
        break;
    default: {  // one equation: 0 u, 1 v, 2 w, 3 p, 4 t; wall data for u and t, delnull for p
        const int meq = eq == 0 || eq == 1 ? 0 : (eq == 2 ? 1 : (eq == 3 ? 2 : 3));
        B.s[0] = band_sys(io.s0, nullptr, A.zscr, meq, eq == 0 || eq == 3 || eq == 4, 0);
        band_launch(st, A, B, 1);
        break;
    }
    }
}

template <int L>
void spec6_seq(int kind, dim3 g, size_t lds, hipStream_t st, const StepArgs &A, const SpecIO<double> &io)
{
    const dim3 b(64 * kColWaves);
    BandArgs B;
    if (kind == 1) {  // temperature: right-hand side, solve, 2/3 rule
        hipLaunchKernelGGL(k_spec6_t_rhs<L>, g, b, lds, st, A, io);
        B.s[0] = band_sys(io.ct, nullptr, A.zscr, 3, 1, 1);
        band_launch(st, A, B, 1);
        return;
    }
    hipLaunchKernelGGL(k_spec6_rhs<L>, g, b, lds, st, A, io);
    B.s[0] = band_sys(io.cu, io.cv, A.zscr, 0, 1, 0);
    B.s[1] = band_sys(io.cw, nullptr, A.zscr + band_off(A), 1, 0, 0);
    band_launch(st, A, B, 2);
    ColIO<double> cio;
    cio.s0 = io.cu; cio.s1 = io.cv; cio.s2 = io.cw; cio.s3 = io.cp;
    cio.sel = io.sel; cio.lo_l = io.lo_l; cio.lo_m = io.lo_m; cio.remap = io.remap;
    hipLaunchKernelGGL(k_z6_div<L>, g, b, lds, st, A, cio, A.zspc, A.idt);
    B.s[0] = band_sys(A.zspc, nullptr, A.zscr, 2, 1, 0);
    band_launch(st, A, B, 1);
    hipLaunchKernelGGL(k_z6_proj<L>, g, b, lds, st, A, cio, A.zspc, 2);
}
}  // namespace

void launch_col6(int kind, int L, dim3 g, size_t lds, hipStream_t st, const StepArgs &A, const ColIO<double> &io, int eq)
{
#define HC_K(L_) case L_: col6_seq<L_>(kind, g, lds, st, A, io, eq); break;
    switch (L) { HC_K(2) HC_K(3) HC_K(4) HC_K(5) HC_K(8) HC_K(16) default: break; }
#undef HC_K
}

void launch_spec_col6(int kind, int L, dim3 g, size_t lds, hipStream_t st, const StepArgs &A, const SpecIO<double> &io)
{
#define HC_K(L_) case L_: spec6_seq<L_>(kind, g, lds, st, A, io); break;
    switch (L) { HC_K(2) HC_K(3) HC_K(4) HC_K(5) HC_K(8) HC_K(16) default: break; }
#undef HC_K
}

hipError_t z6_set_lds(int L, int bytes)
{
    hipError_t e = hipSuccess;
#define HC_A1(K_)                                                                                                   \
    if (e == hipSuccess)                                                                                            \
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(&K_), hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
#define HC_A(L_)                   \
    HC_A1((k_z6_div<L_>))          \
    HC_A1((k_z6_proj<L_>))         \
    HC_A1((k_spec6_rhs<L_>))       \
    HC_A1((k_spec6_t_rhs<L_>))
    switch (L) {
    case 2: HC_A(2) break;
    case 3: HC_A(3) break;
    case 4: HC_A(4) break;
    case 5: HC_A(5) break;
    case 8: HC_A(8) break;
    case 16: HC_A(16) break;
    default: return hipErrorInvalidValue;
    }
#undef HC_A
#undef HC_A1
    return e;
}
