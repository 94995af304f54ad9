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

dim3 stream_grid(const StepArgs &A, int ny = 1) { return dim3((unsigned)((A.nl * A.ny + 255) / 256), (unsigned)ny); }

void col6_seq(int kind, hipStream_t st, const StepArgs &A, const ColIO<double> &io, int eq)
{
    BandArgs B;
    switch (kind) {
    case 0:  // FD substep: u, v and w Helmholtz solves, divergence, pressure, projection and increment
        B.s[0] = band_sys(io.s0, io.s1, A.zscr, 0, 1, 0);
        B.s[1] = band_sys(io.s2, nullptr, A.zscr + band_off(A), 1, 0, 0);
        band_launch(st, A, B, 2);
        hipLaunchKernelGGL(k_z6s_div, stream_grid(A), dim3(256), 0, st, A, io.s0, io.s1, io.s2, io.s3, A.idt);
        B.s[0] = band_sys(io.s3, nullptr, A.zscr, 2, 1, 0);
        band_launch(st, A, B, 1);
        hipLaunchKernelGGL(k_z6s_proj, stream_grid(A), dim3(256), 0, st, A, io.s0, io.s1, io.s2, io.s3, nullptr, 0);
        break;
    case 1:  // temperature
        B.s[0] = band_sys(io.s0, nullptr, A.zscr, 3, 1, 0);
        band_launch(st, A, B, 1);
        break;
    case 2:  // projection onto divergence-free fields (rk_c = 1)
        hipLaunchKernelGGL(k_z6s_div, stream_grid(A), dim3(256), 0, st, A, io.s0, io.s1, io.s2, io.s3, 1.0);
        B.s[0] = band_sys(io.s3, nullptr, A.zscr, 2, 1, 0);
        band_launch(st, A, B, 1);
        hipLaunchKernelGGL(k_z6s_proj, stream_grid(A), dim3(256), 0, st, A, io.s0, io.s1, io.s2, io.s3, nullptr, 1);
        break;
    default: {  // one equation: 0 u, 1 v, 2 w, 3 p, 4 t; wall data for u and t, delnull for p
        const int meq = eq == 0 || eq == 1 ? 0 : (eq == 2 ? 1 : (eq == 3 ? 2 : 3));
        B.s[0] = band_sys(io.s0, nullptr, A.zscr, meq, eq == 0 || eq == 3 || eq == 4, 0);
        band_launch(st, A, B, 1);
        break;
    }
    }
}

void spec6_seq(int kind, hipStream_t st, const StepArgs &A, const SpecIO<double> &io)
{
    BandArgs B;
    if (kind == 1) {  // temperature: right-hand side, solve, 2/3 rule
        hipLaunchKernelGGL(k_z6s_rhs_t, stream_grid(A), dim3(256), 0, st, A, io);
        B.s[0] = band_sys(io.ct, nullptr, A.zscr, 3, 1, 1);
        band_launch(st, A, B, 1);
        return;
    }
    hipLaunchKernelGGL(k_z6s_rhs_uv, stream_grid(A, 2), dim3(256), 0, st, A, io);
    hipLaunchKernelGGL(k_z6s_rhs_w, stream_grid(A), dim3(256), 0, st, A, io);
    B.s[0] = band_sys(io.cu, io.cv, A.zscr, 0, 1, 0);
    B.s[1] = band_sys(io.cw, nullptr, A.zscr + band_off(A), 1, 0, 0);
    band_launch(st, A, B, 2);
    hipLaunchKernelGGL(k_z6s_div, stream_grid(A), dim3(256), 0, st, A, io.cu, io.cv, io.cw, A.zspc, A.idt);
    B.s[0] = band_sys(A.zspc, nullptr, A.zscr, 2, 1, 0);
    band_launch(st, A, B, 1);
    hipLaunchKernelGGL(k_z6s_proj, stream_grid(A), dim3(256), 0, st, A, io.cu, io.cv, io.cw, A.zspc, io.cp, 2);
}
}  // namespace

void launch_col6(int kind, int, dim3, size_t, hipStream_t st, const StepArgs &A, const ColIO<double> &io, int eq)
{
    col6_seq(kind, st, A, io, eq);
}

void launch_spec_col6(int kind, int, dim3, size_t, hipStream_t st, const StepArgs &A, const SpecIO<double> &io)
{
    spec6_seq(kind, st, A, io);
}

// the 6th-order column kernels use no dynamic LDS
hipError_t z6_set_lds(int, int) { return hipSuccess; }
