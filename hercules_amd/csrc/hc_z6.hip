// 6th-order z operators (zorder = 6): interpolants of the right-hand side,
// FD and spectral column kernels.
#include "hc_host.h"

#include "hc_spectral.h"
#include "hc_zorder6.h"

// 6th-order interpolants of the current fields (stored type R) into zbuf,
// and their addresses in the row-kernel arguments B
template <class R>
void zinterp6(hc_ctx *c, StepArgsT<R> &B)
{
    StepArgsT<R> I = B;
    I.zfu = I.zfv = I.zcw = I.zft = nullptr;  // the accessors evaluate the stencils
    R *b[4];
    for (int q = 0; q < 4; ++q) b[q] = reinterpret_cast<R *>(c->zbuf[q]);
    const dim3 grid((unsigned)(((size_t)B.nx * B.ny + 255) / 256), 4, (unsigned)stream_chunks(B.nz));
    hipLaunchKernelGGL((k_zinterp6<R>), grid, dim3(256), 0, c->stream, I, b[0], b[1], b[2], b[3]);
    B.zfu = b[0]; B.zfv = b[1]; B.zcw = b[2]; B.zft = b[3];
}
template void zinterp6<double>(hc_ctx *, StepArgsT<double> &);
template void zinterp6<float>(hc_ctx *, StepArgsT<float> &);

namespace {
// right-hand sides in0/in1 (null: in place), solutions f0/f1
template <class CI>
BandSys<CI> band_sys(const CI *in0, const CI *in1, CI *f0, CI *f1, double *scr, int eq, int fix, int trunc)
{
    BandSys<CI> b;
    b.in[0] = in0 ? in0 : f0;
    b.in[1] = in1 ? in1 : f1;
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
template <class CI>
void band_launch(hipStream_t st, const StepArgs &A, const BandArgs<CI> &B, int nsys)
{
    const int ncol = A.nl * A.ny;
    const dim3 g((unsigned)((ncol + 255) / 256), (unsigned)nsys);
    if (B.s[0].eq == 2) hipLaunchKernelGGL((k_z6_band<7, CI>), g, dim3(256), 0, st, A, B);
    else hipLaunchKernelGGL((k_z6_band<5, CI>), g, dim3(256), 0, st, A, B);
}

// the pressure correction phi: an FP64 half spectrum (A.zspc) in every
// variant, solved in place
void pressure_solve(hipStream_t st, const StepArgs &A)
{
    BandArgs<double2> B;
    B.s[0] = band_sys<double2>(nullptr, nullptr, A.zspc, nullptr, A.zscr, 2, 1, 0);
    band_launch(st, A, B, 1);
}

// U-row scratch of the second system group (the first starts at A.zscr):
// 6 doubles per level and column for half band 5
size_t band_off(const StepArgs &A) { return (size_t)6 * A.nz * A.nl * A.ny; }

// one thread per (l, m) column and z chunk
dim3 stream_grid(const StepArgs &A, int ny = 1)
{
    return dim3((unsigned)((A.nl * A.ny + 255) / 256), (unsigned)ny, (unsigned)stream_chunks(A.nz));
}

template <class R>
void col6_seq(int kind, hipStream_t st, const StepArgs &A, const ColIO<R> &io, int eq)
{
    using CI = cplx<R>;
    BandArgs<CI> B;
    switch (kind) {
    case 0:  // FD substep: u, v and w Helmholtz solves, divergence, pressure, projection and increment (into s3)
        B.s[0] = band_sys<CI>(nullptr, nullptr, io.s0, io.s1, A.zscr, 0, 1, 0);
        B.s[1] = band_sys<CI>(nullptr, nullptr, io.s2, nullptr, A.zscr + band_off(A), 1, 0, 0);
        band_launch(st, A, B, 2);
        hipLaunchKernelGGL((k_z6s_div<CI>), stream_grid(A), dim3(256), 0, st, A, io.s0, io.s1, io.s2, A.zspc, A.idt);
        pressure_solve(st, A);
        hipLaunchKernelGGL((k_z6s_proj<CI>), stream_grid(A), dim3(256), 0, st, A, io.s0, io.s1, io.s2, A.zspc, io.s3, 0);
        break;
    case 1:  // temperature
        B.s[0] = band_sys<CI>(nullptr, nullptr, io.s0, nullptr, A.zscr, 3, 1, 0);
        band_launch(st, A, B, 1);
        break;
    case 2:  // projection onto divergence-free fields (rk_c = 1)
        hipLaunchKernelGGL((k_z6s_div<CI>), stream_grid(A), dim3(256), 0, st, A, io.s0, io.s1, io.s2, A.zspc, 1.0);
        pressure_solve(st, A);
        hipLaunchKernelGGL((k_z6s_proj<CI>), stream_grid(A), dim3(256), 0, st, A, io.s0, io.s1, io.s2, A.zspc,
                           (CI *)nullptr, 1);
        break;
    default: {  // one equation: 0 u, 1 v, 2 w, 3 p, 4 t; wall data for u and t, delnull for p
        const int meq = eq == 0 || eq == 1 ? 0 : (eq == 2 ? 1 : (eq == 3 ? 2 : 3));
        B.s[0] = band_sys<CI>(nullptr, nullptr, io.s0, nullptr, A.zscr, meq, eq == 0 || eq == 3 || eq == 4, 0);
        band_launch(st, A, B, 1);
        break;
    }
    }
}

template <class R>
void spec6_seq(int kind, hipStream_t st, const StepArgs &A, const SpecIO<R> &io)
{
    using CI = cplx<R>;
    BandArgs<CI> B;
    if (kind == 1) {  // temperature: right-hand side, solve, 2/3 rule
        hipLaunchKernelGGL((k_z6s_rhs_t<CI>), stream_grid(A), dim3(256), 0, st, A, io);
        B.s[0] = band_sys<CI>(io.sp[2], nullptr, io.ct, nullptr, A.zscr, 3, 1, 1);
        band_launch(st, A, B, 1);
        return;
    }
    hipLaunchKernelGGL((k_z6s_rhs_uv<CI>), stream_grid(A, 2), dim3(256), 0, st, A, io);
    hipLaunchKernelGGL((k_z6s_rhs_w<CI>), stream_grid(A), dim3(256), 0, st, A, io);
    B.s[0] = band_sys<CI>(io.sp[3], io.sp[4], io.cu, io.cv, A.zscr, 0, 1, 0);
    B.s[1] = band_sys<CI>(io.sp[7], nullptr, io.cw, nullptr, A.zscr + band_off(A), 1, 0, 0);
    band_launch(st, A, B, 2);
    hipLaunchKernelGGL((k_z6s_div<CI>), stream_grid(A), dim3(256), 0, st, A, io.cu, io.cv, io.cw, A.zspc, A.idt);
    pressure_solve(st, A);
    hipLaunchKernelGGL((k_z6s_proj<CI>), stream_grid(A), dim3(256), 0, st, A, io.cu, io.cv, io.cw, A.zspc, io.cp, 2);
}
}  // namespace

template <class R>
void launch_col6(int kind, hipStream_t st, const StepArgs &A, const ColIO<R> &io, int eq)
{
    col6_seq<R>(kind, st, A, io, eq);
}
template <class R>
void launch_spec_col6(int kind, hipStream_t st, const StepArgs &A, const SpecIO<R> &io)
{
    spec6_seq<R>(kind, st, A, io);
}
template void launch_col6<double>(int, hipStream_t, const StepArgs &, const ColIO<double> &, int);
template void launch_col6<float>(int, hipStream_t, const StepArgs &, const ColIO<float> &, int);
template void launch_spec_col6<double>(int, hipStream_t, const StepArgs &, const SpecIO<double> &);
template void launch_spec_col6<float>(int, hipStream_t, const StepArgs &, const SpecIO<float> &);

// the 6th-order column kernels use no dynamic LDS
hipError_t z6_set_lds(int, int) { return hipSuccess; }
