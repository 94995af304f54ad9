// Temperature right-hand side of the FD substep (get_temperature_cnvdiff,
// src/module_cnvdiff.f90:381-473) fused with the x R2C of qt, and the
// interpolants the mean statistics of a step take.
#include "hc_host.h"

// Temperature RHS + x R2C of qt.  Acc: Pt, or PtD in a damped substep.
template <int CDS, int NX, class Acc, int ZO = 2>
__global__ __launch_bounds__(kRhsThreads) void k_rhs_t(StepArgsT<typename Acc::Real> A)
{
    using R = typename Acc::Real;
    using C = cplx<R>;
    constexpr int ROWS = rows_per_block<NX>();
    constexpr int NL = NX / 2;
    __shared__ C sq_[ROWS * NL];
    R *sqd = reinterpret_cast<R *>(sq_);
    const int k = A.klo + (int)blockIdx.y, j0 = blockIdx.x * ROWS, tid = threadIdx.x;
    for (int p = tid; p < ROWS * NX; p += kRhsThreads) {
        int jj = p / NX, i = p - jj * NX, j = j0 + jj;
        R qt = R(0.0);
        if (j < A.ny) {
            Acc q(A, i, j, k);
            R s;
            temperature_point<CDS, ZO>(q, s, qt, A.use_f0 ? A.ss0[q.center()] : R(0));
            A.ss[((size_t)k * A.ny + j) * A.nx + i] = s;
        }
        sqd[jj * NX + i] = qt;
    }
    __syncthreads();
    lds_fft<NL, ROWS, NL, 1, kRhsThreads, -1>(sq_, A.tw_nl, tid);
    for (int p = tid; p < ROWS * NL; p += kRhsThreads) {
        int jj = p / NL, l = p - jj * NL, j = j0 + jj;
        if (j < A.ny) A.s0[((size_t)k * A.ny + j) * NL + l] = r2c_post<NL>(sq_ + jj * NL, l, A.tw_nx, A.norm);
    }
}

// Tiled temperature RHS + x R2C of qt, one point per thread: full x rows
// j0..j0+TY-1 of level k; the t, u, v rows the stencils reach are staged in
// LDS once (all loads issued before the first use), then the same LDS holds
// the qt rows for the FFT.  Undamped substeps, nx <= 256.
template <int CDS, int NX, class R>
__global__ __launch_bounds__(rhs2_threads<NX>(), (rhs2_waves<NX, R>())) void k_rhs_t2(StepArgsT<R> A)
{
    using C = cplx<R>;
    constexpr int TY = rhs2_ty<NX>(), T = rhs2_threads<NX>(), NL = NX / 2;
    using Q = LdT<NX, TY, R>;
    static_assert(Q::NROWS >= TY, "FFT rows must fit in the staging area");
    __shared__ __align__(16) R sm[Q::NROWS * NX];
    __shared__ C tws[NX];  // twiddles of the x transform: half-length, then R2C post
    const int tid = threadIdx.x;
    C tw_a = czero_t<C>(), tw_b = czero_t<C>();
    if (tid < NL) { tw_a = A.tw_nl[tid]; tw_b = A.tw_nx[tid]; }
    const int nty = A.ny / TY, nz = A.nz;
    const int logical = xcd_remap(blockIdx.x, nty * A.nlev);
    const int jt = logical % nty, k = A.klo + logical / nty, j0 = jt * TY;
    const int i = tid % NX, jj = tid / NX, j = j0 + jj;
    constexpr int NST = (Q::NROWS * NX + T - 1) / T;
    R stv[NST];
#pragma unroll
    for (int m = 0; m < NST; ++m) {
        const int q = tid + m * T;
        stv[m] = R(0);
        if ((Q::NROWS * NX) % T == 0 || q < Q::NROWS * NX) {
            const int r = q / NX, x = q - r * NX;
            const R *f = r < Q::OU ? A.t : r < Q::OV ? A.u : A.v;
            const int jr = r < Q::OU ? j0 - 3 + r : r < Q::OV ? j0 + (r - Q::OU) : j0 - 2 + (r - Q::OV);
            stv[m] = f[((unsigned)k * (unsigned)A.ny + (unsigned)wrapi(jr, A.ny)) * (unsigned)NX + (unsigned)x];
        }
    }
    const size_t plane = (size_t)A.ny * NX, ck = (size_t)k * plane + (size_t)j * NX + i;
    const R tkm = k > 0 ? A.t[ck - plane] : R(0), tkp = k < nz - 1 ? A.t[ck + plane] : R(0);
    const R wk = k < nz - 1 ? A.w[ck] : R(0), wkm = k > 0 ? A.w[ck - plane] : R(0);
    const R s0 = A.use_f0 ? A.ss0[ck] : R(0);
#pragma unroll
    for (int m = 0; m < NST; ++m) {
        const int q = tid + m * T;
        if ((Q::NROWS * NX) % T == 0 || q < Q::NROWS * NX) sm[q] = stv[m];
    }
    if (tid < NL) { tws[tid] = tw_a; tws[NL + tid] = tw_b; }
    __syncthreads();
    R qt;
    {
        Q q(A, sm, i, jj, j, k);
        const R tc = sm[(Q::OT + jj + 3) * NX + i];
        q.t_km = k == 0 ? R(2.0) * A.tbot - tc : tkm;
        q.t_kp = k == nz - 1 ? (A.tbc_top ? tc + A.gdz_top : R(2.0) * A.ttop - tc) : tkp;
        q.w_k = wk;
        q.w_km = wkm;
        R s;
        temperature_point<CDS>(q, s, qt, s0);
        A.ss[ck] = s;
    }
    __syncthreads();  // staging no longer read: reuse it for the qt rows
    sm[jj * NX + i] = qt;
    __syncthreads();
    C *sq_ = reinterpret_cast<C *>(sm);
    lds_fft<NL, TY, NL, 1, T, -1>(sq_, tws, tid);
    for (int p = tid; p < TY * NL; p += T) {
        const int row = p / NL, l = p - row * NL;
        A.s0[((size_t)k * A.ny + j0 + row) * NL + l] = r2c_post<NL>(sq_ + row * NL, l, tws + NL, A.norm);
    }
}

// ---- calculate_mean_fields (src/module_tools.f90:362-404), FD, and
// calculate_mean_fields_spec (src/module_spectral.f90:1012-1052).
// The FD sums take u and v from the x / y interpolants uhx, vhy of the
// fields at the start of the step's last substep (get_interp_fields_uvwhxyz
// at src/module_navier_stokes.f90:274 is the last to set them before the
// sums at :358-361); Acc: PtT, or PtD in a damped substep.
template <int CDS, class Acc>
__global__ __launch_bounds__(256) void k_mean_interp(StepArgsT<typename Acc::Real> A, double *uhx, double *vhy)
{
    const size_t n = (size_t)A.nx * A.ny * A.nz, plane = (size_t)A.nx * A.ny;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(q / plane), j = (int)((q - (size_t)k * plane) / A.nx), i = (int)(q % A.nx);
        const Acc g(A, i, j, k);
        uhx[q] = (double)g.template uhx<CDS>(0, 0, 0);
        vhy[q] = (double)g.template vhy<CDS>(0, 0, 0);
    }
}


// The temperature right-hand side of one substep into B.ss and the x
// spectrum of qt (B.s0), with the updated velocities (t read damped).  In a
// damped substep the levels that read a damped t (kdamp0-1 and up) go
// through the accessor kernel, the levels below through the tiled kernel.
template <class R>
int rhs_temperature(hc_ctx *c, StepArgsT<R> &B)
{
    const hc_params &P = c->prm;
    const int rows = rows_for(c->nx);
    const bool damp = P.isdamp == 1, z6 = P.zorder == 4 || P.zorder == 6;
    const int ksplit = damp ? std::max(0, c->kdamp0 - 1) : c->nz;
    B.dmask = 16;  // t read damped, the updated velocities as they are
    if (z6) {
        B.dmask = 0;
        const dim3 grow((c->ny + rows - 1) / rows, c->nz);
#define HC_RTZ(NX_)                                                                                                  \
    if (P.cds == 1) hipLaunchKernelGGL((k_rhs_t<1, NX_, PtT<R>, 6>), grow, dim3(kRhsThreads), 0, c->stream, B);       \
    else hipLaunchKernelGGL((k_rhs_t<2, NX_, PtT<R>, 6>), grow, dim3(kRhsThreads), 0, c->stream, B);
        HC_DISPATCH_N(c->nx, HC_RTZ)
#undef HC_RTZ
        return 0;
    }
    if constexpr (std::is_same<R, double>::value) {
        if (damp) {
            B.klo = ksplit;
            B.nlev = c->nz - ksplit;
            const dim3 gd((c->ny + rows - 1) / rows, B.nlev);
#define HC_RT(NX_)                                                                                      \
    if (P.cds == 1) hipLaunchKernelGGL((k_rhs_t<1, NX_, PtD>), gd, dim3(kRhsThreads), 0, c->stream, B); \
    else hipLaunchKernelGGL((k_rhs_t<2, NX_, PtD>), gd, dim3(kRhsThreads), 0, c->stream, B);
            HC_DISPATCH_N(c->nx, HC_RT)
#undef HC_RT
        }
    }
    if (ksplit > 0) {
        B.klo = 0;
        B.nlev = ksplit;
        const dim3 grow((c->ny + rows - 1) / rows, ksplit);
#define HC_RT(NX_)                                                                                               \
    if constexpr (NX_ <= 256) {                                                                                  \
        dim3 g1((unsigned)(c->ny / rhs2_ty<NX_>() * ksplit));                                                    \
        constexpr int T_ = rhs2_threads<NX_>();                                                                  \
        if (P.cds == 1) hipLaunchKernelGGL((k_rhs_t2<1, NX_, R>), g1, dim3(T_), 0, c->stream, B);                \
        else hipLaunchKernelGGL((k_rhs_t2<2, NX_, R>), g1, dim3(T_), 0, c->stream, B);                           \
    } else {                                                                                                     \
        if (P.cds == 1) hipLaunchKernelGGL((k_rhs_t<1, NX_, PtT<R>>), grow, dim3(kRhsThreads), 0, c->stream, B); \
        else hipLaunchKernelGGL((k_rhs_t<2, NX_, PtT<R>>), grow, dim3(kRhsThreads), 0, c->stream, B);            \
    }
        HC_DISPATCH_N(c->nx, HC_RT)
#undef HC_RT
    }
    B.klo = 0;
    B.nlev = c->nz;
    return 0;
}

// The interpolants the statistics of this step take: the fields at the start
// of the last substep (damped as the stencils read them).
template <class R>
int mean_interp(hc_ctx *c, StepArgsT<R> &B)
{
    const hc_params &P = c->prm;
    const dim3 gm(1024), bm(256);
    if constexpr (std::is_same<R, double>::value) {
        if (P.isdamp == 1) {
            if (P.cds == 1) hipLaunchKernelGGL((k_mean_interp<1, PtD>), gm, bm, 0, c->stream, B, c->mean_uhx, c->mean_vhy);
            else hipLaunchKernelGGL((k_mean_interp<2, PtD>), gm, bm, 0, c->stream, B, c->mean_uhx, c->mean_vhy);
            return 0;
        }
    }
    if (P.cds == 1) hipLaunchKernelGGL((k_mean_interp<1, PtT<R>>), gm, bm, 0, c->stream, B, c->mean_uhx, c->mean_vhy);
    else hipLaunchKernelGGL((k_mean_interp<2, PtT<R>>), gm, bm, 0, c->stream, B, c->mean_uhx, c->mean_vhy);
    return 0;
}

template int rhs_temperature<double>(hc_ctx *, StepArgsT<double> &);
template int rhs_temperature<float>(hc_ctx *, StepArgsT<float> &);
template int mean_interp<double>(hc_ctx *, StepArgsT<double> &);
template int mean_interp<float>(hc_ctx *, StepArgsT<float> &);
