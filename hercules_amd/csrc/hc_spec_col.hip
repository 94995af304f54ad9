// Column kernels of the spectral path: right-hand-side assembly, solves,
// projection and truncation per (l, m) column (k_column_spec, k_column_spec_t).
#include "hc_host.h"

#include "hc_spectral.h"

// Spectral column pass (kind 0: k_column_spec, 1: k_column_spec_t) over
// the spectra of V; with FP32 storage the low-wavenumber blocks are solved
// in FP64 first, as launch_columns does.
template <int L, class RIO, class RC>
static void spec_col_kernel(int kind, dim3 g, size_t lds, hipStream_t st, const StepArgs &A, const SpecIO<RIO> &io)
{
    const dim3 b(64 * kColWaves);
    if (kind == 0) hipLaunchKernelGGL((k_column_spec<L, RIO, RC>), g, b, lds, st, A, io);
    else hipLaunchKernelGGL((k_column_spec_t<L, RIO, RC>), g, b, lds, st, A, io);
}
template <class R>
int launch_spec_columns(hc_ctx *c, int kind, StepArgs A, const View<R> &V)
{
    SpecIO<R> io;
    io.cu = V.cu; io.cv = V.cv; io.cw = V.cw; io.cp = V.cp; io.ct = V.ct;
    io.cff = V.cff; io.cgg = V.cgg; io.chh = V.chh; io.css = V.css;
    for (int q = 0; q < 8; ++q) io.sp[q] = V.sp[q];
    io.lo_l = c->lo_l; io.lo_m = c->lo_m; io.remap = c->tune.col_remap;
    if (A.zorder == 6) {  // 6th/4th-order z operators: every column in FP64 arithmetic
        launch_spec_col6<R>(kind, c->stream, A, io);
        return 0;
    }
    const int nfields = kind == 0 ? 4 : 1;
    A.ntile = col_tiles(c, nfields, sizeof(cplx<R>));
    const size_t lds = col_lds(c, nfields, sizeof(cplx<R>));
    const dim3 gfull((c->nl + kColWaves - 1) / kColWaves, c->ny);
    if constexpr (std::is_same<R, double>::value) {
        io.sel = 0;
#define HC_CK(L_) spec_col_kernel<L_, double, double>(kind, gfull, lds, c->stream, A, io);
        HC_DISPATCH_L(c->Lcol, HC_CK)
#undef HC_CK
    } else {
        if (c->lo_l > 0) {
            io.sel = 1;
            const dim3 glo(c->lo_l, 2 * c->lo_m + 1);
#define HC_CK(L_) spec_col_kernel<L_, float, double>(kind, glo, lds, c->stream, A, io);
            HC_DISPATCH_L(c->Lcol, HC_CK)
#undef HC_CK
        }
        io.sel = c->lo_l > 0 ? 2 : 0;
#define HC_CK(L_) spec_col_kernel<L_, float, float>(kind, gfull, lds, c->stream, A, io);
        HC_DISPATCH_L(c->Lcol, HC_CK)
#undef HC_CK
    }
    return 0;
}

template int launch_spec_columns<double>(hc_ctx *, int, StepArgs, const View<double> &);
template int launch_spec_columns<float>(hc_ctx *, int, StepArgs, const View<float> &);

hipError_t spec_set_lds(int L, int bytes)
{
    hipError_t e = hipSuccess;
#define HC_A1(K_)                                                                                                   \
    if (e == hipSuccess)                                                                                            \
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(&K_), hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
#define HC_A(L_)                                \
    HC_A1((k_column_spec<L_, double, double>))  \
    HC_A1((k_column_spec<L_, float, float>))    \
    HC_A1((k_column_spec<L_, float, double>))
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
