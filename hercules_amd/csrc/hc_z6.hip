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

void launch_col6(int kind, int L, dim3 g, size_t lds, hipStream_t st, const StepArgs &A, const ColIO<double> &io, int eq)
{
    const dim3 b(64 * kColWaves);
#define HC_K(L_)                                                                                  \
    case L_:                                                                                      \
        switch (kind) {                                                                           \
        case 0: hipLaunchKernelGGL((k_column_mom6<L_>), g, b, lds, st, A, io); break;             \
        case 1: hipLaunchKernelGGL((k_column_t6<L_>), g, b, lds, st, A, io); break;               \
        case 2: hipLaunchKernelGGL((k_column_project6<L_>), g, b, lds, st, A, io); break;         \
        default: hipLaunchKernelGGL((k_column_solve6<L_>), g, b, lds, st, A, io, eq); break;      \
        }                                                                                         \
        break;
    switch (L) { HC_K(2) HC_K(3) HC_K(4) HC_K(5) HC_K(8) HC_K(16) default: break; }
#undef HC_K
}

void launch_spec_col6(int kind, int L, dim3 g, size_t lds, hipStream_t st, const StepArgs &A, const SpecIO<double> &io)
{
    const dim3 b(64 * kColWaves);
#define HC_K(L_)                                                                          \
    case L_:                                                                              \
        if (kind == 0) hipLaunchKernelGGL((k_column_spec6<L_>), g, b, lds, st, A, io);    \
        else hipLaunchKernelGGL((k_column_spec_t6<L_>), g, b, lds, st, A, io);            \
        break;
    switch (L) { HC_K(2) HC_K(3) HC_K(4) HC_K(5) HC_K(8) HC_K(16) default: break; }
#undef HC_K
}

hipError_t z6_set_lds(int L, int bytes)
{
    hipError_t e = hipSuccess;
#define HC_A1(K_)                                                                                                   \
    if (e == hipSuccess)                                                                                            \
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(&K_), hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
#define HC_A(L_)                       \
    HC_A1((k_column_mom6<L_>))         \
    HC_A1((k_column_t6<L_>))           \
    HC_A1((k_column_project6<L_>))     \
    HC_A1((k_column_solve6<L_>))       \
    HC_A1((k_column_spec6<L_>))        \
    HC_A1((k_column_spec_t6<L_>))
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
