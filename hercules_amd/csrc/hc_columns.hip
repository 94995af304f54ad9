// Column kernels of the 2nd-order z operators: the fused momentum column
// pass, the temperature solve, the projection of initial states and the
// standalone poisson_solver_fft, FP64 or with FP32 storage.
#include "hc_host.h"

// One fused column pass of the substep: Helmholtz solves of u*, v*, w*, the
// divergence source, the pressure Poisson solve, the projection and the
// pressure increment.  RIO: stored precision of the spectra; RC: the
// arithmetic of the solves.
template <int L, class RIO, class RC>
__global__ __launch_bounds__(64 * kColWaves) void k_column_mom(StepArgs A, ColIO<RIO> io)
{
    using CI = cplx<RIO>;
    using CC = cplx<RC>;
    extern __shared__ double2 tile_raw[];
    CI *tile = reinterpret_cast<CI *>(tile_raw);
    const ColPos pos = col_pos(io, A.ny);
    if (pos.skip) return;  // block-uniform, before any barrier
    // nl is a multiple of kColWaves: every wavefront owns a column
    const int l = pos.bx * kColWaves + (threadIdx.x >> 6), jm = pos.jm;
    const ColMap m = col_map(A.nz);
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    const bool is00 = (l == 0 && jm == 0);
    const size_t plane = (size_t)ny * nl, col0 = (size_t)jm * nl + (size_t)pos.bx * kColWaves;
    CC duv[2][L], dw[1][L], dp[1][L];
    // all three inputs in one phase when the tile holds them (A.ntile sub-tiles)
    const size_t sub = (size_t)nz * kColWaves;
    if (A.ntile >= 3) {
        const CI *srcs[3] = {io.s0, io.s1, io.s2};
        col_fill_n<L, 3>(srcs, tile, nz, plane, col0);
        __syncthreads();
        col_take<L>(tile, m, duv[0]);
        col_take<L>(tile + sub, m, duv[1]);
        col_take<L>(tile + 2 * sub, m, dw[0]);
        __syncthreads();
    } else {
        col_load<L>(io.s0, tile, nz, plane, col0, m, duv[0]);
        col_load<L>(io.s1, tile, nz, plane, col0, m, duv[1]);
        col_load<L>(io.s2, tile, nz, plane, col0, m, dw[0]);
    }
    RC a[L], b[L], c[L];
    // u and v share the matrix (assign_abc(aa_u,...) for both, :291-299)
    assemble<L>(A, 0, l, jm, is00, m, a, b, c);
    if (is00) fixups<L>(A, 0, m, duv[0]);
    wave_tridiag<L, 2>(m.cnt, a, b, c, duv);
    // w (:302-304): kend3 row is trivial and its right-hand side is unused
#pragma unroll
    for (int r = 0; r < L; ++r)
        if (r < m.cnt && m.k0 + r == nz - 1) dw[0][r] = czero_t<CC>();
    assemble<L>(A, 1, l, jm, is00, m, a, b, c);
    wave_tridiag<L, 1>(m.cnt, a, b, c, dw);
    // divergence source /dt (get_p_eqn_src, :127-176) in spectral form
    col_divergence<L>(A, m, l, jm, duv[0], duv[1], dw[0], RC(A.idt), dp[0]);
    // pressure Poisson (:312-314)
    assemble<L>(A, 2, l, jm, is00, m, a, b, c);
    if (is00) fixups<L>(A, 2, m, dp[0]);
    wave_tridiag_refined<L, 1>(m.cnt, a, b, c, dp);
    // projection (calculate_uvw, :659-734) and pressure increment
    // (calculate_new_p, :739-808)
    const RC cdt = RC(A.rc * A.dt);
    col_project<L>(A, m, l, jm, cdt, dp[0], duv[0], duv[1], dw[0]);
    const RC slap = RC(A.symx[SY_L * nl + l] + A.symy[SY_L * ny + jm]);
    CC pprev = shfl_up_c(last_of<L>(dp[0], m.cnt), 1);
    CC pnext = shfl_dn_c(dp[0][0], 1);
    const RC cnu = RC(A.rc * 0.5 * A.dt * A.nu);
    CC dpn[L];
#pragma unroll
    for (int r = 0; r < L; ++r) {
        dpn[r] = czero_t<CC>();
        if (r < m.cnt) {
            int k = m.k0 + r;
            CC ph = dp[0][r];
            CC pm = r == 0 ? (m.lane == 0 ? ph : pprev) : dp[0][r > 0 ? r - 1 : 0];
            CC pn = col_next<L>(m, nz, dp[0], r, pnext);
            RC idzt = RC(A.lev[LV_IDZT * nz + k]), idzb = RC(A.lev[LV_IDZB * nz + k]), idz = RC(A.lev[LV_IDZ * nz + k]);
            CC lz;
            lz.x = ((pn.x - ph.x) * idzt - (ph.x - pm.x) * idzb) * idz;
            lz.y = ((pn.y - ph.y) * idzt - (ph.y - pm.y) * idzb) * idz;
            CC lap = mk2<CC>(slap * ph.x + lz.x, slap * ph.y + lz.y);
            dpn[r] = mk2<CC>(ph.x - cnu * lap.x, ph.y - cnu * lap.y);
        }
    }
    if (A.ntile >= 4) {
        col_put<L>(tile, m, duv[0]);
        col_put<L>(tile + sub, m, duv[1]);
        col_put<L>(tile + 2 * sub, m, dw[0]);
        col_put<L>(tile + 3 * sub, m, dpn);
        __syncthreads();
        col_drain(io.s0, tile, nz, plane, col0);
        col_drain(io.s1, tile + sub, nz, plane, col0);
        col_drain(io.s2, tile + 2 * sub, nz, plane, col0);
        col_drain(io.s3, tile + 3 * sub, nz, plane, col0);
    } else {
        col_store<L>(io.s0, tile, nz, plane, col0, m, duv[0]);
        col_store<L>(io.s1, tile, nz, plane, col0, m, duv[1]);
        col_store<L>(io.s2, tile, nz, plane, col0, m, dw[0]);
        col_store<L>(io.s3, tile, nz, plane, col0, m, dpn);
    }
}

// Temperature Helmholtz solve in place on s0 (:335-337).
template <int L, class RIO, class RC>
__global__ __launch_bounds__(64 * kColWaves) void k_column_t(StepArgs A, ColIO<RIO> io)
{
    using CI = cplx<RIO>;
    extern __shared__ double2 tile_raw[];
    CI *tile = reinterpret_cast<CI *>(tile_raw);
    const ColPos pos = col_pos(io, A.ny);
    if (pos.skip) return;
    const int l = pos.bx * kColWaves + (threadIdx.x >> 6), jm = pos.jm;
    const ColMap m = col_map(A.nz);
    const bool is00 = (l == 0 && jm == 0);
    const size_t plane = (size_t)A.ny * A.nl, col0 = (size_t)jm * A.nl + (size_t)pos.bx * kColWaves;
    cplx<RC> d[1][L];
    col_load<L>(io.s0, tile, A.nz, plane, col0, m, d[0]);
    RC a[L], b[L], c[L];
    assemble<L>(A, 3, l, jm, is00, m, a, b, c);
    if (is00) fixups<L>(A, 3, m, d[0]);
    wave_tridiag<L, 1>(m.cnt, a, b, c, d);
    col_store<L>(io.s0, tile, A.nz, plane, col0, m, d[0]);
}

// Projection of (u, v, w) spectra in s0, s1, s2 onto discretely
// divergence-free fields with the pressure operator at rk_c = 1:
// phi = L^{-1} div u, u -= G phi (the operators of get_p_eqn_src,
// poisson_solver_fft and calculate_uvw).  Used to prepare initial states.
template <int L, class RIO, class RC>
__global__ __launch_bounds__(64 * kColWaves) void k_column_project(StepArgs A, ColIO<RIO> io)
{
    using CI = cplx<RIO>;
    using CC = cplx<RC>;
    extern __shared__ double2 tile_raw[];
    CI *tile = reinterpret_cast<CI *>(tile_raw);
    const ColPos pos = col_pos(io, A.ny);
    if (pos.skip) return;
    const int l = pos.bx * kColWaves + (threadIdx.x >> 6), jm = pos.jm;
    const ColMap m = col_map(A.nz);
    const bool is00 = (l == 0 && jm == 0);
    const size_t plane = (size_t)A.ny * A.nl, col0 = (size_t)jm * A.nl + (size_t)pos.bx * kColWaves;
    CC u[1][L], v[1][L], w[1][L], ph[1][L];
    col_load<L>(io.s0, tile, A.nz, plane, col0, m, u[0]);
    col_load<L>(io.s1, tile, A.nz, plane, col0, m, v[0]);
    col_load<L>(io.s2, tile, A.nz, plane, col0, m, w[0]);
#pragma unroll
    for (int r = 0; r < L; ++r)
        if (r < m.cnt && m.k0 + r == A.nz - 1) w[0][r] = czero_t<CC>();
    col_divergence<L>(A, m, l, jm, u[0], v[0], w[0], RC(1), ph[0]);
    RC a[L], b[L], c[L];
    assemble<L>(A, 2, l, jm, is00, m, a, b, c);
    if (is00) fixups<L>(A, 2, m, ph[0]);
    wave_tridiag_refined<L, 1>(m.cnt, a, b, c, ph);
    col_project<L>(A, m, l, jm, RC(1), ph[0], u[0], v[0], w[0]);
    col_store<L>(io.s0, tile, A.nz, plane, col0, m, u[0]);
    col_store<L>(io.s1, tile, A.nz, plane, col0, m, v[0]);
    col_store<L>(io.s2, tile, A.nz, plane, col0, m, w[0]);
}

// Standalone poisson_solver_fft for one equation, in place on s0.
// eq: 0 u (ubnd), 1 v, 2 w (zstagger), 3 p (delnull), 4 t (tbnd).
template <int L, class RIO, class RC>
__global__ __launch_bounds__(64 * kColWaves) void k_column_solve(StepArgs A, ColIO<RIO> io, int eq)
{
    using CI = cplx<RIO>;
    extern __shared__ double2 tile_raw[];
    CI *tile = reinterpret_cast<CI *>(tile_raw);
    const ColPos pos = col_pos(io, A.ny);
    if (pos.skip) return;
    const int l = pos.bx * kColWaves + (threadIdx.x >> 6), jm = pos.jm;
    const ColMap m = col_map(A.nz);
    const bool is00 = (l == 0 && jm == 0);
    const size_t plane = (size_t)A.ny * A.nl, col0 = (size_t)jm * A.nl + (size_t)pos.bx * kColWaves;
    cplx<RC> d[1][L];
    col_load<L>(io.s0, tile, A.nz, plane, col0, m, d[0]);
#pragma unroll
    for (int r = 0; r < L; ++r)
        if (eq == 2 && r < m.cnt && m.k0 + r == A.nz - 1) d[0][r] = czero_t<cplx<RC>>();
    const int meq = eq == 0 || eq == 1 ? 0 : (eq == 2 ? 1 : (eq == 3 ? 2 : 3));
    RC a[L], b[L], c[L];
    assemble<L>(A, meq, l, jm, is00, m, a, b, c);
    if (is00 && eq != 1 && eq != 2) fixups<L>(A, meq, m, d[0]);
    if (meq == 2) wave_tridiag_refined<L, 1>(m.cnt, a, b, c, d);
    else wave_tridiag<L, 1>(m.cnt, a, b, c, d);
    col_store<L>(io.s0, tile, A.nz, plane, col0, m, d[0]);
}

// Column kernels: kind 0 k_column_mom, 1 k_column_t, 2 k_column_project,
// 3 k_column_solve (equation eq).
template <int L, class RIO, class RC>
static void col_kernel(int kind, dim3 g, size_t lds, hipStream_t st, const StepArgs &A, const ColIO<RIO> &io, int eq)
{
    const dim3 b(64 * kColWaves);
    switch (kind) {
    case 0: hipLaunchKernelGGL((k_column_mom<L, RIO, RC>), g, b, lds, st, A, io); break;
    case 1: hipLaunchKernelGGL((k_column_t<L, RIO, RC>), g, b, lds, st, A, io); break;
    case 2: hipLaunchKernelGGL((k_column_project<L, RIO, RC>), g, b, lds, st, A, io); break;
    default: hipLaunchKernelGGL((k_column_solve<L, RIO, RC>), g, b, lds, st, A, io, eq); break;
    }
}
// One column pass over spectra s[0..3] of V.  FP64 storage: every block in
// FP64.  Mixed: the low-wavenumber blocks in FP64, then the rest in FP32.
template <class R>
int launch_columns(hc_ctx *c, int kind, StepArgs A, const View<R> &V, int nfields, int eq)
{
    ColIO<R> io;
    io.s0 = V.s[0]; io.s1 = V.s[1]; io.s2 = V.s[2]; io.s3 = V.s[3];
    io.lo_l = c->lo_l; io.lo_m = c->lo_m;
    io.remap = c->tune.col_remap;
    if (A.zorder == 6) {  // 6th/4th-order z operators: every column in FP64 arithmetic
        launch_col6<R>(kind, c->stream, A, io, eq);
        return 0;
    }
    A.ntile = col_tiles(c, nfields, sizeof(cplx<R>));
    const size_t lds = col_lds(c, nfields, sizeof(cplx<R>));
    const dim3 gfull((c->nl + kColWaves - 1) / kColWaves, c->ny);
    if constexpr (std::is_same<R, double>::value) {
        io.sel = 0;
#define HC_CK(L_) col_kernel<L_, double, double>(kind, gfull, lds, c->stream, A, io, eq);
        HC_DISPATCH_L(c->Lcol, HC_CK)
#undef HC_CK
    } else {
        if (c->lo_l > 0) {
            io.sel = 1;
            const dim3 glo(c->lo_l, 2 * c->lo_m + 1);
#define HC_CK(L_) col_kernel<L_, float, double>(kind, glo, lds, c->stream, A, io, eq);
            HC_DISPATCH_L(c->Lcol, HC_CK)
#undef HC_CK
        }
        io.sel = c->lo_l > 0 ? 2 : 0;
#define HC_CK(L_) col_kernel<L_, float, float>(kind, gfull, lds, c->stream, A, io, eq);
        HC_DISPATCH_L(c->Lcol, HC_CK)
#undef HC_CK
    }
    return 0;
}

template int launch_columns<double>(hc_ctx *, int, StepArgs, const View<double> &, int, int);
template int launch_columns<float>(hc_ctx *, int, StepArgs, const View<float> &, int, int);

hipError_t columns_set_lds(int L, int bytes)
{
    hipError_t e = hipSuccess;
#define HC_A1(K_)                                                                                                   \
    if (e == hipSuccess)                                                                                            \
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(&K_), hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
#define HC_A(L_)                               \
    HC_A1((k_column_mom<L_, double, double>))  \
    HC_A1((k_column_mom<L_, float, float>))    \
    HC_A1((k_column_mom<L_, float, double>))
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
