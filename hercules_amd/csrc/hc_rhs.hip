// Momentum right-hand side of the FD substep (get_momentum_cnvdiff,
// src/module_cnvdiff.f90:35-374) fused with the x-direction R2C of qu, qv,
// qw: the two-phase LDS tile (nx <= 256), the staged-row kernel (nx = 512),
// the global-memory kernel of damped substeps and of the 6th-order z
// operators, and the Rayleigh damping means.
#include "hc_host.h"

// Momentum RHS + x R2C of qu,qv,qw.  Block: full rows j0..j0+TY-1 of level
// k, one point per thread; every row the stencils reach is staged in LDS
// once (coalesced), then the same LDS holds the q rows for the FFT.
template <int NX>
constexpr int rhs_ty()
{
    return NX == 16 ? 16 : NX == 32 ? 32 : NX == 64 ? 16 : NX == 128 ? 8 : NX == 192 ? 4 : NX == 256 ? 4 : 1;
}

template <int NX>
constexpr int rhs_threads() { return NX * rhs_ty<NX>() <= 512 ? NX * rhs_ty<NX>() : NX * rhs_ty<NX>() / 2; }

template <int CDS, int NX, class R>
__global__ __launch_bounds__(rhs_threads<NX>()) void k_rhs_mom(StepArgsT<R> A)
{
    using C = cplx<R>;
    constexpr int TY = rhs_ty<NX>();
    constexpr int T = rhs_threads<NX>(), PP = NX * TY / T, NL = NX / 2;
    using Tl = RhsTile<NX, TY>;
    static_assert(Tl::NROWS * NX >= 3 * TY * NX, "FFT rows must fit in the staging area");
    __shared__ __align__(16) R sm[Tl::NROWS * NX];
    const int tid = threadIdx.x;
    const int nty = A.ny / TY;
    const int logical = xcd_remap(blockIdx.x, nty * A.nlev);
    const int jt = logical % nty, k = A.klo + logical / nty, j0 = jt * TY;
    for (int q = tid; q < Tl::NROWS * NX; q += T) {
        int r = q / NX, i = q - r * NX;
        sm[q] = rhs_stage_value<NX, TY>(A, r, i, j0, k);
    }
    const R dpx = A.dp_opt == 2 ? R(A.scal[0]) : A.dpx;
    PointVals<R> pv[PP];
#pragma unroll
    for (int pp = 0; pp < PP; ++pp) {
        const int pt = tid + pp * T;
        pv[pp].load(A, pt % NX, j0 + pt / NX, k);
    }
    __syncthreads();
    R qs[PP][3];
#pragma unroll
    for (int pp = 0; pp < PP; ++pp) {
        const int pt = tid + pp * T, i = pt % NX, jj = pt / NX, j = j0 + jj;
        Ld<NX, TY, R> q(A, sm, i, jj, j, k, pv[pp]);
        R f, g, h;
        momentum_point<CDS>(q, dpx, f, g, h, qs[pp][0], qs[pp][1], qs[pp][2]);
        const size_t c = ((size_t)k * A.ny + j) * NX + i;
        A.ff[c] = f;
        A.gg[c] = g;
        A.hh[c] = h;
    }
    __syncthreads();  // staging no longer read: reuse it for the q rows
#pragma unroll
    for (int pp = 0; pp < PP; ++pp) {
        const int pt = tid + pp * T, i = pt % NX, jj = pt / NX;
        sm[(0 * TY + jj) * NX + i] = qs[pp][0];
        sm[(1 * TY + jj) * NX + i] = qs[pp][1];
        sm[(2 * TY + jj) * NX + i] = qs[pp][2];
    }
    __syncthreads();
    C *sq_ = reinterpret_cast<C *>(sm);
    lds_fft<NL, 3 * TY, NL, 1, T, -1>(sq_, A.tw_nl, tid);
    C *outs[3] = {A.s0, A.s1, A.s2};
    for (int p = tid; p < 3 * TY * NL; p += T) {
        int row = p / NL, l = p - row * NL, fi = row / TY, r = row - fi * TY;
        outs[fi][((size_t)k * A.ny + j0 + r) * NL + l] = r2c_post<NL>(sq_ + row * NL, l, A.tw_nx, A.norm);
    }
}


template <int CDS, int NX, class R>
__global__ __launch_bounds__(rhs2_threads<NX>(), (rhs2_waves<NX, R>())) void k_rhs_mom2(StepArgsT<R> A)
{
    using C = cplx<R>;
    constexpr int TY = rhs2_ty<NX>();
    constexpr int T = rhs2_threads<NX>(), NL = NX / 2;
    using Tl = Tile2<NX, TY>;
    static_assert(Tl::NRAW >= 3 * TY, "FFT rows must fit in the staging area");
    constexpr int NFWV = (TY + 3) * NX, PF = (NFWV + T - 1) / T;
    __shared__ __align__(16) R sm[Tl::NROWS * NX];
    __shared__ C tws[NX];  // twiddles of the x transform: half-length, then R2C post
    const int tid = threadIdx.x;
    C tw_a = czero_t<C>(), tw_b = czero_t<C>();
    if (tid < NL) { tw_a = A.tw_nl[tid]; tw_b = A.tw_nx[tid]; }
    const int nty = A.ny / TY;
    // tile order: groups of G consecutive y tiles, levels fastest-changing
    // across groups, so that an XCD's L2 holds both the y halos and the
    // neighbouring levels of the tiles in flight
    const int G = A.rgroup > 0 ? A.rgroup : (nty % 16 == 0) ? 16 : (nty % 8 == 0) ? 8 : (nty % 4 == 0) ? 4 : 1;
    const int logical = xcd_remap(blockIdx.x, nty * A.nlev);
    const int grp = logical / (G * A.nlev), rem = logical - grp * (G * A.nlev);
    const int k = A.klo + rem / G, jt = grp * G + rem % G, j0 = jt * TY;
    const int i = tid % NX, jj = tid / NX, j = j0 + jj;
    const bool top = k == A.nz - 1;
    // staging: every thread issues all of its loads before the first use
    constexpr int NST = (Tl::NRAW * NX + T - 1) / T;
    R stv[NST];
#pragma unroll
    for (int m = 0; m < NST; ++m) {
        const int q = tid + m * T;
        stv[m] = R(0.0);
        if ((Tl::NRAW * NX) % T == 0 || q < Tl::NRAW * NX) {
            const int r = q / NX, x = q - r * NX;
            const R *f = r < Tl::OV ? A.u : r < Tl::OW ? A.v : r < Tl::OP ? A.w : A.p;
            const int jr = r < Tl::OP ? j0 - 3 + r % Tl::RU : j0 - 1 + (r - Tl::OP);
            stv[m] = f[((unsigned)k * (unsigned)A.ny + (unsigned)wrapi(jr, A.ny)) * (unsigned)NX + (unsigned)x];
        }
    }
    PointVals2<R> pv;
    pv.load(A, i, j, k);
#pragma unroll
    for (int m = 0; m < NST; ++m) {
        const int q = tid + m * T;
        if ((Tl::NRAW * NX) % T == 0 || q < Tl::NRAW * NX) sm[q] = stv[m];
    }
    // v(k+1) under the why*vhz rows, loaded with the staging
    R vkp[PF];
#pragma unroll
    for (int m = 0; m < PF; ++m) {
        const int e = tid + m * T;
        vkp[m] = R(0.0);
        if ((NFWV % T == 0 || e < NFWV) && !top) {
            const int r = e / NX, x = e - r * NX;
            vkp[m] = A.v[((size_t)(k + 1) * A.ny + wrapi(j0 - 2 + r, A.ny)) * NX + x];
        }
    }
    const R dpx = A.dp_opt == 2 ? R(A.scal[0]) : A.dpx;
    const R um = A.u_mrf;
    const R lt = top ? R(0.0) : A.lev[LV_LT * A.nz + k];
    if (tid < NL) { tws[tid] = tw_a; tws[NL + tid] = tw_b; }
    __syncthreads();
    // ---- phase 1: interpolants and flux products of level k
    {
        const R *ru = sm + (Tl::OU + jj + 3) * NX, *rw = sm + (Tl::OW + jj + 3) * NX;
        const R uhx = ixm<CDS, NX>(ru, i);
        const R whx = top ? R(0.0) : ixp<CDS, NX>(rw, i);
        const R uc = ru[i];
        const R uhz = top ? (A.ochannel == 0 ? -um : uc) : pv.u_kp * lt + uc * (R(1.0) - lt);
        sm[(Tl::OUHX + jj) * NX + i] = uhx;
        sm[(Tl::OWHX + jj) * NX + i] = whx;
        sm[(Tl::OFUW + jj) * NX + i] = uhz * whx;
    }
    for (int e = tid; e < Tl::NH1 * NX; e += T) {
        const int r = e / NX, x = e - r * NX;
        if (r < TY + 1) {  // vhx, rows j0-1+r
            sm[(Tl::OVHX + r) * NX + x] = ixp<CDS, NX>(sm + (Tl::OV + r + 2) * NX, x);
        } else {           // vhy, rows j0-1+r'
            const int r2 = r - (TY + 1);
            sm[(Tl::OVHY + r2) * NX + x] = iym<CDS, NX>(sm + (Tl::OV + r2 + 2) * NX + x);
        }
    }
    for (int e = tid; e < (TY + 3) * NX; e += T) {  // uhy*vhx, rows j0-2+r
        const int r = e / NX, x = e - r * NX;
        const R uhy = iyp<CDS, NX>(sm + (Tl::OU + r + 1) * NX + x);
        const R vhx = ixp<CDS, NX>(sm + (Tl::OV + r + 1) * NX, x);
        sm[(Tl::OFUV + r) * NX + x] = uhy * vhx;
    }
#pragma unroll
    for (int m = 0; m < PF; ++m) {  // why*vhz, rows j0-2+r
        const int e = tid + m * T;
        if (NFWV % T == 0 || e < NFWV) {
            const int r = e / NX, x = e - r * NX;
            const R vc = sm[(Tl::OV + r + 1) * NX + x];
            const R why = top ? R(0.0) : iyp<CDS, NX>(sm + (Tl::OW + r + 1) * NX + x);
            const R vhz = top ? (A.ochannel == 0 ? R(0.0) : vc) : vkp[m] * lt + vc * (R(1.0) - lt);
            sm[(Tl::OFWV + r) * NX + x] = vhz * why;
        }
    }
    __syncthreads();
    // ---- phase 2
    R qu, qv, qw;
    {
        R f, g, h;
        momentum_tile<CDS, NX, TY>(A, sm, i, jj, k, pv, dpx, f, g, h, qu, qv, qw);
        const size_t c = ((size_t)k * A.ny + j) * NX + i;
        A.ff[c] = f;
        A.gg[c] = g;
        A.hh[c] = h;
    }
    __syncthreads();  // staging no longer read: reuse it for the q rows
    sm[(0 * TY + jj) * NX + i] = qu;
    sm[(1 * TY + jj) * NX + i] = qv;
    sm[(2 * TY + jj) * NX + i] = qw;
    __syncthreads();
    C *sq_ = reinterpret_cast<C *>(sm);
    lds_fft<NL, 3 * TY, NL, 1, T, -1>(sq_, tws, tid);
    for (int p = tid; p < 3 * TY * NL; p += T) {
        int row = p / NL, l = p - row * NL, fi = row / TY, r = row - fi * TY;
        C *o = fi == 0 ? A.s0 : fi == 1 ? A.s1 : A.s2;
        o[((size_t)k * A.ny + j0 + r) * NL + l] = r2c_post<NL>(sq_ + row * NL, l, tws + NL, A.norm);
    }
}

// Momentum right-hand side on a TX x TY tile of level k (x tiles fastest,
// then y tiles, then levels, in XCD-contiguous order), without the x
// transform: the same phases, terms and wall rules as k_rhs_mom2 on LDS rows
// of S = TX + 6 points (columns i0-3 .. i0+TX+2; a derived row holds the
// columns that phase 2 reads).  ff, gg, hh are stored; qu, qv, qw go as
// physical rows into the storage of their spectra (a row of nx values takes
// the bytes of its nx/2 modes), which launch_r2c_n then transforms in place.
// The small LDS tile lets several blocks share a CU, so that one block's
// loads overlap another's stencils.
// DAMP (levels that read the sponge of a damped substep, FP64): every value
// is read as rayleigh_damping leaves it (src/module_tools.f90:705-749): a
// level k >= kdamp0 damped at the domain's interior (i, j), undamped at a
// wrapped x / y position (the stale halo) and in a z ghost.  That equals the
// reference for the points two or more columns and rows from the domain's
// x / y edges, whose interpolants all sit at interior centres;
// k_rhs_mom_edge then recomputes the points nearer the edges.
template <int CDS, int TX, int TY, class R, bool DAMP = false>
__global__ __launch_bounds__(TX * TY, 4) void k_rhs_mom_xt(StepArgsT<R> A)
{
    constexpr int T = TX * TY, S = TX + 6, HX = 3;
    using Tl = Tile2<S, TY>;
    constexpr int NST = (Tl::NRAW * S + T - 1) / T;
    constexpr int NUK = TY * (TX + 3), PU = (NUK + T - 1) / T;   // uhx/whx/fuw columns 1 .. TX+3
    constexpr int NVK = (TY + 3) * TX, PV = (NVK + T - 1) / T;   // fwv rows j0-2 .. j0+TY, own columns
    __shared__ __align__(16) R sm[Tl::NROWS * S];
    const int tid = threadIdx.x, nx = A.nx, ny = A.ny;
    const int ntx = nx / TX, nty = ny / TY;
    // tile order: groups of G rows of y tiles swept level by level, x tiles
    // fastest, so that an XCD's L2 holds the y halos and the neighbouring
    // levels of the tiles in flight
    int G = A.rgroup > 0 ? A.rgroup : 8;
    if (nty % G) G = nty % 8 == 0 ? 8 : nty % 4 == 0 ? 4 : 1;
    const int per = G * ntx * A.nlev;
    const int logical = xcd_remap(blockIdx.x, ntx * nty * A.nlev);
    const int grp = logical / per, rem = logical - grp * per, r2 = rem % (G * ntx);
    const int k = A.klo + rem / (G * ntx), jt = grp * G + r2 / ntx, it = r2 % ntx;
    const int i0 = it * TX, j0 = jt * TY;
    const int il = tid % TX, jj = tid / TX;
    const int i = i0 + il, j = j0 + jj, xl = il + HX;
    const bool top = k == A.nz - 1;
    auto gx = [&](int x) -> unsigned { return (unsigned)wrapi(i0 - HX + x, nx); };
    // field fid of level kk as the damped substep reads it (interior point)
    auto dmp = [&](int fid, int kk, R v) __attribute__((always_inline)) {
        if constexpr (DAMP) {
            if (kk >= A.kdamp0 && ((A.dmask >> fid) & 1)) {
                const R s = A.dmp_s[kk], m = A.dmp_m[fid * A.nz + kk];
                v = v - s * (v - m);
            }
        }
        return v;
    };
    // staging: every thread issues all of its loads before the first use
    R stv[NST];
#pragma unroll
    for (int m = 0; m < NST; ++m) {
        const int q = tid + m * T;
        stv[m] = R(0.0);
        if ((Tl::NRAW * S) % T == 0 || q < Tl::NRAW * S) {
            const int r = q / S, x = q - r * S;
            const R *f = r < Tl::OV ? A.u : r < Tl::OW ? A.v : r < Tl::OP ? A.w : A.p;
            const int jr = r < Tl::OP ? j0 - 3 + r % Tl::RU : j0 - 1 + (r - Tl::OP);
            stv[m] = f[((unsigned)k * (unsigned)ny + (unsigned)wrapi(jr, ny)) * (unsigned)nx + gx(x)];
        }
    }
    PointVals2<R> pv;
    pv.load(A, i, j, k);
    // u(k+1) under the uhz*whx columns and v(k+1) under the why*vhz rows
    R ukp[PU], vkp[PV];
#pragma unroll
    for (int m = 0; m < PU; ++m) {
        const int e = tid + m * T;
        ukp[m] = R(0.0);
        if ((NUK % T == 0 || e < NUK) && !top) {
            const int r = e / (TX + 3), x = 1 + e % (TX + 3);
            ukp[m] = A.u[((unsigned)(k + 1) * (unsigned)ny + (unsigned)(j0 + r)) * (unsigned)nx + gx(x)];
        }
    }
#pragma unroll
    for (int m = 0; m < PV; ++m) {
        const int e = tid + m * T;
        vkp[m] = R(0.0);
        if ((NVK % T == 0 || e < NVK) && !top) {
            const int r = e / TX, x = e % TX;
            vkp[m] = A.v[((unsigned)(k + 1) * (unsigned)ny + (unsigned)wrapi(j0 - 2 + r, ny)) * (unsigned)nx +
                         (unsigned)(i0 + x)];
        }
    }
    if constexpr (DAMP) {
        // the staged rows: damped at the domain's interior positions only
#pragma unroll
        for (int m = 0; m < NST; ++m) {
            const int q = tid + m * T, r = q / S, x = q - r * S;
            const int fid = r < Tl::OV ? 0 : r < Tl::OW ? 1 : r < Tl::OP ? 2 : 3;
            const int jr = r < Tl::OP ? j0 - 3 + r % Tl::RU : j0 - 1 + (r - Tl::OP), xg = i0 - HX + x;
            if (xg >= 0 && xg < nx && jr >= 0 && jr < ny) stv[m] = dmp(fid, k, stv[m]);
        }
#pragma unroll
        for (int m = 0; m < PU; ++m) {
            const int e = tid + m * T, xg = i0 - HX + 1 + e % (TX + 3);
            if (!top && xg >= 0 && xg < nx) ukp[m] = dmp(0, k + 1, ukp[m]);
        }
#pragma unroll
        for (int m = 0; m < PV; ++m) {
            const int jr = j0 - 2 + (tid + m * T) / TX;
            if (!top && jr >= 0 && jr < ny) vkp[m] = dmp(1, k + 1, vkp[m]);
        }
        // the point's z neighbours (wall ghosts stay as loaded)
        if (k >= 1) {
            pv.u_km = dmp(0, k - 1, pv.u_km);
            pv.v_km = dmp(1, k - 1, pv.v_km);
#pragma unroll
            for (int d = 0; d < 4; ++d) pv.wkx[d] = dmp(2, k - 1, pv.wkx[d]);
#pragma unroll
            for (int d = 0; d < 3; ++d) pv.wky[d] = dmp(2, k - 1, pv.wky[d]);
        }
        if (!top) {
            pv.u_kp = dmp(0, k + 1, pv.u_kp);
            pv.v_kp = dmp(1, k + 1, pv.v_kp);
            if (k + 1 < A.nz - 1) pv.w_kp = dmp(2, k + 1, pv.w_kp);
            pv.p_kp = dmp(3, k + 1, pv.p_kp);
            if (A.isscalar) pv.t_kp = dmp(4, k + 1, pv.t_kp);
        }
        if (A.isscalar) pv.t_k = dmp(4, k, pv.t_k);
    }
#pragma unroll
    for (int m = 0; m < NST; ++m) {
        const int q = tid + m * T;
        if ((Tl::NRAW * S) % T == 0 || q < Tl::NRAW * S) sm[q] = stv[m];
    }
    const R dpx = A.dp_opt == 2 ? R(A.scal[0]) : A.dpx;
    const R um = A.u_mrf;
    const R lt = top ? R(0.0) : A.lev[LV_LT * A.nz + k];
    __syncthreads();
    // ---- phase 1: interpolants and flux products of level k
    for (int e = tid; e < TY * (TX + 3); e += T) {  // uhx, rows j0 .. j0+TY-1, columns 2 .. TX+4
        const int r = e / (TX + 3), x = 2 + e % (TX + 3);
        sm[(Tl::OUHX + r) * S + x] = ixm_d<CDS>(sm + (Tl::OU + r + 3) * S, x);
    }
#pragma unroll
    for (int m = 0; m < PU; ++m) {  // whx and uhz*whx, rows j0 .. j0+TY-1, columns 1 .. TX+3
        const int e = tid + m * T;
        if (NUK % T == 0 || e < NUK) {
            const int r = e / (TX + 3), x = 1 + e % (TX + 3);
            const R whx = top ? R(0.0) : ixp_d<CDS>(sm + (Tl::OW + r + 3) * S, x);
            const R uc = sm[(Tl::OU + r + 3) * S + x];
            const R uhz = top ? (A.ochannel == 0 ? -um : uc) : ukp[m] * lt + uc * (R(1.0) - lt);
            sm[(Tl::OWHX + r) * S + x] = whx;
            sm[(Tl::OFUW + r) * S + x] = uhz * whx;
        }
    }
    for (int e = tid; e < (TY + 1) * (TX + 3); e += T) {  // vhx, rows j0-1 .. j0+TY-1, columns 1 .. TX+3
        const int r = e / (TX + 3), x = 1 + e % (TX + 3);
        sm[(Tl::OVHX + r) * S + x] = ixp_d<CDS>(sm + (Tl::OV + r + 2) * S, x);
    }
    for (int e = tid; e < (TY + 3) * TX; e += T) {  // vhy, rows j0-1 .. j0+TY+1, own columns
        const int r = e / TX, x = HX + e % TX;
        sm[(Tl::OVHY + r) * S + x] = iym<CDS, S>(sm + (Tl::OV + r + 2) * S + x);
    }
    for (int e = tid; e < (TY + 3) * (TX + 3); e += T) {  // uhy*vhx, rows j0-2 .. j0+TY, columns 1 .. TX+3
        const int r = e / (TX + 3), x = 1 + e % (TX + 3);
        const R uhy = iyp<CDS, S>(sm + (Tl::OU + r + 1) * S + x);
        const R vhx = ixp_d<CDS>(sm + (Tl::OV + r + 1) * S, x);
        sm[(Tl::OFUV + r) * S + x] = uhy * vhx;
    }
#pragma unroll
    for (int m = 0; m < PV; ++m) {  // why*vhz, rows j0-2 .. j0+TY, own columns
        const int e = tid + m * T;
        if (NVK % T == 0 || e < NVK) {
            const int r = e / TX, x = HX + e % TX;
            const R vc = sm[(Tl::OV + r + 1) * S + x];
            const R why = top ? R(0.0) : iyp<CDS, S>(sm + (Tl::OW + r + 1) * S + x);
            const R vhz = top ? (A.ochannel == 0 ? R(0.0) : vc) : vkp[m] * lt + vc * (R(1.0) - lt);
            sm[(Tl::OFWV + r) * S + x] = vhz * why;
        }
    }
    __syncthreads();
    // ---- phase 2
    R f, g, h, qu, qv, qw;
    momentum_tile<CDS, S, TY, R, false>(A, sm, xl, jj, k, pv, dpx, f, g, h, qu, qv, qw);
    const size_t c = ((size_t)k * ny + j) * nx + i;
    A.ff[c] = f;
    A.gg[c] = g;
    A.hh[c] = h;
    reinterpret_cast<R *>(A.s0)[c] = qu;
    reinterpret_cast<R *>(A.s1)[c] = qv;
    reinterpret_cast<R *>(A.s2)[c] = qw;
}

// The points of a damped substep's sponge levels within two columns or rows
// of the domain's x / y edges (x in 0,1,nx-2,nx-1 or y in 0,1,ny-2,ny-1):
// their stencils read stale halos and interpolants evaluated at wrapped
// centres, which the accessor PtD reproduces.  Writes ff, gg, hh and the
// physical qu, qv, qw over the values k_rhs_mom_xt<..., DAMP> left there.
template <int CDS>
__global__ __launch_bounds__(256) void k_rhs_mom_edge(StepArgs A)
{
    const int nx = A.nx, ny = A.ny, per = 4 * nx + 4 * (ny - 4);
    const int t = (int)(blockIdx.x * 256 + threadIdx.x);
    if (t >= per * A.nlev) return;
    const int k = A.klo + t / per, e = t % per;
    int i, j;
    if (e < 4 * nx) {
        const int r = e / nx;
        j = r < 2 ? r : ny - 4 + r;
        i = e - r * nx;
    } else {
        const int u = e - 4 * nx, q = u % 4;
        j = 2 + u / 4;
        i = q < 2 ? q : nx - 4 + q;
    }
    PtD q(A, i, j, k);
    double f, g, h, qu, qv, qw;
    const double dpx = A.dp_opt == 2 ? A.scal[0] : A.dpx;
    momentum_point<CDS, 2>(q, dpx, f, g, h, qu, qv, qw);
    const size_t c = ((size_t)k * ny + j) * nx + i;
    A.ff[c] = f;
    A.gg[c] = g;
    A.hh[c] = h;
    reinterpret_cast<double *>(A.s0)[c] = qu;
    reinterpret_cast<double *>(A.s1)[c] = qv;
    reinterpret_cast<double *>(A.s2)[c] = qw;
}

// Tile shapes of k_rhs_mom_xt (launch-shape knob rhs_xt; 0 is the fused
// whole-row kernel)
constexpr int kRhsXtTX[5] = {0, 32, 64, 32, 64};
constexpr int kRhsXtTY[5] = {0, 8, 8, 16, 4};
bool rhs_xt_fits(int shape, int nx, int ny)
{
    return shape >= 1 && shape <= 4 && nx % kRhsXtTX[shape] == 0 && ny % kRhsXtTY[shape] == 0;
}

// Momentum RHS + x R2C of qu,qv,qw through a global-memory accessor: PtD in
// a damped substep, Pt with the 6th-order z operators (ZO = 6).  Block:
// rows_per_block full rows of level k.
template <int CDS, int NX, class Acc = PtD, int ZO = 2>
__global__ __launch_bounds__(kRhsThreads) void k_rhs_mom_damped(StepArgsT<typename Acc::Real> A)
{
    using R = typename Acc::Real;
    using C = cplx<R>;
    constexpr int ROWS = rows_per_block<NX>();
    constexpr int NL = NX / 2;
    __shared__ C sq_[3 * ROWS * NL];
    R *sqd = reinterpret_cast<R *>(sq_);
    const int k = A.klo + (int)blockIdx.y, j0 = blockIdx.x * ROWS, tid = threadIdx.x;
    const R dpx = A.dp_opt == 2 ? R(A.scal[0]) : A.dpx;
    for (int p = tid; p < ROWS * NX; p += kRhsThreads) {
        int jj = p / NX, i = p - jj * NX, j = j0 + jj;
        R qu = R(0.0), qv = R(0.0), qw = R(0.0);
        if (j < A.ny) {
            Acc q(A, i, j, k);
            R f, g, h;
            momentum_point<CDS, ZO>(q, dpx, f, g, h, qu, qv, qw);
            const size_t c = ((size_t)k * A.ny + j) * NX + i;
            A.ff[c] = f;
            A.gg[c] = g;
            A.hh[c] = h;
        }
        sqd[(0 * ROWS + jj) * NX + i] = qu;
        sqd[(1 * ROWS + jj) * NX + i] = qv;
        sqd[(2 * ROWS + jj) * NX + i] = qw;
    }
    __syncthreads();
    lds_fft<NL, 3 * ROWS, NL, 1, kRhsThreads, -1>(sq_, A.tw_nl, tid);
    C *outs[3] = {A.s0, A.s1, A.s2};
    for (int p = tid; p < 3 * ROWS * NL; p += kRhsThreads) {
        int row = p / NL, l = p - row * NL, fi = row / ROWS, jj = row - fi * ROWS, j = j0 + jj;
        if (j < A.ny) outs[fi][((size_t)k * A.ny + j) * NL + l] = r2c_post<NL>(sq_ + row * NL, l, A.tw_nx, A.norm);
    }
}

// Plane means of the damped levels, one block per (level, field):
// the mean_local of rayleigh_damping (src/module_tools.f90:718-732).
struct DampFields {
    const double *f[5];
};
__global__ __launch_bounds__(1024) void k_damp_means(DampFields F, double *dmp_m, int nx, int ny, int nz, int kdamp0)
{
    __shared__ double red[1024];
    const int kk = kdamp0 + blockIdx.x, fid = blockIdx.y;
    const double *plane = F.f[fid] + (size_t)kk * nx * ny;
    double s = 0.0;
    for (int q = threadIdx.x; q < nx * ny; q += 1024) s += plane[q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) dmp_m[fid * nz + kk] = red[0] / nx / ny;
}

// In-place damping of one field's interior (the temperature when no
// temperature equation overwrites it later in the substep).
__global__ __launch_bounds__(256) void k_damp_apply(double *f, const double *dmp_s, const double *mean, int nx,
                                                     int ny, int kdamp0, int nlev)
{
    const size_t n = (size_t)nx * ny * nlev;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int kk = kdamp0 + (int)(q / ((size_t)nx * ny));
        const size_t c = (size_t)kdamp0 * nx * ny + q;
        const double v = f[c], s = dmp_s[kk], m = mean[kk];
        f[c] = v - s * (v - m);
    }
}


// The momentum right-hand side of one substep into B.ff, gg, hh and the x
// spectra of qu, qv, qw (B.s0..s2).  A damped substep (isdamp = 1) first
// forms the plane means of the undamped fields (and damps a passive t in
// place); the levels that read a damped value (kdamp0-1 and up: the
// stencils reach one level down) then go through the accessor kernel that
// applies the damping as it reads, and the levels below through the tiled
// kernels of an undamped substep.
template <class R>
int rhs_momentum(hc_ctx *c, const StepArgs &A, StepArgsT<R> &B)
{
    const hc_params &P = c->prm;
    const int rows = rows_for(c->nx);
    const bool damp = P.isdamp == 1, z6 = P.zorder == 4 || P.zorder == 6;
    const int ksplit = damp ? std::max(0, c->kdamp0 - 1) : c->nz;
    if constexpr (std::is_same<R, double>::value) {
        if (damp) {
            // rayleigh_damping of u, v, w, p, t at the start of the substep
            // (src/module_navier_stokes.f90:258-266)
            DampFields F = {{c->u, c->v, c->w, c->p, c->t}};
            hipLaunchKernelGGL(k_damp_means, dim3(P.nzdamp + 1, 5), dim3(1024), 0, c->stream, F, c->dmp_m, c->nx, c->ny,
                               c->nz, c->kdamp0);
            if (P.isscalar == 0) {
                hipLaunchKernelGGL(k_damp_apply, dim3(256), dim3(256), 0, c->stream, c->t, c->dmp_s,
                                   c->dmp_m + 4 * c->nz, c->nx, c->ny, c->kdamp0, P.nzdamp + 1);
            }
            B.dmask = 31;
            B.klo = ksplit;
            B.nlev = c->nz - ksplit;
            // the sponge levels: x-y tiles reading damped values, then the
            // points near the x / y edges through the accessor, then the x
            // transform in place; the accessor kernel alone where no tile
            // shape divides the grid
            const int sh = c->tune.rhs_xt > 0 ? c->tune.rhs_xt : 1;
            if (rhs_xt_fits(sh, c->nx, c->ny) && c->nx >= 4 && c->ny >= 4) {
                const dim3 g((unsigned)(c->nx / kRhsXtTX[sh] * (c->ny / kRhsXtTY[sh]) * B.nlev));
#define HC_XTD(TX_, TY_)                                                                                                 \
    if (P.cds == 1) hipLaunchKernelGGL((k_rhs_mom_xt<1, TX_, TY_, R, true>), g, dim3(TX_ * TY_), 0, c->stream, B);       \
    else hipLaunchKernelGGL((k_rhs_mom_xt<2, TX_, TY_, R, true>), g, dim3(TX_ * TY_), 0, c->stream, B);
                switch (sh) {
                case 1: HC_XTD(32, 8) break;
                case 2: HC_XTD(64, 8) break;
                case 3: HC_XTD(32, 16) break;
                default: HC_XTD(64, 4) break;
                }
#undef HC_XTD
                const int per = 4 * c->nx + 4 * (c->ny - 4);
                const dim3 ge((unsigned)(((size_t)per * B.nlev + 255) / 256));
                if (P.cds == 1) hipLaunchKernelGGL((k_rhs_mom_edge<1>), ge, dim3(256), 0, c->stream, B);
                else hipLaunchKernelGGL((k_rhs_mom_edge<2>), ge, dim3(256), 0, c->stream, B);
                const R *src[3] = {reinterpret_cast<const R *>(B.s0), reinterpret_cast<const R *>(B.s1),
                                   reinterpret_cast<const R *>(B.s2)};
                cplx<R> *dst[3] = {B.s0, B.s1, B.s2};
                if (int e = launch_r2c_n<R>(c, 3, src, dst, ksplit, B.nlev)) return e;
            } else {
                const dim3 gd((c->ny + rows - 1) / rows, B.nlev);
#define HC_RHSD(NX_)                                                                                       \
    if (P.cds == 1) hipLaunchKernelGGL((k_rhs_mom_damped<1, NX_>), gd, dim3(kRhsThreads), 0, c->stream, B); \
    else hipLaunchKernelGGL((k_rhs_mom_damped<2, NX_>), gd, dim3(kRhsThreads), 0, c->stream, B);
                HC_DISPATCH_N(c->nx, HC_RHSD)
#undef HC_RHSD
            }
        }
    }
    if (z6 && !damp) {  // 6th/4th-order z operators: the global-memory point kernel
        B.dmask = 0;
        zinterp6<R>(c, B);
        const dim3 grow((c->ny + rows - 1) / rows, c->nz);
#define HC_RHSZ(NX_)                                                                                                   \
    if (P.cds == 1) hipLaunchKernelGGL((k_rhs_mom_damped<1, NX_, PtT<R>, 6>), grow, dim3(kRhsThreads), 0, c->stream, B); \
    else hipLaunchKernelGGL((k_rhs_mom_damped<2, NX_, PtT<R>, 6>), grow, dim3(kRhsThreads), 0, c->stream, B);
        HC_DISPATCH_N(c->nx, HC_RHSZ)
#undef HC_RHSZ
    }
    int err = 0;
    if (!z6 && ksplit > 0) {
        B.klo = 0;
        B.nlev = ksplit;
        if (rhs_xt_fits(c->tune.rhs_xt, c->nx, c->ny)) {
            const int sh = c->tune.rhs_xt;
            const dim3 g((unsigned)(c->nx / kRhsXtTX[sh] * (c->ny / kRhsXtTY[sh]) * ksplit));
#define HC_XT(TX_, TY_)                                                                                      \
    if (P.cds == 1) hipLaunchKernelGGL((k_rhs_mom_xt<1, TX_, TY_, R>), g, dim3(TX_ * TY_), 0, c->stream, B); \
    else hipLaunchKernelGGL((k_rhs_mom_xt<2, TX_, TY_, R>), g, dim3(TX_ * TY_), 0, c->stream, B);
            switch (sh) {
            case 1: HC_XT(32, 8) break;
            case 2: HC_XT(64, 8) break;
            case 3: HC_XT(32, 16) break;
            default: HC_XT(64, 4) break;
            }
#undef HC_XT
            const R *src[3] = {reinterpret_cast<const R *>(B.s0), reinterpret_cast<const R *>(B.s1),
                               reinterpret_cast<const R *>(B.s2)};
            cplx<R> *dst[3] = {B.s0, B.s1, B.s2};
            err = launch_r2c_n<R>(c, 3, src, dst, 0, ksplit);
        } else {
#define HC_RHS(NX_)                                                                                     \
    if constexpr (NX_ <= 256) {                                                                          \
        constexpr int TY_ = rhs2_ty<NX_>();                                                              \
        dim3 g1((unsigned)(c->ny / TY_ * ksplit));                                                       \
        constexpr int T_ = rhs2_threads<NX_>();                                                          \
        if (P.cds == 1) hipLaunchKernelGGL((k_rhs_mom2<1, NX_, R>), g1, dim3(T_), 0, c->stream, B);      \
        else hipLaunchKernelGGL((k_rhs_mom2<2, NX_, R>), g1, dim3(T_), 0, c->stream, B);                 \
    } else {                                                                                             \
        constexpr int TY_ = rhs_ty<NX_>();                                                               \
        dim3 g1((unsigned)(c->ny / TY_ * ksplit));                                                       \
        constexpr int T_ = rhs_threads<NX_>();                                                           \
        if (P.cds == 1) hipLaunchKernelGGL((k_rhs_mom<1, NX_, R>), g1, dim3(T_), 0, c->stream, B);       \
        else hipLaunchKernelGGL((k_rhs_mom<2, NX_, R>), g1, dim3(T_), 0, c->stream, B);                  \
    }
            HC_DISPATCH_N(c->nx, HC_RHS)
#undef HC_RHS
        }
    }
    B.klo = 0;
    B.nlev = c->nz;
    return err;
}

template int rhs_momentum<double>(hc_ctx *, const StepArgs &, StepArgsT<double> &);
template int rhs_momentum<float>(hc_ctx *, const StepArgs &, StepArgsT<float> &);
