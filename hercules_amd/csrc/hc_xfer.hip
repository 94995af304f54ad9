// x and y transforms of the half spectra: the plain x R2C of one field, the
// inverse x transform (C2R) with store / accumulate / damped accumulate, and
// the in-place y FFT of up to four spectra (mpi_2d_fft,
// src/module_tools.f90:518-700, on one GPU).
#include "hc_host.h"

// x R2C of up to four real fields, scaled by norm.  Blocks of one field
// (blockIdx.z) loop over its tiles of ROWS rows with a grid stride, two tiles
// per trip with two register buffers: the next tile's rows are loaded before
// the current tile's transform.  src may be dst (an in-place transform): a
// tile is read and written only by its own block (a fetch past the last tile
// rereads the last one and is never used).
template <int NX, class R>
__global__ __launch_bounds__(kRhsThreads) void k_x_r2c(R2CFields<R> F, int ny, int nk, const cplx<R> *tw_nl,
                                                       const cplx<R> *tw_nx, R norm)
{
    using C = cplx<R>;
    constexpr int ROWS = rows_per_block<NX>();
    constexpr int NL = NX / 2;
    constexpr int NIN = (ROWS * NX + kRhsThreads - 1) / kRhsThreads;
    __shared__ C sq_[ROWS * NL];
    __shared__ C tws[2 * NL];  // half-length twiddles, then the R2C post-twiddles
    R *sqd = reinterpret_cast<R *>(sq_);
    const int tid = threadIdx.x, G = gridDim.x;
    const R *src = F.src[blockIdx.z];
    C *dst = F.dst[blockIdx.z];
    const int njt = (ny + ROWS - 1) / ROWS, ntiles = njt * nk;
    for (int q = tid; q < NL; q += kRhsThreads) { tws[q] = tw_nl[q]; tws[NL + q] = tw_nx[q]; }
    auto fetch = [&](int tile, R (&x)[NIN]) __attribute__((always_inline)) {
        const int t = tile < ntiles ? tile : ntiles - 1;
        const int k = t / njt, j0 = (t - k * njt) * ROWS;
#pragma unroll
        for (int m = 0; m < NIN; ++m) {
            const int p = (ROWS * NX) % kRhsThreads == 0 ? tid + m * kRhsThreads
                                                          : min(tid + m * kRhsThreads, ROWS * NX - 1);
            const int jj = p / NX, i = p - jj * NX, j = min(j0 + jj, ny - 1);
            x[m] = src[((size_t)k * ny + j) * NX + i];
        }
    };
    auto work = [&](int tile, const R (&x)[NIN]) __attribute__((always_inline)) {
        const int k = tile / njt, j0 = (tile - k * njt) * ROWS;
        __syncthreads();  // twiddles stored; the previous tile's rows read
#pragma unroll
        for (int m = 0; m < NIN; ++m) {
            const int p = tid + m * kRhsThreads;
            if ((ROWS * NX) % kRhsThreads == 0 || p < ROWS * NX) sqd[p] = j0 + p / NX < ny ? x[m] : R(0.0);
        }
        __syncthreads();
        lds_fft<NL, ROWS, NL, 1, kRhsThreads, -1>(sq_, tws, tid);
        for (int p = tid; p < ROWS * NL; p += kRhsThreads) {
            const int jj = p / NL, l = p - jj * NL, j = j0 + jj;
            if (j < ny) dst[((size_t)k * ny + j) * NL + l] = r2c_post<NL>(sq_ + jj * NL, l, tws + NL, norm);
        }
    };
    R xa[NIN], xb[NIN];
    fetch(blockIdx.x, xa);
    for (int tile = blockIdx.x; tile < ntiles; tile += 2 * G) {
        fetch(tile + G, xb);
        work(tile, xa);
        if (tile + G >= ntiles) break;
        fetch(tile + 2 * G, xa);
        work(tile + G, xb);
    }
}

// Field of a row: wave-uniform when a wavefront's 64 lanes share the row
// (the row length is a multiple of 64), and then taken from a scalar register
// so that the field's pointer is selected per wavefront.
template <int LEN, int ROWS>
HC_DEV int c2r_field(int p)
{
    const int f = p / LEN / ROWS;
    if constexpr (LEN % 64 == 0) return __builtin_amdgcn_readfirstlane(f);
    return f;
}

// The spectra rows and, for the accumulated fields, the old values of one
// tile (ROWS rows of level k, F fields).  Every call issues the same loads,
// so that the compiler's wait for one buffer never covers the loads of the
// other: a tile past the last one reads the last tile (never used), a field
// that is only stored reads its first entry in place of the old row, and xt
// holds the mirrored entry (NL - l) mod NL as loaded (its conjugate is taken
// where it is used).
template <int NX, int F, class R, int ROWS, int NIN, int NOUT>
HC_DEV void c2r_fetch(const C2RArgsT<R> &A, int tile, int njt, int ntiles, int tid, cplx<R> (&xl)[NIN],
                      cplx<R> (&xt)[NIN], R (&old)[NOUT])
{
    using C = cplx<R>;
    constexpr int NL = NX / 2, B = F * ROWS;
    const int t = tile < ntiles ? tile : ntiles - 1;
    const int k = t / njt, j0 = (t - k * njt) * ROWS;
#pragma unroll
    for (int m = 0; m < NIN; ++m) {
        const int p = (B * NL) % kRhsThreads == 0 ? tid + m * kRhsThreads : min(tid + m * kRhsThreads, B * NL - 1);
        const int row = p / NL, l = p - row * NL, f = c2r_field<NL, ROWS>(p), jj = row - f * ROWS;
        const int j = min(j0 + jj, A.ny - 1);
        const C *s = A.src[f] + ((size_t)k * A.ny + j) * NL;
        xl[m] = s[l];
        xt[m] = s[l == 0 ? 0 : NL - l];
    }
#pragma unroll
    for (int m = 0; m < NOUT; ++m) {
        const int p = (B * NX) % kRhsThreads == 0 ? tid + m * kRhsThreads : min(tid + m * kRhsThreads, B * NX - 1);
        const int row = p / NX, i = p - row * NX, f = c2r_field<NX, ROWS>(p), jj = row - f * ROWS;
        const int j = min(j0 + jj, A.ny - 1);
        const size_t c = A.mode[f] != 0 ? ((size_t)k * A.ny + j) * NX + i : 0;
        old[m] = A.dst[f][c];
    }
}

// One tile of the inverse x transform from fetched rows: pre-twiddle into
// LDS, half-length inverse FFT, store or accumulate.
template <int NX, int F, class R, int ROWS, int NIN, int NOUT>
HC_DEV void c2r_tile(const C2RArgsT<R> &A, int tile, int njt, int tid, const cplx<R> (&xl)[NIN],
                     const cplx<R> (&xt)[NIN], const R (&old)[NOUT], cplx<R> *sq_, const cplx<R> *tws)
{
    using C = cplx<R>;
    constexpr int NL = NX / 2, B = F * ROWS;
    const int k = tile / njt, j0 = (tile - k * njt) * ROWS;
    __syncthreads();  // twiddles stored; the previous tile's rows read
#pragma unroll
    for (int m = 0; m < NIN; ++m) {
        const int p = tid + m * kRhsThreads;
        if ((B * NL) % kRhsThreads == 0 || p < B * NL) {
            const int l = p % NL;
            const C t = l != 0 ? cconj(xt[m]) : czero_t<C>();
            sq_[p] = c2r_pre(xl[m], t, tws[NL + l]);
        }
    }
    __syncthreads();
    lds_fft<NL, B, NL, 1, kRhsThreads, +1>(sq_, tws, tid);
    const R *sqd = reinterpret_cast<const R *>(sq_);
#pragma unroll
    for (int m = 0; m < NOUT; ++m) {
        const int p = tid + m * kRhsThreads;
        const int row = p / NX, i = p - row * NX, f = c2r_field<NX, ROWS>(p), jj = row - f * ROWS, j = j0 + jj;
        if (((B * NX) % kRhsThreads == 0 || p < B * NX) && j < A.ny) {
            size_t c = ((size_t)k * A.ny + j) * NX + i;
            R x = sqd[row * NX + i];
            if (A.mode[f] == 1) {
                A.dst[f][c] = old[m] + x;
            } else if (A.mode[f] == 2) {
                R o = old[m];
                if (k >= A.kdamp0) o = o - A.dmp_s[k] * (o - A.dmp_mean[k]);
                A.dst[f][c] = o + x;
            } else {
                A.dst[f][c] = x;
            }
        }
    }
}

// Blocks loop over the tiles with a grid stride, two tiles per trip with two
// register buffers: the next tile's spectra and old values are loaded before
// the current tile's transform, so that the loads of one tile overlap the LDS
// work of the previous one.
template <int NX, int F, class R, int PTS = kC2RPoints>
__global__ __launch_bounds__(kRhsThreads) void k_x_c2r(C2RArgsT<R> A)
{
    using C = cplx<R>;
    constexpr int ROWS = c2r_rows<NX, F, PTS>();
    constexpr int NL = NX / 2;
    constexpr int B = F * ROWS;
    __shared__ C sq_[B * NL];
    __shared__ C tws[2 * NL];  // half-length twiddles, then the C2R pre-twiddles
    const int tid = threadIdx.x;
    const int njt = (A.ny + ROWS - 1) / ROWS, ntiles = njt * A.nz, G = gridDim.x;
    if (tid < NL) { tws[tid] = A.tw_nl[tid]; tws[NL + tid] = A.tw_nx[tid]; }
    constexpr int NIN = (B * NL + kRhsThreads - 1) / kRhsThreads;
    constexpr int NOUT = (B * NX + kRhsThreads - 1) / kRhsThreads;
    C xa[NIN], ta[NIN], xb[NIN], tb[NIN];
    R oa[NOUT], ob[NOUT];
    c2r_fetch<NX, F, R, ROWS>(A, blockIdx.x, njt, ntiles, tid, xa, ta, oa);
    for (int tile = blockIdx.x; tile < ntiles; tile += 2 * G) {
        c2r_fetch<NX, F, R, ROWS>(A, tile + G, njt, ntiles, tid, xb, tb, ob);
        c2r_tile<NX, F, R, ROWS>(A, tile, njt, tid, xa, ta, oa, sq_, tws);
        if (tile + G >= ntiles) break;
        c2r_fetch<NX, F, R, ROWS>(A, tile + 2 * G, njt, ntiles, tid, xa, ta, oa);
        c2r_tile<NX, F, R, ROWS>(A, tile + G, njt, tid, xb, tb, ob, sq_, tws);
    }
}

// TL consecutive l per block: 8 double2 or 16 float2, one 128-byte line per row.
template <int NY, int SIGN, class R, int TL = kYTile>
__global__ __launch_bounds__(256) void k_yfft(YArgsT<R> A)
{
    using C = cplx<R>;
    constexpr int TP = TL + 1;
    __shared__ C tile[NY * TP];
    __shared__ C twy[NY];
    C *fld = A.f[blockIdx.z];
    const int l0 = blockIdx.x * TL, k = blockIdx.y, tid = threadIdx.x;
    for (int q = tid; q < NY; q += 256) twy[q] = A.tw_ny[q];
    constexpr int NIT = (NY * TL + 255) / 256;
    C v[NIT];
#pragma unroll
    for (int m = 0; m < NIT; ++m) {
        const int p = tid + m * 256, j = p / TL, ll = p - j * TL;
        v[m] = p < NY * TL ? fld[((size_t)k * NY + j) * A.nl + l0 + ll] : czero_t<C>();
    }
#pragma unroll
    for (int m = 0; m < NIT; ++m) {
        const int p = tid + m * 256, j = p / TL, ll = p - j * TL;
        if (p < NY * TL) tile[j * TP + ll] = (SIGN > 0 && j == NY / 2) ? czero_t<C>() : v[m];
    }
    __syncthreads();
    lds_fft<NY, TL, 1, TP, 256, SIGN>(tile, twy, tid);
    for (int p = tid; p < NY * TL; p += 256) {
        int j = p / TL, ll = p - j * TL;
        C v = tile[j * TP + ll];
        if (SIGN < 0 && j == NY / 2) v = czero_t<C>();
        fld[((size_t)k * NY + j) * A.nl + l0 + ll] = v;
    }
}

template <class C>
int launch_yfft(hc_ctx *c, int sign, int nf, C *f0, typename nd<C>::type *f1, typename nd<C>::type *f2,
                       typename nd<C>::type *f3)
{
    using R = real_of<C>;
    YArgsT<R> Y;
    Y.f[0] = f0; Y.f[1] = f1; Y.f[2] = f2; Y.f[3] = f3;
    Y.nl = c->nl; Y.nz = c->nz; Y.tw_ny = view<R>(c).tw_ny;
    // FP32 spectra: 16 l per block (one cache line per row) where nl allows;
    // the autotuner may choose either width for either precision
    const int tl = c->tune.ytile ? c->tune.ytile : (sizeof(R) == 4 ? 2 * kYTile : kYTile);
    const bool wide = tl == 2 * kYTile && c->nl % (2 * kYTile) == 0;
    dim3 g(c->nl / (wide ? 2 * kYTile : kYTile), c->nz, nf);
#define HC_Y(NY_)                                                                                          \
    if (wide) {                                                                                            \
        if (sign < 0) hipLaunchKernelGGL((k_yfft<NY_, -1, R, 2 * kYTile>), g, dim3(256), 0, c->stream, Y); \
        else hipLaunchKernelGGL((k_yfft<NY_, +1, R, 2 * kYTile>), g, dim3(256), 0, c->stream, Y);          \
    } else {                                                                                               \
        if (sign < 0) hipLaunchKernelGGL((k_yfft<NY_, -1, R>), g, dim3(256), 0, c->stream, Y);             \
        else hipLaunchKernelGGL((k_yfft<NY_, +1, R>), g, dim3(256), 0, c->stream, Y);                      \
    }
    HC_DISPATCH_N(c->ny, HC_Y)
#undef HC_Y
    return 0;
}

template <class R>
int launch_c2r(hc_ctx *c, int nf, C2RArgsT<R> &X)
{
    if (nf < 1 || nf > 4) return set_err(HC_ERR_ARG, "launch_c2r: %d fields", nf);
    const View<R> V = view<R>(c);
    X.ny = c->ny; X.nz = c->nz; X.tw_nl = V.tw_nl; X.tw_nx = V.tw_nx;
    const int pts = c->tune.c2r_pts;
    int rows = pts / c->nx / nf > 0 ? pts / c->nx / nf : 1;
    // grid-stride blocks: as many as the CUs hold at once (8 per CU), at most one per tile
    const long ntiles = (long)((c->ny + rows - 1) / rows) * c->nz;
    dim3 g((unsigned)std::min<long>(ntiles, 8L * c->ncu));
#define HC_C1(NX_, P_)                                                                                      \
    if (nf == 4) hipLaunchKernelGGL((k_x_c2r<NX_, 4, R, P_>), g, dim3(kRhsThreads), 0, c->stream, X);      \
    else if (nf == 3) hipLaunchKernelGGL((k_x_c2r<NX_, 3, R, P_>), g, dim3(kRhsThreads), 0, c->stream, X); \
    else if (nf == 2) hipLaunchKernelGGL((k_x_c2r<NX_, 2, R, P_>), g, dim3(kRhsThreads), 0, c->stream, X); \
    else hipLaunchKernelGGL((k_x_c2r<NX_, 1, R, P_>), g, dim3(kRhsThreads), 0, c->stream, X);
#define HC_C(NX_)                                 \
    if (pts == 512) { HC_C1(NX_, 512) }           \
    else if (pts == 2048) { HC_C1(NX_, 2048) }    \
    else { HC_C1(NX_, 1024) }
    HC_DISPATCH_N(c->nx, HC_C)
#undef HC_C
#undef HC_C1
    return 0;
}

// x R2C of nf (1..4) fields in one launch, levels k0 .. k0+nk-1 (nk < 0:
// to the top)
template <class R>
int launch_r2c_n(hc_ctx *c, int nf, const R *const *src, cplx<R> *const *dst, int k0, int nk)
{
    if (nf < 1 || nf > 4) return set_err(HC_ERR_ARG, "launch_r2c_n: %d fields", nf);
    if (nk < 0) nk = c->nz - k0;
    if (nk <= 0) return 0;
    const View<R> V = view<R>(c);
    R2CFields<R> F;
    const size_t pr = (size_t)k0 * c->nx * c->ny, pc = (size_t)k0 * c->nl * c->ny;
    for (int q = 0; q < 4; ++q) {
        F.src[q] = q < nf ? src[q] + pr : nullptr;
        F.dst[q] = q < nf ? dst[q] + pc : nullptr;
    }
    const int rows = rows_for(c->nx);
    // grid-stride blocks: as many per field as the CUs hold at once (8 per
    // CU over the fields), at most one per tile
    const long ntiles = (long)((c->ny + rows - 1) / rows) * nk;
    dim3 g((unsigned)std::min<long>(ntiles, std::max<long>(1, 8L * c->ncu / nf)), 1, nf);
    R norm = R(1.0 / ((double)c->nx * (double)c->ny));
#define HC_R(NX_) hipLaunchKernelGGL((k_x_r2c<NX_, R>), g, dim3(kRhsThreads), 0, c->stream, F, c->ny, nk, \
                                     (const cplx<R> *)V.tw_nl, (const cplx<R> *)V.tw_nx, norm);
    HC_DISPATCH_N(c->nx, HC_R)
#undef HC_R
    return 0;
}
template <class R>
int launch_r2c(hc_ctx *c, const R *src, cplx<R> *dst)
{
    const R *s1[1] = {src};
    cplx<R> *d1[1] = {dst};
    return launch_r2c_n<R>(c, 1, s1, d1);
}

template int launch_yfft<double2>(hc_ctx *, int, int, double2 *, double2 *, double2 *, double2 *);
template int launch_yfft<float2>(hc_ctx *, int, int, float2 *, float2 *, float2 *, float2 *);
template int launch_c2r<double>(hc_ctx *, int, C2RArgsT<double> &);
template int launch_c2r<float>(hc_ctx *, int, C2RArgsT<float> &);
template int launch_r2c<double>(hc_ctx *, const double *, double2 *);
template int launch_r2c_n<double>(hc_ctx *, int, const double *const *, double2 *const *, int, int);
template int launch_r2c_n<float>(hc_ctx *, int, const float *const *, float2 *const *, int, int);
template int launch_r2c<float>(hc_ctx *, const float *, float2 *);
