// The divergence screen of time_advancement_fd / _sp: screen_div_error
// (src/module_io.f90:85-206), called after every substep with isfinal = 1
// (src/module_navier_stokes.f90:344-350) and at start-up with isfinal = 0.
//
// The reference scans the points in (k, j, i) order and keeps
//   div_avg      = sum |d| / (nx ny nz)
//   div_max      : starts at 0; at each point with |d| > div_max, div_max = d
//   div_hist_max : the same rule, kept across calls, updated when isfinal = 1
// and stops on the first NaN in u, v or w (screen_div_error_spec,
// src/module_spectral.f90:899-1007, stores |d|: plain maxima).  div_max and div_hist_max are
// therefore SIGNED and depend on the scan order: after a negative d is stored
// the next point always replaces it.  The screen reproduces that sequence
// exactly in two launches:
//   k_div_rows    all points in parallel: per block (16 rows of one level)
//                 the sum of |d| and the NaN flag, per row and per block the
//                 largest |d|;
//   k_div_signed  one wavefront walks the blocks in scan order and opens only
//                 those whose largest |d| can change a state (|d| > state, or
//                 any state < 0), then their rows, then their points, taking
//                 the first qualifying point of each 64 by a ballot.  Every
//                 state change happens at a point it visits in order, so the
//                 result equals the sequential scan; the untouched blocks cost
//                 one load of their maximum.
#include "hc_host.h"

#include <mutex>

namespace {
constexpr int kRowsPerBlock = 16;
constexpr int kScrLen = 32;

__global__ void k_scr_zero(double *f, int n)
{
    if ((int)threadIdx.x < n) f[threadIdx.x] = 0.0;
}

template <int CDS, class R>
__global__ __launch_bounds__(256) void k_div_rows(StepArgsT<R> F, StepArgs D, double *rowmax, double *blk)
{
    __shared__ double rs[4], rm[4];
    __shared__ int rn[4];
    const int k = blockIdx.y, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int gx = gridDim.x, b = k * gx + blockIdx.x;
    double s = 0.0, bmax = 0.0;
    int nan = 0;
    for (int r = wv; r < kRowsPerBlock; r += 4) {
        const int j = blockIdx.x * kRowsPerBlock + r;
        if (j >= D.ny) break;
        double m = 0.0;
        for (int i = lane; i < D.nx; i += 64) {
            const double d = div_at<CDS, R>(F, D, i, j, k);
            s += fabs(d);
            m = fmax(m, fabs(d));
            const size_t o = ((size_t)k * D.ny + j) * D.nx + i;
            const R u0 = F.u[o], v0 = F.v[o], w0 = F.w[o];
            if (u0 != u0 || v0 != v0 || w0 != w0) nan = 1;
        }
        for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
        if (lane == 0) rowmax[(size_t)k * D.ny + j] = m;
        bmax = fmax(bmax, m);
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        nan |= __shfl_xor(nan, o, 64);
    }
    if (lane == 0) { rs[wv] = s; rm[wv] = bmax; rn[wv] = nan; }
    __syncthreads();
    if (threadIdx.x == 0) {
        blk[3 * (size_t)b] = (rs[0] + rs[1]) + (rs[2] + rs[3]);
        blk[3 * (size_t)b + 1] = fmax(fmax(rm[0], rm[1]), fmax(rm[2], rm[3]));
        blk[3 * (size_t)b + 2] = (double)(rn[0] | rn[1] | rn[2] | rn[3]);
    }
}

// State of the scan: s1 = div_max, s2 = div_hist_max (updated when fin).
// screen_div_error_spec stores |d| instead of d (src/module_spectral.f90:931-946),
// so with absval the states are plain running maxima.
struct ScanState {
    double s1, s2;
    bool fin, absval;
    // can a value of magnitude a change a state?
    HC_DEV bool hit(double a) const { return a > s1 || (fin && a > s2); }
    HC_DEV void take(double d)
    {
        const double a = fabs(d), keep = absval ? a : d;
        if (a > s1) s1 = keep;
        if (fin && a > s2) s2 = keep;
    }
};

// one row of points in scan order (i ascending), 64 at a time
template <int CDS, class R>
HC_DEV void scan_row(const StepArgsT<R> &F, const StepArgs &D, int j, int k, ScanState &S, int lane)
{
    for (int i0 = 0; i0 < D.nx; i0 += 64) {
        const int i = i0 + lane;
        const double d = i < D.nx ? div_at<CDS, R>(F, D, i, j, k) : 0.0;
        const double a = i < D.nx ? fabs(d) : -1.0;
        int pos = -1;
        while (true) {
            const unsigned long long m = __ballot(lane > pos && S.hit(a));
            if (!m) break;
            const int L = __ffsll((long long)m) - 1;
            S.take(__shfl(d, L, 64));
            pos = L;
        }
    }
}

template <int CDS, class R>
__global__ __launch_bounds__(64) void k_div_signed(StepArgsT<R> F, StepArgs D, const double *rowmax, const double *blk,
                                                   int nblk, int gx, double *state, double *rec, int fin)
{
    // state[0]: div_hist_max before the call (updated in place when fin)
    const int lane = threadIdx.x;
    // fixed-order sum of the block sums, NaN flag
    double s = 0.0;
    int nan = 0;
    for (int b = lane; b < nblk; b += 64) {
        s += blk[3 * (size_t)b];
        nan |= blk[3 * (size_t)b + 2] != 0.0;
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        nan |= __shfl_xor(nan, o, 64);
    }
    ScanState S{0.0, state[0], fin != 0, CDS == 3};
    if (!nan) {
        constexpr int U = 8;
        for (int b0 = 0; b0 < nblk; b0 += 64 * U) {
            double bm[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int b = b0 + u * 64 + lane;
                bm[u] = b < nblk ? blk[3 * (size_t)b + 1] : -1.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int pos = -1;
                while (true) {
                    const unsigned long long m = __ballot(lane > pos && bm[u] >= 0.0 && S.hit(bm[u]));
                    if (!m) break;
                    const int L = __ffsll((long long)m) - 1;
                    const int b = b0 + u * 64 + L, k = b / gx, j0 = (b - k * gx) * kRowsPerBlock;
                    // the rows of block b in order
                    const int nr = min(kRowsPerBlock, D.ny - j0);
                    const double rmx = lane < nr ? rowmax[(size_t)k * D.ny + j0 + lane] : -1.0;
                    int rpos = -1;
                    while (true) {
                        const unsigned long long rm = __ballot(lane > rpos && rmx >= 0.0 && S.hit(rmx));
                        if (!rm) break;
                        const int r = __ffsll((long long)rm) - 1;
                        scan_row<CDS, R>(F, D, j0 + r, k, S, lane);
                        rpos = r;
                    }
                    pos = L;
                }
            }
        }
    }
    if (lane == 0) {
        rec[0] = s / ((double)D.nx * D.ny * D.nz);
        rec[1] = S.s1;
        rec[2] = nan ? state[0] : S.s2;
        rec[3] = nan;
        if (fin && !nan) state[0] = S.s2;
    }
}

__global__ void k_scr_set(double *f, double v)
{
    if (threadIdx.x == 0) f[0] = v;
}

// Partials of one screen: per-row maxima (ny nz), per-block sums, maxima and
// NaN flags (3 gx nz), where gx = ceil(ny / 16).
struct ScreenBufs {
    double *rows, *blk, *state, *rec;
};
inline size_t screen_rows_len(int ny, int nz) { return (size_t)ny * nz; }
inline size_t screen_blk_len(int ny, int nz) { return 3 * (size_t)((ny + kRowsPerBlock - 1) / kRowsPerBlock) * nz; }

template <int CDS, class R>
void launch_screen(hc_ctx *c, const StepArgsT<R> &F, const StepArgs &D, const ScreenBufs &B, int fin)
{
    const int gx = (c->ny + kRowsPerBlock - 1) / kRowsPerBlock, nblk = gx * c->nz;
    hipLaunchKernelGGL((k_div_rows<CDS, R>), dim3(gx, c->nz), dim3(256), 0, c->stream, F, D, B.rows, B.blk);
    hipLaunchKernelGGL((k_div_signed<CDS, R>), dim3(1), dim3(64), 0, c->stream, F, D, B.rows, B.blk, nblk, gx,
                       B.state, B.rec, fin);
}

// Queue a screen of the context's current fields with the partials in B.
int queue_screen(hc_ctx *c, const ScreenBufs &B, int fin)
{
    StepArgs A = base_args(c);
    if (c->prm.cds == 3)
        if (int e = spec_divergence(c)) return e;
    if (c->mixed) {
        StepArgsT<float> F = rows_args<float>(A, c->m32);
        F.aux = c->m32.ptmp;
        if (c->prm.cds == 1) launch_screen<1, float>(c, F, A, B, fin);
        else if (c->prm.cds == 2) launch_screen<2, float>(c, F, A, B, fin);
        else launch_screen<3, float>(c, F, A, B, fin);
    } else {
        A.aux = c->ptmp;
        if (c->prm.cds == 1) launch_screen<1, double>(c, A, A, B, fin);
        else if (c->prm.cds == 2) launch_screen<2, double>(c, A, A, B, fin);
        else launch_screen<3, double>(c, A, A, B, fin);
    }
    HIPCHK(hipGetLastError());
    return 0;
}
}  // namespace

int screen_alloc(hc_ctx *c)
{
    if (c->scr) return 0;
    const int gx = (c->ny + kRowsPerBlock - 1) / kRowsPerBlock;
    int e = dalloc(c, (void **)&c->scr, kScrLen * sizeof(double));
    if (!e) e = dalloc(c, (void **)&c->scr_rows, (size_t)c->ny * c->nz * sizeof(double));
    if (!e) e = dalloc(c, (void **)&c->scr_blk, 3 * (size_t)gx * c->nz * sizeof(double));
    if (!e) {
        // div_hist_max starts at 0 (module_io's save variable)
        hipLaunchKernelGGL(k_scr_zero, dim3(1), dim3(64), 0, c->stream, c->scr, kScrLen);
        HIPCHK(hipGetLastError());
    }
    return e;
}

// Queue the screen of the current fields into record slot `slot` (4 doubles
// at c->scr + 4 + 4 slot: div_avg, div_max, div_hist_max, NaN flag).
int screen_queue(hc_ctx *c, int slot, int fin)
{
    if (int e = screen_alloc(c)) return e;
    return queue_screen(c, ScreenBufs{c->scr_rows, c->scr_blk, c->scr, c->scr + 4 + 4 * slot}, fin);
}

// ---------------------------------------------------------------------------
// The screen without allocation: its partials live in `scratch`, a
// file-scope device array that every GPU the library runs on holds until the
// process ends.  Host code never passes the array's name to a kernel (that is
// a host-side stand-in, not a device address): the device address is looked
// up once per device with hipGetSymbolAddress and kept in g_scratch.
namespace {
constexpr size_t kScratchDoubles = (size_t)1 << 19;  // 4 MiB
constexpr int kMaxDevices = 64;
__device__ __align__(256) double scratch[kScratchDoubles];
double *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_ptr(int device, double **out)
{
    if (device < 0 || device >= kMaxDevices) return set_err(HC_ERR_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(HC_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch[device] = (double *)p;
    }
    *out = g_scratch[device];
    return 0;
}
}  // namespace

extern "C" int hc_screen_div_error_noalloc(hc_ctx *c, int isfinal, double *div_avg, double *div_max,
                                           double *div_hist_max)
{
    if (!c || !div_hist_max) return set_err(HC_ERR_ARG, "null argument");
    if (isfinal != 0 && isfinal != 1) return set_err(HC_ERR_ARG, "isfinal must be 0 or 1");
    const size_t nrows = screen_rows_len(c->ny, c->nz), nblk = screen_blk_len(c->ny, c->nz);
    const size_t need = 8 + nrows + nblk;
    if (need > kScratchDoubles)
        return set_err(HC_ERR_ARG, "screen partials of %zu doubles (ny=%d, nz=%d) exceed the scratch of %zu doubles",
                       need, c->ny, c->nz, kScratchDoubles);
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != c->device) return set_err(HC_ERR_ARG, "the context lives on device %d, current is %d", c->device, dev);
    double *ptr = nullptr;
    if (int e = scratch_ptr(dev, &ptr)) return e;
    // scratch layout: [0] div_hist_max in/out, [4..7] the record, then the
    // row maxima and the block partials (both fully rewritten by each call)
    const ScreenBufs B{ptr + 8, ptr + 8 + nrows, ptr, ptr + 4};
    hipLaunchKernelGGL(k_scr_set, dim3(1), dim3(64), 0, c->stream, ptr, *div_hist_max);
    HIPCHK(hipGetLastError());
    if (int e = queue_screen(c, B, isfinal)) return e;
    double rec[4];
    HIPCHK(hipMemcpyAsync(rec, ptr + 4, sizeof rec, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (div_avg) *div_avg = rec[0];
    if (div_max) *div_max = rec[1];
    *div_hist_max = rec[2];
    if (rec[3] != 0.0) return set_err(HC_ERR_NAN, "NaN error!");
    return 0;
}
