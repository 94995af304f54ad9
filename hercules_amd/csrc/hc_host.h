// Host-side state of one channel (hc_ctx) and the interfaces between the
// translation units of the library: every kernel family is compiled in its
// own unit (hc_xfer.hip, hc_rhs.hip, hc_rhs_t.hip, hc_columns.hip,
// hc_spec.hip, hc_spec_col.hip, hc_z6.hip, hc_adj.hip) and launched through
// the host functions declared here; hc_api.hip holds the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/hercules_amd.h"
#include "hc_columns.h"

// error handling: the message of the last failing call on this thread
int set_err(int code, const char *fmt, ...);
#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return set_err(HC_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// Low-wavenumber blocks of the mixed variant solved in FP64: l < kColWaves *
// HC_LOWK_L and |ky index| <= HC_LOWK_M.
#ifndef HC_LOWK_L
#define HC_LOWK_L 1
#endif
#ifndef HC_LOWK_M
#define HC_LOWK_M 8
#endif
constexpr double kPi = 3.1415926535897932;  // src/module_parameters.f90:32
constexpr double kRkc1[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};  // :210-212
constexpr double kRkc2[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};
constexpr double kRkc3[3] = {8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0};

// Device buffers of one stored precision: fields, right-hand-side
// ping-pongs, the four half spectra, and the tables the row kernels read.
template <class R>
struct View {
    R *u, *v, *w, *p, *t;
    R *fb[2][3];
    R *sb[2];
    cplx<R> *s[4];
    R *lev, *symx, *symy;
    cplx<R> *tw_nl, *tw_nx, *tw_ny;
    // spectral path (cds = 3): field and right-hand-side spectra, product
    // spectra (sp[0..3] are s[0..3]), a physical scratch field
    cplx<R> *cu, *cv, *cw, *cp, *ct, *cff, *cgg, *chh, *css, *sp[8];
    R *ptmp;
};

struct hc_ctx {
    hc_params prm;
    // mixed precision (prm.precision = 1): fields, spectra and row kernels in
    // FP32 (m32); column blocks with l < 8 lo_l and |m| <= lo_m solved in FP64
    bool mixed = false;
    View<float> m32 = {};
    // launch shapes chosen by hc_autotune (the defaults otherwise)
    struct {
        int rgroup = 0;     // RHS tile order group (0: largest of 16, 8, 4, 1 dividing ny/TY)
        int ytile = 0;      // y transform columns per block (0: 8 for FP64, 16 for FP32)
        int c2r_pts = 1024; // inverse-x points per field per block
        int col_remap = 0;  // XCD-contiguous column block order
        int rhs_xt = 0;     // FD momentum right-hand side: 0 whole rows with the x R2C fused, 1-4 x-y tiles
                            // (32x8, 64x8, 32x16, 64x4) and a separate in-place R2C
    } tune;
    int lo_l = 0, lo_m = 0;
    int nx, ny, nz, nl, Lcol;
    double dx, dy;
    int device;
    int ncu = 256;  // compute units of the device
    hipStream_t stream;
    double *u, *v, *w, *p, *t;
    double *fb[2][3];  // ff, gg, hh ping-pong
    double *sb[2];     // ss ping-pong
    int fcur, scur;
    double2 *s[4];
    double *lev, *symx, *symy, *scal, *part, *stage;
    // 6th-order z operators (zorder = 6): stencil starts and weights, and
    // the interpolants the right-hand-side kernels read (k_zinterp6)
    int *zst = nullptr;
    double *zwt = nullptr;
    double *zbuf[4] = {nullptr, nullptr, nullptr, nullptr};
    // direct banded solves: band rows of ZO_C2_DD, ZO_C2_DN, ZO_F2, ZO_P
    // (kBandRow per level) and the U-row scratch (12 nz nl ny doubles)
    double *zband = nullptr, *zscr = nullptr;
    std::vector<double> h_zband;
    // calculate_mean_fields: running sums (HC_MEAN_COUNT fields), the x / y
    // interpolants of u, v taken at the start of a step's last substep (FD),
    // and the number of steps summed
    double *mean[HC_MEAN_COUNT] = {};
    double *mean_uhx = nullptr, *mean_vhy = nullptr;
    bool mean_on = false;
    long mean_steps = 0;
    std::vector<int> h_zst;
    std::vector<double> h_zwt;
    double *dmp_s, *dmp_m;  // Rayleigh damping: sigz per level, plane means per (field, level)
    double2 *spund;         // spectral damping: undamped spectra of u, v, w, p, t on levels >= kdamp0
    // spectral path (cds = 3): field and right-hand-side spectra, product
    // spectra (sp[0..3] are s[0..3]), a physical scratch field
    double2 *cu = nullptr, *cv = nullptr, *cw = nullptr, *cp = nullptr, *ct = nullptr;
    double2 *cff = nullptr, *cgg = nullptr, *chh = nullptr, *css = nullptr, *sp[8] = {};
    double *ptmp = nullptr;
    bool spec_dirty = true;  // physical fields changed since the spectra were formed
    int kdamp0;             // first damped level (nz when isdamp = 0)
    double2 *tw_nl, *tw_nx, *tw_ny;
    double aU0, cUe, aT0, cTe;
    double time_sim, got;
    int stps;
    std::vector<double> h_lev;
    std::vector<void *> allocs;
    // discrete adjoint: adjoints of u, v, w, p; of f, g, h (ping-pong); of
    // the Helmholtz sources q (physical scratch)
    double *adj[4] = {};
    double *adf[2][3] = {};
    double *adq[3] = {};
    double *adfb[3] = {}, *adface[8] = {};  // scratch of the transposed right-hand side
    int adcur = 0;
    bool big_lds = false;  // column kernels may use more than 64 KiB of dynamic LDS
    // optional per-kernel timing with HIP events on the context stream
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_kid;  // kernel id of each recorded (start, stop) pair
    size_t ev_used = 0;
    double prof_ms[HC_KERNEL_COUNT] = {};
    long prof_n[HC_KERNEL_COUNT] = {};
    // divergence screen (screen_div_error after every substep when screen =
    // 1): scr[0] div_hist_max, scr[4 + 4 s ..] the record of slot s (div_avg,
    // div_max, div_hist_max, NaN flag; slot 0 on request, 1..3 the substeps);
    // scr_rows / scr_blk the per-row and per-block partials
    int screen = 0;
    double *scr = nullptr, *scr_rows = nullptr, *scr_blk = nullptr;
};

// Steps taken for tuning or for a gradient's forward run leave the
// statistics and the divergence screen of the context untouched.
struct QuietSteps {
    hc_ctx *c;
    bool mean;
    int screen;
    explicit QuietSteps(hc_ctx *c_) : c(c_), mean(c_->mean_on), screen(c_->screen)
    {
        c->mean_on = false;
        c->screen = 0;
    }
    ~QuietSteps()
    {
        c->mean_on = mean;
        c->screen = screen;
    }
};

// Record a start (begin=true) or stop event for kernel `kid` when profiling.
inline void prof_mark(hc_ctx *c, int kid, bool begin)
{
    if (!c->prof) return;
    if (begin) {
        if (c->ev_used + 2 > c->ev_pool.size()) {
            size_t add = c->ev_pool.size() < 256 ? 256 : c->ev_pool.size();
            for (size_t q = 0; q < add; ++q) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) { c->prof = false; return; }
                c->ev_pool.push_back(e);
            }
        }
        c->ev_kid.push_back(kid);
        (void)hipEventRecord(c->ev_pool[c->ev_used], c->stream);
    } else {
        (void)hipEventRecord(c->ev_pool[c->ev_used + 1], c->stream);
        c->ev_used += 2;
    }
}
#define PROF_BEGIN(c, kid) prof_mark(c, kid, true)
#define PROF_END(c, kid) prof_mark(c, kid, false)

inline bool fft_len_ok(int n)
{
    return n == 16 || n == 32 || n == 48 || n == 64 || n == 96 || n == 128 || n == 192 || n == 256 || n == 384 ||
           n == 512;
}
// rows of a right-hand-side tile for nx (rhs2_ty for nx <= 256, rhs_ty above):
// ny must be a multiple of it
inline int rhs_tile_rows(int nx)
{
    return nx == 16 ? 16 : nx == 32 ? 32 : nx == 48 ? 16 : nx == 64 ? 16 : nx == 96 ? 8 : nx == 128 ? 8 : nx <= 256 ? 4 : 1;
}

inline int col_levels(int nz)
{
    int L = (nz + 63) / 64;
    if (L <= 2) return 2;
    if (L <= 3) return 3;
    if (L <= 4) return 4;
    if (L <= 5) return 5;
    if (L <= 8) return 8;
    if (L <= 16) return 16;
    return -1;
}

int dalloc(hc_ctx *c, void **p, size_t bytes);
// Kernel arguments of the context's current state (FP64 tables and fields).
StepArgs base_args(hc_ctx *c);

template <class R> View<R> view(hc_ctx *c);
template <> inline View<double> view<double>(hc_ctx *c)
{
    View<double> V;
    V.u = c->u; V.v = c->v; V.w = c->w; V.p = c->p; V.t = c->t;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 3; ++b) V.fb[a][b] = c->fb[a][b];
    V.sb[0] = c->sb[0]; V.sb[1] = c->sb[1];
    for (int q = 0; q < 4; ++q) V.s[q] = c->s[q];
    V.lev = c->lev; V.symx = c->symx; V.symy = c->symy;
    V.tw_nl = c->tw_nl; V.tw_nx = c->tw_nx; V.tw_ny = c->tw_ny;
    V.cu = c->cu; V.cv = c->cv; V.cw = c->cw; V.cp = c->cp; V.ct = c->ct;
    V.cff = c->cff; V.cgg = c->cgg; V.chh = c->chh; V.css = c->css;
    for (int q = 0; q < 8; ++q) V.sp[q] = c->sp[q];
    V.ptmp = c->ptmp;
    return V;
}
template <> inline View<float> view<float>(hc_ctx *c) { return c->m32; }

// Row-kernel arguments in the stored precision R: the pointers of V and the
// scalars of A rounded to R (the FP64 device scalars stay shared).
template <class R>
inline StepArgsT<R> rows_args(const StepArgs &A, const View<R> &V)
{
    StepArgsT<R> B;
    std::memset(&B, 0, sizeof B);
    B.u = V.u; B.v = V.v; B.w = V.w; B.p = V.p; B.t = V.t;
    B.s0 = V.s[0]; B.s1 = V.s[1]; B.s2 = V.s[2]; B.s3 = V.s[3];
    B.lev = V.lev; B.symx = V.symx; B.symy = V.symy;
    B.tw_nl = V.tw_nl; B.tw_nx = V.tw_nx; B.tw_ny = V.tw_ny;
    B.cu = V.cu; B.cv = V.cv; B.cw = V.cw; B.cp = V.cp; B.ct = V.ct;
    B.cff = V.cff; B.cgg = V.cgg; B.chh = V.chh; B.css = V.css;
    B.scal = A.scal;
    B.nx = A.nx; B.ny = A.ny; B.nz = A.nz; B.nl = A.nl;
    B.ochannel = A.ochannel; B.dp_opt = A.dp_opt; B.isscalar = A.isscalar; B.use_f0 = A.use_f0;
    const double sc[] = {A.dx, A.dy, A.dx2, A.dy2, A.dt, A.nu, A.alpha, A.got, A.fc, A.ug, A.vg, A.u_mrf, A.t_ref,
                         A.tbot, A.ttop, A.ra, A.rb, A.rc, A.dpx, A.norm, A.c24x, A.c12x2, A.c24y, A.c12y2, A.idx,
                         A.idx2, A.idy, A.idy2, A.idt, A.aU0, A.cUe, A.aT0, A.cTe, A.coe, A.gdz_top};
    R *dst[] = {&B.dx, &B.dy, &B.dx2, &B.dy2, &B.dt, &B.nu, &B.alpha, &B.got, &B.fc, &B.ug, &B.vg, &B.u_mrf, &B.t_ref,
                &B.tbot, &B.ttop, &B.ra, &B.rb, &B.rc, &B.dpx, &B.norm, &B.c24x, &B.c12x2, &B.c24y, &B.c12y2, &B.idx,
                &B.idx2, &B.idy, &B.idy2, &B.idt, &B.aU0, &B.cUe, &B.aT0, &B.cTe, &B.coe, &B.gdz_top};
    static_assert(sizeof sc / sizeof sc[0] == sizeof dst / sizeof dst[0], "scalar lists differ");
    for (size_t q = 0; q < sizeof sc / sizeof sc[0]; ++q) *dst[q] = R(sc[q]);
    B.kdamp0 = A.kdamp0; B.dmask = A.dmask; B.tbc_top = A.tbc_top;
    B.klo = A.klo; B.nlev = A.nlev;
    B.issk = A.issk; B.lc = A.lc; B.mc = A.mc;
    B.ntile = A.ntile;
    B.rgroup = A.rgroup;
    // the high-order z tables (FP64 in every variant)
    B.zorder = A.zorder; B.zn1 = A.zn1; B.zn2 = A.zn2;
    B.zst = A.zst; B.zwt = A.zwt; B.zband = A.zband; B.zscr = A.zscr; B.zspc = A.zspc;
    B.spund = A.spund;
    return B;
}
template <>
inline StepArgsT<double> rows_args<double>(const StepArgs &A, const View<double> &V)
{
    StepArgs B = A;
    B.u = V.u; B.v = V.v; B.w = V.w; B.p = V.p; B.t = V.t;
    B.s0 = V.s[0]; B.s1 = V.s[1]; B.s2 = V.s[2]; B.s3 = V.s[3];
    return B;
}

// ---- template dispatch ----------------------------------------------------
#define HC_DISPATCH_N(n, MACRO)                                                          \
    switch (n) {                                                                          \
    case 16: MACRO(16); break;                                                            \
    case 32: MACRO(32); break;                                                            \
    case 48: MACRO(48); break;                                                            \
    case 64: MACRO(64); break;                                                            \
    case 96: MACRO(96); break;                                                            \
    case 128: MACRO(128); break;                                                          \
    case 192: MACRO(192); break;                                                          \
    case 256: MACRO(256); break;                                                          \
    case 384: MACRO(384); break;                                                          \
    case 512: MACRO(512); break;                                                          \
    default: return set_err(HC_ERR_UNSUPPORTED, "unsupported transform length %d", n);   \
    }
#define HC_DISPATCH_L(L_, MACRO)                                                         \
    switch (L_) {                                                                         \
    case 2: MACRO(2); break;                                                              \
    case 3: MACRO(3); break;                                                              \
    case 4: MACRO(4); break;                                                              \
    case 5: MACRO(5); break;                                                              \
    case 8: MACRO(8); break;                                                              \
    case 16: MACRO(16); break;                                                            \
    default: return set_err(HC_ERR_UNSUPPORTED, "unsupported nz");                       \
    }

inline int rows_for(int nx) { return nx >= kRowPoints ? 1 : kRowPoints / nx; }

// LDS of the column kernels: `nfields` sub-tiles of nz x kColWaves complex
// (elem bytes each) when they fit in kColLds bytes, else one.
constexpr size_t kColLds = 160 * 1024;
inline int col_tiles(const hc_ctx *c, int nfields, size_t elem = sizeof(double2))
{
    const size_t one = (size_t)c->nz * kColWaves * elem;
    return (c->big_lds && one * nfields <= kColLds) ? nfields : 1;
}
inline size_t col_lds(const hc_ctx *c, int nfields = 1, size_t elem = sizeof(double2))
{
    return (size_t)col_tiles(c, nfields, elem) * c->nz * kColWaves * elem;
}

template <class T> struct nd { using type = T; };


// ---- launchers of the kernel families --------------------------------------
// x / y transforms (hc_xfer.hip)
template <class C>
int launch_yfft(hc_ctx *c, int sign, int nf, C *f0, typename nd<C>::type *f1, typename nd<C>::type *f2,
                typename nd<C>::type *f3);
template <class R> int launch_c2r(hc_ctx *c, int nf, C2RArgsT<R> &X);
template <class R> int launch_r2c(hc_ctx *c, const R *src, cplx<R> *dst);
template <class R>
int launch_r2c_n(hc_ctx *c, int nf, const R *const *src, cplx<R> *const *dst, int k0 = 0, int nk = -1);
// whether tile shape 1-4 of the x-y tiled momentum right-hand side divides the grid
bool rhs_xt_fits(int shape, int nx, int ny);

// Momentum right-hand side of one substep with the x R2C of qu, qv, qw
// (hc_rhs.hip): Rayleigh damping means first when isdamp = 1, then the
// tiled, damped or 6th-order kernel.
template <class R> int rhs_momentum(hc_ctx *c, const StepArgs &A, StepArgsT<R> &B);
// Temperature right-hand side with the x R2C of qt, and the interpolants of
// the mean statistics (hc_rhs_t.hip).
template <class R> int rhs_temperature(hc_ctx *c, StepArgsT<R> &B);
template <class R> int mean_interp(hc_ctx *c, StepArgsT<R> &B);

// Column passes over the spectra s[0..3] of V (hc_columns.hip): kind 0
// k_column_mom, 1 k_column_t, 2 k_column_project, 3 k_column_solve (eq).
template <class R> int launch_columns(hc_ctx *c, int kind, StepArgs A, const View<R> &V, int nfields, int eq = 0);
hipError_t columns_set_lds(int L, int bytes);

// Spectral path (hc_spec.hip, hc_spec_col.hip)
int spec_init(hc_ctx *c);
template <class R> int launch_spec_columns(hc_ctx *c, int kind, StepArgs A, const View<R> &V);
int substep_sp(hc_ctx *c, int t_its);
template <class R> int spec_inverse(hc_ctx *c, int nf, const cplx<R> *const *spec, R *const *phys);
// screen_div_error_spec: the spectral divergence of (c_u, c_v, w) in
// physical space, into the scratch field ptmp of the stored precision
int spec_divergence(hc_ctx *c);
// FP64 spectral helpers of the per-stage entry points (hc_spec.hip): the
// products of the momentum (which = 0) or temperature (1) right-hand side and
// their spectra in sp[], and real_2_cmplx of one field
int spec_products(hc_ctx *c, const StepArgs &A, int which);
int spec_forward64(hc_ctx *c, const double *phys, double2 *spec);
int spec_damp_one(hc_ctx *c, double2 *f);
hipError_t spec_set_lds(int L, int bytes);

// 6th- and 4th-order z operators (hc_z6.hip): the interpolant pre-pass of
// the right-hand side, and the column passes (kind as in launch_columns /
// launch_spec_columns) over spectra stored as R
template <class RIO> struct SpecIO;  // hc_spectral.h
template <class R> void zinterp6(hc_ctx *c, StepArgsT<R> &B);
template <class R> void launch_col6(int kind, hipStream_t st, const StepArgs &A, const ColIO<R> &io, int eq);
template <class R> void launch_spec_col6(int kind, hipStream_t st, const StepArgs &A, const SpecIO<R> &io);
hipError_t z6_set_lds(int L, int bytes);

// Discrete adjoint of one FD substep (hc_adj.hip)
int adjoint_substep(hc_ctx *c, int t_its);

// small kernels (hc_api.hip)
void dpx_update_launch(hc_ctx *c);
// calculate_ubar of an FP64 interior field into the device scalar ubar
int ubar_of(hc_ctx *c, const double *u);
// FP64 field transfers between a caller's array (host or device, `layout`)
// and a device interior array
int put_f64(hc_ctx *c, double *dev, const double *src, int layout);
int get_f64(hc_ctx *c, const double *dev, double *dst, int layout);

// Divergence screen (hc_screen.hip): queue screen_div_error of the current
// fields into record slot `slot`; fin = 1 updates div_hist_max.
int screen_alloc(hc_ctx *c);
int screen_queue(hc_ctx *c, int slot, int fin);
