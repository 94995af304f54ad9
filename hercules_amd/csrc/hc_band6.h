// Direct solves of the 6th-order implicit z systems (zorder = 6), one thread
// per (l, m) column.
//
// Every system of the step is real and banded,
//   u, v, t:  (1 - rc hxy) I - rc kappa Z      Z = ZO_C2_* on the levels
//   w:        (1 - rc hxy) I - rc kappa Z      Z = ZO_F2 on the interior faces
//   p:        rc (hxy I + ZO_P)                ZO_P = ZO_D o ZO_G
// and differs between columns only through the horizontal symbol hxy.  Z has
// half band 5 (7-point stencils, shifted inwards at the walls) and ZO_P half
// band 7.  Gaussian elimination without pivoting is stable for all of them
// (growth factor below 1.7 on the channel grids), so each thread factors its
// column's matrix row by row while eliminating the right-hand sides:
//   forward  k = 0..n-1: row k of the matrix from the band table, eliminated
//            against the previous PB rows of U kept in registers; U's row
//            (PB entries right of the diagonal and 1/diagonal) goes to a
//            scratch array, the eliminated right-hand side back in place;
//   backward k = n-1..0: x(k) from U's row and the PB values of x above it.
// Consecutive threads own consecutive l, so every load and store of a level
// is one coalesced line across the wavefront.  The (0,0) mode adds the wall
// data of u and t to the first right-hand side (the Crank-Nicolson implicit
// half) and replaces the first pressure row by the delnull row, as the
// iterative solves did.
#pragma once


// CI: the stored complex type of the right-hand sides (double2, or float2 in
// the mixed variant); the factorisation and the solve run in FP64.
template <class CI>
struct BandSys {
    const CI *in[2];  // right-hand sides (k, j, l)
    CI *f[2];         // solutions (may be in[]); hold the eliminated right-hand sides in between
    double *scr;    // U rows: PB + 1 doubles per level and column, level-major
    int nr;         // right-hand sides sharing the matrix: 1 or 2
    int eq;         // 0 u/v, 1 w, 2 p, 3 t
    int fix;        // apply the (0,0)-mode wall data / delnull row
    int trunc;      // zero the columns outside the 2/3 rule on output
};
template <class CI>
struct BandArgs {
    BandSys<CI> s[2];  // one system per blockIdx.y
};

// Wall data of the (0,0) mode for row k of the u (eq 0) or t (eq 3)
// equation: rc kappa times the weight of the wall datum times its value
// (bottom: -u_mrf or tbot; top: -u_mrf for a closed channel, ttop or the
// gradient for t), i.e. the implicit half of the Crank-Nicolson wall terms.
HC_DEV double z6_wall_rhs(const StepArgs &A, int eq, int op, double ck, int k)
{
    const int nz = A.nz;
    double bot, top;
    bool has_top;
    if (eq == 0) {
        bot = -A.u_mrf;
        top = -A.u_mrf;
        has_top = A.ochannel == 0;
    } else {
        bot = A.tbot;
        top = A.tbc_top ? A.gdz_top * A.lev[LV_IDZT * nz + nz - 1] : A.ttop;
        has_top = true;
    }
    const size_t e = (size_t)op * nz + k;
    const int s = A.zst[e];
    const double *w = A.zwt + e * kZW;
    double v = 0.0;
    if (s == 0) v += ck * w[0] * bot;
    const int jt = nz + 1 - s;
    if (has_top && jt < 7) v += ck * w[jt] * top;
    return v;
}

template <int PB, class CI>
__global__ __launch_bounds__(256) void k_z6_band(StepArgs A, BandArgs<CI> B)
{
    const BandSys<CI> S = B.s[blockIdx.y];
    const int nl = A.nl, ny = A.ny, nz = A.nz;
    const int ncol = nl * ny;
    const int col = blockIdx.x * 256 + threadIdx.x;
    if (col >= ncol) return;
    const int l = col % nl, jm = col / nl;
    const size_t plane = (size_t)ncol;
    const bool face = S.eq == 1, pres = S.eq == 2;
    const int n = face ? nz - 1 : nz;
    int band, op;
    double hxy, kap;
    switch (S.eq) {
    case 0:
        band = A.ochannel ? ZB_C2_DN : ZB_C2_DD;
        op = A.ochannel ? ZO_C2_DN : ZO_C2_DD;
        hxy = A.symx[SY_HU * nl + l] + A.symy[SY_HU * ny + jm];
        kap = 0.5 * A.dt * A.nu;
        break;
    case 1:
        band = ZB_F2;
        op = ZO_F2;
        hxy = A.symx[SY_HU * nl + l] + A.symy[SY_HU * ny + jm];
        kap = 0.5 * A.dt * A.nu;
        break;
    case 2:
        band = ZB_P;
        op = ZO_P;
        hxy = A.symx[SY_PX * nl + l] + A.symy[SY_PX * ny + jm];
        kap = 0.0;
        break;
    default:
        band = A.tbc_top ? ZB_C2_DN : ZB_C2_DD;
        op = A.tbc_top ? ZO_C2_DN : ZO_C2_DD;
        hxy = A.symx[SY_HT * nl + l] + A.symy[SY_HT * ny + jm];
        kap = 0.5 * A.dt * A.alpha;
        break;
    }
    const double rc = A.rc, ck = rc * kap;
    const double cdiag = pres ? rc * hxy : 1.0 - rc * hxy, coff = pres ? rc : -ck;
    const bool is00 = S.fix && l == 0 && jm == 0;
    const double *bt = A.zband + (size_t)band * nz * kBandRow + kBandMax;
    const int nr = S.nr;
    double U[PB][PB], dv[PB];
    double2 y[2][PB];
#pragma unroll
    for (int t = 0; t < PB; ++t) {
        dv[t] = 0.0;
        y[0][t] = y[1][t] = make_double2(0.0, 0.0);
#pragma unroll
        for (int c = 0; c < PB; ++c) U[t][c] = 0.0;
    }
    // The right-hand sides (forward) and the U rows with the partial
    // solutions (backward) of level k are loaded PF levels ahead into rings
    // indexed by the unrolled loop position, in the stored type, so that a
    // column's chain of levels waits less on its loads (a launch has about
    // one wavefront per SIMD).  With FP64 spectra the kernel moves its bytes
    // near the HBM rate already, and the one-level ring keeps its registers
    // (two wavefronts per SIMD); FP32 spectra take four levels.
    constexpr int PF = sizeof(CI) == 8 ? (PB >= 7 ? 2 : 4) : 1;
    const CI zero = czero_t<CI>();
    CI rq[PF][2];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        const size_t o = (size_t)u * plane + col;
        rq[u][0] = u < n ? S.in[0][o] : zero;
        rq[u][1] = (u < n && nr > 1) ? S.in[1][o] : zero;
    }
    // ---- forward elimination
    for (int k0 = 0; k0 < n; k0 += PF) {
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int k = k0 + u;
        if (k >= n) break;
        double a[2 * PB + 1];
#pragma unroll
        for (int d = -PB; d <= PB; ++d) a[d + PB] = coff * bt[(size_t)k * kBandRow + d];
        a[PB] += cdiag;
        const size_t o = (size_t)k * plane + col;
        double2 r[2];
        r[0] = ccast<double2>(rq[u][0]);
        r[1] = ccast<double2>(rq[u][1]);
        if (k + PF < n) {
            const size_t on = (size_t)(k + PF) * plane + col;
            rq[u][0] = S.in[0][on];
            if (nr > 1) rq[u][1] = S.in[1][on];
        }
        if (is00) {
            if (pres) {
                if (k == 0) {  // delnull: b0 phi(0) = 0 (src/module_poisson_solver.f90:558-564)
#pragma unroll
                    for (int d = 0; d <= 2 * PB; ++d) a[d] = 0.0;
                    a[PB] = rc * (hxy + A.lev[LV_DP00 * nz]);
                    r[0] = make_double2(0.0, 0.0);
                }
            } else if (S.eq == 0 || S.eq == 3) {
                r[0].x += z6_wall_rhs(A, S.eq, op, ck, k);
            }
        }
#pragma unroll
        for (int t = 0; t < PB; ++t) {
            const double lt = a[t] * dv[t];
#pragma unroll
            for (int c = 1; c <= PB; ++c) a[t + c] -= lt * U[t][c - 1];
            r[0].x -= lt * y[0][t].x;
            r[0].y -= lt * y[0][t].y;
            r[1].x -= lt * y[1][t].x;
            r[1].y -= lt * y[1][t].y;
        }
        const double dk = 1.0 / a[PB];
        double *srow = S.scr + (size_t)k * (PB + 1) * plane + col;
#pragma unroll
        for (int c = 1; c <= PB; ++c) srow[(size_t)(c - 1) * plane] = a[PB + c];
        srow[(size_t)PB * plane] = dk;
        S.f[0][o] = ccast<CI>(r[0]);
        if (nr > 1) S.f[1][o] = ccast<CI>(r[1]);
#pragma unroll
        for (int t = 0; t < PB - 1; ++t) {
            dv[t] = dv[t + 1];
            y[0][t] = y[0][t + 1];
            y[1][t] = y[1][t + 1];
#pragma unroll
            for (int c = 0; c < PB; ++c) U[t][c] = U[t + 1][c];
        }
        dv[PB - 1] = dk;
        y[0][PB - 1] = r[0];
        y[1][PB - 1] = r[1];
#pragma unroll
        for (int c = 0; c < PB; ++c) U[PB - 1][c] = a[PB + 1 + c];
      }
    }
    // ---- back substitution
    const int mw = jm < ny / 2 ? jm : jm - ny;
    const bool zero_out = S.trunc && !(l < A.lc && mw < A.mc && mw > -A.mc);
    double2 xs[2][PB];  // x at k+1 .. k+PB
#pragma unroll
    for (int c = 0; c < PB; ++c) xs[0][c] = xs[1][c] = make_double2(0.0, 0.0);
    if (face) {  // w at the top wall face
        const size_t o = (size_t)(nz - 1) * plane + col;
        S.f[0][o] = zero;
        if (nr > 1) S.f[1][o] = zero;
    }
    double ub[PF][PB + 1];
    CI fq[PF][2];
    auto ld_back = [&](int k, double (&uu)[PB + 1], CI (&ff)[2]) __attribute__((always_inline)) {
        const size_t o = (size_t)k * plane + col;
        const double *srow = S.scr + (size_t)k * (PB + 1) * plane + col;
#pragma unroll
        for (int c = 0; c <= PB; ++c) uu[c] = srow[(size_t)c * plane];
        ff[0] = S.f[0][o];
        ff[1] = nr > 1 ? S.f[1][o] : zero;
    };
#pragma unroll
    for (int u = 0; u < PF; ++u)
        if (n - 1 - u >= 0) ld_back(n - 1 - u, ub[u], fq[u]);
    for (int k0 = n - 1; k0 >= 0; k0 -= PF) {
#pragma unroll
      for (int q = 0; q < PF; ++q) {
        const int k = k0 - q;
        if (k < 0) break;
        const size_t o = (size_t)k * plane + col;
        double u[PB];
#pragma unroll
        for (int c = 0; c < PB; ++c) u[c] = ub[q][c];
        const double dk = ub[q][PB];
        double2 r[2];
        r[0] = ccast<double2>(fq[q][0]);
        r[1] = ccast<double2>(fq[q][1]);
        if (k - PF >= 0) ld_back(k - PF, ub[q], fq[q]);
#pragma unroll
        for (int c = 0; c < PB; ++c) {
            r[0].x -= u[c] * xs[0][c].x;
            r[0].y -= u[c] * xs[0][c].y;
            r[1].x -= u[c] * xs[1][c].x;
            r[1].y -= u[c] * xs[1][c].y;
        }
        r[0] = make_double2(r[0].x * dk, r[0].y * dk);
        r[1] = make_double2(r[1].x * dk, r[1].y * dk);
#pragma unroll
        for (int c = PB - 1; c > 0; --c) {
            xs[0][c] = xs[0][c - 1];
            xs[1][c] = xs[1][c - 1];
        }
        xs[0][0] = r[0];
        xs[1][0] = r[1];
        S.f[0][o] = zero_out ? zero : ccast<CI>(r[0]);
        if (nr > 1) S.f[1][o] = zero_out ? zero : ccast<CI>(r[1]);
      }
    }
}

// ---------------------------------------------------------------------------
// Explicit column operations of the 6th-order step, one thread per (l, m)
// column streaming along z.  A thread keeps a window of kWin levels of the
// fields that an operator reaches (levels k-7 .. k+7 around the row k being
// formed) and applies band rows (ZB_* tables, read as block-uniform values)
// to it; the fields read at the row itself pass through a delay of kBandMax
// levels.  Loads run kPF levels ahead of the window through a staging ring
// indexed by the unrolled loop position, so the wavefront never waits on a
// load issued less than kPF levels earlier.
constexpr int kWin = 2 * kBandMax + 1;

// A thread forms the rows kb .. ke-1 of its column's z chunk (blockIdx.z of
// gridDim.z equal chunks), so that a launch has several wavefronts per SIMD.
// Iterations i = kb - kBandMax, ...: window field w enters level i, point
// fields enter level i - kBandMax, and row k = i - kBandMax (kb <= k < ke) is
// formed by row(k, win, pts) with win[kBandMax + t] = level k + t (zero
// outside 0 .. wlev[f] - 1; a chunk reads kBandMax levels beyond each end).
// Fields are stored as CI (double2 or float2) and held in FP64.  A kernel
// writes only its point fields' rows or fields it does not read.
constexpr int kStreamRowsPerChunk = 32;
inline int stream_chunks(int nz) { return nz / kStreamRowsPerChunk < 1 ? 1 : (nz / kStreamRowsPerChunk > 8 ? 8 : nz / kStreamRowsPerChunk); }
template <int N, class CI = double2>
struct FieldSet {
    const CI *p[N > 0 ? N : 1];
    int n[N > 0 ? N : 1];  // levels of a window field
};
template <int NW, int NP, int PF, class CW, class CP, class Row>
HC_DEV void z6_stream(size_t plane, size_t col, int nrows, const FieldSet<NW, CW> ws, const FieldSet<NP, CP> ps,
                      Row &&row)
{
    const auto &wlev = ws.n;
    const double2 zero = make_double2(0.0, 0.0);
    const CW zw_ = czero_t<CW>();
    const CP zp_ = czero_t<CP>();
    const int csz = (nrows + (int)gridDim.z - 1) / (int)gridDim.z;
    const int kb = (int)blockIdx.z * csz, ke = kb + csz < nrows ? kb + csz : nrows;
    const int ib = kb - kBandMax;
    // the staging ring holds the stored type: widening to FP64 happens when a
    // level leaves the ring, so no conversion waits on a load in flight
    double2 win[NW][kWin];
    CW fw[NW][PF];
    CP fp[PF][NP > 0 ? NP : 1];
#pragma unroll
    for (int f = 0; f < NW; ++f)
#pragma unroll
        for (int t = 0; t < kWin; ++t) win[f][t] = zero;
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        const int li = ib + u, lp = li - kBandMax;
#pragma unroll
        for (int f = 0; f < NW; ++f) fw[f][u] = (li >= 0 && li < wlev[f]) ? ws.p[f][(size_t)li * plane + col] : zw_;
#pragma unroll
        for (int p = 0; p < NP; ++p) fp[u][p] = (lp >= kb && lp < ke) ? ps.p[p][(size_t)lp * plane + col] : zp_;
    }
    for (int i0 = ib; i0 < ke + kBandMax; i0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int i = i0 + u;
            double2 pts[NP > 0 ? NP : 1];
#pragma unroll
            for (int f = 0; f < NW; ++f) {
#pragma unroll
                for (int t = 0; t < kWin - 1; ++t) win[f][t] = win[f][t + 1];
                win[f][kWin - 1] = ccast<double2>(fw[f][u]);
            }
#pragma unroll
            for (int p = 0; p < NP; ++p) pts[p] = ccast<double2>(fp[u][p]);
            // the loads of iteration i + PF
            const int li = i + PF, lp = li - kBandMax;
#pragma unroll
            for (int f = 0; f < NW; ++f)
                fw[f][u] = (li >= 0 && li < wlev[f]) ? ws.p[f][(size_t)li * plane + col] : zw_;
#pragma unroll
            for (int p = 0; p < NP; ++p) fp[u][p] = (lp >= kb && lp < ke) ? ps.p[p][(size_t)lp * plane + col] : zp_;
            const int k = i - kBandMax;
            if (k >= kb && k < ke) row(k, win, pts);
        }
    }
}

// sum_t band(k, t) x(k + t) over the stored band
HC_DEV double2 band_dot(const double *brow, const double2 (&x)[kWin])
{
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int t = 0; t < kWin; ++t) {
        const double b = brow[t];
        acc.x += b * x[t].x;
        acc.y += b * x[t].y;
    }
    return acc;
}
HC_DEV const double *band_row(const StepArgs &A, int band, int k)
{
    return A.zband + ((size_t)band * A.nz + k) * kBandRow;
}

struct ColIdx {
    int l, jm;
    size_t col, plane;
};
HC_DEV bool col_idx(const StepArgs &A, ColIdx &c)
{
    const int ncol = A.nl * A.ny;
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= ncol) return false;
    c.l = q % A.nl;
    c.jm = q / A.nl;
    c.col = (size_t)q;
    c.plane = (size_t)ncol;
    return true;
}

// Divergence D_x u + D_y v + D_z w of a column times `scale` into the FP64
// spectrum dst (get_p_eqn_src / get_div_spec with ZO_D over the interior
// faces).
template <class CI>
__global__ __launch_bounds__(256) void k_z6s_div(StepArgs A, const CI *u, const CI *v, const CI *w, double2 *dst,
                                                 double scale)
{
    ColIdx c;
    if (!col_idx(A, c)) return;
    const int nl = A.nl, ny = A.ny, nz = A.nz;
    const double2 sdx = make_double2(A.symx[SY_DRE * nl + c.l], A.symx[SY_DIM * nl + c.l]);
    const double2 sdy = make_double2(A.symy[SY_DRE * ny + c.jm], A.symy[SY_DIM * ny + c.jm]);
    const FieldSet<1, CI> ws = {{w}, {nz - 1}};
    const FieldSet<2, CI> ps = {{u, v}, {0, 0}};
    z6_stream<1, 2, 8>(c.plane, c.col, nz, ws, ps, [&](int k, const double2 (&win)[1][kWin], const double2 (&p)[2]) __attribute__((always_inline)) {
        const double2 dz = band_dot(band_row(A, ZB_D, k), win[0]);
        const double2 d = cadd(cadd(cmul(sdx, p[0]), cmul(sdy, p[1])), dz);
        dst[(size_t)k * c.plane + c.col] = make_double2(d.x * scale, d.y * scale);
    });
}

// Projection u, v, w -= cdt grad phi (calculate_uvw, ZO_G at the interior
// faces), phi an FP64 spectrum.  mode 0 (FD substep, cdt = rc dt): the
// pressure increment phi - rc dt nu / 2 (hxy_L + ZO_P) phi (calculate_new_p)
// goes to `pinc`; mode 1 (projection of a state, cdt = 1); mode 2 (spectral
// substep, cdt = rc dt): the increment is added to c_p in `pinc` and the
// columns outside the 2/3 rule are zeroed.
template <class CI>
__global__ __launch_bounds__(256) void k_z6s_proj(StepArgs A, CI *u, CI *v, CI *w, const double2 *phi, CI *pinc,
                                                  int mode)
{
    ColIdx c;
    if (!col_idx(A, c)) return;
    const int nl = A.nl, ny = A.ny, nz = A.nz;
    const double2 sgx = make_double2(A.symx[SY_GRE * nl + c.l], A.symx[SY_GIM * nl + c.l]);
    const double2 sgy = make_double2(A.symy[SY_GRE * ny + c.jm], A.symy[SY_GIM * ny + c.jm]);
    const double slap = A.symx[SY_L * nl + c.l] + A.symy[SY_L * ny + c.jm];
    const double cdt = mode == 1 ? 1.0 : A.rc * A.dt, cnu = A.rc * 0.5 * A.dt * A.nu;
    const int mw = c.jm < ny / 2 ? c.jm : c.jm - ny;
    const bool kept = mode != 2 || (c.l < A.lc && mw < A.mc && mw > -A.mc);
    const CI zero = czero_t<CI>();
    const FieldSet<1> ws = {{phi}, {nz}};
    // the fourth point field is read only in mode 2 (c_p); u stands in otherwise
    const FieldSet<4, CI> ps = {{u, v, w, mode == 2 ? pinc : u}, {0, 0, 0, 0}};
    z6_stream<1, 4, 6>(c.plane, c.col, nz, ws, ps, [&](int k, const double2 (&win)[1][kWin], const double2 (&p)[4]) __attribute__((always_inline)) {
        const size_t o = (size_t)k * c.plane + c.col;
        const double2 ph = win[0][kBandMax];
        double2 uu = csub(p[0], cscale(cmul(sgx, ph), cdt));
        double2 vv = csub(p[1], cscale(cmul(sgy, ph), cdt));
        double2 ww = p[2];
        if (k < nz - 1) {
            const double2 g = band_dot(band_row(A, ZB_G, k), win[0]);
            ww = make_double2(ww.x - cdt * g.x, ww.y - cdt * g.y);
        }
        u[o] = kept ? ccast<CI>(uu) : zero;
        v[o] = kept ? ccast<CI>(vv) : zero;
        w[o] = kept ? ccast<CI>(ww) : zero;
        if (mode == 1) return;
        const double2 pz = band_dot(band_row(A, ZB_P, k), win[0]);
        const double2 inc = make_double2(ph.x - cnu * (slap * ph.x + pz.x), ph.y - cnu * (slap * ph.y + pz.y));
        if (mode == 0) pinc[o] = ccast<CI>(inc);
        else pinc[o] = kept ? ccast<CI>(cadd(p[3], inc)) : zero;
    });
}

// Right-hand sides of the spectral substep's Helmholtz systems
// (get_momentum_cnvdiff_spec, src/module_spectral.f90:35-504, with the
// 6th-order z operators) for u, v (blockIdx.y = 0, 1; into sp[3], sp[4]) or
// w (k_z6s_rhs_w; into sp[7]):
//   q = c + hxy c + dif Z c - cdt grad_e p + rb dt f_old + ra dt f
//   f = -u_mrf i kx c - coe (i kx X + i ky Y) + coe R
// with X, Y, R the transformed products of the component; f replaces f_old.
// The (0,0) mode adds the explicit half of the wall data of u.
template <class CI>
__global__ __launch_bounds__(256) void k_z6s_rhs_uv(StepArgs A, SpecIO<real_of<CI>> io)
{
    ColIdx c;
    if (!col_idx(A, c)) return;
    const int e = blockIdx.y;
    const int nl = A.nl, ny = A.ny, nz = A.nz;
    const double kx = A.symx[SY_GIM * nl + c.l], ky = A.symy[SY_GIM * ny + c.jm];
    const double hxy = A.rc * (A.symx[SY_HU * nl + c.l] + A.symy[SY_HU * ny + c.jm]);
    const double dif = A.rc * 0.5 * A.dt * A.nu, coe = A.coe, um = A.u_mrf, cdt = A.rc * A.dt;
    const double rb = A.rb * A.dt, ra = A.ra * A.dt, kk = e == 0 ? kx : ky;
    const bool is00 = e == 0 && c.l == 0 && c.jm == 0;
    const int band = A.ochannel ? ZB_C2_DN : ZB_C2_DD, op = A.ochannel ? ZO_C2_DN : ZO_C2_DD;
    // q goes over R_u / R_v (read at the same row by the same thread): c_u,
    // c_v stay intact for the neighbouring z chunks' windows
    CI *cs = e == 0 ? io.cu : io.cv, *fo = e == 0 ? io.cff : io.cgg, *qo = e == 0 ? io.sp[3] : io.sp[4];
    const FieldSet<1, CI> ws = {{cs}, {nz}};
    const FieldSet<5, CI> ps = {{io.sp[e == 0 ? 0 : 1], io.sp[e == 0 ? 1 : 2], io.sp[e == 0 ? 3 : 4], io.cp, fo},
                                {0, 0, 0, 0, 0}};
    z6_stream<1, 5, 4>(c.plane, c.col, nz, ws, ps, [&](int k, const double2 (&win)[1][kWin], const double2 (&p)[5]) __attribute__((always_inline)) {
        const double2 x = win[0][kBandMax];
        const double2 t = band_dot(band_row(A, band, k), win[0]);
        double2 q = make_double2(x.x + hxy * x.x + dif * t.x, x.y + hxy * x.y + dif * t.y);
        if (is00) q.x += z6_wall_rhs(A, 0, op, dif, k);
        double2 f = times_i(x, -um * kx);
        f = csub(f, times_i(p[0], coe * kx));
        f = csub(f, times_i(p[1], coe * ky));
        f = make_double2(f.x + coe * p[2].x, f.y + coe * p[2].y);
        q = csub(q, times_i(p[3], cdt * kk));
        q = make_double2(q.x + rb * p[4].x + ra * f.x, q.y + rb * p[4].y + ra * f.y);
        const size_t o = (size_t)k * c.plane + c.col;
        fo[o] = ccast<CI>(f);
        qo[o] = ccast<CI>(q);
    });
}
template <class CI>
__global__ __launch_bounds__(256) void k_z6s_rhs_w(StepArgs A, SpecIO<real_of<CI>> io)
{
    ColIdx c;
    if (!col_idx(A, c)) return;
    const int nl = A.nl, ny = A.ny, nz = A.nz;
    const double kx = A.symx[SY_GIM * nl + c.l], ky = A.symy[SY_GIM * ny + c.jm];
    const double hxy = A.rc * (A.symx[SY_HU * nl + c.l] + A.symy[SY_HU * ny + c.jm]);
    const double dif = A.rc * 0.5 * A.dt * A.nu, coe = A.coe, um = A.u_mrf, cdt = A.rc * A.dt;
    const double rb = A.rb * A.dt, ra = A.ra * A.dt;
    const FieldSet<2, CI> ws = {{io.cw, io.cp}, {nz - 1, nz}};
    const FieldSet<4, CI> ps = {{io.sp[5], io.sp[6], io.sp[7], io.chh}, {0, 0, 0, 0}};
    z6_stream<2, 4, 4>(c.plane, c.col, nz, ws, ps, [&](int k, const double2 (&win)[2][kWin], const double2 (&p)[4]) __attribute__((always_inline)) {
        const size_t o = (size_t)k * c.plane + c.col;
        const double2 x = win[0][kBandMax];
        double2 f = times_i(x, -um * kx);
        f = csub(f, times_i(p[0], coe * kx));
        f = csub(f, times_i(p[1], coe * ky));
        f = make_double2(f.x + coe * p[2].x, f.y + coe * p[2].y);
        io.chh[o] = ccast<CI>(f);
        if (k == nz - 1) return;  // the top wall face: no equation
        const double2 t = band_dot(band_row(A, ZB_F2, k), win[0]);
        const double2 g = band_dot(band_row(A, ZB_G, k), win[1]);
        double2 q = make_double2(x.x + hxy * x.x + dif * t.x - cdt * g.x, x.y + hxy * x.y + dif * t.y - cdt * g.y);
        q = make_double2(q.x + rb * p[3].x + ra * f.x, q.y + rb * p[3].y + ra * f.y);
        io.sp[7][o] = ccast<CI>(q);  // over R_w, as for u and v
    });
}

// Right-hand side of the spectral temperature system
// (get_temperature_cnvdiff_spec with the 6th-order z operator) into sp[2];
// the new s replaces the old one in c_ss; the (0,0) mode adds the explicit
// half of the wall data.
template <class CI>
__global__ __launch_bounds__(256) void k_z6s_rhs_t(StepArgs A, SpecIO<real_of<CI>> io)
{
    ColIdx c;
    if (!col_idx(A, c)) return;
    const int nl = A.nl, ny = A.ny, nz = A.nz;
    const double kx = A.symx[SY_GIM * nl + c.l], ky = A.symy[SY_GIM * ny + c.jm];
    const double hxy = A.rc * (A.symx[SY_HT * nl + c.l] + A.symy[SY_HT * ny + c.jm]);
    const double dif = A.rc * 0.5 * A.dt * A.alpha, um = A.u_mrf;
    const double rb = A.rb * A.dt, ra = A.ra * A.dt;
    const bool is00 = c.l == 0 && c.jm == 0;
    const int band = A.tbc_top ? ZB_C2_DN : ZB_C2_DD, op = A.tbc_top ? ZO_C2_DN : ZO_C2_DD;
    const FieldSet<1, CI> ws = {{io.ct}, {nz}};
    const FieldSet<4, CI> ps = {{io.sp[0], io.sp[1], io.sp[2], io.css}, {0, 0, 0, 0}};
    z6_stream<1, 4, 6>(c.plane, c.col, nz, ws, ps, [&](int k, const double2 (&win)[1][kWin], const double2 (&p)[4]) __attribute__((always_inline)) {
        const double2 x = win[0][kBandMax];
        const double2 t = band_dot(band_row(A, band, k), win[0]);
        double2 d = make_double2(x.x + hxy * x.x + dif * t.x, x.y + hxy * x.y + dif * t.y);
        if (is00) d.x += z6_wall_rhs(A, 3, op, dif, k);
        double2 f = times_i(x, -um * kx);
        f = csub(f, times_i(p[0], kx));
        f = csub(f, times_i(p[1], ky));
        f = cadd(f, p[2]);
        d = make_double2(d.x + rb * p[3].x + ra * f.x, d.y + rb * p[3].y + ra * f.y);
        const size_t o = (size_t)k * c.plane + c.col;
        io.css[o] = ccast<CI>(f);
        io.sp[2][o] = ccast<CI>(d);  // over the z flux term
    });
}
