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


struct BandSys {
    double2 *f[2];  // right-hand sides (k, j, l), solved in place
    double *scr;    // U rows: PB + 1 doubles per level and column, level-major
    int nr;         // right-hand sides sharing the matrix: 1 or 2
    int eq;         // 0 u/v, 1 w, 2 p, 3 t
    int fix;        // apply the (0,0)-mode wall data / delnull row
    int trunc;      // zero the columns outside the 2/3 rule on output
};
struct BandArgs {
    BandSys s[2];  // one system per blockIdx.y
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

template <int PB>
__global__ __launch_bounds__(256) void k_z6_band(StepArgs A, BandArgs B)
{
    const BandSys S = B.s[blockIdx.y];
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
    // ---- forward elimination
    for (int k = 0; k < n; ++k) {
        double a[2 * PB + 1];
#pragma unroll
        for (int d = -PB; d <= PB; ++d) a[d + PB] = coff * bt[(size_t)k * kBandRow + d];
        a[PB] += cdiag;
        const size_t o = (size_t)k * plane + col;
        double2 r[2];
        r[0] = S.f[0][o];
        r[1] = nr > 1 ? S.f[1][o] : make_double2(0.0, 0.0);
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
        S.f[0][o] = r[0];
        if (nr > 1) S.f[1][o] = r[1];
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
    // ---- back substitution
    const int mw = jm < ny / 2 ? jm : jm - ny;
    const bool zero_out = S.trunc && !(l < A.lc && mw < A.mc && mw > -A.mc);
    double2 xs[2][PB];  // x at k+1 .. k+PB
#pragma unroll
    for (int c = 0; c < PB; ++c) xs[0][c] = xs[1][c] = make_double2(0.0, 0.0);
    if (face) {  // w at the top wall face
        const size_t o = (size_t)(nz - 1) * plane + col;
        S.f[0][o] = make_double2(0.0, 0.0);
        if (nr > 1) S.f[1][o] = make_double2(0.0, 0.0);
    }
    for (int k = n - 1; k >= 0; --k) {
        const size_t o = (size_t)k * plane + col;
        const double *srow = S.scr + (size_t)k * (PB + 1) * plane + col;
        double u[PB];
#pragma unroll
        for (int c = 0; c < PB; ++c) u[c] = srow[(size_t)c * plane];
        const double dk = srow[(size_t)PB * plane];
        double2 r[2];
        r[0] = S.f[0][o];
        r[1] = nr > 1 ? S.f[1][o] : make_double2(0.0, 0.0);
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
        S.f[0][o] = zero_out ? make_double2(0.0, 0.0) : r[0];
        if (nr > 1) S.f[1][o] = zero_out ? make_double2(0.0, 0.0) : r[1];
    }
}
