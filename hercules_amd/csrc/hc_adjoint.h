// Discrete adjoint of the finite-difference substep (cds = 1, 2; no
// temperature, no damping).  The substep is
//   f  = F(u, v, w)                                 (convective terms, drive)
//   q  = u + dt (ra f + rb f_old) + rc dt nu/2 Lap u - rc dt G p
//   [U**, P_inc] = M FFT(q)    per (l, m) column: Helmholtz solves, divergence,
//                              pressure Poisson, projection, pressure increment
//   u' = IFFT(U**),  p' = p + IFFT(P_inc),  f_old' = f
// (src/module_navier_stokes.f90:247-355, src/module_cnvdiff.f90:35-374).
// Its adjoint runs the same transforms: for real translation-invariant maps
// IFFT M FFT, the transpose is IFFT M^H FFT with M^H the conjugate transpose
// of each column's map (k_adj_column).  The right-hand side is quadratic in
// (u, v, w); its transpose at the forward state is formed in three passes:
// the total adjoints of f, g, h (k_adj_fbar), the adjoint of every face
// interpolant at every point (k_adj_faces), and for each point the gather of
// every face interpolant, diffusion and gradient term that reads it
// (k_adj_gather).  Affine constants (wall values, u_mrf, the drive) drop out.
#pragma once

// Forward row (a, b, c) of level k of one column equation, as assemble()
// forms it.  eq: 0 u/v, 1 w, 2 p.
HC_DEV void row_fwd(const StepArgs &A, int eq, double hxy, bool is00, int k, double &a, double &b, double &c)
{
    const int nz = A.nz;
    const double rc = A.rc;
    int ta, tc, td, td00;
    switch (eq) {
    case 0: ta = LV_AU; tc = LV_CU; td = LV_DU; td00 = LV_DU00; break;
    case 1: ta = LV_AW; tc = LV_CW; td = LV_DW; td00 = LV_DW00; break;
    default: ta = LV_AP; tc = LV_CP; td = LV_DP; td00 = LV_DP00; break;
    }
    if (eq == 1 && k == nz - 1) { a = 0.0; b = 1.0; c = 0.0; return; }
    a = rc * A.lev[ta * nz + k];
    c = rc * A.lev[tc * nz + k];
    const double dg = A.lev[(is00 ? td00 : td) * nz + k];
    if (eq == 2) {
        b = rc * (hxy + dg);
        if (is00 && k == 0) { a = 0.0; c = 0.0; }
    } else {
        b = 1.0 + rc * (-hxy + dg);
    }
}

// Rows of the transposed system: a'(k) = c(k-1), b'(k) = b(k), c'(k) = a(k+1).
template <int L>
HC_DEV void assemble_T(const StepArgs &A, int eq, int l, int jm, bool is00, const ColMap &m, double (&a)[L],
                       double (&b)[L], double (&c)[L])
{
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    const int sy = eq == 2 ? SY_PX : SY_HU;
    const double hxy = A.symx[sy * nl + l] + A.symy[sy * ny + jm];
#pragma unroll
    for (int r = 0; r < L; ++r) {
        a[r] = 0.0; b[r] = 1.0; c[r] = 0.0;
        if (r < m.cnt) {
            const int k = m.k0 + r;
            double ta, tb, tc;
            row_fwd(A, eq, hxy, is00, k, ta, tb, tc);
            b[r] = tb;
            if (k >= 1) { row_fwd(A, eq, hxy, is00, k - 1, ta, tb, tc); a[r] = tc; }
            if (k <= nz - 2) { row_fwd(A, eq, hxy, is00, k + 1, ta, tb, tc); c[r] = ta; }
        }
    }
}

// value of level r+1 of this lane's column (or the next lane's first)
template <int L>
HC_DEV double2 lvl_up(const ColMap &m, const double2 (&v)[L], int r, double2 next)
{
    double2 cand = czero();
#pragma unroll
    for (int q = 0; q < L; ++q)
        if (q == r + 1) cand = v[q];
    return r + 1 < m.cnt ? cand : next;
}
template <int L>
HC_DEV double2 lvl_dn(const double2 (&v)[L], int r, double2 prev)
{
    double2 cand = czero();
#pragma unroll
    for (int q = 0; q < L; ++q)
        if (q == r - 1) cand = v[q];
    return r == 0 ? prev : cand;
}

// M^H of one column: in s0..s3 the adjoints of U**, V**, W**, P_inc; out in
// s0..s2 the adjoints of the Helmholtz sources Qu, Qv, Qw.
template <int L>
__global__ __launch_bounds__(64 * kColWaves) void k_adj_column(StepArgs A)
{
    extern __shared__ double2 tile_raw[];
    double2 *tile = tile_raw;
    const int l = blockIdx.x * kColWaves + (threadIdx.x >> 6), jm = blockIdx.y;
    const ColMap m = col_map(A.nz);
    const int nz = A.nz, nl = A.nl, ny = A.ny;
    const bool is00 = (l == 0 && jm == 0);
    const size_t plane = (size_t)ny * nl, col0 = (size_t)jm * nl + (size_t)blockIdx.x * kColWaves;
    double2 ub[2][L], wb[1][L], pb[L];
    col_load<L>(A.s0, tile, nz, plane, col0, m, ub[0]);
    col_load<L>(A.s1, tile, nz, plane, col0, m, ub[1]);
    col_load<L>(A.s2, tile, nz, plane, col0, m, wb[0]);
    col_load<L>(A.s3, tile, nz, plane, col0, m, pb);
    const double cdt = A.rc * A.dt, cnu = A.rc * 0.5 * A.dt * A.nu;
    const double2 sgx = make_double2(A.symx[SY_GRE * nl + l], A.symx[SY_GIM * nl + l]);
    const double2 sgy = make_double2(A.symy[SY_GRE * ny + jm], A.symy[SY_GIM * ny + jm]);
    const double2 sdx = make_double2(A.symx[SY_DRE * nl + l], A.symx[SY_DIM * nl + l]);
    const double2 sdy = make_double2(A.symy[SY_DRE * ny + jm], A.symy[SY_DIM * ny + jm]);
    const double slap = A.symx[SY_L * nl + l] + A.symy[SY_L * ny + jm];
    auto lev = [&](int t, int k) { return A.lev[t * nz + k]; };
    // 1. adjoint of phi: projection u** = u* - cdt G phi and the increment
    //    P_inc = phi - cnu (slap + Lz) phi (Neumann ghosts)
    const double2 pprev = shfl_up_c(last_of<L>(pb, m.cnt), 1), pnext = shfl_dn_c(pb[0], 1);
    const double2 wprev = shfl_up_c(last_of<L>(wb[0], m.cnt), 1);
    double2 ph[1][L];
#pragma unroll
    for (int r = 0; r < L; ++r) {
        ph[0][r] = czero();
        if (r >= m.cnt) continue;
        const int k = m.k0 + r;
        double2 t = cadd(cmul(cconj(sgx), ub[0][r]), cmul(cconj(sgy), ub[1][r]));
        t = cscale(t, -cdt);
        if (k <= nz - 2) t = caxpy_neg(t, -cdt * lev(LV_IDZT, k), wb[0][r]);
        if (k >= 1) t = caxpy_neg(t, cdt * lev(LV_IDZT, k - 1), lvl_dn<L>(wb[0], r, wprev));
        double dg = 0.0;
        if (k <= nz - 2) dg -= lev(LV_IDZT, k);
        if (k >= 1) dg -= lev(LV_IDZB, k);
        dg *= lev(LV_IDZ, k);
        double2 lz = cscale(pb[r], dg);
        if (k >= 1) lz = caxpy_neg(lz, -lev(LV_IDZT, k - 1) * lev(LV_IDZ, k - 1), lvl_dn<L>(pb, r, pprev));
        if (k <= nz - 2) lz = caxpy_neg(lz, -lev(LV_IDZB, k + 1) * lev(LV_IDZ, k + 1), lvl_up<L>(m, pb, r, pnext));
        t = cadd(t, pb[r]);
        t = caxpy_neg(t, cnu * slap, pb[r]);
        t = caxpy_neg(t, cnu, lz);
        ph[0][r] = t;
    }
    // 2. adjoint of the divergence source: transposed (refined) pressure
    //    solve, then the delnull zeroing of level 0 of the (0,0) source
    double a[L], b[L], c[L];
    assemble_T<L>(A, 2, l, jm, is00, m, a, b, c);
    wave_tridiag_refined<L, 1>(m.cnt, a, b, c, ph);
    if (is00 && m.lane == 0) ph[0][0] = czero();
    // 3. adjoint of u*, v*, w*: add the transposed divergence / dt
    const double2 snext = shfl_dn_c(ph[0][0], 1);
#pragma unroll
    for (int r = 0; r < L; ++r) {
        if (r >= m.cnt) continue;
        const int k = m.k0 + r;
        const double2 s = cscale(ph[0][r], A.idt);
        ub[0][r] = cadd(ub[0][r], cmul(cconj(sdx), s));
        ub[1][r] = cadd(ub[1][r], cmul(cconj(sdy), s));
        wb[0][r] = caxpy_neg(wb[0][r], -lev(LV_IDZ, k), s);
        if (k + 1 <= nz - 1)
            wb[0][r] = caxpy_neg(wb[0][r], lev(LV_IDZ, k + 1) * A.idt, lvl_up<L>(m, ph[0], r, snext));
    }
    // 4. transposed Helmholtz solves; the w source's top level is discarded
    //    by the forward solve, so its adjoint is zero
    assemble_T<L>(A, 0, l, jm, is00, m, a, b, c);
    wave_tridiag<L, 2>(m.cnt, a, b, c, ub);
    assemble_T<L>(A, 1, l, jm, is00, m, a, b, c);
    wave_tridiag<L, 1>(m.cnt, a, b, c, wb);
#pragma unroll
    for (int r = 0; r < L; ++r)
        if (r < m.cnt && m.k0 + r == nz - 1) wb[0][r] = czero();
    col_store<L>(A.s0, tile, nz, plane, col0, m, ub[0]);
    col_store<L>(A.s1, tile, nz, plane, col0, m, ub[1]);
    col_store<L>(A.s2, tile, nz, plane, col0, m, wb[0]);
}

// Adjoint fields of one substep in physical space.
struct AdjIO {
    const double *qu, *qv, *qw;   // adjoints of the Helmholtz sources q
    const double *fu, *fv, *fw;   // adjoints of this substep's f (next substep's f_old)
    double *au, *av, *aw, *ap;    // out: adjoints of the substep's u, v, w; p in/out
    double *ou, *ov, *ow;         // out: adjoints of f_old
    double *fb[3];                // scratch: total adjoints of f, g, h
    double *face[8];              // scratch: adjoints of the face interpolants
};
enum { FB_UHX = 0, FB_UHY, FB_VHX, FB_VHY, FB_WHX, FB_WHY, FB_UHZ, FB_VHZ };

// Stencil tables: interpolation to the face "behind" a point (I-, e.g. uhx
// from u), to the face "ahead" (I+, e.g. vhx from v), and the derivative
// weights of quantities at S offsets (uux, vvy, dpdx) and at P offsets
// (uvy, vux, vx, wux, wvy).
template <int CDS> struct AdjStencil;
template <> struct AdjStencil<1> {
    static constexpr int NI = 2, ND = 2;
    static constexpr int tm[2] = {0, -1};
    static constexpr int tp[2] = {1, 0};
    static constexpr double wi[2] = {0.5, 0.5};
    static constexpr int sS[2] = {1, 0};
    static constexpr int sP[2] = {0, -1};
    static constexpr double wd[2] = {1.0, -1.0};
};
template <> struct AdjStencil<2> {
    static constexpr int NI = 4, ND = 4;
    static constexpr int tm[4] = {1, 0, -1, -2};
    static constexpr int tp[4] = {2, 1, 0, -1};
    static constexpr double wi[4] = {-1.0 / 16.0, 9.0 / 16.0, 9.0 / 16.0, -1.0 / 16.0};
    static constexpr int sS[4] = {2, 1, 0, -1};
    static constexpr int sP[4] = {1, 0, -1, -2};
    static constexpr double wd[4] = {-1.0, 27.0, -27.0, 1.0};
};

// Periodic-wrap field access and the stencils of the transposed right-hand
// side at the forward state of A.
template <int CDS>
struct AdjOps {
    using St = AdjStencil<CDS>;
    const StepArgs &A;
    int nx, ny, nz;
    HC_DEV AdjOps(const StepArgs &a) : A(a), nx(a.nx), ny(a.ny), nz(a.nz) {}
    HC_DEV size_t at(int i, int j, int k) const { return ((size_t)k * ny + wrapi(j, ny)) * nx + wrapi(i, nx); }
    HC_DEV double lev(int t, int k) const { return A.lev[t * nz + k]; }
    HC_DEV double f(const double *p, int i, int j, int k) const { return p[at(i, j, k)]; }
    HC_DEV double U(int i, int j, int k) const { return A.u[at(i, j, k)]; }
    HC_DEV double V(int i, int j, int k) const { return A.v[at(i, j, k)]; }
    HC_DEV double W(int i, int j, int k) const { return (k < 0 || k >= nz - 1) ? 0.0 : A.w[at(i, j, k)]; }
    // I- and I+ of a field along axis ax (0 x, 1 y)
    template <class F> HC_DEV double Im(F g, int i, int j, int k, int ax) const
    {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < St::NI; ++q) s += St::wi[q] * (ax == 0 ? g(i + St::tm[q], j, k) : g(i, j + St::tm[q], k));
        return s;
    }
    template <class F> HC_DEV double Ip(F g, int i, int j, int k, int ax) const
    {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < St::NI; ++q) s += St::wi[q] * (ax == 0 ? g(i + St::tp[q], j, k) : g(i, j + St::tp[q], k));
        return s;
    }
    // transposes of I- and I+
    template <class F> HC_DEV double ImT(F g, int i, int j, int k, int ax) const
    {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < St::NI; ++q) s += St::wi[q] * (ax == 0 ? g(i - St::tm[q], j, k) : g(i, j - St::tm[q], k));
        return s;
    }
    template <class F> HC_DEV double IpT(F g, int i, int j, int k, int ax) const
    {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < St::NI; ++q) s += St::wi[q] * (ax == 0 ? g(i - St::tp[q], j, k) : g(i, j - St::tp[q], k));
        return s;
    }
    // adjoint of a quantity at S / P offsets: sc * sum_s w_s fbar(point - s)
    HC_DEV double DS(const double *p, int i, int j, int k, int ax, double sc) const
    {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < St::ND; ++q) s += St::wd[q] * (ax == 0 ? f(p, i - St::sS[q], j, k) : f(p, i, j - St::sS[q], k));
        return s * sc;
    }
    HC_DEV double DP(const double *p, int i, int j, int k, int ax, double sc) const
    {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < St::ND; ++q) s += St::wd[q] * (ax == 0 ? f(p, i - St::sP[q], j, k) : f(p, i, j - St::sP[q], k));
        return s * sc;
    }
    // z flux divergence adjoint at face k: cell k (+1/dz), cell k+1 (-1/dz)
    HC_DEV double EZ(const double *p, int i, int j, int k) const
    {
        return f(p, i, j, k) * lev(LV_IDZ, k) - (k + 1 <= nz - 1 ? f(p, i, j, k + 1) * lev(LV_IDZ, k + 1) : 0.0);
    }
};

// Pass 1: total adjoints of f, g, h (this substep's f feeds q with ra and
// the stored f_old' of the next substep).
__global__ __launch_bounds__(256) void k_adj_fbar(StepArgs A, AdjIO io)
{
    const size_t plane = (size_t)A.nx * A.ny, n = plane * A.nz, top = plane * (A.nz - 1);
    const double dra = A.dt * A.ra, drb = A.dt * A.rb;
    (void)drb;  // the f_old adjoints rb dt q are written by k_adj_gather
    for (size_t pt = blockIdx.x * (size_t)blockDim.x + threadIdx.x; pt < n; pt += (size_t)gridDim.x * blockDim.x) {
        const double qu = io.qu[pt], qv = io.qv[pt], qw = pt < top ? io.qw[pt] : 0.0;
        io.fb[0][pt] = io.fu[pt] + dra * qu;
        io.fb[1][pt] = io.fv[pt] + dra * qv;
        io.fb[2][pt] = pt < top ? io.fw[pt] + dra * qw : 0.0;
    }
}

// Pass 2: adjoints of the face interpolants at every point (sign: f = -uux
// - ...).  Faces in z (uhz, vhz) and the w faces (whx, why) exist on levels
// 0..nz-2; the top level is written 0.
template <int CDS>
__global__ __launch_bounds__(256) void k_adj_faces(StepArgs A, AdjIO io)
{
    const AdjOps<CDS> o(A);
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const size_t n = (size_t)nx * ny * nz;
    const double cx = CDS == 1 ? A.idx : A.c24x, cy = CDS == 1 ? A.idy : A.c24y;
    const double um = A.u_mrf, fc = A.dp_opt == 3 ? A.fc : 0.0;
    const double *FU = io.fb[0], *FV = io.fb[1], *FW = io.fb[2];
    auto Uf = [&](int i, int j, int k) { return o.U(i, j, k); };
    auto Vf = [&](int i, int j, int k) { return o.V(i, j, k); };
    auto Wf = [&](int i, int j, int k) { return o.W(i, j, k); };
    for (size_t pt = blockIdx.x * (size_t)blockDim.x + threadIdx.x; pt < n; pt += (size_t)gridDim.x * blockDim.x) {
        const unsigned p32 = (unsigned)pt, row = p32 / (unsigned)nx;
        const int i = (int)(p32 - row * (unsigned)nx), j = (int)(row % (unsigned)ny), k = (int)(row / (unsigned)ny);
        const double uhx = o.Im(Uf, i, j, k, 0), uhy = o.Ip(Uf, i, j, k, 1);
        const double vhx = o.Ip(Vf, i, j, k, 0), vhy = o.Im(Vf, i, j, k, 1);
        const double dpu_y = o.DP(FU, i, j, k, 1, cy), dpv_x = o.DP(FV, i, j, k, 0, cx);
        double b_uhy = -vhx * (dpu_y + dpv_x);
        double b_vhx = -uhy * dpu_y - (uhy + um) * dpv_x;
        if (fc != 0.0) {
            b_uhy -= 0.5 * fc * (o.f(FV, i, j, k) + o.f(FV, i + 1, j, k));
            b_vhx += 0.5 * fc * (o.f(FU, i, j, k) + o.f(FU, i, j + 1, k));
        }
        io.face[FB_UHX][pt] = -(2.0 * uhx + um) * o.DS(FU, i, j, k, 0, cx);
        io.face[FB_UHY][pt] = b_uhy;
        io.face[FB_VHX][pt] = b_vhx;
        io.face[FB_VHY][pt] = -2.0 * vhy * o.DS(FV, i, j, k, 1, cy);
        double bwx = 0.0, bwy = 0.0, buz = 0.0, bvz = 0.0;
        if (k <= nz - 2) {
            const double lt = o.lev(LV_LT, k);
            const double uhz = o.U(i, j, k + 1) * lt + o.U(i, j, k) * (1.0 - lt);
            const double vhz = o.V(i, j, k + 1) * lt + o.V(i, j, k) * (1.0 - lt);
            const double whx = o.Ip(Wf, i, j, k, 0), why = o.Ip(Wf, i, j, k, 1);
            const double ezu = o.EZ(FU, i, j, k), ezv = o.EZ(FV, i, j, k);
            const double dpw_x = o.DP(FW, i, j, k, 0, cx), dpw_y = o.DP(FW, i, j, k, 1, cy);
            bwx = -uhz * ezu - (uhz + um) * dpw_x;
            bwy = -vhz * ezv - vhz * dpw_y;
            buz = -whx * (ezu + dpw_x);
            bvz = -why * (ezv + dpw_y);
        }
        io.face[FB_WHX][pt] = bwx;
        io.face[FB_WHY][pt] = bwy;
        io.face[FB_UHZ][pt] = buz;
        io.face[FB_VHZ][pt] = bvz;
    }
}

// Pass 3: gather of the transposed interpolations, explicit diffusion,
// pressure gradients and the w self-advection at every point.
template <int CDS>
__global__ __launch_bounds__(256) void k_adj_gather(StepArgs A, AdjIO io)
{
    const AdjOps<CDS> o(A);
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const size_t pl = (size_t)nx * ny, n = pl * nz;
    const double cx = CDS == 1 ? A.idx : A.c24x, cy = CDS == 1 ? A.idy : A.c24y;
    const double cnu = A.rc * 0.5 * A.dt * A.nu, crp = A.rc * A.dt;
    const double *FW = io.fb[2];
    auto face = [&](int id) { return [&, id](int i, int j, int k) { return o.f(io.face[id], i, j, k); }; };
    // symmetric horizontal second differences
    auto Lh = [&](const double *g, int i, int j, int k) {
        auto F = [&](int x, int y) { return g[o.at(x, y, k)]; };
        if (CDS == 1)
            return (F(i + 1, j) - 2.0 * F(i, j) + F(i - 1, j)) * A.idx2 + (F(i, j + 1) - 2.0 * F(i, j) + F(i, j - 1)) * A.idy2;
        return (-F(i + 2, j) + 16.0 * F(i + 1, j) - 30.0 * F(i, j) + 16.0 * F(i - 1, j) - F(i - 2, j)) * A.c12x2 +
               (-F(i, j + 2) + 16.0 * F(i, j + 1) - 30.0 * F(i, j) + 16.0 * F(i, j - 1) - F(i, j - 2)) * A.c12y2;
    };
    for (size_t pt = blockIdx.x * (size_t)blockDim.x + threadIdx.x; pt < n; pt += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(pt % nx), j = (int)((pt / nx) % ny), k = (int)(pt / pl);
        const double qu = io.qu[pt], qv = io.qv[pt], qw = k < nz - 1 ? io.qw[pt] : 0.0;
        const double idz = o.lev(LV_IDZ, k), idzt = o.lev(LV_IDZT, k), idzb = o.lev(LV_IDZB, k);
        // ---- u and v: identity, explicit diffusion (transposed z operator with
        // the no-slip / free-slip ghosts), interpolant transposes
        double dgz = -(idzt + idzb);
        if (k == 0) dgz = -(idzt + 2.0 * idzb);
        if (k == nz - 1) dgz = A.ochannel == 0 ? -(2.0 * idzt + idzb) : -idzb;
        auto LzT = [&](const double *g) {
            double s = g[pt] * dgz * idz;
            if (k >= 1) s += g[pt - pl] * o.lev(LV_IDZT, k - 1) * o.lev(LV_IDZ, k - 1);
            if (k <= nz - 2) s += g[pt + pl] * o.lev(LV_IDZB, k + 1) * o.lev(LV_IDZ, k + 1);
            return s;
        };
        double au = qu + cnu * (Lh(io.qu, i, j, k) + LzT(io.qu));
        double av = qv + cnu * (Lh(io.qv, i, j, k) + LzT(io.qv));
        au += o.ImT(face(FB_UHX), i, j, k, 0) + o.IpT(face(FB_UHY), i, j, k, 1);
        av += o.IpT(face(FB_VHX), i, j, k, 0) + o.ImT(face(FB_VHY), i, j, k, 1);
        if (k <= nz - 2) {
            const double lt = o.lev(LV_LT, k);
            au += io.face[FB_UHZ][pt] * (1.0 - lt);
            av += io.face[FB_VHZ][pt] * (1.0 - lt);
        }
        if (k >= 1) {
            const double lt = o.lev(LV_LT, k - 1);
            au += io.face[FB_UHZ][pt - pl] * lt;
            av += io.face[FB_VHZ][pt - pl] * lt;
        }
        // ---- w (levels 0..nz-2; w at the top face is fixed)
        double aw = 0.0;
        if (k <= nz - 2) {
            double s = qw * (-(o.lev(LV_IDZ, k + 1) + idz) * idzt);
            if (k >= 1) s += io.qw[pt - pl] * idz * o.lev(LV_IDZT, k - 1);
            if (k + 1 <= nz - 2) s += io.qw[pt + pl] * o.lev(LV_IDZ, k + 1) * o.lev(LV_IDZT, k + 1);
            aw = qw + cnu * (Lh(io.qw, i, j, k) + s);
            aw += o.IpT(face(FB_WHX), i, j, k, 0) + o.IpT(face(FB_WHY), i, j, k, 1);
            // wwz = ((w(k)+w(k+1))^2/4 - (w(k)+w(k-1))^2/4) / dz_t(k)
            const double w0 = o.W(i, j, k);
            const double c1 = 0.5 * (w0 + o.W(i, j, k + 1)), c2 = 0.5 * (w0 + o.W(i, j, k - 1));
            double r = FW[pt] * (c1 - c2) * idzt;
            if (k >= 1) r += FW[pt - pl] * 0.5 * (o.W(i, j, k - 1) + w0) * o.lev(LV_IDZT, k - 1);
            if (k + 1 <= nz - 2) r -= FW[pt + pl] * 0.5 * (o.W(i, j, k + 1) + w0) * o.lev(LV_IDZT, k + 1);
            aw -= r;
        }
        // ---- p: the pressure gradients of q
        double ap = io.ap[pt] - crp * (o.DS(io.qu, i, j, k, 0, cx) + o.DS(io.qv, i, j, k, 1, cy));
        if (k >= 1) ap -= crp * io.qw[pt - pl] * o.lev(LV_IDZT, k - 1);
        if (k <= nz - 2) ap += crp * qw * idzt;
        io.au[pt] = au;
        io.av[pt] = av;
        io.aw[pt] = aw;
        io.ap[pt] = ap;
        // adjoints of f_old (this substep's q took rb dt f_old)
        const double drb = A.dt * A.rb;
        io.ou[pt] = drb * qu;
        io.ov[pt] = drb * qv;
        io.ow[pt] = drb * qw;
    }
}
