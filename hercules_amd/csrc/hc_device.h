// Device building blocks for the HERCULES hot path on CDNA4 (gfx950):
//   * batched in-place Stockham FFTs over LDS (radix 4/2/3),
//   * a wave64 partitioned tridiagonal solver (one system per wavefront,
//     levels split across lanes, interface system by parallel cyclic
//     reduction over lane shuffles).
#pragma once
#include <hip/hip_runtime.h>

#define HC_DEV __device__ __forceinline__

// Complex arithmetic over the HIP vector types double2 and float2.
template <class R> struct CplxOf;
template <> struct CplxOf<double> { using type = double2; };
template <> struct CplxOf<float> { using type = float2; };
template <class R> using cplx = typename CplxOf<R>::type;
template <class C> struct RealOf;
template <> struct RealOf<double2> { using type = double; };
template <> struct RealOf<float2> { using type = float; };
template <class C> using real_of = typename RealOf<C>::type;

template <class C> HC_DEV C mk2(real_of<C> x, real_of<C> y)
{
    C r;
    r.x = x;
    r.y = y;
    return r;
}
template <class C> HC_DEV C cadd(C a, C b) { return mk2<C>(a.x + b.x, a.y + b.y); }
template <class C> HC_DEV C csub(C a, C b) { return mk2<C>(a.x - b.x, a.y - b.y); }
template <class C> HC_DEV C cmul(C a, C b) { return mk2<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <class C> HC_DEV C cscale(C a, real_of<C> s) { return mk2<C>(a.x * s, a.y * s); }
template <class C> HC_DEV C cconj(C a) { return mk2<C>(a.x, -a.y); }
template <class C> HC_DEV C czero_t() { return mk2<C>(real_of<C>(0), real_of<C>(0)); }
HC_DEV double2 czero() { return make_double2(0.0, 0.0); }
// a - s*b for real s
template <class C> HC_DEV C caxpy_neg(C a, real_of<C> s, C b) { return mk2<C>(a.x - s * b.x, a.y - s * b.y); }
// complex value converted to another precision
template <class C, class S> HC_DEV C ccast(S a) { return mk2<C>(real_of<C>(a.x), real_of<C>(a.y)); }

// 1/x for finite, normal, nonzero x: the hardware reciprocal estimate
// (v_rcp_f64) refined by two Newton-Raphson steps r <- r + r(1 - x r),
// which leaves it within an ulp or two of the correctly rounded value at a
// third of the cost of an IEEE division.
HC_DEV double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}
// FP32: v_rcp_f32 (1 ulp) and one Newton-Raphson step.
HC_DEV float rcp32(float x)
{
    float r = __builtin_amdgcn_rcpf(x);
    float e = __builtin_fmaf(-x, r, 1.0f);
    return __builtin_fmaf(r, e, r);
}
HC_DEV double rcpr(double x) { return rcp64(x); }
HC_DEV float rcpr(float x) { return rcp32(x); }

// ---------------------------------------------------------------------------
// FFT plans: N = 4^a 2^b 3^c, stages in that order (radix 4 first).
constexpr int fft_nstages(int n)
{
    int s = 0;
    while (n % 4 == 0) { n /= 4; ++s; }
    while (n % 2 == 0) { n /= 2; ++s; }
    while (n % 3 == 0) { n /= 3; ++s; }
    return n == 1 ? s : -1;
}
constexpr int fft_radix(int n, int st)
{
    int s = 0;
    while (n % 4 == 0) { if (s == st) return 4; n /= 4; ++s; }
    while (n % 2 == 0) { if (s == st) return 2; n /= 2; ++s; }
    while (n % 3 == 0) { if (s == st) return 3; n /= 3; ++s; }
    return 1;
}
constexpr int fft_ns(int n, int st)  // product of the radices before stage st
{
    int p = 1;
    for (int s = 0; s < st; ++s) p *= fft_radix(n, s);
    return p;
}

// In-register R-point DFT, sign SIGN (-1 forward, +1 backward), unnormalised.
template <int R, int SIGN, class C>
HC_DEV void dft_small(C *v)
{
    using Re = real_of<C>;
    if constexpr (R == 2) {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        C a2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        // a3 = (SIGN i) * d
        C a3 = SIGN < 0 ? mk2<C>(d.y, -d.x) : mk2<C>(-d.y, d.x);
        v[0] = cadd(a0, a2);
        v[2] = csub(a0, a2);
        v[1] = cadd(a1, a3);
        v[3] = csub(a1, a3);
    } else if constexpr (R == 3) {
        const Re h = Re(0.86602540378443864676);  // sqrt(3)/2
        const Re half = Re(0.5);
        C t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
        C m = mk2<C>(v[0].x - half * t1.x, v[0].y - half * t1.y);
        // s = (SIGN i) * h * t2
        C s = SIGN < 0 ? mk2<C>(h * t2.y, -h * t2.x) : mk2<C>(-h * t2.y, h * t2.x);
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, s);
        v[2] = csub(m, s);
    }
}

// One Stockham stage over B transforms of length N held in LDS, element
// (b, n) at lds[b*BS + n*ES].  All loads of the stage happen before the
// barrier and all stores after it, so one buffer suffices.  tw[e] =
// exp(-2 pi i e / N).  C is double2 or float2.
template <int N, int B, int BS, int ES, int T, int SIGN, int ST, class C>
HC_DEV void fft_stages(C *lds, const C *__restrict__ tw, int tid)
{
    if constexpr (ST < fft_nstages(N)) {
        constexpr int R = fft_radix(N, ST);
        constexpr int NS = fft_ns(N, ST);
        constexpr int M = N / R;
        constexpr int NB = B * M;
        constexpr int PER = (NB + T - 1) / T;
        C v[PER][R];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int bf = tid + q * T;
            if (NB % T == 0 || bf < NB) {
                int b = bf / M, j = bf - b * M;
#pragma unroll
                for (int r = 0; r < R; ++r) v[q][r] = lds[b * BS + (j + r * M) * ES];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int bf = tid + q * T;
            if (NB % T == 0 || bf < NB) {
                int b = bf / M, j = bf - b * M;
                int e = j % NS;
                if constexpr (NS > 1) {
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        C w = tw[r * e * (N / (NS * R))];
                        if (SIGN > 0) w.y = -w.y;
                        v[q][r] = cmul(v[q][r], w);
                    }
                }
                dft_small<R, SIGN>(v[q]);
                int base = (j / NS) * NS * R + e;
#pragma unroll
                for (int r = 0; r < R; ++r) lds[b * BS + (base + r * NS) * ES] = v[q][r];
            }
        }
        __syncthreads();
        fft_stages<N, B, BS, ES, T, SIGN, ST + 1>(lds, tw, tid);
    }
}

template <int N, int B, int BS, int ES, int T, int SIGN, class C>
HC_DEV void lds_fft(C *lds, const C *__restrict__ tw, int tid)
{
    static_assert(fft_nstages(N) > 0, "FFT length must be 2^a 3^b");
    fft_stages<N, B, BS, ES, T, SIGN, 0>(lds, tw, tid);
}

// ---------------------------------------------------------------------------
// Wave64 partitioned tridiagonal solver.
//
// One system a_k x_{k-1} + b_k x_k + c_k x_{k+1} = d_k of n levels per
// wavefront.  Lane s owns `cnt` consecutive levels (cnt <= L) held in
// registers; lanes are filled in order, so lanes with cnt == 0 come last.
// The first level of each lane is its interface unknown X_s; the remaining
// levels are eliminated locally (Thomas, pivots never shared) in terms of
// X_s and X_{s+1}, giving a tridiagonal system in X over the lanes that is
// solved by parallel cyclic reduction through lane shuffles, after which the
// interior is recovered.  Requires a_0 = 0 and c_{n-1} = 0.  NR complex
// right-hand sides share the matrix; d is overwritten by the solution.
// Re (double or float) is the arithmetic of the whole solve.
template <class Re> HC_DEV Re shfl_up1(Re v, int d) { return __shfl_up(v, d, 64); }
template <class Re> HC_DEV Re shfl_dn1(Re v, int d) { return __shfl_down(v, d, 64); }
template <class C> HC_DEV C shfl_up_c(C v, int d) { return mk2<C>(__shfl_up(v.x, d, 64), __shfl_up(v.y, d, 64)); }
template <class C> HC_DEV C shfl_dn_c(C v, int d) { return mk2<C>(__shfl_down(v.x, d, 64), __shfl_down(v.y, d, 64)); }

template <int L, int NR, class Re>
HC_DEV void wave_tridiag(int cnt, const Re (&a)[L], const Re (&b)[L], const Re (&c)[L], cplx<Re> (&d)[NR][L])
{
    using C = cplx<Re>;
    const Re zero = Re(0), one = Re(1);
    const int lane = threadIdx.x & 63;
    Re cp[L], g[L], h[L];
    // forward sweep over the interior levels 1..cnt-1
#pragma unroll
    for (int r = 0; r < L; ++r) { cp[r] = zero; g[r] = zero; h[r] = zero; }
#pragma unroll
    for (int r = 1; r < L; ++r) {
        if (r < cnt) {
            Re den = b[r] - (r > 1 ? a[r] * cp[r - 1] : zero);
            Re iv = rcpr(den);
            cp[r] = c[r] * iv;
            Re gr = r == 1 ? -a[1] : -a[r] * g[r - 1];
            g[r] = gr * iv;
            Re hr = (r == cnt - 1 ? -c[r] : zero) - (r > 1 ? a[r] * h[r - 1] : zero);
            h[r] = hr * iv;
#pragma unroll
            for (int q = 0; q < NR; ++q) {
                C y = d[q][r];
                if (r > 1) y = caxpy_neg(y, a[r], d[q][r - 1]);
                d[q][r] = cscale(y, iv);
            }
        }
    }
    // backward sweep
#pragma unroll
    for (int r = L - 2; r >= 1; --r) {
        if (r < cnt - 1) {
            g[r] -= cp[r] * g[r + 1];
            h[r] -= cp[r] * h[r + 1];
#pragma unroll
            for (int q = 0; q < NR; ++q) d[q][r] = caxpy_neg(d[q][r], cp[r], d[q][r + 1]);
        }
    }
    // interface quantities: first interior (y1,g1,h1) and last level (yl,gl,hl)
    Re g1 = zero, h1 = one, gl = one, hl = zero;
    C y1[NR], yl[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) { y1[q] = czero_t<C>(); yl[q] = czero_t<C>(); }
    if (cnt >= 2) {
        g1 = g[1]; h1 = h[1];
#pragma unroll
        for (int q = 0; q < NR; ++q) y1[q] = d[q][1];
#pragma unroll
        for (int r = 1; r < L; ++r)
            if (r == cnt - 1) {
                gl = g[r]; hl = h[r];
#pragma unroll
                for (int q = 0; q < NR; ++q) yl[q] = d[q][r];
            }
    }
    // values of the previous lane's last level
    Re Gp = shfl_up1(gl, 1), Hp = shfl_up1(hl, 1);
    C Yp[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) Yp[q] = shfl_up_c(yl[q], 1);
    if (lane == 0) {
        Gp = zero; Hp = zero;
#pragma unroll
        for (int q = 0; q < NR; ++q) Yp[q] = czero_t<C>();
    }
    Re al, be, ga;
    C de[NR];
    if (cnt >= 1) {
        al = a[0] * Gp;
        be = b[0] + a[0] * Hp + c[0] * g1;
        ga = c[0] * h1;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            C t = caxpy_neg(d[q][0], a[0], Yp[q]);
            de[q] = caxpy_neg(t, c[0], y1[q]);
        }
    } else {
        al = zero; be = one; ga = zero;
#pragma unroll
        for (int q = 0; q < NR; ++q) de[q] = czero_t<C>();
    }
    // parallel cyclic reduction over the 64 lanes
#pragma unroll
    for (int st = 1; st < 64; st <<= 1) {
        Re al_lo = shfl_up1(al, st), be_lo = shfl_up1(be, st), ga_lo = shfl_up1(ga, st);
        Re al_hi = shfl_dn1(al, st), be_hi = shfl_dn1(be, st), ga_hi = shfl_dn1(ga, st);
        C de_lo[NR], de_hi[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            de_lo[q] = shfl_up_c(de[q], st);
            de_hi[q] = shfl_dn_c(de[q], st);
        }
        if (lane < st) {
            al_lo = zero; be_lo = one; ga_lo = zero;
#pragma unroll
            for (int q = 0; q < NR; ++q) de_lo[q] = czero_t<C>();
        }
        if (lane + st >= 64) {
            al_hi = zero; be_hi = one; ga_hi = zero;
#pragma unroll
            for (int q = 0; q < NR; ++q) de_hi[q] = czero_t<C>();
        }
        Re k1 = al * rcpr(be_lo), k2 = ga * rcpr(be_hi);
        Re nal = -al_lo * k1, nga = -ga_hi * k2;
        Re nbe = be - ga_lo * k1 - al_hi * k2;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            C t = caxpy_neg(de[q], k1, de_lo[q]);
            de[q] = caxpy_neg(t, k2, de_hi[q]);
        }
        al = nal; be = nbe; ga = nga;
    }
    Re ib = rcpr(be);
    C X[NR], Xn[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        X[q] = cscale(de[q], ib);
        Xn[q] = shfl_dn_c(X[q], 1);
        if (lane == 63) Xn[q] = czero_t<C>();
    }
    // recover the interior
#pragma unroll
    for (int r = 1; r < L; ++r) {
        if (r < cnt) {
#pragma unroll
            for (int q = 0; q < NR; ++q) {
                C y = d[q][r];
                y.x += g[r] * X[q].x + h[r] * Xn[q].x;
                y.y += g[r] * X[q].y + h[r] * Xn[q].y;
                d[q][r] = y;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) d[q][0] = X[q];
}

// The same solve followed by one step of fixed-precision iterative
// refinement: r = d - T x is formed with the neighbouring lanes' values and
// solved again.  PCR's residual on the stiff stretched-grid Neumann systems
// is 1e3-1e4 times that of a sequential Thomas sweep (its forward error is
// comparable); one refinement step brings the residual back to the Thomas
// level, which is what the divergence of the projected velocity measures.
template <int L, int NR, class Re>
HC_DEV void wave_tridiag_refined(int cnt, const Re (&a)[L], const Re (&b)[L], const Re (&c)[L], cplx<Re> (&d)[NR][L])
{
    using C = cplx<Re>;
    const int lane = threadIdx.x & 63;
    C d0[NR][L];
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
        for (int r = 0; r < L; ++r) d0[q][r] = d[q][r];
    wave_tridiag<L, NR>(cnt, a, b, c, d);
    C res[NR][L];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        C last = czero_t<C>();
#pragma unroll
        for (int r = 0; r < L; ++r)
            if (r == cnt - 1) last = d[q][r];
        C xprev = shfl_up_c(last, 1);
        C xnext = shfl_dn_c(d[q][0], 1);
        if (lane == 0) xprev = czero_t<C>();
        if (lane == 63) xnext = czero_t<C>();
#pragma unroll
        for (int r = 0; r < L; ++r) {
            res[q][r] = czero_t<C>();
            if (r < cnt) {
                C xm = r == 0 ? xprev : d[q][r > 0 ? r - 1 : 0];
                C xp = xnext;
#pragma unroll
                for (int s = 0; s < L; ++s)
                    if (s == r + 1 && s < cnt) xp = d[q][s];
                C x0 = d[q][r];
                res[q][r].x = d0[q][r].x - (a[r] * xm.x + b[r] * x0.x + c[r] * xp.x);
                res[q][r].y = d0[q][r].y - (a[r] * xm.y + b[r] * x0.y + c[r] * xp.y);
            }
        }
    }
    wave_tridiag<L, NR>(cnt, a, b, c, res);
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
        for (int r = 0; r < L; ++r)
            if (r < cnt) d[q][r] = cadd(d[q][r], res[q][r]);
}
