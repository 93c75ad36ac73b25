// euler2d_state.hpp -- ideal-gas 2D Euler state algebra for the gfx950 kernels.
//
// Templated on the arithmetic type T (double for the FP64 path, float for the
// FP32 reconstruction / limiter / Roe flux of the mixed path).  Every formula
// keeps the reference's operation order so that, built with
// -ffp-contract=off, the FP64 instantiation reproduces the reference's
// arithmetic bit for bit:
//   state functions          src_2D/Euler2D/Euler2DState.h:1033-1200, 1511-1560,
//                            1748-1860, 2153-2161, 2702-2720, 2968-2986, 3429-3435
//   RoeAverage, Reflect,     src_2D/Euler2D/Euler2DState.cc:654-683, 751-787,
//   BC_Characteristic_Pressure, HartenFix, FluxRoe_x/_n
//                            1383-1465, 1872-1905, 2311-2353, 2446-2489
//   HartenFixPos/Neg, Limiter_Venkatakrishnan
//                            src_2D/CFD/CFD.cc:605-643, 728-752
// min/max are the reference's ternaries (Math.h:73-92), so signed zeros and
// NaNs propagate the same way, except where a comment shows that the IEEE
// minimum or maximum (tmin/tmax, one instruction) gives the same bits.
#pragma once

#include <hip/hip_runtime.h>

#define CFFC_HD __host__ __device__ __forceinline__

#define CFFC_POW_FN __device__ __forceinline__
#define CFFC_POW_TAB static __constant__ const
#include "glibc_pow.h"

namespace cffc {

template <typename T> struct Consts {
    static constexpr T ZERO = T(0), ONE = T(1), TWO = T(2), HALF = T(0.5);
    static constexpr T TOLER = T(0.0000001);
};

// gamma = 1.4; gm1 and gm1i exactly as Euler2DState.cc:13-15 evaluates them
template <typename T> CFFC_HD T gam() { return T(1.40); }
template <typename T> CFFC_HD T gm1() { return T(1.40 - 1.0); }
template <typename T> CFFC_HD T gm1i() { return T(1.0 / (1.40 - 1.0)); }

template <typename T> CFFC_HD T dmax(T x, T y) { return (x > y) ? x : y; }
template <typename T> CFFC_HD T dmin(T x, T y) { return (x < y) ? x : y; }
// IEEE minimum (v_min_f64 / v_min_f32): dmin's value wherever neither
// argument is NaN and the arguments are not zeros of opposite sign
template <typename T> CFFC_HD T tmin(T x, T y) { return fmin(x, y); }
template <> CFFC_HD float tmin<float>(float x, float y) { return fminf(x, y); }
template <typename T> CFFC_HD T tmax(T x, T y) { return fmax(x, y); }
template <> CFFC_HD float tmax<float>(float x, float y) { return fmaxf(x, y); }
// The same IEEE minimum and maximum of two FP64 values as one v_min_f64 /
// v_max_f64 each, for operands read from LDS: fmin/fmax first quiet
// signalling NaNs with an extra instruction per operand, which finite
// operands do not need.
CFFC_HD double vmin64(double x, double y) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(x), "v"(y));
    return r;
#else
    return fmin(x, y);
#endif
}
CFFC_HD double vmax64(double x, double y) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(x), "v"(y));
    return r;
#else
    return fmax(x, y);
#endif
}
template <typename T> CFFC_HD T sqr(T x) { return x * x; }
template <typename T> CFFC_HD T tabs(T x) { return fabs(x); }
template <typename T> CFFC_HD T tsqrt(T x) { return sqrt(x); }
template <> CFFC_HD float tsqrt<float>(float x) { return sqrtf(x); }
template <> CFFC_HD float tabs<float>(float x) { return fabsf(x); }

// Arithmetic policy.  double: exact IEEE division and square root, in the
// reference's order (bitwise path).  float (mixed path, no bitwise
// requirement): hardware reciprocal / square-root approximations.
template <typename T> struct Fast { static constexpr bool on = false; };
template <> struct Fast<float> { static constexpr bool on = true; };

// Correctly rounded FP64 division with the divisor's reciprocal shared by
// every quotient over the same divisor.  The device steps are gfx950's own
// IEEE a / b: r = v_rcp_f64(b) refined by two Newton steps
// r += r(1 - b r), then q = a r, the exact FMA remainder a - b q, the
// corrected quotient q + (a - b q) r and v_div_fixup_f64, which gives zeros,
// infinities and NaNs their IEEE results.  Only the operand scaling
// (v_div_scale_f64 and the scale applied by v_div_fmas_f64) is left out: it
// acts when a/b or 1/b would be subnormal, |a/b| >= 2^767 or a nonzero
// |a| < 2^-969, so for every other operand pair the result is a / b bit for
// bit.  The Newton steps depend on b alone, which is what makes the
// reciprocal shareable.  float: plain division (callers choose their own
// approximations on the mixed path).
template <typename T> struct Recip { T b, r; };
template <typename T> CFFC_HD Recip<T> recip(T b) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (!Fast<T>::on) {
        double r = __builtin_amdgcn_rcp(b);
        double e = fma(-b, r, 1.0);
        r = fma(r, e, r);
        e = fma(-b, r, 1.0);
        r = fma(r, e, r);
        return {b, r};
    }
#endif
    return {b, T(0)};
}
template <typename T> CFFC_HD T rdiv(T a, const Recip<T> &d) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (!Fast<T>::on) {
        double q = a * d.r;
        const double rem = fma(-d.b, q, a);
        q = fma(rem, d.r, q);
        return __builtin_amdgcn_div_fixup(q, d.b, a);
    }
#endif
    return a / d.b;
}
template <typename T> CFFC_HD T qdiv(T a, T b) {
    if constexpr (Fast<T>::on) return a * __builtin_amdgcn_rcpf(b);
    else return rdiv(a, recip(b));
}
// Correctly rounded FP64 square root: gfx950's own IEEE sqrt(x) (v_rsq_f64,
// then Goldschmidt's coupled iteration for g ~ sqrt(x) and h ~ 1/(2 sqrt(x))
// and two FMA corrections g += (x - g^2) h; zeros and +inf returned as
// they are) without the input scaling by 2^256 that only acts below
// x = 2^-767, so for x >= 2^-767 the result is sqrt(x) bit for bit.
template <typename T> CFFC_HD T qsqrt(T a) {
    if constexpr (Fast<T>::on) return __builtin_amdgcn_sqrtf(a);
    else {
#if defined(__HIP_DEVICE_COMPILE__)
        const double y = __builtin_amdgcn_rsq(a);
        double g = a * y, h = y * 0.5;
        const double r = fma(-h, g, 0.5);
        g = fma(g, r, g);
        double d = fma(-g, g, a);
        h = fma(h, r, h);
        g = fma(d, h, g);
        d = fma(-g, g, a);
        g = fma(d, h, g);
        // 0x260: -0, +0, +inf
        return __builtin_amdgcn_class(a, 0x260) ? a : g;
#else
        return sqrt(a);
#endif
    }
}

template <typename T> struct PS { T d, vx, vy, p; };   // Euler2D_pState
template <typename T> struct CS { T d, mx, my, E; };   // Euler2D_cState

template <typename T> CFFC_HD T get(const PS<T> &w, int n) {
    return n == 1 ? w.d : n == 2 ? w.vx : n == 3 ? w.vy : w.p;
}

template <typename T> CFFC_HD PS<T> W_of_U(const CS<T> &U) {
    PS<T> W;
    const Recip<T> rd = recip(U.d);
    W.d = U.d;
    W.vx = rdiv(U.mx, rd);
    W.vy = rdiv(U.my, rd);
    W.p = gm1<T>() * (U.E - rdiv(Consts<T>::HALF * (U.mx * U.mx + U.my * U.my), rd));
    return W;
}
template <typename T> CFFC_HD CS<T> U_of_W(const PS<T> &W) {
    CS<T> U;
    U.d = W.d;
    U.mx = W.d * W.vx;
    U.my = W.d * W.vy;
    U.E = W.p * gm1i<T>() + Consts<T>::HALF * W.d * (W.vx * W.vx + W.vy * W.vy);
    return U;
}
template <typename T> CFFC_HD T W_h(const PS<T> &W) {
    return qdiv(gam<T>() * W.p, gm1<T>() * W.d) + Consts<T>::HALF * (W.vx * W.vx + W.vy * W.vy);
}
template <typename T> CFFC_HD T W_H(const PS<T> &W) {
    return gam<T>() * gm1i<T>() * W.p + Consts<T>::HALF * W.d * (W.vx * W.vx + W.vy * W.vy);
}
template <typename T> CFFC_HD T W_a(const PS<T> &W) { return qsqrt(qdiv(gam<T>() * W.p, W.d)); }
// axisymmetric source S(W, X), Euler2DState.h:3179-3182 (y the radius)
template <typename T> CFFC_HD CS<T> axisym_source(const PS<T> &W, T y) {
    CS<T> S;
    const Recip<T> ry = recip(y);
    S.d = rdiv(-W.d * W.vy, ry);
    S.mx = rdiv(-W.d * W.vx * W.vy, ry);
    S.my = rdiv(-W.d * (W.vy * W.vy), ry);
    S.E = rdiv(-W.vy * W_H(W), ry);
    return S;
}
template <typename T> CFFC_HD T W_a2(const PS<T> &W) { return qdiv(gam<T>() * W.p, W.d); }
template <typename T> CFFC_HD T U_e(const CS<T> &U) {
    return qdiv(U.E, U.d) - qdiv(Consts<T>::HALF * (U.mx * U.mx + U.my * U.my), sqr(U.d));
}
template <typename T> CFFC_HD CS<T> W_Fx(const PS<T> &W) {
    CS<T> F;
    F.d = W.d * W.vx;
    F.mx = W.d * sqr(W.vx) + W.p;
    F.my = W.d * W.vx * W.vy;
    F.E = W.vx * W_H(W);
    return F;
}

// HartenFixPos / HartenFixNeg, CFD.cc:605-643.  Both share the corrected
// half-width delta; harten_delta returns it and the callers form
// HALF*(la+|la|+delta) (positive part) or HALF*(la-|la|-delta) (negative).
template <typename T> CFFC_HD T harten_delta(T la, T ll, T lr) {
    typedef Consts<T> K;
    // only compared with |la| and TOLER, or divided by when >= TOLER: the
    // sign of a zero delta is immaterial, so the IEEE maximum serves
    T delta = tmax(K::ZERO, K::TWO * (lr - ll));
    if constexpr (Fast<T>::on) {
        const T fixd = K::HALF * (qdiv(sqr(la), delta) + delta) - tabs(la);
        return (tabs(la) > delta || delta < K::TOLER) ? K::ZERO : fixd;
    } else {
        if (tabs(la) > delta || delta < K::TOLER) return K::ZERO;
        return K::HALF * (qdiv(sqr(la), delta) + delta) - tabs(la);
    }
}
template <typename T> CFFC_HD T harten_pos(T la, T ll, T lr) {
    return Consts<T>::HALF * (la + tabs(la) + harten_delta(la, ll, lr));
}
template <typename T> CFFC_HD T harten_neg(T la, T ll, T lr) {
    return Consts<T>::HALF * (la - tabs(la) - harten_delta(la, ll, lr));
}

// RoeAverage, Euler2DState.cc:654-683
template <typename T> CFFC_HD PS<T> roe_average(const PS<T> &Wl, const PS<T> &Wr) {
    typedef Consts<T> K;
    T hl = W_h(Wl), hr = W_h(Wr);
    T sdl = qsqrt(Wl.d), sdr = qsqrt(Wr.d);
    PS<T> Wa;
    Wa.d = sdl * sdr;
    if constexpr (Fast<T>::on) {
        const T r = __builtin_amdgcn_rcpf(sdl + sdr);
        Wa.vx = (sdl * Wl.vx + sdr * Wr.vx) * r;
        Wa.vy = (sdl * Wl.vy + sdr * Wr.vy) * r;
        const T ha = (sdl * hl + sdr * hr) * r;
        const T aa2 = gm1<T>() * (ha - K::HALF * (sqr(Wa.vx) + sqr(Wa.vy)));
        Wa.p = Wa.d * aa2 * T(1.0 / 1.4);
    } else {
        const Recip<T> rs = recip(sdl + sdr);
        Wa.vx = rdiv(sdl * Wl.vx + sdr * Wr.vx, rs);
        Wa.vy = rdiv(sdl * Wl.vy + sdr * Wr.vy, rs);
        T ha = rdiv(sdl * hl + sdr * hr, rs);
        T aa2 = gm1<T>() * (ha - K::HALF * (sqr(Wa.vx) + sqr(Wa.vy)));
        Wa.p = qdiv(Wa.d * aa2, gam<T>());
    }
    return Wa;
}

// FluxRoe_x, Euler2DState.cc:2311-2353.  The wave loop is unrolled; the
// left/right eigenvectors are lp_x/rc_x (Euler2DState.h:2702-2720,
// 2968-2986) and the wave strengths are the pState inner product
// (Euler2DState.h:1222-1224), all in the reference's order.  The upwind side
// is a select rather than a branch: F = Fx(W_upwind) +/- sum over the waves
// whose split eigenvalue is nonzero (HartenFixNeg for Wa.u >= 0, else
// HartenFixPos, whose results are <= 0 / >= 0 respectively).
template <typename T> CFFC_HD CS<T> flux_roe_x(const PS<T> &Wl, const PS<T> &Wr) {
    typedef Consts<T> K;
    if constexpr (Fast<T>::on) {
        // mixed path: the same upwind form with the eigenvector algebra
        // folded (no multiplications by the zero entries of lp/rc, one
        // reciprocal of the Roe sound speed)
        const PS<T> Wa = roe_average(Wl, Wr);
        const T dd = Wr.d - Wl.d, du = Wr.vx - Wl.vx, dv = Wr.vy - Wl.vy, dp = Wr.p - Wl.p;
        const T a2a = W_a2(Wa);
        const T cl = W_a(Wl), cr = W_a(Wr), ca = qsqrt(a2a);
        const T rca = __builtin_amdgcn_rcpf(ca), rca2 = rca * rca;
        const T ha = W_h(Wa);
        const T la1 = Wa.vx - ca, la4 = Wa.vx + ca;
        const bool left = (Wa.vx >= K::ZERO);
        const T d1 = harten_delta(la1, Wl.vx - cl, Wr.vx - cr);
        const T d4 = harten_delta(la4, Wl.vx + cl, Wr.vx + cr);
        const T w1 = left ? K::HALF * (la1 - tabs(la1) - d1) : K::HALF * (la1 + tabs(la1) + d1);
        const T w2 = left ? K::HALF * (Wa.vx - tabs(Wa.vx)) : K::HALF * (Wa.vx + tabs(Wa.vx));
        const T w4 = left ? K::HALF * (la4 - tabs(la4) - d4) : K::HALF * (la4 + tabs(la4) + d4);
        const PS<T> Wu = left ? Wl : Wr;
        CS<T> F = W_Fx(Wu);
        const T sg = left ? T(1) : T(-1);
        const T pa = K::HALF * rca2 * dp, ua = K::HALF * Wa.d * rca * du;
        const T a1 = sg * w1 * (pa - ua);             // lp1 . dW
        const T a4 = sg * w4 * (pa + ua);             // lp4 . dW
        const T a2 = sg * w2 * (dd - rca2 * dp);      // lp2 . dW (1/a^2 = 1/c^2)
        const T a3 = sg * w2 * dv;                    // lp3 . dW
        const T uc = Wa.vx * ca;
        F.d += a1 + a2 + a4;
        F.mx += a1 * la1 + a2 * Wa.vx + a4 * la4;
        F.my += (a1 + a2 + a4) * Wa.vy + a3 * Wa.d;
        F.E += a1 * (ha - uc) + a2 * (K::HALF * (Wa.vx * Wa.vx + Wa.vy * Wa.vy)) + a3 * (Wa.d * Wa.vy) +
               a4 * (ha + uc);
        return F;
    }
    const PS<T> Wa = roe_average(Wl, Wr);
    const PS<T> dW = {Wr.d - Wl.d, Wr.vx - Wl.vx, Wr.vy - Wl.vy, Wr.p - Wl.p};
    // a() = sqrt(g*p/d) and a2() = g*p/d share the quotient; h() is needed by
    // rc_x for waves 1 and 4, computed once
    const T a2a = W_a2(Wa);
    const T cl = W_a(Wl), cr = W_a(Wr), ca = qsqrt(a2a);
    const T ha = W_h(Wa);
    const T la1 = Wa.vx - ca, la4 = Wa.vx + ca;
    const bool left = (Wa.vx >= K::ZERO);
    const T d1 = harten_delta(la1, Wl.vx - cl, Wr.vx - cr);
    const T d4 = harten_delta(la4, Wl.vx + cl, Wr.vx + cr);
    const T w1 = left ? K::HALF * (la1 - tabs(la1) - d1) : K::HALF * (la1 + tabs(la1) + d1);
    const T w2 = left ? K::HALF * (Wa.vx - tabs(Wa.vx)) : K::HALF * (Wa.vx + tabs(Wa.vx));
    const T w4 = left ? K::HALF * (la4 - tabs(la4) - d4) : K::HALF * (la4 + tabs(la4) + d4);
    const PS<T> Wu = left ? Wl : Wr;
    CS<T> F = W_Fx(Wu);
    const T sg = left ? T(1) : T(-1);   // Flux += (left) / Flux -= (right): exact sign flip
    // waves 1 and 4 share lp4 = HALF/c^2 and, up to sign, lp2 = -+HALF*d/c:
    // (-HALF*d)/c is -(HALF*d/c) bit for bit (a negated product, then a
    // negated quotient), so each is divided once
    T lp2w4 = K::ZERO, lp4w14 = K::ZERO;
    if (w1 != K::ZERO || w4 != K::ZERO) {
        lp2w4 = qdiv(K::HALF * Wa.d, ca);
        lp4w14 = qdiv(K::HALF, sqr(ca));
    }
    // wave 1: lp = (0, -HALF*d/c, 0, HALF/c^2), rc = (1, u-c, v, h-u*c)
    if (Fast<T>::on || w1 != K::ZERO) {
        const T lp2 = -lp2w4, lp4 = lp4w14;
        const T a = sg * (w1 * (K::ZERO * dW.d + (lp2 * dW.vx + K::ZERO * dW.vy) + lp4 * dW.p));
        F.d += a * K::ONE;
        F.mx += a * (Wa.vx - ca);
        F.my += a * Wa.vy;
        F.E += a * (ha - Wa.vx * ca);
    }
    // waves 2 and 3 share the eigenvalue u; lp2 = (1,0,0,-1/a2), rc2 =
    // (1,u,v,|v|^2/2); lp3 = (0,0,1,0), rc3 = (0,0,d,d*v)
    if (Fast<T>::on || w2 != K::ZERO) {
        const T lp4 = qdiv(-K::ONE, a2a);
        const T a = sg * (w2 * (K::ONE * dW.d + (K::ZERO * dW.vx + K::ZERO * dW.vy) + lp4 * dW.p));
        F.d += a * K::ONE;
        F.mx += a * Wa.vx;
        F.my += a * Wa.vy;
        F.E += a * (K::HALF * (Wa.vx * Wa.vx + Wa.vy * Wa.vy));
        const T b = sg * (w2 * (K::ZERO * dW.d + (K::ZERO * dW.vx + K::ONE * dW.vy) + K::ZERO * dW.p));
        F.d += b * K::ZERO;
        F.mx += b * K::ZERO;
        F.my += b * Wa.d;
        F.E += b * (Wa.d * Wa.vy);
    }
    // wave 4: lp = (0, HALF*d/c, 0, HALF/c^2), rc = (1, u+c, v, h+u*c)
    if (Fast<T>::on || w4 != K::ZERO) {
        const T lp2 = lp2w4, lp4 = lp4w14;
        const T a = sg * (w4 * (K::ZERO * dW.d + (lp2 * dW.vx + K::ZERO * dW.vy) + lp4 * dW.p));
        F.d += a * K::ONE;
        F.mx += a * (Wa.vx + ca);
        F.my += a * Wa.vy;
        F.E += a * (ha + Wa.vx * ca);
    }
    return F;
}

// FluxRoe_n, Euler2DState.cc:2446-2489
template <typename T> CFFC_HD CS<T> flux_roe_n(const PS<T> &Wl, const PS<T> &Wr, T nx, T ny) {
    PS<T> a = {Wl.d, Wl.vx * nx + Wl.vy * ny, -Wl.vx * ny + Wl.vy * nx, Wl.p};
    PS<T> b = {Wr.d, Wr.vx * nx + Wr.vy * ny, -Wr.vx * ny + Wr.vy * nx, Wr.p};
    CS<T> Fr = flux_roe_x(a, b);
    CS<T> F = {Fr.d, Fr.mx * nx - Fr.my * ny, Fr.mx * ny + Fr.my * nx, Fr.E};
    return F;
}

// FluxHLLE_x / FluxHLLE_n, Euler2DState.cc:2921-2963, 3046-3095: wave speeds
// from the left, right and Roe-average eigenvalues u -/+ a, then
// (sR*Fl - sL*Fr + sL*sR*(Ur - Ul)) / (sR - sL).  FP64 only (the coarse
// levels of the multigrid solver).
template <typename T> CFFC_HD CS<T> flux_hlle_x(const PS<T> &Wl, const PS<T> &Wr) {
    const PS<T> Wa = roe_average(Wl, Wr);
    const CS<T> Ul = U_of_W(Wl), Ur = U_of_W(Wr);
    const T sl = dmin(Wl.vx - W_a(Wl), Wa.vx - W_a(Wa));
    const T sr = dmax(Wr.vx + W_a(Wr), Wa.vx + W_a(Wa));
    if (sl >= T(0)) return W_Fx(Wl);
    if (sr <= T(0)) return W_Fx(Wr);
    const CS<T> Fl = W_Fx(Wl), Fr = W_Fx(Wr);
    const T slr = sl * sr;
    const Recip<T> rden = recip(sr - sl);
    CS<T> F;
    F.d = rdiv((sr * Fl.d - sl * Fr.d) + slr * (Ur.d - Ul.d), rden);
    F.mx = rdiv((sr * Fl.mx - sl * Fr.mx) + slr * (Ur.mx - Ul.mx), rden);
    F.my = rdiv((sr * Fl.my - sl * Fr.my) + slr * (Ur.my - Ul.my), rden);
    F.E = rdiv((sr * Fl.E - sl * Fr.E) + slr * (Ur.E - Ul.E), rden);
    return F;
}
template <typename T> CFFC_HD CS<T> flux_hlle_n(const PS<T> &Wl, const PS<T> &Wr, T nx, T ny) {
    PS<T> a = {Wl.d, Wl.vx * nx + Wl.vy * ny, -Wl.vx * ny + Wl.vy * nx, Wl.p};
    PS<T> b = {Wr.d, Wr.vx * nx + Wr.vy * ny, -Wr.vx * ny + Wr.vy * nx, Wr.p};
    CS<T> Fr = flux_hlle_x(a, b);
    CS<T> F = {Fr.d, Fr.mx * nx - Fr.my * ny, Fr.mx * ny + Fr.my * nx, Fr.E};
    return F;
}

// Reflect, Euler2DState.cc:751-787
template <typename T> CFFC_HD PS<T> reflect(const PS<T> &W, T nx, T ny) {
    T ur = W.vx * nx + W.vy * ny;
    T vr = -W.vx * ny + W.vy * nx;
    ur = -ur;
    PS<T> r = {W.d, ur * nx - vr * ny, ur * ny + vr * nx, W.p};
    return r;
}

// pow of the characteristic boundary state (Euler2DState.cc:1432).  FP64:
// glibc's own method and roundings (glibc_pow.h), so the boundary density
// has the reference's bits; FP32 (mixed path): the device powf.
__device__ __forceinline__ double bc_pow(double x, double y) { return cffc_pow_glibc(x, y); }
__device__ __forceinline__ float bc_pow(float x, float y) { return powf(x, y); }

// BC_Characteristic_Pressure, Euler2DState.cc:1383-1465
template <typename T> CFFC_HD PS<T> bc_characteristic_pressure(const PS<T> &Wi, const PS<T> &Wo, T nx, T ny) {
    typedef Consts<T> K;
    PS<T> Wir = {Wi.d, Wi.vx * nx + Wi.vy * ny, -Wi.vx * ny + Wi.vy * nx, Wi.p};
    PS<T> Wor = {Wo.d, Wo.vx * nx + Wo.vy * ny, -Wo.vx * ny + Wo.vy * nx, Wo.p};
    T mi = Wir.vx / W_a(Wir);
    T db, pb, ab, ubr, vbr;
    if (mi >= K::ONE) {
        db = Wir.d; ubr = Wir.vx; vbr = Wir.vy; pb = Wir.p;
    } else if (mi >= K::ZERO) {
        pb = Wor.p;
        db = Wir.d * bc_pow(pb / Wir.p, K::ONE / gam<T>());
        ab = tsqrt(gam<T>() * pb / db);
        ubr = Wir.vx + K::TWO * (W_a(Wir) - ab) * gm1i<T>();
        vbr = Wir.vy;
    } else if (mi >= -K::ONE) {
        pb = Wor.p;
        db = Wor.d;
        vbr = Wor.vy;
        ab = tsqrt(gam<T>() * pb / db);
        ubr = Wir.vx + K::TWO * (W_a(Wir) - ab) * gm1i<T>();
    } else {
        db = Wor.d; ubr = Wor.vx; vbr = Wor.vy; pb = Wor.p;
    }
    PS<T> r = {db, ubr * nx - vbr * ny, ubr * ny + vbr * nx, pb};
    return r;
}

// Limiter_Venkatakrishnan for one quadrature value, CFD.cc:728-752.  The
// caller folds phi = min(phi, f) over the four face midpoints.  For y >= 2
// the rational term f(y) = (y^2+2y)/(y^2+y+2) is >= 1 and cannot lower
// phi <= 1.  When a multiply-only test proves y >= 2.001 (a margin far above
// rounding) and y is nowhere near overflow, both divisions are skipped: the
// result is bit-identical to evaluating the reference's expression.
template <typename T> CFFC_HD T venkat_fold(T phi, T uq, T u0, T u0Min, T u0Max) {
    typedef Consts<T> K;
    const T du = uq - u0;
    if constexpr (Fast<T>::on) {
        // mixed path: branch-free.  y >= 0 always (num and du share their
        // sign); capping y at 1e6 keeps y^2 finite, and f(1e6) rounds to 1,
        // which cannot lower phi <= 1 (nor can any f(y >= 2)).  Inside the
        // dead band the quotient is discarded by the select.
        const T num = du > T(0) ? u0Max - u0 : u0Min - u0;
        const T y = fminf(num * __builtin_amdgcn_rcpf(du), T(1.0e6));
        const T f = (y * y + K::TWO * y) * __builtin_amdgcn_rcpf(y * y + y + K::TWO);
        return fminf(phi, fabsf(du) > K::TOLER ? f : K::ONE);
    }
    T num;
    if (du > K::TOLER) {
        num = u0Max - u0;  // y = num/du, both >= 0
        if (num >= T(2.002) * du && num <= T(1.0e12) * du) return phi;
    } else if (du < -K::TOLER) {
        num = u0Min - u0;  // y = num/du, both <= 0
        if (num <= T(2.002) * du && num >= T(1.0e12) * du) return phi;
    } else {
        return vmin64(phi, K::ONE);
    }
    // phi and f(y) are never -0 or NaN (phi starts at 1, f(+-0) = +0; y is
    // finite since |du| > TOLER and |num| << 1e147), so the one-instruction
    // minimum equals the reference's compare-and-select
    const T y = qdiv(num, du);
    return vmin64(phi, qdiv(sqr(y) + K::TWO * y, sqr(y) + y + K::TWO));
}

}  // namespace cffc
