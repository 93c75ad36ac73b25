/* glibc_pow.h -- pow(x, y) for finite x > 0 (normal) and finite y, returning
 * the same bits as x86-64 glibc's pow() (2.28 and later, the FMA variant the
 * library selects on CPUs with FMA and AVX2).
 *
 * Why: the reference's characteristic pressure boundary state evaluates
 * rho_b = rho_i * pow(p_b / p_i, 1/gamma) (src_2D/Euler2D/Euler2DState.cc:1432).
 * glibc's pow is accurate to about 0.52 ulp, not correctly rounded, so a
 * device pow with any other error pattern differs from it in the last bit
 * on a fraction of inputs, and the limiter's |u_q - u0| > TOLER test turns a
 * last-bit difference into a different trajectory.  Bit identity needs the
 * same method with the same roundings.
 *
 * Method (published with ARM Optimized Routines, shipped by glibc as
 * sysdeps/ieee754/dbl-64/e_pow.c): pow(x, y) = exp(y * log(x)), with
 *   log(x): x = 2^k z, z in [0x1.69555p-1, 0x1.69555p0); a 128-entry table
 *           gives 1/c (few bits, so r = z/c - 1 is exact with one FMA) and
 *           log(c) as hi + tail; log(x) = k ln2 + log(c) + log1p(r), with
 *           log1p(r) = r - r^2/2 + r^3 q(r) (degree-8 polynomial), carried as
 *           a double-double (hi, lo);
 *   exp(e): e = k ln2/128 + r; 2^(k/128) from a 128-entry table as
 *           H (1 + T); exp(r) - 1 by a degree-5 polynomial;
 *           result = s + s * tmp, s = H * 2^(k div 128).
 * The FMA build of glibc fuses every multiply whose single use is an add
 * (GCC's default contraction); the fma() calls below sit exactly there.
 * The tables are generated from their construction rule by
 * gen_glibc_pow_tables.py.  Only the general path is implemented: callers
 * pass x in the normal positive range and a y for which |y log x| < 512
 * (the boundary state's pressure ratio and 1/gamma), where glibc takes no
 * special-case branch.
 *
 * Plain C99 / C++: the host test (gcc) includes it with the default
 * qualifiers, the HIP kernels define CFFC_POW_FN / CFFC_POW_TAB first.
 */
#ifndef CFFC_GLIBC_POW_H
#define CFFC_GLIBC_POW_H

#include <stdint.h>
#ifndef __HIPCC__
#include <math.h>
#include <string.h>
#endif

#ifndef CFFC_POW_FN
#define CFFC_POW_FN static inline
#endif
#ifndef CFFC_POW_TAB
#define CFFC_POW_TAB static const
#endif

#include "glibc_pow_tables.h"

CFFC_POW_FN double cffc_pow_bits_to_double(uint64_t u) {
#ifdef __HIPCC__
    return __builtin_bit_cast(double, u);
#else
    double d;
    memcpy(&d, &u, sizeof d);
    return d;
#endif
}
CFFC_POW_FN uint64_t cffc_pow_double_to_bits(double d) {
#ifdef __HIPCC__
    return __builtin_bit_cast(uint64_t, d);
#else
    uint64_t u;
    memcpy(&u, &d, sizeof u);
    return u;
#endif
}

/* log1p polynomial, scaled as evaluated: c0 = -1/2, then -2, -2, 4, 4, -8, -8
 * times the minimax coefficients on |r| <= 0x1.6bp-8 */
#define CFFC_POW_A0 (-0x1p-1)
#define CFFC_POW_A1 (0x1.555555555556p-2 * -2)
#define CFFC_POW_A2 (-0x1.0000000000006p-2 * -2)
#define CFFC_POW_A3 (0x1.999999959554ep-3 * 4)
#define CFFC_POW_A4 (-0x1.555555529a47ap-3 * 4)
#define CFFC_POW_A5 (0x1.2495b9b4845e9p-3 * -8)
#define CFFC_POW_A6 (-0x1.0002b8b263fc3p-3 * -8)
#define CFFC_POW_LN2HI 0x1.62e42fefa3800p-1
#define CFFC_POW_LN2LO 0x1.ef35793c76730p-45
/* exp step: 128/ln2, -ln2/128 split in two, the round-to-integer shift and the
 * exp(r) - 1 polynomial coefficients of r^2 .. r^5 */
#define CFFC_POW_INVLN2N (0x1.71547652b82fep0 * 128)
#define CFFC_POW_NEGLN2HIN (-0x1.62e42fefa0000p-8)
#define CFFC_POW_NEGLN2LON (-0x1.cf79abc9e3b3ap-47)
#define CFFC_POW_SHIFT 0x1.8p52
#define CFFC_POW_C2 0x1.ffffffffffdbdp-2
#define CFFC_POW_C3 0x1.555555555543cp-3
#define CFFC_POW_C4 0x1.55555cf172b91p-5
#define CFFC_POW_C5 0x1.1111167a4d017p-7

CFFC_POW_FN double cffc_pow_glibc(double x, double y) {
    /* ---- log(x) as hi + lo ---- */
    const uint64_t ix = cffc_pow_double_to_bits(x);
    const uint64_t off = ix - 0x3fe6955500000000ULL;
    const int i = (int)((off >> 45) & 127);
    const double kd = (double)((int64_t)off >> 52);
    const double z = cffc_pow_bits_to_double(ix - (off & (0xfffULL << 52)));
    const double invc = cffc_pow_log_tab[i][0];
    const double logc = cffc_pow_log_tab[i][1];
    const double logctail = cffc_pow_log_tab[i][2];
    const double r = fma(z, invc, -1.0);                 /* exact */
    const double t1 = fma(kd, CFFC_POW_LN2HI, logc);     /* exact */
    const double t2 = t1 + r;
    const double lo1 = fma(kd, CFFC_POW_LN2LO, logctail);
    const double lo2 = t1 - t2 + r;
    const double ar = CFFC_POW_A0 * r;
    const double ar2 = r * ar;
    const double ar3 = r * ar2;
    const double hi0 = t2 + ar2;
    const double lo3 = fma(ar, r, -ar2);
    const double lo4 = t2 - hi0 + ar2;
    const double q = fma(ar2, fma(ar2, fma(r, CFFC_POW_A6, CFFC_POW_A5), fma(r, CFFC_POW_A4, CFFC_POW_A3)),
                         fma(r, CFFC_POW_A2, CFFC_POW_A1));
    const double lo = fma(ar3, q, lo1 + lo2 + lo3 + lo4);
    const double lhi = hi0 + lo;
    const double llo = hi0 - lhi + lo;
    /* ---- y * log(x) as ehi + elo ---- */
    const double ehi = y * lhi;
    const double elo = fma(y, llo, fma(y, lhi, -ehi));
    /* ---- exp(ehi + elo) ---- */
    double kd2 = fma(CFFC_POW_INVLN2N, ehi, CFFC_POW_SHIFT);
    const uint64_t ki = cffc_pow_double_to_bits(kd2);
    kd2 -= CFFC_POW_SHIFT;
    double re = fma(kd2, CFFC_POW_NEGLN2LON, fma(kd2, CFFC_POW_NEGLN2HIN, ehi));
    re += elo;
    const uint64_t idx = 2 * (ki & 127);
    const double tail = cffc_pow_bits_to_double(cffc_pow_exp_tab[idx]);
    const double scale = cffc_pow_bits_to_double(cffc_pow_exp_tab[idx + 1] + (ki << 45));
    const double r2 = re * re;
    const double tmp = fma(r2 * r2, fma(re, CFFC_POW_C5, CFFC_POW_C4),
                           fma(r2, fma(re, CFFC_POW_C3, CFFC_POW_C2), tail + re));
    return fma(scale, tmp, scale);
}

#endif /* CFFC_GLIBC_POW_H */
