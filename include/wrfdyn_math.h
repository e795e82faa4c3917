/*
 * Deterministic FP64 elementary functions for the dycore's equation of state
 * (pow), the log-pressure specific volume (log) and the Rayleigh damping
 * profile (sin).
 *
 * The same source is compiled for the GPU kernels (hipcc) and for host C code
 * (gcc), so a given input gives the same bits on both: every step is a plain
 * IEEE add/multiply/divide (build with -ffp-contract=off) or an explicit,
 * correctly rounded fma(); frexp/ldexp/rint are exact.  Accuracy is about one
 * ulp, comparable to a platform libm.
 *
 * Methods (textbook, written from the mathematics):
 *  - log: x = m 2^e, m in [1/sqrt2, sqrt2); ln m = 2 atanh(s), s = (m-1)/(m+1),
 *    with s carried as a double-double and the atanh series tail in double;
 *    e ln2 added with a two-part ln2.
 *  - exp: Cody-Waite reduction r = t - k ln2 (two-part ln2), Taylor series of
 *    e^r - 1 to degree 14 on |r| <= ln2/2, then scaling by 2^k.
 *  - pow(x, y) = exp(y ln x) with y ln x formed as a double-double.
 *  - sin: Cody-Waite reduction by pi/2 (three-part pi/2), Taylor series of
 *    sin / cos on |r| <= pi/4.
 * Domain: pow for finite x > 0 with |y ln x| < 700; sin for |x| < 1e5.
 */
#ifndef WRFDYN_MATH_H
#define WRFDYN_MATH_H

#include <math.h>

#if defined(__HIPCC__)
#define WD_MATH_FN __host__ __device__ static inline
#else
#define WD_MATH_FN static inline
#endif

/* Cody-Waite split constants: the first part is the constant rounded to 33
 * significant bits (so k * part is exact for |k| < 2^20), the next part the
 * remainder rounded likewise, the last part the remainder to 53 bits.
 *   ln2  - (LN2_HI + LN2_LO)           ~ -1.3e-27
 *   pi/2 - (PIO2_1 + PIO2_2 + PIO2_3)  ~  1.0e-37                      */
#define WD_LN2_HI 0x1.62e42ff000000p-1
#define WD_LN2_LO (-0x1.718432a1b0e26p-35)
#define WD_INV_LN2 0x1.71547652b82fep+0
#define WD_PIO2_1 0x1.921fb54400000p+0
#define WD_PIO2_2 0x1.0b4611a600000p-34
#define WD_PIO2_3 0x1.3198a2e037073p-69
#define WD_INV_PIO2 0x1.45f306dc9c883p-1

/* natural log as a double-double hi + lo, x finite and > 0 */
WD_MATH_FN void wd_log_dd(double x, double *hi, double *lo) {
    int e;
    double m = frexp(x, &e); /* x = m 2^e, m in [0.5, 1) */
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    double f = m - 1.0;          /* exact (Sterbenz) */
    double v_hi = 2.0 + f;       /* |f| < 2: Fast2Sum */
    double v_lo = f - (v_hi - 2.0);
    double s_hi = f / v_hi;
    double r = fma(-s_hi, v_hi, f);
    r = r - s_hi * v_lo;
    double s_lo = r / v_hi;
    /* ln m = 2 s + 2 s^3 (1/3 + t/5 + t^2/7 + ...), t = s^2 */
    double t = s_hi * s_hi;
    double p = 1.0 / 29.0;
    p = p * t + 1.0 / 27.0;
    p = p * t + 1.0 / 25.0;
    p = p * t + 1.0 / 23.0;
    p = p * t + 1.0 / 21.0;
    p = p * t + 1.0 / 19.0;
    p = p * t + 1.0 / 17.0;
    p = p * t + 1.0 / 15.0;
    p = p * t + 1.0 / 13.0;
    p = p * t + 1.0 / 11.0;
    p = p * t + 1.0 / 9.0;
    p = p * t + 1.0 / 7.0;
    p = p * t + 1.0 / 5.0;
    p = p * t + 1.0 / 3.0;
    double tail = (2.0 * s_hi) * t * p;
    double lm_hi = 2.0 * s_hi;                 /* exact */
    double lm_lo = 2.0 * s_lo + tail;
    /* + e ln2 */
    double de = (double)e;
    double a = de * WD_LN2_HI;                 /* exact */
    double sum = a + lm_hi;                    /* TwoSum */
    double bb = sum - a;
    double err = (a - (sum - bb)) + (lm_hi - bb);
    double lo2 = err + (lm_lo + de * WD_LN2_LO);
    double h = sum + lo2;                      /* renormalise */
    *lo = lo2 - (h - sum);
    *hi = h;
}

/* e^(hi + lo), |hi| < 700 */
WD_MATH_FN double wd_exp_dd(double hi, double lo) {
    double k = rint(hi * WD_INV_LN2);
    double r = fma(-k, WD_LN2_HI, hi);
    r = fma(-k, WD_LN2_LO, r);
    r = r + lo;
    /* e^r - 1 = r + r^2 (1/2 + r/6 + ...) */
    double q = 1.0 / 87178291200.0;            /* 1/14! */
    q = q * r + 1.0 / 6227020800.0;
    q = q * r + 1.0 / 479001600.0;
    q = q * r + 1.0 / 39916800.0;
    q = q * r + 1.0 / 3628800.0;
    q = q * r + 1.0 / 362880.0;
    q = q * r + 1.0 / 40320.0;
    q = q * r + 1.0 / 5040.0;
    q = q * r + 1.0 / 720.0;
    q = q * r + 1.0 / 120.0;
    q = q * r + 1.0 / 24.0;
    q = q * r + 1.0 / 6.0;
    q = q * r + 0.5;
    double em1 = r + (r * r) * q;
    return ldexp(1.0 + em1, (int)k);
}

/* ln(x) for finite x > 0: the double-double log rounded to double */
WD_MATH_FN double wd_log(double x) {
    double lh, ll;
    wd_log_dd(x, &lh, &ll);
    return lh + ll;
}

/* x^y for finite x > 0 */
WD_MATH_FN double wd_pow(double x, double y) {
    double lh, ll;
    wd_log_dd(x, &lh, &ll);
    double p_hi = y * lh;
    double p_lo = fma(y, lh, -p_hi) + y * ll;
    double t_hi = p_hi + p_lo;
    double t_lo = p_lo - (t_hi - p_hi);
    return wd_exp_dd(t_hi, t_lo);
}

/* sin(x), |x| < 1e5 */
WD_MATH_FN double wd_sin(double x) {
    double k = rint(x * WD_INV_PIO2);
    double r = fma(-k, WD_PIO2_1, x);
    r = fma(-k, WD_PIO2_2, r);
    r = fma(-k, WD_PIO2_3, r);
    int q = ((int)k) & 3;
    double t = r * r;
    double s, c;
    /* sin r = r + r^3 (-1/3! + t/5! - ...) */
    double ps = -1.0 / 121645100408832000.0;   /* -1/19! */
    ps = ps * t + 1.0 / 355687428096000.0;     /* 1/17! */
    ps = ps * t - 1.0 / 1307674368000.0;       /* -1/15! */
    ps = ps * t + 1.0 / 6227020800.0;
    ps = ps * t - 1.0 / 39916800.0;
    ps = ps * t + 1.0 / 362880.0;
    ps = ps * t - 1.0 / 5040.0;
    ps = ps * t + 1.0 / 120.0;
    ps = ps * t - 1.0 / 6.0;
    s = r + (r * t) * ps;
    /* cos r = 1 - t/2 + t^2 (1/4! - t/6! + ...) */
    /* pc = -(1/4! - t/6! + t^2/8! - ...), Horner from the t^8 term */
    double pc = -1.0 / 2432902008176640000.0; /* -1/20! */
    pc = pc * t + 1.0 / 6402373705728000.0;    /* 1/18! */
    pc = pc * t - 1.0 / 20922789888000.0;      /* -1/16! */
    pc = pc * t + 1.0 / 87178291200.0;         /* 1/14! */
    pc = pc * t - 1.0 / 479001600.0;           /* -1/12! */
    pc = pc * t + 1.0 / 3628800.0;
    pc = pc * t - 1.0 / 40320.0;
    pc = pc * t + 1.0 / 720.0;
    pc = pc * t - 1.0 / 24.0;
    double hz = 0.5 * t;
    double w = 1.0 - hz;
    c = w + (((1.0 - w) - hz) + (t * t) * (-pc));
    if (q == 0) return s;
    if (q == 1) return c;
    if (q == 2) return -s;
    return -c;
}

#endif
