/* kd_math.h -- exp, log and pow in FP64 from IEEE +, -, *, /, fma and exact
 * scaling only, so a host build (C or C++) and the gfx950 device build
 * return the same bits for the same arguments.
 *
 * The hot path's transcendentals (the entropy S(rho, u) of calc_Sfromu,
 * physics.c:907-917, and of the entropy inversion, u2p.c:914-974; the
 * Planck source B = sigma T^4 / pi of calc_Gi, rad.c:1787-1888; the M1 slow
 * root bound of u2prad.c) call these.  A device libm and the host libm may
 * round a few arguments differently; the same functions on both sides make
 * a device run and its CPU check agree bit for bit.
 *
 * Methods (from the published algorithms, tables computed by
 * gen_kd_math_tables.py with 60-digit decimal arithmetic):
 *  - exp: P. T. P. Tang, "Table-driven implementation of the exponential
 *    function in IEEE floating-point arithmetic", ACM TOMS 15 (1989):
 *    x = (64 m + j) ln2/64 + r, |r| <= ln2/128, exp x = 2^m 2^(j/64) e^r
 *    with 2^(j/64) as a double-double, e^r - 1 a degree-6 polynomial and
 *    the leading products and sums formed exactly (fma, two-sum).
 *  - log: P. T. P. Tang, "Table-driven implementation of the logarithm
 *    function in IEEE floating-point arithmetic", ACM TOMS 16 (1990):
 *    x = 2^e m, m in [1, 2), c = 1/(1 + (j + 1/2)/128) for the interval j
 *    of m, log x = e ln2 - log c + log1p(m c - 1), the product m c - 1
 *    formed exactly as a double-double, the table -log c as double-double
 *    and log1p by its series to the 8th power (|m c - 1| < 2^-8); within
 *    2^-7 of 1, log1p(x - 1) by its series to the 11th power.
 *  - pow(x, y) = exp(y log x) with log x as a double-double and y log x
 *    formed as a double-double, C99 Annex F special cases.
 * The result before the final rounding is within about 2^-66 relative of
 * the exact value, so results are correctly rounded except within that
 * distance of a rounding boundary (measured against 60-digit references in
 * tests/test_kd_math.py).  Results that underflow
 * to subnormals are rounded twice (at most 1 ulp of the subnormal).
 */
#ifndef KD_MATH_H
#define KD_MATH_H
#include <math.h>
#include "kd_math_tables.h"

#if defined(__HIP__)
#define KD_MATHFN __host__ __device__ static inline
#else
#define KD_MATHFN static inline
#endif

#ifdef __cplusplus
static constexpr double kd_exp_tab[128] = KD_EXP_TAB;
static constexpr double kd_log_c[128] = KD_LOG_C;
static constexpr double kd_log_tab[256] = KD_LOG_TAB;
#else
static const double kd_exp_tab[128] = KD_EXP_TAB;
static const double kd_log_c[128] = KD_LOG_C;
static const double kd_log_tab[256] = KD_LOG_TAB;
#endif

/* s + e = a + b exactly (Knuth's two-sum) */
KD_MATHFN double kd_two_sum(double a, double b, double *e)
{
  double s = a + b;
  double bb = s - a;
  *e = (a - (s - bb)) + (b - bb);
  return s;
}

/* e^(h + l), |l| <= ulp(h): 2^m 2^(j/64) e^r */
KD_MATHFN double kd_exp_dd(double h, double l)
{
  if (h != h) return h;
  if (h > 709.782712893384) return HUGE_VAL;
  if (h < -745.1332191019412) return 0.;
  /* k = nearest integer to h 64/ln2 (the shifter rounds to even) */
  const double shift = 0x1.8p52;
  double kd = (h * KD_INV_L64 + shift) - shift;
  int k = (int)kd;
  /* r = h - k ln2/64 + l = rh + rl: k L64_HI is exact (36-bit head,
     |k| < 2^17) and h - k L64_HI is exact (Sterbenz) */
  double rh = h - kd * KD_L64_HI;
  double rl = l - kd * KD_L64_LO;
  double re;
  double r = kd_two_sum(rh, rl, &re);
  int j = k & 63;
  int m = (k - j) / 64;
  /* e^r - 1 = r + q, q = r^2/2 + ... + r^6/720, |r| < 0.0055 */
  double q = r * r * (0.5 + r * (1. / 6. + r * (1. / 24. + r * (1. / 120. + r * (1. / 720.)))));
  double th = kd_exp_tab[2 * j], tl = kd_exp_tab[2 * j + 1];
  /* 2^(j/64) e^r = th + th r + (tl + th (re + q) + tl r), th r exact as t1 + t2 */
  double t1 = th * r;
  double t2 = fma(th, r, -t1);
  double e1;
  double s = kd_two_sum(th, t1, &e1);
  double v = s + (e1 + (t2 + (tl + (th * (re + q) + tl * r))));
  /* 2^m in two exact steps where one would leave the exponent range */
  if (m < -1000) return ldexp(ldexp(v, m + 1000), -1000);
  if (m > 1000) return ldexp(ldexp(v, m - 1000), 1000);
  return ldexp(v, m);
}

KD_MATHFN double kd_exp(double x) { return kd_exp_dd(x, 0.); }

/* log x = hi + lo for finite x > 0 (double-double, about 2^-68 absolute
   for |log x| < 1 and 2^-60 relative otherwise) */
KD_MATHFN void kd_log_dd(double x, double *hi, double *lo)
{
  if (x > 1. - 0x1p-7 && x < 1. + 0x1p-7) {
    /* log1p(r), r = x - 1 exact: r - r^2/2 (exact parts) + series to r^11 */
    double r = x - 1.;
    double r2h = r * r;
    double r2l = fma(r, r, -r2h);
    double e0;
    double s0 = kd_two_sum(r, -0.5 * r2h, &e0);
    double t = r2h * r * (1. / 3. + r * (-0.25 + r * (0.2 + r * (-1. / 6. + r * (1. / 7. + r * (-0.125 +
                   r * (1. / 9. + r * (-0.1 + r * (1. / 11.)))))))));
    double rest = e0 + (t - 0.5 * r2l);
    double h = s0 + rest;
    *lo = rest - (h - s0);
    *hi = h;
    return;
  }
  int e = 0;
  double m = frexp(x, &e); /* x = m 2^e, m in [0.5, 1) */
  m = m * 2.;
  e = e - 1; /* m in [1, 2) */
  int j = (int)((m - 1.) * 128.);
  if (j > 127) j = 127;
  double c = kd_log_c[j];
  /* m c = ph + pl exactly; ph - 1 is exact (ph in [0.98, 1.02]) */
  double ph = m * c;
  double pl = fma(m, c, -ph);
  double ra = ph - 1.;
  double r = ra + pl;
  double rb = pl - (r - ra); /* r + rb = ra + pl exactly */
  /* log1p(r) - r = -r^2/2 + r^3/3 - ... + r^8/8 */
  double t = r * r * (-0.5 + r * (1. / 3. + r * (-0.25 + r * (0.2 + r * (-1. / 6. + r * (1. / 7. + r * (-0.125)))))));
  /* sum the leading parts with two-sums, collect the small ones */
  double a = (double)e * KD_LN2_HI; /* exact */
  double err1, err2;
  double s1 = kd_two_sum(a, kd_log_tab[2 * j], &err1);
  double s2 = kd_two_sum(s1, r, &err2);
  double small = (double)e * KD_LN2_LO + kd_log_tab[2 * j + 1] + rb + t + err1 + err2;
  double h = s2 + small;
  *lo = small - (h - s2);
  *hi = h;
}

KD_MATHFN double kd_log(double x)
{
  if (x != x) return x;
  if (x < 0.) return NAN;
  if (x == 0.) return -HUGE_VAL;
  if (x == HUGE_VAL) return x;
  double h, l;
  kd_log_dd(x, &h, &l);
  return h + l;
}

/* y is an integer: 1 odd, 2 even, 0 not an integer */
KD_MATHFN int kd_int_kind(double y)
{
  if (y != y || fabs(y) == HUGE_VAL) return 0;
  if (floor(y) != y) return 0;
  if (fabs(y) >= 0x1p53) return 2;
  double h = y * 0.5;
  return floor(h) == h ? 2 : 1;
}

KD_MATHFN double kd_pow(double x, double y)
{
  if (y == 0.) return 1.;
  if (x == 1.) return 1.;
  if (x != x || y != y) return x + y;
  const int yk = kd_int_kind(y);
  const double ax = fabs(x);
  if (fabs(y) == HUGE_VAL) {
    if (ax == 1.) return 1.;
    if ((ax < 1.) == (y < 0.)) return HUGE_VAL;
    return 0.;
  }
  if (x == 0.) {
    double r = y < 0. ? HUGE_VAL : 0.;
    return (yk == 1 && signbit(x)) ? -r : r;
  }
  if (ax == HUGE_VAL) {
    double r = y < 0. ? 0. : HUGE_VAL;
    return (x < 0. && yk == 1) ? -r : r;
  }
  if (x < 0. && yk == 0) return NAN;
  double lh, ll;
  kd_log_dd(ax, &lh, &ll);
  double ph = y * lh;
  double pl = fma(y, lh, -ph) + y * ll;
  double h = ph + pl;
  double l = pl - (h - ph);
  double r = kd_exp_dd(h, l);
  return (x < 0. && yk == 1) ? -r : r;
}

#endif
