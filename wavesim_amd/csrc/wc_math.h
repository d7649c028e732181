// Portable, IEEE-only elementary functions used identically on the host and on gfx950.
//
// wavecore calls the platform libm for `hypot` (the breaking switch, crates/wavecore/src/
// dispersion.rs:148, and Manning friction, solver.rs:565) and `powf(4/3)` (Manning,
// forcing.rs:295). A GPU libm differs from glibc in the last bit now and then, and a bit
// flipped in one cell is enough to make two long runs part at 1e-12. These routines use
// only +, -, *, /, sqrt, frexp and ldexp, all exact or correctly rounded in IEEE-754
// arithmetic, so the same source gives the same bits on any conforming target as long as
// nothing fuses a*b+c into an FMA: every file that includes this header is compiled with
// -ffp-contract=off, and the routines below never rely on a fused form either.
//
// Both are accurate to well under one unit in the last place and agree with glibc 2.35's
// correctly rounded results except in rare hard-to-round cases (measured by the tests).
#pragma once

#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define WC_HD __host__ __device__
#else
#define WC_HD
#endif

// Physical constants and thresholds of the scheme (crates/wavecore/src/solver.rs:26-32,
// dispersion.rs:41, forcing.rs:286, grid.rs:4).
#define WC_GHOST 3
#define WC_G 9.81
#define WC_H_DRY 1e-8
#define WC_H_SECOND 1e-3
#define WC_H_FACE_MIN 1e-4
#define WC_H_DISPERSION_MIN 0.05
#define WC_H_FRICTION_MIN 1e-3

// Rust's f64::max / f64::min: a NaN operand yields the other one. Rust leaves the sign of a
// zero tie unspecified; here a tie always returns the first operand, on every target.
WC_HD static inline double wc_max(double a, double b) {
  if (isnan(a)) return b;
  if (isnan(b)) return a;
  return a < b ? b : a;
}
WC_HD static inline double wc_min(double a, double b) {
  if (isnan(a)) return b;
  if (isnan(b)) return a;
  return b < a ? b : a;
}
// par.rs max_or_nan: NaN if either is NaN, else the larger.
WC_HD static inline double wc_max_or_nan(double a, double b) {
  if (isnan(a) || isnan(b)) return NAN;
  return wc_max(a, b);
}
// Rust's f64::clamp: NaN passes through.
WC_HD static inline double wc_clamp(double x, double lo, double hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
// dispersion.rs:108-111.
WC_HD static inline double wc_smoothstep(double lo, double hi, double x) {
  const double t = wc_clamp((x - lo) / (hi - lo), 0.0, 1.0);
  return t * t * (3.0 - 2.0 * t);
}
// The breaking switch's "off" (dispersion.rs:140-150): the larger of the erosion smoothstep of
// the depth ratio and the steepness smoothstep of hypot(ex, ey), hypot being the portable one
// below. hypot is the costly part, and most cells do not need its last bits: where the larger
// slope component alone already puts hypot past the steepness band (hypot >= max(|ex|, |ey|),
// and a correctly rounded hypot stays within 2^-50 of it), the smoothstep is exactly 1; where
// 1.5 times it stays below the band (hypot <= sqrt 2 max(|ex|, |ey|)), exactly 0; and where the
// erosion smoothstep is 1 the larger of the two is 1 whatever the other. Those cells skip hypot;
// the result has the bits of the full expression (wc_switch_off_full).
WC_HD static inline double wc_hypot(double a, double b);
WC_HD static inline double wc_switch_off_full(double er_lo, double er_hi, double sl_lo,
                                              double sl_hi, double ratio, double ex, double ey) {
  return wc_max(wc_smoothstep(er_lo, er_hi, ratio), wc_smoothstep(sl_lo, sl_hi, wc_hypot(ex, ey)));
}
WC_HD static inline double wc_switch_off(double er_lo, double er_hi, double sl_lo, double sl_hi,
                                         double ratio, double ex, double ey) {
  const double er = wc_smoothstep(er_lo, er_hi, ratio);
  const double ax = fabs(ex), ay = fabs(ey);
  const double mx = ax > ay ? ax : ay;
  if (sl_lo < sl_hi && mx <= 0x1p1000) {  // finite slopes (NaN fails the test)
    if (er == 1.0) return er;
    if (mx >= sl_hi * (1.0 + 0x1p-40)) return wc_max(er, 1.0);
    if (1.5 * mx <= sl_lo) return wc_max(er, 0.0);
  }
  return wc_max(er, wc_smoothstep(sl_lo, sl_hi, wc_hypot(ex, ey)));
}
// solver.rs:90-92: velocity of a cell, 0 where it is dry.
WC_HD static inline double wc_velocity(double h, double hm) { return h > WC_H_DRY ? hm / h : 0.0; }

// Veltkamp split of x into two halves whose products are exact (|x| < 2^995).
WC_HD static inline void wc_split(double x, double* hi, double* lo) {
  const double c = 134217729.0 * x;  // 2^27 + 1
  const double h = c - (c - x);
  *hi = h;
  *lo = x - h;
}

// Dekker's exact product without FMA: a*b = p + e exactly.
WC_HD static inline void wc_two_prod(double a, double b, double* p, double* e) {
  double ah, al, bh, bl;
  const double x = a * b;
  wc_split(a, &ah, &al);
  wc_split(b, &bh, &bl);
  *p = x;
  *e = ((ah * bh - x) + ah * bl + al * bh) + al * bl;
}

// Knuth's exact sum: a+b = s + e exactly.
WC_HD static inline void wc_two_sum(double a, double b, double* s, double* e) {
  const double x = a + b;
  const double bv = x - a;
  const double av = x - bv;
  *s = x;
  *e = (a - av) + (b - bv);
}

// sqrt(a^2 + b^2) without undue overflow or underflow, within a hair of correct rounding.
// IEEE special cases as C99: an infinite argument gives +inf even beside a NaN.
WC_HD static inline double wc_hypot(double a, double b) {
  a = fabs(a);
  b = fabs(b);
  if (isinf(a) || isinf(b)) return INFINITY;
  if (isnan(a) || isnan(b)) return NAN;
  double hi = a > b ? a : b;
  double lo = a > b ? b : a;
  if (hi == 0.0) return 0.0;
  // Scale by an exact power of two so the squares neither overflow nor lose the low part.
  int scale = 0;
  if (hi > 0x1p+450) {
    scale = 600;
  } else if (hi < 0x1p-450) {
    scale = -600;
  }
  if (scale != 0) {
    hi = ldexp(hi, -scale);
    lo = ldexp(lo, -scale);
  }
  double p1, e1, p2, e2, s, t;
  wc_two_prod(hi, hi, &p1, &e1);
  wc_two_prod(lo, lo, &p2, &e2);
  wc_two_sum(p1, p2, &s, &t);
  t = t + (e1 + e2);
  const double big = s + t;
  const double small = t - (big - s);
  const double r = sqrt(big);
  double r2, re;
  wc_two_prod(r, r, &r2, &re);
  const double d = ((big - r2) - re) + small;
  double out = r + d / (2.0 * r);
  if (scale != 0) out = ldexp(out, scale);
  return out;
}

// Natural logarithm to about 1e-9 relative, used only for the tiny correction in wc_pow43.
WC_HD static inline double wc_log_rough(double x) {
  int e;
  double m = frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
  if (m < 0.70710678118654752) {
    m = 2.0 * m;
    e -= 1;
  }
  // ln(m) = 2 atanh(z), z = (m - 1)/(m + 1), |z| < 0.172
  const double z = (m - 1.0) / (m + 1.0);
  const double z2 = z * z;
  double s = 1.0 / 13.0;
  s = 1.0 / 11.0 + z2 * s;
  s = 1.0 / 9.0 + z2 * s;
  s = 1.0 / 7.0 + z2 * s;
  s = 1.0 / 5.0 + z2 * s;
  s = 1.0 / 3.0 + z2 * s;
  s = 1.0 + z2 * s;
  return 2.0 * z * s + (double)e * 0.69314718055994530942;
}

// h^y with y = 4/3 rounded to double (as `h.powf(4.0 / 3.0)` evaluates it), for finite h > 0.
// The cube root comes from Newton's method on the mantissa, is refined once in double-double,
// and the product h * cbrt(h) is corrected for y differing from 4/3 by 2^-52/3.
WC_HD static inline double wc_pow43(double h) {
  if (!(h > 0.0) || isinf(h)) {
    if (h == 0.0) return 0.0;
    if (isinf(h) && h > 0.0) return INFINITY;
    return NAN;
  }
  int e;
  const double m0 = frexp(h, &e);  // h = m0 2^e
  int q = e >= 0 ? e / 3 : -((-e + 2) / 3);
  const int r = e - 3 * q;  // 0, 1 or 2
  const double m = ldexp(m0, r);  // [0.5, 4)
  // Newton for y^3 = m from a linear start; 7 steps are far past double precision.
  double y = 0.75 + 0.25 * m;
  for (int it = 0; it < 7; ++it) {
    y = y - (y * y * y - m) / (3.0 * y * y);
  }
  const double c = ldexp(y, q);
  // One double-double correction: cbrt(h) = c + cl.
  double c2, c2e, c3, c3e;
  wc_two_prod(c, c, &c2, &c2e);
  wc_two_prod(c2, c, &c3, &c3e);
  const double c3l = c3e + c2e * c;
  const double d = (h - c3) - c3l;
  const double cl = d / (3.0 * c2);
  // h^(4/3) = h c + h cl, then times h^(y - 4/3) = 1 - (2^-52/3) ln h to first order.
  double ph, pe;
  wc_two_prod(h, c, &ph, &pe);
  double pl = pe + h * cl;
  const double corr = -(0x1p-52 / 3.0) * wc_log_rough(h);
  pl = pl + ph * corr;
  return ph + pl;
}
