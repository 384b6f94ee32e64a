// Correctly rounded x**y for the FP64 path, in double-double arithmetic.
//
// The reference evaluates REAL**REAL through the C library's pow
// (Schiller-Naumann drag Re**0.687, src/constitutive_2d.f90:5378; hindered
// settling (1 - a/a_max)**4.65, :4850-4851; entrainment Ri**2.4, :5049).  A
// device pow with a 1-ulp error bound gives a different last bit in a
// fraction of the calls, and the settling iteration's 1e-6 stop and the
// Newton tolerances amplify such bits along a run.  pow_cr returns the
// correctly rounded result (round to nearest) of the exact x**y instead:
//   log x = e ln2 + 2 atanh(s),  s = (m - 1)/(m + 1),  m in [sqrt(1/2), sqrt(2)),
//           the atanh series summed to s**43 in double-double (|s| <= 0.1716);
//   exp z = 2**k (1 + p),  p = expm1(r / 256) by its Taylor series to r**10,
//           squared back 8 times as p <- 2p + p**2;  ln2 split in three parts.
// The relative error of the double-double result is below 2**-95, so the
// final rounding is correct unless x**y lies within 2**-42 ulp of a rounding
// boundary.  Every function here compiles for the host as well as the device
// (the host build is what the CPU tests check against a quad-precision pow).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace imex {
namespace ddm {

struct dd {
  double hi, lo;
};

__host__ __device__ __forceinline__ dd two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  return dd{s, e};
}

// requires |a| >= |b| (or a == 0)
__host__ __device__ __forceinline__ dd fast_two_sum(double a, double b) {
  const double s = a + b;
  return dd{s, b - (s - a)};
}

__host__ __device__ __forceinline__ dd two_prod(double a, double b) {
  const double p = a * b;
  return dd{p, fma(a, b, -p)};
}

__host__ __device__ __forceinline__ dd add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  const dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_two_sum(s.hi, s.lo);
}

__host__ __device__ __forceinline__ dd mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return fast_two_sum(p.hi, p.lo);
}

__host__ __device__ __forceinline__ dd mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  p.lo += a.lo * b;
  return fast_two_sum(p.hi, p.lo);
}

__host__ __device__ __forceinline__ dd div(dd a, dd b) {
  const double q1 = a.hi / b.hi;
  dd r = add(a, mul_d(b, -q1));
  const double q2 = r.hi / b.hi;
  r = add(r, mul_d(b, -q2));
  const double q3 = r.hi / b.hi;
  return add(fast_two_sum(q1, q2), dd{q3, 0.0});
}

// 1/(2k+1), k = 0..21, as exact-to-2**-106 pairs
__host__ __device__ __forceinline__ dd inv_odd(int k) {
  switch (k) {
    case 0: return dd{0x1.0000000000000p+0, 0x0.0p+0};
    case 1: return dd{0x1.5555555555555p-2, 0x1.5555555555555p-56};
    case 2: return dd{0x1.999999999999ap-3, -0x1.999999999999ap-57};
    case 3: return dd{0x1.2492492492492p-3, 0x1.2492492492492p-57};
    case 4: return dd{0x1.c71c71c71c71cp-4, 0x1.c71c71c71c71cp-58};
    case 5: return dd{0x1.745d1745d1746p-4, -0x1.745d1745d1746p-59};
    case 6: return dd{0x1.3b13b13b13b14p-4, -0x1.3b13b13b13b14p-58};
    case 7: return dd{0x1.1111111111111p-4, 0x1.1111111111111p-60};
    case 8: return dd{0x1.e1e1e1e1e1e1ep-5, 0x1.e1e1e1e1e1e1ep-61};
    case 9: return dd{0x1.af286bca1af28p-5, 0x1.af286bca1af28p-59};
    case 10: return dd{0x1.8618618618618p-5, 0x1.8618618618618p-59};
    case 11: return dd{0x1.642c8590b2164p-5, 0x1.642c8590b2164p-60};
    case 12: return dd{0x1.47ae147ae147bp-5, -0x1.eb851eb851eb8p-61};
    case 13: return dd{0x1.2f684bda12f68p-5, 0x1.2f684bda12f68p-59};
    case 14: return dd{0x1.1a7b9611a7b96p-5, 0x1.1a7b9611a7b96p-61};
    case 15: return dd{0x1.0842108421084p-5, 0x1.0842108421084p-60};
    case 16: return dd{0x1.f07c1f07c1f08p-6, -0x1.f07c1f07c1f08p-61};
    case 17: return dd{0x1.d41d41d41d41dp-6, 0x1.0750750750750p-60};
    case 18: return dd{0x1.bacf914c1bad0p-6, -0x1.bacf914c1bad0p-60};
    case 19: return dd{0x1.a41a41a41a41ap-6, 0x1.0690690690690p-60};
    case 20: return dd{0x1.8f9c18f9c18fap-6, -0x1.f3831f3831f38p-61};
    default: return dd{0x1.7d05f417d05f4p-6, 0x1.7d05f417d05f4p-62};
  }
}

// 1/n!, n = 2..10
__host__ __device__ __forceinline__ dd inv_fact(int n) {
  switch (n) {
    case 2: return dd{0x1.0000000000000p-1, 0x0.0p+0};
    case 3: return dd{0x1.5555555555555p-3, 0x1.5555555555555p-57};
    case 4: return dd{0x1.5555555555555p-5, 0x1.5555555555555p-59};
    case 5: return dd{0x1.1111111111111p-7, 0x1.1111111111111p-63};
    case 6: return dd{0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65};
    case 7: return dd{0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73};
    case 8: return dd{0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76};
    case 9: return dd{0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73};
    default: return dd{0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76};
  }
}

// ln 2 = LN2_1 + LN2_2 + LN2_3 to about 2**-160
constexpr double LN2_1 = 0x1.62e42fefa39efp-1;
constexpr double LN2_2 = 0x1.abc9e3b39803fp-56;
constexpr double LN2_3 = 0x1.7b57a079a1934p-111;
constexpr double SQRT_HALF = 0x1.6a09e667f3bcdp-1;

// natural logarithm of a positive finite x
__host__ __device__ __forceinline__ dd log_dd(double x) {
  int e = 0;
  double m = frexp(x, &e);  // x = m 2**e, m in [0.5, 1)
  if (m < SQRT_HALF) {
    m = 2.0 * m;
    e -= 1;
  }
  const double f = m - 1.0;  // exact (Sterbenz)
  const dd s = div(dd{f, 0.0}, two_sum(m, 1.0));
  const dd s2 = mul(s, s);
  dd acc = inv_odd(21);
#pragma unroll
  for (int k = 20; k >= 0; k--) acc = add(inv_odd(k), mul(s2, acc));
  dd lm = mul(s, acc);
  lm.hi = 2.0 * lm.hi;
  lm.lo = 2.0 * lm.lo;
  const double ed = (double)e;
  dd el = add(two_prod(ed, LN2_1), two_prod(ed, LN2_2));
  el = add(el, dd{ed * LN2_3, 0.0});
  return add(el, lm);
}

// exp of a double-double argument, rounded to the nearest double
__host__ __device__ __forceinline__ double exp_dd(dd z) {
  if (z.hi > 709.8) return HUGE_VAL;
  if (z.hi < -745.2) return 0.0;
  const double kd = floor(z.hi * 0x1.71547652b82fep+0 + 0.5);  // z / ln2
  dd r = add(z, two_prod(-kd, LN2_1));
  r = add(r, two_prod(-kd, LN2_2));
  r = add(r, dd{-kd * LN2_3, 0.0});
  r.hi = r.hi * 0x1.0p-8;  // exact scaling
  r.lo = r.lo * 0x1.0p-8;
  dd acc = inv_fact(10);
#pragma unroll
  for (int n = 9; n >= 2; n--) acc = add(inv_fact(n), mul(r, acc));
  dd p = add(r, mul(mul(r, r), acc));  // expm1(r)
#pragma unroll
  for (int i = 0; i < 8; i++) p = add(dd{2.0 * p.hi, 2.0 * p.lo}, mul(p, p));
  const dd res = add(dd{1.0, 0.0}, p);
  const int k = (int)kd;
  if (k > -1022) return ldexp(res.hi, k);
  // (sub)normal result below 2**-1021: round the pair once on the grid of
  // 2**-1074 -- scale so that grid is the integers (exact), round hi to an
  // integer, then let the remainder hi - f + lo decide the final step
  const double hs = ldexp(res.hi, k + 1074), ls = ldexp(res.lo, k + 1074);
  double f = rint(hs);
  const double rem = (hs - f) + ls;
  if (rem > 0.5 || (rem == 0.5 && fmod(f, 2.0) != 0.0)) f += 1.0;
  else if (rem < -0.5 || (rem == -0.5 && fmod(f, 2.0) != 0.0)) f -= 1.0;
  return ldexp(f, -1074);
}

}  // namespace ddm

// x**y correctly rounded; the special cases follow C's pow
__host__ __device__ __forceinline__ double pow_cr(double x, double y) {
  if (y == 0.0 || x == 1.0) return 1.0;
  if (x == 0.0) return (y > 0.0) ? 0.0 : HUGE_VAL;
  if (!(x > 0.0) || !(x < HUGE_VAL) || !(fabs(y) < HUGE_VAL)) return pow(x, y);
  return ddm::exp_dd(ddm::mul_d(ddm::log_dd(x), y));
}

}  // namespace imex
