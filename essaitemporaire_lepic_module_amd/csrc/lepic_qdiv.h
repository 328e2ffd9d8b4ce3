// lepic_qdiv.h -- x / h, correctly rounded, for a divisor fixed per launch.
//
// The reference divides every particle coordinate by the grid step
// (int(x/h)+1, (ix*h-x)/h, floor(x/h)).  An IEEE FP64 division costs a
// reciprocal seed, five FMAs and two fix-up instructions per lane; with the
// reciprocal y = RN(1/h) computed once, the quotient needs one multiply and
// two FMA corrections:
//   q0 = RN(x*y)                      within 2 ulp of x/h
//   q1 = RN(q0 + y*RN(x - h*q0))      faithful (within 1 ulp)
//   q2 = RN(q1 + y*(x - h*q1))        x - h*q1 is exact (q1 faithful), and by
//                                     Markstein's theorem (y within 1/2 ulp of
//                                     1/h, q1 within 1 ulp of x/h) q2 = RN(x/h)
// (P. Markstein, IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of
// Floating-Point Arithmetic, ch. 4).  The theorem assumes no underflow or
// overflow, so inputs outside [1e-280, 1e280] in magnitude (and 0, inf, NaN)
// take the plain division; results are then identical to x / h for every x.
#pragma once

#include <cmath>

#ifndef __HIPCC__
#define LEPIC_HD
#else
#define LEPIC_HD __host__ __device__
#endif

struct QDiv {
  double h;   // divisor
  double y;   // RN(1/h)
};

inline QDiv qdiv_make(double h) {
  QDiv d;
  d.h = h;
  d.y = 1.0 / h;
  return d;
}

// True when qdiv is exact for every x: the divisor and its reciprocal are
// normal numbers well inside the exponent range.
inline bool qdiv_divisor_ok(double h) {
  const double a = std::fabs(h);
  return a >= 1.0e-200 && a <= 1.0e200;
}

LEPIC_HD inline double qdiv(double x, const QDiv &d) {
  const double ax = std::fabs(x);
  if (!(ax >= 1.0e-280 && ax <= 1.0e280)) return x / d.h;
  double q = x * d.y;
  double e = std::fma(-q, d.h, x);
  q = std::fma(e, d.y, q);
  e = std::fma(-q, d.h, x);
  return std::fma(e, d.y, q);
}

// Coordinates for which every quotient a particle kernel forms -- x/h and
// (i*h - x)/h for a cell index i -- is exact on the short path: 0, or a
// magnitude in [1e-280, 1e279].  (i*h - x is then 0 or, with the divisor
// inside [1e-200, 1e200], at least ulp(i*h) >= 1e-216 and far below 1e280.)
// For x = -0 the quotient comes out +0 instead of -0; every use converts it
// to an integer or adds it to a nonzero term.
LEPIC_HD inline bool qdiv_coord_ok(double x) {
  const double ax = std::fabs(x);
  return x == 0.0 || (ax >= 1.0e-280 && ax <= 1.0e279);
}

// The same test for a particle's three coordinates at once, stricter only in
// sending a particle with a coordinate exactly 0 to the plain division:
// every magnitude in [1e-280, 1e279].  (A NaN coordinate may pass; both paths
// then give NaN.)
LEPIC_HD inline bool qdiv_coords_ok(double x, double y, double z) {
  const double ax = std::fabs(x), ay = std::fabs(y), az = std::fabs(z);
  const double mx = std::fmax(std::fmax(ax, ay), az);
  const double mn = std::fmin(std::fmin(ax, ay), az);
  return mn >= 1.0e-280 && mx <= 1.0e279;
}

// qdiv with the range test made once per particle: ok from qdiv_coord_ok of
// the coordinates the quotients come from (or qdiv_coords_ok of all three).
LEPIC_HD inline double qdiv_pre(double x, const QDiv &d, bool ok) {
  if (!ok) return x / d.h;
  double q = x * d.y;
  double e = std::fma(-q, d.h, x);
  q = std::fma(e, d.y, q);
  e = std::fma(-q, d.h, x);
  return std::fma(e, d.y, q);
}
