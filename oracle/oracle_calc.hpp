// oracle_calc.hpp -- TEST INFRASTRUCTURE (the oracle), never the product.
//
// CPU restatement of the reference's per-dial math in src/Utils/include/.
// Each function repeats the reference's arithmetic in the same operation
// order, so that on x86-64 (SSE2, no FMA contraction) it reproduces the
// reference bit for bit.  That claim is pinned by golden vectors produced
// from the reference headers themselves (oracle/_ref, tests/golden).
#pragma once

#include <cmath>

namespace oracle {

// Index clamp used by the monotonic and compact splines: the three knot
// intervals around ix are limited to [0, dim-2].
inline int clamp_interval(int i, int dim) {
  if (i < 0) i = 0;
  if (i > dim - 2) i = dim - 2;
  return i;
}

// Cubic Hermite on the unit interval in Horner form:
// p(fx) = (((2a - 2b + mb + ma) fx + 3b - 3a - mb - 2ma) fx + ma) fx + a.
// The sums are evaluated left to right exactly as the reference writes them
// (CalculateMonotonicSpline.h:128-131, CalculateUniformSpline.h:70-73).
inline double hermite_horner(double a, double b, double ma, double mb,
                             double fx) {
  double c3 = 2.0 * a - 2.0 * b + mb + ma;
  double c2 = c3 * fx + 3.0 * b - 3.0 * a - mb - 2.0 * ma;
  double c1 = c2 * fx + ma;
  return c1 * fx + a;
}

inline double clamp_value(double v, double lo, double hi) {
  if (v < lo) v = lo;
  if (v > hi) v = hi;
  return v;
}

// CalculateMonotonicSpline.h:49-138 (FRITSCH_CARLSON defined).
// data = [x0, step, y0..y_{dim-1}]; dim = number of knots.
inline double calc_monotonic(double x, double lo, double hi, const double* data,
                             int dim) {
  const double low = data[0];
  const double step = data[1];
  const double xx = (x - low) / step;
  const int ix = (int)((xx < 0) ? xx - 1 : xx);
  const int i21 = clamp_interval(ix - 1, dim);
  const int i32 = clamp_interval(ix, dim);
  const int i43 = clamp_interval(ix + 1, dim);
  const double p2 = data[2 + i32];
  const double p3 = data[2 + i32 + 1];
  const double fx = xx - i32;
  const double d21 = data[2 + i21 + 1] - data[2 + i21];
  const double d32 = p3 - p2;
  const double d43 = data[2 + i43 + 1] - data[2 + i43];
  double m2 = 0.5 * (d21 + d32);
  double m3 = 0.5 * (d32 + d43);
  // Fritsch & Carlson, SIAM J. Numer. Anal. 17 (1980): flat/cusp -> 0 slope,
  // |m| limited to three times the smaller neighbouring secant.
  if (d32 * d21 <= 0.0) m2 = 0.0;
  if (d43 * d32 <= 0.0) m3 = 0.0;
  const double a21 = (d21 < 0) ? -d21 : d21;
  const double a32 = (d32 < 0) ? -d32 : d32;
  const double a43 = (d43 < 0) ? -d43 : d43;
  const double lim2 = 3.0 * ((a21 < a32) ? a21 : a32);
  const double lim3 = 3.0 * ((a32 < a43) ? a32 : a43);
  if (m2 > lim2) m2 = lim2;
  if (m2 < -lim2) m2 = -lim2;
  if (m3 > lim3) m3 = lim3;
  if (m3 < -lim3) m3 = -lim3;
  return clamp_value(hermite_horner(p2, p3, m2, m3, fx), lo, hi);
}

// CalculateCompactSpline.h:48-115: the Catmull-Rom variant, no limiter.
inline double calc_compact(double x, double lo, double hi, const double* data,
                           int dim) {
  const double low = data[0];
  const double step = data[1];
  const double xx = (x - low) / step;
  const int ix = (int)((xx < 0) ? xx - 1 : xx);
  const int i21 = clamp_interval(ix - 1, dim);
  const int i32 = clamp_interval(ix, dim);
  const int i43 = clamp_interval(ix + 1, dim);
  const double p2 = data[2 + i32];
  const double p3 = data[2 + i32 + 1];
  const double fx = xx - i32;
  const double d21 = data[2 + i21 + 1] - data[2 + i21];
  const double d32 = p3 - p2;
  const double d43 = data[2 + i43 + 1] - data[2 + i43];
  const double m2 = 0.5 * (d21 + d32);
  const double m3 = 0.5 * (d32 + d43);
  return clamp_value(hermite_horner(p2, p3, m2, m3, fx), lo, hi);
}

// CalculateUniformSpline.h:46-80.  data = [x0, step, (y, slope)...];
// dim = total number of doubles (2n+2).
inline double calc_uniform(double x, double lo, double hi, const double* data,
                           int dim) {
  const double step = data[1];
  const double xx = (x - data[0]) / step;
  int ix = (int)xx;
  if (ix < 0) ix = 0;
  if (2 * ix + 7 > dim) ix = (dim - 2) / 2 - 2;
  const double fx = xx - ix;
  const double p1 = data[2 + 2 * ix];
  const double m1 = data[2 + 2 * ix + 1] * step;
  const double p2 = data[2 + 2 * ix + 2];
  const double m2 = data[2 + 2 * ix + 3] * step;
  return clamp_value(hermite_horner(p1, p2, m1, m2, fx), lo, hi);
}

// CalculateGeneralSpline.h:51-197, default BINARY_IF search.
// data = [x0, step, (y, slope, x)...]; dim = total doubles (3n+2).
inline double calc_general(double x, double lo, double hi, const double* data,
                           int dim) {
  const int last = (dim - 2) / 3 - 1;  // number of intervals
  int ix = 0;
  for (int off = 16; off > 0; off >>= 1) {
    if ((ix + off < last) && (x > data[2 + 3 * (ix + off) + 2])) ix += off;
  }
  const double x1 = data[2 + 3 * ix + 2];
  const double x2 = data[2 + 3 * (ix + 1) + 2];
  const double step = x2 - x1;
  const double fx = (x - x1) / step;
  const double p1 = data[2 + 3 * ix];
  const double m1 = data[2 + 3 * ix + 1] * step;
  const double p2 = data[2 + 3 * (ix + 1)];
  const double m2 = data[2 + 3 * (ix + 1) + 1] * step;
  return clamp_value(hermite_horner(p1, p2, m1, m2, fx), lo, hi);
}

// CalculateGraph.h:43-88.  data = [(y, x)...]; dim = total doubles (2n).
inline double calc_graph(double x, double lo, double hi, const double* data,
                         int dim) {
  if (dim < 4) return data[0];  // single point: constant, unclamped
  const int knots = dim / 2;
  int ix = 0;
  for (int off = 1 << (31 - __builtin_clz(knots)); off > 0; off >>= 1) {
    if ((ix + off < knots) && (x > data[2 * (ix + off) + 1])) ix += off;
  }
  if (ix + 1 >= knots) ix--;  // extrapolate past the last knot linearly
  const double p1 = data[2 * ix];
  const double x1 = data[2 * ix + 1];
  const double p2 = data[2 * (ix + 1)];
  const double x2 = data[2 * (ix + 1) + 1];
  const double step = x2 - x1;
  const double fx = (x - x1) / step;
  const double m = p2 - p1;
  return clamp_value(p1 + fx * m, lo, hi);
}

}  // namespace oracle
