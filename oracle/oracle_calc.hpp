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

// 2-D dials.  data = the reference's getDialData() layout
// [nx, ny, x knots (nx), y knots (ny), values (nx x ny, x-major)]
// (Bilinear.cpp:5-32, Bicubic.cpp:5-32).

// BilinearIndex / BicubicIndex (CalculateBilinearInterpolation.h:36-45,
// CalculateBicubicSpline.h): a four-step search over offsets 8, 4, 2, 1 (it
// never returns more than 15).  The reference multiplies each comparison by
// (i + off < na), reading a[i + off] even past the array; the comparison is
// skipped here when it cannot count, which gives the same index.
inline int grid_index(double v, const double* a, int na) {
  int i = (8 < na && !(v < a[8])) ? 8 : 0;
  i += (i + 4 < na && !(v < a[i + 4])) ? 4 : 0;
  i += (i + 2 < na && !(v < a[i + 2])) ? 2 : 0;
  i += (i + 1 < na && !(v < a[i + 1])) ? 1 : 0;
  return i;
}

// CalculateBilinearInterpolation.h:47-152.
inline double bilinear_cint(double x, double x0, double p0, double x1, double p1) {
  return p0 + (p1 - p0) * (x - x0) / (x1 - x0);
}

inline double calc_bilinear(double x, double y, double lo, double hi, const double* data) {
  const int nx = int(data[0]), ny = int(data[1]);
  const double* xx = data + 2;
  const double* yy = xx + nx;
  const double* k = yy + ny;
  const int ix = grid_index(x, xx, nx - 1);
  const int iy = grid_index(y, yy, ny - 1);
  const double v0 = bilinear_cint(x, xx[ix], k[ix * ny + iy], xx[ix + 1], k[(ix + 1) * ny + iy]);
  const double v1 =
      bilinear_cint(x, xx[ix], k[ix * ny + iy + 1], xx[ix + 1], k[(ix + 1) * ny + iy + 1]);
  double val = bilinear_cint(y, yy[iy], v0, yy[iy + 1], v1);
  if (val < lo) val = lo;
  if (val > hi) val = hi;
  return val;
}

// CalculateBicubicSpline.h: the Catmull-Rom Hermite through four points in
// the Horner form the reference compiles (BICUBIC_USE_HORNER_FACTORIZATION).
inline double bicubic_cint(double x, double x0, double p0, double x1, double p1, double x2,
                           double p2, double x3, double p3) {
  const double step = x2 - x1;
  const double fx = (x - x1) / step;
  double m1 = step * (p2 - p0) / (x2 - x0);
  double m2 = step * (p3 - p1) / (x3 - x1);
  if (fx < 0.0) m2 = m1;
  if (fx > 1.0) m1 = m2;
  return ((((-2.0 * p2 + 2.0 * p1 + m2 + m1) * fx + 3 * p2 - 3 * p1 - m2 - 2 * m1) * fx + m1) *
              fx +
          p1);
}

// CalculateBicubicSpline.h: a 4 x 4 neighbourhood of (ix, iy); missing rows
// are extrapolated linearly in y from the two nearest, missing columns
// linearly in x from the two nearest (of the y-extrapolated values at the
// corners), then four x interpolations and one y interpolation.
inline double calc_bicubic(double x, double y, double lo, double hi, const double* data) {
  const int nx = int(data[0]), ny = int(data[1]);
  const double* xx = data + 2;
  const double* yy = xx + nx;
  const double* k = yy + ny;
  const int ix = grid_index(x, xx, nx - 1);
  const int iy = grid_index(y, yy, ny - 1);
  auto row_val = [&](int c, int r) -> double {  // column c in [0, nx)
    if (r < 0) return 2 * k[c * ny + iy] - k[c * ny + iy + 1];
    if (r > ny - 1) return 2 * k[c * ny + iy + 1] - k[c * ny + iy];
    return k[c * ny + r];
  };
  auto val = [&](int c, int r) -> double {
    if (c < 0) return 2 * row_val(ix, r) - row_val(ix + 1, r);
    if (c > nx - 1) return 2 * row_val(ix + 1, r) - row_val(ix, r);
    return row_val(c, r);
  };
  const double a0 = (ix > 0) ? xx[ix - 1] : 2 * xx[ix] - xx[ix + 1];
  const double a1 = xx[ix], a2 = xx[ix + 1];
  const double a3 = (ix < nx - 2) ? xx[ix + 2] : 2 * xx[ix + 1] - xx[ix];
  double v[4];
  for (int j = 0; j < 4; ++j) {
    const int r = iy - 1 + j;
    v[j] = bicubic_cint(x, a0, val(ix - 1, r), a1, val(ix, r), a2, val(ix + 1, r), a3,
                        val(ix + 2, r));
  }
  const double u0 = (iy > 0) ? yy[iy - 1] : 2 * yy[iy] - yy[iy + 1];
  const double u3 = (iy < ny - 2) ? yy[iy + 2] : 2 * yy[iy + 1] - yy[iy];
  double out = bicubic_cint(y, u0, v[0], yy[iy], v[1], yy[iy + 1], v[2], u3, v[3]);
  if (out < lo) out = lo;
  if (out > hi) out = hi;
  return out;
}

}  // namespace oracle
