// dial_eval.hpp -- device evaluation of one dial record staged in LDS.
//
// R is the arithmetic type of the record and of the dial math: double in the
// FP64 mode (operation order identical to the reference's Calculate*.h, the
// translation unit is compiled with -ffp-contract=off so every operation
// rounds as on the host), float in the FP32 mode.
//
// Record words (LDS, tile-relative):
//   MONOTONIC, COMPACT : x0, step, y[n]              (n = knots)
//   UNIFORM            : x0, step, (y, slope)[n]
//   GENERAL            : x0, step, (y, slope, x)[n]
//   GRAPH              : (y, x)[n]
//   SHIFT              : value
//   TABULATED, KRIGED  : packed {u32 table slot, f32 weight} entries, one
//                        64-bit word each in FP64 mode, two 32-bit words each
//                        in FP32 mode (n = 1 for TABULATED, n = k for KRIGED)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gcm.h"

namespace gcm {

template <typename R>
__device__ __forceinline__ int clamp_interval(int i, int dim) {
  if (i < 0) i = 0;
  if (i > dim - 2) i = dim - 2;
  return i;
}

// Cubic Hermite in the Horner form of CalculateMonotonicSpline.h:128-131 and
// CalculateUniformSpline.h:70-73, summed left to right.
template <typename R>
__device__ __forceinline__ R hermite(R a, R b, R ma, R mb, R fx) {
  R c3 = R(2) * a - R(2) * b + mb + ma;
  R c2 = c3 * fx + R(3) * b - R(3) * a - mb - R(2) * ma;
  R c1 = c2 * fx + ma;
  return c1 * fx + a;
}

template <typename R>
__device__ __forceinline__ R clamp_value(R v, R lo, R hi) {
  if (v < lo) v = lo;
  if (v > hi) v = hi;
  return v;
}

// Monotonic (FC limiter) and compact (Catmull-Rom) splines on uniform knots:
// CalculateMonotonicSpline.h:49-138, CalculateCompactSpline.h:48-115.
template <typename R, bool kMonotonic>
__device__ __forceinline__ R spline_uniform_knots(R x, const R* d, int dim) {
  const R low = d[0];
  const R step = d[1];
  const R xx = (x - low) / step;
  const int ix = (int)((xx < 0) ? xx - R(1) : xx);
  const int i21 = clamp_interval<R>(ix - 1, dim);
  const int i32 = clamp_interval<R>(ix, dim);
  const int i43 = clamp_interval<R>(ix + 1, dim);
  const R p2 = d[2 + i32];
  const R p3 = d[2 + i32 + 1];
  const R fx = xx - R(i32);
  const R d21 = d[2 + i21 + 1] - d[2 + i21];
  const R d32 = p3 - p2;
  const R d43 = d[2 + i43 + 1] - d[2 + i43];
  R m2 = R(0.5) * (d21 + d32);
  R m3 = R(0.5) * (d32 + d43);
  if (kMonotonic) {
    if (d32 * d21 <= R(0)) m2 = R(0);
    if (d43 * d32 <= R(0)) m3 = R(0);
    const R a21 = (d21 < 0) ? -d21 : d21;
    const R a32 = (d32 < 0) ? -d32 : d32;
    const R a43 = (d43 < 0) ? -d43 : d43;
    const R lim2 = R(3) * ((a21 < a32) ? a21 : a32);
    const R lim3 = R(3) * ((a32 < a43) ? a32 : a43);
    if (m2 > lim2) m2 = lim2;
    if (m2 < -lim2) m2 = -lim2;
    if (m3 > lim3) m3 = lim3;
    if (m3 < -lim3) m3 = -lim3;
  }
  return clamp_value<R>(hermite<R>(p2, p3, m2, m3, fx), R(-1E20), R(1E20));
}

// CalculateUniformSpline.h:46-80 (dim = 2n + 2 words).
template <typename R>
__device__ __forceinline__ R spline_uniform(R x, const R* d, int dim) {
  const R step = d[1];
  const R xx = (x - d[0]) / step;
  int ix = (int)xx;
  if (ix < 0) ix = 0;
  if (2 * ix + 7 > dim) ix = (dim - 2) / 2 - 2;
  const R fx = xx - R(ix);
  const R p1 = d[2 + 2 * ix];
  const R m1 = d[2 + 2 * ix + 1] * step;
  const R p2 = d[2 + 2 * ix + 2];
  const R m2 = d[2 + 2 * ix + 3] * step;
  return clamp_value<R>(hermite<R>(p1, p2, m1, m2, fx), R(-1E20), R(1E20));
}

// Internal descriptor type codes (never seen by callers): general splines
// and graphs whose knot abscissae are not non-decreasing (or hold a NaN) keep
// the reference's binary search; sorted ones use the counting search below,
// which returns the same segment index.
constexpr int kTypeGeneralUnsorted = 11;
constexpr int kTypeGraphUnsorted = 12;

// #{ j in [first, end) : x > X[stride j] }, with the loads of a batch issued
// together (independent LDS reads instead of a chain of dependent ones).
template <typename R, int kBatch>
__device__ __forceinline__ int count_below(R x, const R* X, int stride, int first, int end) {
  int c = 0;
  for (int j0 = first; j0 < end; j0 += kBatch) {
    R v[kBatch];
#pragma unroll
    for (int q = 0; q < kBatch; ++q) {
      const int j = (j0 + q < end) ? j0 + q : end - 1;
      v[q] = X[stride * j];
    }
#pragma unroll
    for (int q = 0; q < kBatch; ++q) c += (j0 + q < end && x > v[q]) ? 1 : 0;
  }
  return c;
}

// CalculateGeneralSpline.h:51-197.  The reference's BINARY_IF search (offsets
// 16 .. 1) finds the largest j <= 31, j < last, with x > X[j] (or 0); for
// non-decreasing X that is min(#{j in [1, last) : x > X[j]}, 31).
template <typename R, bool kSorted>
__device__ __forceinline__ R spline_general(R x, const R* d, int dim) {
  const int last = (dim - 2) / 3 - 1;
  int ix = 0;
  if (kSorted) {
    ix = count_below<R, 16>(x, d + 4, 3, 1, last);
    if (ix > 31) ix = 31;
  } else {
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) {
      if ((ix + off < last) && (x > d[2 + 3 * (ix + off) + 2])) ix += off;
    }
  }
  const R x1 = d[2 + 3 * ix + 2];
  const R x2 = d[2 + 3 * (ix + 1) + 2];
  const R step = x2 - x1;
  const R fx = (x - x1) / step;
  const R p1 = d[2 + 3 * ix];
  const R m1 = d[2 + 3 * ix + 1] * step;
  const R p2 = d[2 + 3 * (ix + 1)];
  const R m2 = d[2 + 3 * (ix + 1) + 1] * step;
  return clamp_value<R>(hermite<R>(p1, p2, m1, m2, fx), R(-1E20), R(1E20));
}

// CalculateGraph.h:43-88 (dim = 2n words).  The binary search covers every
// knot, so for non-decreasing X it returns #{j in [1, knots) : x > X[j]}.
template <typename R, bool kSorted>
__device__ __forceinline__ R graph_linear(R x, const R* d, int dim) {
  if (dim < 4) return d[0];
  const int knots = dim / 2;
  int ix = 0;
  if (kSorted) {
    ix = count_below<R, 8>(x, d + 1, 2, 1, knots);
  } else {
    for (int off = 1 << (31 - __builtin_clz(knots)); off > 0; off >>= 1) {
      if ((ix + off < knots) && (x > d[2 * (ix + off) + 1])) ix += off;
    }
  }
  if (ix + 1 >= knots) ix--;
  const R p1 = d[2 * ix];
  const R x1 = d[2 * ix + 1];
  const R p2 = d[2 * (ix + 1)];
  const R x2 = d[2 * (ix + 1) + 1];
  const R step = x2 - x1;
  const R fx = (x - x1) / step;
  const R m = p2 - p1;
  return clamp_value<R>(p1 + fx * m, R(-1E20), R(1E20));
}

// 2-D dials on a grid: d = [nx, ny, x knots (nx), y knots (ny), values
// (nx x ny, x-major)], the payload of Bilinear.cpp:5-32 / Bicubic.cpp:5-32.
// The four-step index search of CalculateBilinearInterpolation.h:36-45 (the
// same in CalculateBicubicSpline.h): offsets 8, 4, 2, 1, never past 15; the
// comparison is skipped when it cannot count (the reference multiplies it by
// zero after reading a[i + off]), which gives the same index without reading
// past the knots.
template <typename R>
__device__ __forceinline__ int grid_index(R v, const R* a, int na) {
  int i = (8 < na && !(v < a[8])) ? 8 : 0;
  i += (i + 4 < na && !(v < a[i + 4])) ? 4 : 0;
  i += (i + 2 < na && !(v < a[i + 2])) ? 2 : 0;
  i += (i + 1 < na && !(v < a[i + 1])) ? 1 : 0;
  return i;
}

template <typename R>
__device__ __forceinline__ R lerp_cint(R x, R x0, R p0, R x1, R p1) {
  return p0 + (p1 - p0) * (x - x0) / (x1 - x0);
}

// CalculateBilinearInterpolation.h:47-152.
template <typename R>
__device__ __forceinline__ R grid_bilinear(R x, R y, const R* d) {
  const int nx = (int)d[0], ny = (int)d[1];
  const R* xx = d + 2;
  const R* yy = xx + nx;
  const R* k = yy + ny;
  const int ix = grid_index<R>(x, xx, nx - 1);
  const int iy = grid_index<R>(y, yy, ny - 1);
  const R v0 = lerp_cint<R>(x, xx[ix], k[ix * ny + iy], xx[ix + 1], k[(ix + 1) * ny + iy]);
  const R v1 = lerp_cint<R>(x, xx[ix], k[ix * ny + iy + 1], xx[ix + 1], k[(ix + 1) * ny + iy + 1]);
  return clamp_value<R>(lerp_cint<R>(y, yy[iy], v0, yy[iy + 1], v1), R(-1E20), R(1E20));
}

// The Catmull-Rom Hermite through four points, Horner form of
// CalculateBicubicSpline.h (BICUBIC_USE_HORNER_FACTORIZATION).
template <typename R>
__device__ __forceinline__ R cubic_cint(R x, R x0, R p0, R x1, R p1, R x2, R p2, R x3, R p3) {
  const R step = x2 - x1;
  const R fx = (x - x1) / step;
  R m1 = step * (p2 - p0) / (x2 - x0);
  R m2 = step * (p3 - p1) / (x3 - x1);
  if (fx < R(0)) m2 = m1;
  if (fx > R(1)) m1 = m2;
  return ((((R(-2) * p2 + R(2) * p1 + m2 + m1) * fx + R(3) * p2 - R(3) * p1 - m2 - R(2) * m1) * fx +
           m1) * fx + p1);
}

// CalculateBicubicSpline.h: the 4 x 4 neighbourhood of (ix, iy); rows past
// the grid are extrapolated linearly in y from the two nearest rows, columns
// past it linearly in x from the two nearest columns (of those values).
template <typename R>
__device__ __forceinline__ R grid_value(const R* k, int ny, int nx, int ix, int iy, int c, int r) {
  auto row = [&](int cc) -> R {
    if (r < 0) return R(2) * k[cc * ny + iy] - k[cc * ny + iy + 1];
    if (r > ny - 1) return R(2) * k[cc * ny + iy + 1] - k[cc * ny + iy];
    return k[cc * ny + r];
  };
  if (c < 0) return R(2) * row(ix) - row(ix + 1);
  if (c > nx - 1) return R(2) * row(ix + 1) - row(ix);
  return row(c);
}

template <typename R>
__device__ __forceinline__ R grid_bicubic(R x, R y, const R* d) {
  const int nx = (int)d[0], ny = (int)d[1];
  const R* xx = d + 2;
  const R* yy = xx + nx;
  const R* k = yy + ny;
  const int ix = grid_index<R>(x, xx, nx - 1);
  const int iy = grid_index<R>(y, yy, ny - 1);
  const R a0 = (ix > 0) ? xx[ix - 1] : R(2) * xx[ix] - xx[ix + 1];
  const R a1 = xx[ix], a2 = xx[ix + 1];
  const R a3 = (ix < nx - 2) ? xx[ix + 2] : R(2) * xx[ix + 1] - xx[ix];
  // one row at a time (not unrolled: fewer values live at once)
  R v0 = 0, v1 = 0, v2 = 0, v3 = 0;
#pragma unroll 1
  for (int j = 0; j < 4; ++j) {
    const int r = iy - 1 + j;
    const R vj = cubic_cint<R>(x, a0, grid_value<R>(k, ny, nx, ix, iy, ix - 1, r), a1,
                               grid_value<R>(k, ny, nx, ix, iy, ix, r), a2,
                               grid_value<R>(k, ny, nx, ix, iy, ix + 1, r), a3,
                               grid_value<R>(k, ny, nx, ix, iy, ix + 2, r));
    v0 = (j == 0) ? vj : v0;
    v1 = (j == 1) ? vj : v1;
    v2 = (j == 2) ? vj : v2;
    v3 = (j == 3) ? vj : v3;
  }
  const R v[4] = {v0, v1, v2, v3};
  const R u0 = (iy > 0) ? yy[iy - 1] : R(2) * yy[iy] - yy[iy + 1];
  const R u3 = (iy < ny - 2) ? yy[iy + 2] : R(2) * yy[iy + 1] - yy[iy];
  return clamp_value<R>(cubic_cint<R>(y, u0, v[0], yy[iy], v[1], yy[iy + 1], v[2], u3, v[3]),
                        R(-1E20), R(1E20));
}

// A 2-D dial (Bilinear.cpp:40-64, Bicubic.cpp:40-64): x, y are the mirrored
// values of its two inputs; both are clamped to the knot range when
// extrapolation is off.
template <typename R>
__device__ __forceinline__ double eval_dial2d(int type, const R* d, double x_in, double y_in,
                                              int flags) {
  R x = (R)x_in, y = (R)y_in;
  if (flags & GCM_INPUT_NO_EXTRAPOLATION) {
    const int nx = (int)d[0], ny = (int)d[1];
    const R* xx = d + 2;
    const R* yy = xx + nx;
    if (x < xx[0]) x = xx[0];
    if (x > xx[nx - 1]) x = xx[nx - 1];
    if (y < yy[0]) y = yy[0];
    if (y > yy[ny - 1]) y = yy[ny - 1];
  }
  return (double)(type == GCM_DIAL_BICUBIC ? grid_bicubic<R>(x, y, d) : grid_bilinear<R>(x, y, d));
}

// Packed table entry {u32 slot, f32 weight}.
template <typename R>
struct TableEntry;

template <>
struct TableEntry<double> {
  static __device__ __forceinline__ void get(const double* d, int j,
                                             uint32_t* slot, float* w) {
    const uint64_t bits = __double_as_longlong(d[j]);
    *slot = (uint32_t)bits;
    *w = __uint_as_float((uint32_t)(bits >> 32));
  }
};

template <>
struct TableEntry<float> {
  static __device__ __forceinline__ void get(const float* d, int j,
                                             uint32_t* slot, float* w) {
    *slot = __float_as_uint(d[2 * j]);
    *w = d[2 * j + 1];
  }
};

// Tabulated and Kriged dials in two halves, so that a dial's table loads are
// in flight while its thread evaluates another dial.  table_issue starts the
// loads of the first kTabBatch table values the dial reads (Tabulated: the
// entries at slot and slot + 1; Kriged: its first kTabBatch entries);
// table_finish forms the response from them in the reference's order and
// with its types (Tabulated.h:27-28: the float fraction weights the lower
// entry; Kriged.cpp:14-21: the float weights summed in entry order), loading
// any further Kriged entries itself.  The result equals eval_dial's.
constexpr int kTabBatch = 8;

template <typename R>
__device__ __forceinline__ void table_issue(int type, int n, const R* d,
                                            const double* __restrict__ tab,
                                            double (&tv)[kTabBatch]) {
  uint32_t s[kTabBatch];
  int m;
  if (type == GCM_DIAL_TABULATED) {
    float f;
    TableEntry<R>::get(d, 0, &s[0], &f);
    s[1] = s[0] + 1;
#pragma unroll
    for (int q = 2; q < kTabBatch; ++q) s[q] = s[0];
    m = 2;
  } else {
    m = n < kTabBatch ? n : kTabBatch;
#pragma unroll
    for (int q = 0; q < kTabBatch; ++q) {
      float w;
      s[q] = 0;
      if (q < m) TableEntry<R>::get(d, q, &s[q], &w);
    }
  }
#pragma unroll
  for (int q = 0; q < kTabBatch; ++q) tv[q] = (q < m) ? tab[s[q]] : 0.0;
}

template <typename R>
__device__ __forceinline__ double table_finish(int type, int n, const R* d,
                                               const double* __restrict__ tab,
                                               const double (&tv)[kTabBatch]) {
  if (type == GCM_DIAL_TABULATED) {
    uint32_t slot;
    float f;
    TableEntry<R>::get(d, 0, &slot, &f);
    if (sizeof(R) == 8) {
      const double fd = (double)f;
      return fd * tv[0] + (1.0 - fd) * tv[1];
    }
    return (double)(f * (float)tv[0] + (1.0f - f) * (float)tv[1]);
  }
  double v = 0.0;
  float vf = 0.0f;
#pragma unroll
  for (int q = 0; q < kTabBatch; ++q) {
    if (q < n) {
      uint32_t slot;
      float w;
      TableEntry<R>::get(d, q, &slot, &w);
      if (sizeof(R) == 8) v += (double)w * tv[q];
      else vf += w * (float)tv[q];
    }
  }
  for (int j0 = kTabBatch; j0 < n; j0 += kTabBatch) {
    double tw[kTabBatch];
    float wv[kTabBatch];
#pragma unroll
    for (int q = 0; q < kTabBatch; ++q) {
      uint32_t slot = 0;
      float w = 0.0f;
      if (j0 + q < n) TableEntry<R>::get(d, j0 + q, &slot, &w);
      wv[q] = w;
      tw[q] = (j0 + q < n) ? tab[slot] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kTabBatch; ++q) {
      if (j0 + q < n) {
        if (sizeof(R) == 8) v += (double)wv[q] * tw[q];
        else vf += wv[q] * (float)tw[q];
      }
    }
  }
  return sizeof(R) == 8 ? v : (double)vf;
}

// Evaluate one dial.  x is the mirrored input value; flags are the input's
// GCM_INPUT_* flags; the point's table value at slot s is tab[s] (the
// point's own table vector), or with kInterleaved tab[s * k_stride] (the
// point-interleaved tables of a K-point launch, k_stride = K).  Returns the raw response, before the supervisor
// clamp.
template <typename R, bool kInterleaved = false>
__device__ __forceinline__ double eval_dial(int type, int n, const R* d,
                                           double x_in, int flags,
                                           const double* __restrict__ tab,
                                           int64_t k_stride = 1) {
  const int64_t tstride = kInterleaved ? k_stride : 1;
  R x = (R)x_in;
  const bool no_extrap = (flags & GCM_INPUT_NO_EXTRAPOLATION) != 0;
  switch (type) {
    case GCM_DIAL_NORM:
      return x_in;  // Norm.h:21: the input value itself
    case GCM_DIAL_SHIFT:
      return (double)d[0];
    case GCM_DIAL_MONOTONIC:
    case GCM_DIAL_COMPACT: {
      if (no_extrap) {  // MonotonicSpline.cpp:78-81
        const R lo = d[0], hi = d[0] + R(n - 1) * d[1];
        if (x <= lo) x = lo;
        else if (x >= hi) x = hi;
      }
      if (type == GCM_DIAL_MONOTONIC)
        return (double)spline_uniform_knots<R, true>(x, d, n);
      return (double)spline_uniform_knots<R, false>(x, d, n);
    }
    case GCM_DIAL_UNIFORM: {
      if (no_extrap) {
        const R lo = d[0], hi = d[0] + R(n - 1) * d[1];
        if (x <= lo) x = lo;
        else if (x >= hi) x = hi;
      }
      return (double)spline_uniform<R>(x, d, 2 * n + 2);
    }
    case GCM_DIAL_GENERAL:
    case kTypeGeneralUnsorted: {
      if (no_extrap) {
        const R lo = d[0], hi = d[2 + 3 * (n - 1) + 2];
        if (x <= lo) x = lo;
        else if (x >= hi) x = hi;
      }
      if (type == GCM_DIAL_GENERAL) return (double)spline_general<R, true>(x, d, 3 * n + 2);
      return (double)spline_general<R, false>(x, d, 3 * n + 2);
    }
    case GCM_DIAL_GRAPH:
    case kTypeGraphUnsorted: {
      if (no_extrap) {  // Graph.cpp:48-51
        if (x <= d[1]) return (double)d[0];
        else if (x >= d[2 * n - 1]) return (double)d[2 * n - 2];
      }
      if (type == GCM_DIAL_GRAPH) return (double)graph_linear<R, true>(x, d, 2 * n);
      return (double)graph_linear<R, false>(x, d, 2 * n);
    }
    case GCM_DIAL_TABULATED: {  // Tabulated.h:27-28
      uint32_t slot;
      float f;
      TableEntry<R>::get(d, 0, &slot, &f);
      if (sizeof(R) == 8) {
        const double fd = (double)f;
        if constexpr (kInterleaved)
          return fd * tab[slot * tstride] + (1.0 - fd) * tab[(slot + 1) * tstride];
        else
          return fd * tab[slot] + (1.0 - fd) * tab[slot + 1];
      }
      if constexpr (kInterleaved)
        return (double)(f * (float)tab[slot * tstride] + (1.0f - f) * (float)tab[(slot + 1) * tstride]);
      else
        return (double)(f * (float)tab[slot] + (1.0f - f) * (float)tab[slot + 1]);
    }
    case GCM_DIAL_KRIGED: {  // Kriged.cpp:14-21
      // The table gathers are issued eight at a time (they hit L2 or HBM);
      // the sum is still accumulated in entry order.
      double v = 0.0;
      float vf = 0.0f;
      for (int j0 = 0; j0 < n; j0 += 8) {
        double tv[8];
        float wv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          uint32_t slot = 0;
          float w = 0.0f;
          if (j0 + q < n) TableEntry<R>::get(d, j0 + q, &slot, &w);
          wv[q] = w;
          if constexpr (kInterleaved)
            tv[q] = (j0 + q < n) ? tab[(int64_t)slot * tstride] : 0.0;
          else
            tv[q] = (j0 + q < n) ? tab[slot] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          if (j0 + q < n) {
            if (sizeof(R) == 8) v += (double)wv[q] * tv[q];
            else vf += wv[q] * (float)tv[q];
          }
        }
      }
      return sizeof(R) == 8 ? v : (double)vf;
    }
    default:
      return __builtin_nan("");
  }
}

}  // namespace gcm
