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
constexpr int kTypeGeneralUnsorted = 9;
constexpr int kTypeGraphUnsorted = 10;

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

// Evaluate one dial.  x is the mirrored input value; flags are the input's
// GCM_INPUT_* flags; tab points at this point's table vector.  Returns the
// raw response, before the supervisor clamp.
template <typename R>
__device__ __forceinline__ double eval_dial(int type, int n, const R* d,
                                           double x_in, int flags,
                                           const double* __restrict__ tab) {
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
        return fd * tab[slot] + (1.0 - fd) * tab[slot + 1];
      }
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
