// ref_calc.cpp -- TEST INFRASTRUCTURE: a driver that compiles the reference's
// own header-only dial math (src/Utils/include/Calculate*.h, read in place
// under the reference tree, never copied) into oracle/_ref/libref_calc.so so
// that tests/golden/make_golden.py can record golden vectors from it.
// Built by `make ref` in this directory; see oracle/Makefile.
#include "CalculateMonotonicSpline.h"
#include "CalculateUniformSpline.h"
#include "CalculateGeneralSpline.h"
#include "CalculateCompactSpline.h"
#include "CalculateGraph.h"
#include "CalculateBilinearInterpolation.h"
#include "CalculateBicubicSpline.h"

#include "../include/gcm.h"

extern "C" double ref_calc(int type, double x, double lo, double hi,
                           const double* data, int dim) {
  switch (type) {
    case GCM_DIAL_MONOTONIC: return CalculateMonotonicSpline(x, lo, hi, data, dim);
    case GCM_DIAL_UNIFORM: return CalculateUniformSpline(x, lo, hi, data, dim);
    case GCM_DIAL_GENERAL: return CalculateGeneralSpline(x, lo, hi, data, dim);
    case GCM_DIAL_COMPACT: return CalculateCompactSpline(x, lo, hi, data, dim);
    case GCM_DIAL_GRAPH: return CalculateGraph(x, lo, hi, data, dim);
    default: return __builtin_nan("");
  }
}

// 2-D dials: data is the getDialData() layout [nx, ny, x knots, y knots,
// values], unpacked as Bilinear::evalResponse / Bicubic::evalResponse do.
extern "C" double ref_calc2d(int type, double x, double y, double lo, double hi,
                             const double* data) {
  const int nx = int(data[0]), ny = int(data[1]);
  const double* xx = data + 2;
  const double* yy = xx + nx;
  const double* knots = yy + ny;
  if (type == GCM_DIAL_BILINEAR)
    return CalculateBilinearInterpolation(x, y, lo, hi, knots, nx, ny, xx, nx, yy, ny);
  if (type == GCM_DIAL_BICUBIC)
    return CalculateBicubicSpline(x, y, lo, hi, knots, nx, ny, xx, nx, yy, ny);
  return __builtin_nan("");
}
