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
