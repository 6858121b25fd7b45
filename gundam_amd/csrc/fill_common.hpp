// fill_common.hpp -- definitions shared by the device code of gcm.hip (one
// point per lane of work: fill_kernel) and gcm_points.hip (one lane per
// point: fill_points_kernel): the device layout of events, tiles and input
// records, the fill arguments, the LDS-DMA helpers, and the per-bin
// likelihood.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/gcm.h"

namespace gcm {

int fail(int code, const std::string& msg);  // gcm.hip

#define GCM_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(GCM_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

constexpr int kWave = 64;
constexpr int kMaxTileEvents = 64;
// Launches of at least kPointsMin points evaluate with one lane per point
// (fill_points_kernel); each tile's events are split into kRanges contiguous
// ranges, one per evaluator wave, each with its own partial slots.  10
// ranges: with up to two producers, two blocks of 12 waves fit a CU at six
// waves per SIMD (measured against 8-14 ranges at 5-8 waves per SIMD).
constexpr int kPointsMin = 16;
constexpr int kRanges = 10;
constexpr int kDefaultBlock = 256;
constexpr int kDefaultTileWords = 2048;  // 16 KiB of records per tile
constexpr int kDefaultBlocksPerCU = 3;
constexpr int kFinalizeBlock = 256;
constexpr int kTotalBlock = 1024;  // threads of the per-point LLH sum
constexpr int kFinalizeMaxBlocks = 8192;  // waves in flight for the per-bin pass
constexpr size_t kInputsLdsMax = 16384;  // stage input records in LDS up to this
constexpr size_t kLdsPerCU = 160 * 1024;   // gfx950 LDS per compute unit
constexpr size_t kHeaderBytes = 32;      // tile header copy in each LDS buffer
constexpr int kMaxDynamicLds = 160 * 1024;  // gfx950 LDS per CU
constexpr size_t kReadyBytes = 16;  // the fill kernel's per-buffer ready flags

// One binned event, in bin-sorted order: base weight, global bin, and its
// dials as (first event-major slot in the tile, count).
struct EventRec {
  double base;
  int32_t bin;
  uint32_t dials;  // slot | count << 16
};
static_assert(sizeof(EventRec) == 16, "event record is 16 bytes");

// A tile's dials are type-sorted: the table dials (Tabulated, Kriged) in
// [0, n_tab), whose records are the first tab_words record words, the other
// 1-D dials in [n_tab, grid_begin), the 2-D dials in [grid_begin, n_dials).
struct TileHdr {
  uint32_t first_event;
  uint32_t first_dial;
  uint64_t rec_begin;  // in record words of the precision's type
  uint16_t rec_words;  // <= tile_words x 8 / sizeof(R) <= 16384
  uint16_t n_tab;      // table dials
  uint16_t n_events;
  uint16_t n_dials;
  uint32_t first_partial;
  uint16_t grid_begin;  // first 2-D dial
  uint16_t tab_words;   // record words of the table dials
};
static_assert(sizeof(TileHdr) == 32, "tile header is 32 bytes");

// Per-point record of one dial input: the mirrored value and the response
// supervisor bounds (unset bounds are -+inf).
struct InputRec {
  double x;
  double lo;
  double hi;
  int32_t flags;
  int32_t pad;
};
static_assert(sizeof(InputRec) == 32, "input record is 32 bytes");

// Response supervisor and flags shared by the inputs of one class.  Dial
// descriptors carry their input's class (4 bits), so that when the inputs'
// x values are staged in LDS an evaluator lane reads x and its class record
// from LDS and nothing from global memory; class kEscapeClass (workloads
// with more than kInputClasses - 1 distinct (lo, hi, flags)) and the
// unstaged case read the input's full record from global memory.
struct ClassRec {
  double lo;
  double hi;
  int32_t flags;
  int32_t pad;
  double pad2;
};
static_assert(sizeof(ClassRec) == 32, "class record is 32 bytes");
constexpr int kInputClasses = 16;
constexpr int kEscapeClass = kInputClasses - 1;
constexpr size_t kClassBytes = kInputClasses * sizeof(ClassRec);

// Knuth two-sum: s + e == a + b exactly.
__device__ __forceinline__ void two_sum(double a, double b, double& s,
                                        double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

struct FillArgs {
  const ClassRec* cls;  // [kInputClasses]
  const TileHdr* tiles;
  int64_t n_tiles;
  const uint64_t* desc;
  const void* rec;
  const EventRec* ev;
  const int32_t* ev_result;
  const InputRec* xin;
  int n_inputs;
  const double* tab;
  int64_t table_len;
  double cap;
  int k_points;
  double* psum;
  double* psum2;
  double* pcomp;
  double* pcomp2;
  int64_t n_partials;
  double* weights_out;
  int* flags;
  int rec_bytes;   // LDS record capacity per buffer (multiple of 16 bytes)
  int max_dials;   // LDS descriptor/response capacity (multiple of 128)
  int inputs_lds;  // input records staged in LDS (K * n_inputs * 32 bytes)
  int tab_waves;   // evaluator waves that take only the table dials (0: none)
};

// Cache policy of the tile stream: non-temporal, so the streamed records do
// not evict the per-point lookup tables and input records from L2.
#ifndef GCM_STREAM_POLICY
#define GCM_STREAM_POLICY 2
#endif
constexpr int kStreamPolicy = GCM_STREAM_POLICY;

typedef __attribute__((address_space(3))) void* lds_void_ptr;
typedef __attribute__((address_space(1))) void* global_void_ptr;

// Block barrier that waits only for this wave's LDS traffic (not for its
// global stores or, in the producer, for anything it has not waited on).
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Producer wave: copy `bytes` (a multiple of 16) from global memory straight
// into LDS with 16-byte-per-lane DMA loads; lane i of a chunk lands at
// dst + chunk + 16 i.  src and dst are wave-uniform: each load takes its
// address as a scalar base plus the lane's fixed offset, so whole chunks cost
// no vector instruction; only the last, partial chunk masks lanes off.
__device__ __forceinline__ void dma_to_lds(const unsigned char* src,
                                           unsigned char* dst, int bytes,
                                           int lane) {
  const uint32_t off = (uint32_t)lane * 16u;
  int j = 0;
  for (; j + 1024 <= bytes; j += 1024)
    __builtin_amdgcn_global_load_lds((global_void_ptr)(src + j + off), (lds_void_ptr)(dst + j),
                                     16, 0, kStreamPolicy);
  if (j < bytes && j + (int)off < bytes)
    __builtin_amdgcn_global_load_lds((global_void_ptr)(src + j + off), (lds_void_ptr)(dst + j),
                                     16, 0, kStreamPolicy);
}

// A tile header with every field in scalar registers (the producer's DMA
// addresses are then scalar).
__device__ __forceinline__ TileHdr uniform_header(const TileHdr& h) {
  TileHdr u;
  u.first_event = __builtin_amdgcn_readfirstlane(h.first_event);
  u.first_dial = __builtin_amdgcn_readfirstlane(h.first_dial);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)h.rec_begin);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(h.rec_begin >> 32));
  u.rec_begin = ((uint64_t)hi << 32) | lo;
  u.rec_words = (uint16_t)__builtin_amdgcn_readfirstlane(h.rec_words);
  u.n_tab = (uint16_t)__builtin_amdgcn_readfirstlane(h.n_tab);
  u.n_events = (uint16_t)__builtin_amdgcn_readfirstlane(h.n_events);
  u.n_dials = (uint16_t)__builtin_amdgcn_readfirstlane(h.n_dials);
  u.first_partial = __builtin_amdgcn_readfirstlane(h.first_partial);
  u.grid_begin = (uint16_t)__builtin_amdgcn_readfirstlane(h.grid_begin);
  u.tab_words = (uint16_t)__builtin_amdgcn_readfirstlane(h.tab_words);
  return u;
}

// s_waitcnt vmcnt(m) for the largest m <= min(n, 47): waits until at most n
// of this wave's vector-memory loads are outstanding (or fewer).
__device__ __forceinline__ void wait_vm_at_most(int n) {
  switch (n < 0 ? 0 : n > 47 ? 47 : n) {
#define GCM_VM_CASE(m) \
    case m: asm volatile("s_waitcnt vmcnt(" #m ")" ::: "memory"); break;
    GCM_VM_CASE(0) GCM_VM_CASE(1) GCM_VM_CASE(2) GCM_VM_CASE(3) GCM_VM_CASE(4)
    GCM_VM_CASE(5) GCM_VM_CASE(6) GCM_VM_CASE(7) GCM_VM_CASE(8) GCM_VM_CASE(9)
    GCM_VM_CASE(10) GCM_VM_CASE(11) GCM_VM_CASE(12) GCM_VM_CASE(13) GCM_VM_CASE(14)
    GCM_VM_CASE(15) GCM_VM_CASE(16) GCM_VM_CASE(17) GCM_VM_CASE(18) GCM_VM_CASE(19)
    GCM_VM_CASE(20) GCM_VM_CASE(21) GCM_VM_CASE(22) GCM_VM_CASE(23) GCM_VM_CASE(24)
    GCM_VM_CASE(25) GCM_VM_CASE(26) GCM_VM_CASE(27) GCM_VM_CASE(28) GCM_VM_CASE(29)
    GCM_VM_CASE(30) GCM_VM_CASE(31) GCM_VM_CASE(32) GCM_VM_CASE(33) GCM_VM_CASE(34)
    GCM_VM_CASE(35) GCM_VM_CASE(36) GCM_VM_CASE(37) GCM_VM_CASE(38) GCM_VM_CASE(39)
    GCM_VM_CASE(40) GCM_VM_CASE(41) GCM_VM_CASE(42) GCM_VM_CASE(43) GCM_VM_CASE(44)
    GCM_VM_CASE(45) GCM_VM_CASE(46) GCM_VM_CASE(47)
#undef GCM_VM_CASE
    default: break;
  }
}

// DMA instructions dma_to_lds issues for `bytes`.
__device__ __forceinline__ int dma_instrs(int bytes) { return (bytes + 1023) / 1024; }

template <typename R>
__device__ __forceinline__ void dma_tile(const TileHdr& h, const unsigned char* rec,
                                         const unsigned char* desc,
                                         const EventRec* ev, unsigned char* buf,
                                         int rec_bytes, int max_dials, int lane) {
  dma_to_lds(rec + h.rec_begin * sizeof(R), buf,
             (int)((h.rec_words * sizeof(R) + 15) / 16 * 16), lane);
  dma_to_lds(desc + (size_t)h.first_dial * 8, buf + rec_bytes,
             (h.n_dials + 1) / 2 * 16, lane);
  dma_to_lds(reinterpret_cast<const unsigned char*>(ev + h.first_event),
             buf + rec_bytes + (size_t)max_dials * 8, h.n_events * 16, lane);
}

// The tile in two parts: first the header and the table dials' records and
// descriptors (what the table waves need a step early), then the rest.  The
// parts are cut on 16-byte boundaries; the up to 8 bytes on either side of a
// cut are copied by both parts (the same bytes).  Returns the number of DMA
// instructions of the second part.
template <typename R>
__device__ __forceinline__ int dma_tile_split(const TileHdr& h, const TileHdr* hdr_src,
                                              const unsigned char* rec,
                                              const unsigned char* desc, const EventRec* ev,
                                              unsigned char* buf, int rec_bytes, int max_dials,
                                              size_t hdr_off, int lane) {
  const unsigned char* r0 = rec + h.rec_begin * sizeof(R);
  const unsigned char* d0 = desc + (size_t)h.first_dial * 8;
  const int tab_bytes = (int)(h.tab_words * sizeof(R));
  const int rec_all = (int)((h.rec_words * sizeof(R) + 15) / 16 * 16);
  const int desc_all = (h.n_dials + 1) / 2 * 16;
  const int tr = (tab_bytes + 15) / 16 * 16, tr0 = tab_bytes / 16 * 16;
  const int td = (h.n_tab * 8 + 15) / 16 * 16, td0 = h.n_tab * 8 / 16 * 16;
  dma_to_lds(reinterpret_cast<const unsigned char*>(hdr_src), buf + hdr_off,
             (int)sizeof(TileHdr), lane);
  dma_to_lds(r0, buf, tr, lane);
  dma_to_lds(d0, buf + rec_bytes, td, lane);
  dma_to_lds(r0 + tr0, buf + tr0, rec_all - tr0, lane);
  dma_to_lds(d0 + td0, buf + rec_bytes + td0, desc_all - td0, lane);
  dma_to_lds(reinterpret_cast<const unsigned char*>(ev + h.first_event),
             buf + rec_bytes + (size_t)max_dials * 8, h.n_events * 16, lane);
  return dma_instrs(rec_all - tr0) + dma_instrs(desc_all - td0) + dma_instrs(h.n_events * 16);
}

// The tile header travels with the tile (evaluators and the event wave read
// it from LDS instead of waiting on a global load after each barrier).
__device__ __forceinline__ void dma_header(const TileHdr* hdr, unsigned char* buf,
                                           size_t hdr_off, int lane) {
  dma_to_lds(reinterpret_cast<const unsigned char*>(hdr), buf + hdr_off,
             (int)sizeof(TileHdr), lane);
}

struct LlhParams {
  int32_t type;
  int32_t use_poisson;
  double frac_limit;
  double expected_min;
  int32_t lsq_poisson;
  int32_t no_update_weights;
  const double* nominal_err;     // [n_bins] or null
  const double* extra_fraction;  // [n_bins] or null
  const uint8_t* ignore;         // [n_bins] or null
};

// Per-bin -2 ln L of bin b from (data, sum w, sqrt(sum w^2)) as
// JointProbabilityBase::eval(SamplePair, bin) hands them over
// (JointProbabilityBase.h:37-43).  Where the reference throws (LogThrowIf)
// the result is NaN.
__device__ inline double llh_bin(const LlhParams& c, int b, double data, double pred,
                          double err) {
  const double inf = __builtin_huge_val();
  const double nan = __builtin_nan("");
  const double dmin = 2.2250738585072014e-308;  // numeric_limits<double>::min
  const double deps = 2.220446049250313e-16;    // numeric_limits<double>::epsilon
  // JointProbabilityBase.h:74-77: bins flagged zero at the prior
  if (c.ignore && c.ignore[b]) return 0.0;
  switch (c.type) {
    case GCM_LLH_POISSON:  // PoissonLogLikelihood.h:19-53
      if (pred <= 0) return inf;
      if (data <= 0) return 2.0 * pred;
      return 2.0 * (pred - data + data * log(data / pred));
    case GCM_LLH_BARLOW_OA2021: {  // BarlowBeestonBanff2022.h:92-294
      double mcuncert = (c.no_update_weights && c.nominal_err) ? c.nominal_err[b] : err;
      mcuncert *= mcuncert;
      if (!isfinite(mcuncert) || mcuncert < 0.0) return inf;
      if (pred <= c.expected_min) {
        if (pred <= 0.0 && data > 0.0) return inf;
        return 0.0;
      }
      double newmc = pred;
      double penalty = 0.0;
      if (mcuncert > 0.0) {
        const double fractional = fmin(sqrt(mcuncert) / pred, c.frac_limit);
        const double temp = pred * fractional * fractional - 1;
        const double temp2 = temp * temp + 4 * data * fractional * fractional;
        if (temp2 < 0) return inf;
        const double beta = (-1 * temp + sqrt(temp2)) / 2.;
        newmc = fmax(pred * beta, dmin);
        penalty = (beta - 1) * (beta - 1) / (2 * fractional * fractional);
      }
      double stat = 0;
      if (data <= deps) {
        stat = newmc;
      } else if (newmc > dmin) {
        stat = newmc - data + data * (log(data) - log(newmc));
      } else {
        stat = dmin - data + data * (log(data) - log(dmin));
      }
      double chisq = 2.0 * stat;
      if (!c.use_poisson) chisq += 2.0 * penalty;
      return chisq;
    }
    case GCM_LLH_BARLOW_BEESTON: {  // BarlowBeeston.h:33-62 (err as given)
      const double dataVal = fmax(data, 0.0);
      const double predVal = fmax(pred, sqrt(dmin));
      const double mcUncert = fmax(err, dmin);
      const double rel_var = mcUncert / (predVal * predVal);
      const double bq = (predVal * rel_var) - 1;
      const double cq = 4 * dataVal * rel_var;
      const double beta = (-bq + sqrt(bq * bq + cq)) / 2.0;
      const double mc_hat = predVal * beta;
      double chi2 = 0.0;
      if (dataVal <= 0.0) {
        chi2 = 2 * mc_hat;
        chi2 += (beta - 1) * (beta - 1) / rel_var;
      } else {
        chi2 = 2 * (mc_hat - dataVal);
        if (dataVal > 0.0) chi2 += 2 * dataVal * (log(dataVal) - log(mc_hat));
        chi2 += (beta - 1) * (beta - 1) / rel_var;
      }
      return chi2;
    }
    case GCM_LLH_BARLOW_OA2020:      // BarlowBeestonBanff2020.h:18-85
    case GCM_LLH_BARLOW_OA2021_SFGD: {  // BarlowBeestonBanff2022Sfgd.h:20-134
      double fractional;
      if (c.type == GCM_LLH_BARLOW_OA2020) {
        fractional = sqrt(err) / pred;
      } else {
        const double sfg = c.extra_fraction ? err + pred * c.extra_fraction[b] : err;
        fractional = sfg / pred;
      }
      const double temp = pred * fractional * fractional - 1;
      const double temp2 = temp * temp + 4 * data * fractional * fractional;
      if (temp2 < 0) return nan;
      const double beta = (-1 * temp + sqrt(temp2)) / 2.;
      const double newmc = pred * beta;
      const double penalty = (beta - 1) * (beta - 1) / (2 * fractional * fractional);
      double stat = 0;
      if (data == 0) stat = newmc;
      else if (newmc > 0) stat = newmc - data + data * log(data / newmc);
      double chisq = 0.;
      if (c.type == GCM_LLH_BARLOW_OA2020 && isnan(penalty)) return inf;
      if (pred > 0.0 && data > 0.0) chisq += 2.0 * (stat + penalty);
      else if (pred > 0.0) chisq += 2.0 * (stat + penalty);
      if (c.type == GCM_LLH_BARLOW_OA2020 && isnan(chisq)) return nan;
      return chisq;
    }
    case GCM_LLH_CHI_SQUARED: {  // ChiSquared.h:19-29
      if (pred == 0) return inf;
      const double d = pred - data;
      return d * d / pred;
    }
    case GCM_LLH_LEAST_SQUARES: {  // LeastSquares.h:45-52
      double v = data - pred;
      v = v * v;
      if (c.lsq_poisson && data > 1.0) v /= 0.5 * data;
      return v;
    }
    default:
      return nan;
  }
}

// ---- the K-point path (gcm_points.hip) ----------------------------------
struct PointsArgs {
  const double* xT;         // [n_inputs][K]
  const double* tabT;       // [table_len][K]
  const uint32_t* tile_pr;  // [n_tiles] first range partial slot of each tile
  double* prs;              // [n_pr][K] sums of w
  double* prs2;             // [n_pr][K] sums of w^2
  double* prc;              // [n_pr][K] compensations (FP32 mode)
  double* prc2;
};

// The per-bin pass of a K-point launch in two levels, so that no wave adds
// more than `chunk` slots one after another (a bin's slot count grows with
// its events): level 1 sums runs of at most `chunk` consecutive slots of one
// bin, and finishes a bin that is a single run; level 2 adds a longer bin's
// run sums in run order and finishes it.  Every bin has at least one run
// (an empty bin a run of no slots), so every bin is written.
struct FinalizePlan {
  const uint4* runs;  // [n_runs] {first slot, slots, bin, 0 | 1 + item}
  int64_t n_runs;
  const uint4* long_bins;  // [n_long] {bin, first item, items, 0}
  int64_t n_long;
  double* ips;  // [items][K] the run sums of the long bins
  double* ips2;
  double* ipc;  // compensations (FP32 mode)
  double* ipc2;
};

// LDS bytes of one fill_points_kernel block for the tile layout in a.
size_t points_lds_bytes(const FillArgs& a, int nb);
// Allow fill_points_kernel its dynamic LDS (once per process is enough).
int points_set_lds_attrs();
// Enqueue fill_points_kernel over the tiles (a persistent grid of the blocks
// that fit on n_cu CUs at once); finalize_points_kernel over the bins;
// transpose_tables_kernel (tab [K][len] -> tabT [len][K]).
int launch_points(const FillArgs& a, const PointsArgs& pa, bool fp32, bool grid, int nb,
                  int n_cu, size_t lds, hipStream_t s);
int launch_finalize_points(const PointsArgs& pa, const FinalizePlan& fp, bool fp32, int n_bins,
                           int k_points, const double* data, const LlhParams& cfg,
                           double* hist, double* hist2, double* llh_terms, hipStream_t s);
int launch_transpose_tables(const double* tab, int64_t len, int k_points, double* tabT,
                            hipStream_t s);

}  // namespace gcm
