// gcm.hip -- MI355X (gfx950) event-reweighting cache: device layout, HIP
// kernels and the C ABI declared in include/gcm.h.
//
// Device layout (built once by gcm_build, resident in HBM):
//   * Binned events are stably sorted by global bin, so each bin's events are
//     contiguous and keep the caller's order inside the bin.
//   * The sorted events are cut into tiles of whole events (<= 64 events,
//     <= max_dials dials, records <= tile_bytes).  A tile's records are one
//     contiguous, 16-byte aligned run of words, ordered type-major so the
//     lanes of a wave evaluate the same dial type; every dial carries a 64-bit
//     descriptor {input, type, knots, event-major slot, record offset}.
//   * Per tile, the distinct bins of its events are "segments"; each segment
//     owns one partial-sum slot.  Slots are numbered tile by tile, so the
//     slots of one bin are contiguous (bin_pbeg).
//
// One evaluation (gcm_fill) is three launches on one stream:
//   prologue  : mirror each dial input (DialInputBuffer.cpp:56-70) and the
//               prior penalty, per parameter point.
//   fill      : persistent blocks stream tiles HBM -> LDS with 16-byte loads,
//               evaluate every dial from LDS, clamp (DialResponseSupervisor),
//               form each event's product in dial-list order, cap, times the
//               base weight (EventDialCache.cpp:156-171), then a wave-64
//               segmented scan by bin writes one (sum w, sum w^2) partial per
//               segment.  No atomics on the data path.
//   finalize  : per bin, sum the partial slots in slot order (deterministic),
//               evaluate the bin's LLH, reduce per block, and the last block
//               of each point sums the block results in order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gcm.h"
#include "../../include/gcm_table_update.h"
#include "dial_eval.hpp"
#include "fill_common.hpp"

namespace gcm {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}


// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

constexpr int kPenaltyChunk = 256;
constexpr int kTableUpdateBlock = 128;  // threads per (job, point) table update

// The value input i hands its dials: the parameter, mirrored into
// [lo, lo + range] when the input has mirror edges (DialInputBuffer.cpp:56-70).
__device__ __forceinline__ double input_value(const double* th, const int32_t* in_par,
                                              const double* in_mlo, const double* in_mrange,
                                              int i) {
  double v = th[in_par[i]];
  const double lo = in_mlo[i];
  if (!isnan(lo)) {
    const double range = in_mrange[i];
    v = fabs(fmod(v - lo, 2 * range));
    if (v > range) {
      v -= 2 * range;
      v = -v;
    }
    v += lo;
  }
  return v;
}

// Per point k (one block each, 256 threads): the point's parameters read
// from the caller's pinned staging buffer theta_src (when given) into theta,
// then the mirrored input values (and, for K-point launches of
// fill_points_kernel, their input-major copy xT), the inputs of the device
// table-update jobs (tab_in[k][j] = value of input tab_idx[j]) and the prior
// penalty.
__global__ void prologue_kernel(const double* theta_src, double* theta, int n_par,
                                const int32_t* __restrict__ in_par,
                                const double* __restrict__ in_mlo,
                                const double* __restrict__ in_mrange,
                                const double* __restrict__ in_min,
                                const double* __restrict__ in_max,
                                const int32_t* __restrict__ in_flags,
                                int n_inputs, InputRec* __restrict__ xin,
                                const int32_t* __restrict__ tab_idx, int n_tab_idx,
                                double* __restrict__ tab_in,
                                const double* __restrict__ prior_mean,
                                const double* __restrict__ prior_sigma,
                                double* __restrict__ penalty, int* __restrict__ flags,
                                double* __restrict__ xT) {
  const int k = blockIdx.x;
  double* th = theta + (size_t)k * n_par;
  if (theta_src) {  // host memory: system-scope loads, so no stale cached copy
    const double* src = theta_src + (size_t)k * n_par;
    for (int p = threadIdx.x; p < n_par; p += blockDim.x)
      th[p] = __hip_atomic_load(src + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __syncthreads();
  }
  // the NaN-weight flag reports this evaluation only (cleared in-stream, so
  // that graph replays and fills that are never synchronised start clean)
  if (k == 0 && threadIdx.x == 0) *flags = 0;
  for (int j = threadIdx.x; j < n_tab_idx; j += blockDim.x)
    tab_in[(size_t)k * n_tab_idx + j] = input_value(th, in_par, in_mlo, in_mrange, tab_idx[j]);
  for (int i = threadIdx.x; i < n_inputs; i += blockDim.x) {
    const double v = input_value(th, in_par, in_mlo, in_mrange, i);
    InputRec r;
    r.x = v;
    r.lo = in_min[i];
    r.hi = in_max[i];
    r.flags = in_flags[i];
    r.pad = 0;
    xin[(size_t)k * n_inputs + i] = r;
    if (xT) xT[(size_t)i * gridDim.x + k] = v;  // input-major copy for fill_points_kernel
  }
  // LikelihoodInterface.cpp:186-206 with a diagonal covariance: the terms are
  // formed in parallel and summed in parameter order by one thread.
  __shared__ double s_z[kPenaltyChunk];
  double pen = 0;
  if (prior_sigma) {
    for (int p0 = 0; p0 < n_par; p0 += kPenaltyChunk) {
      const int p = p0 + threadIdx.x;
      if (threadIdx.x < kPenaltyChunk && p < n_par) {
        const double s = prior_sigma[p];
        double z2 = 0.0;  // no term: adding +0.0 leaves the sum unchanged
        if (s > 0) {
          const double z = (th[p] - prior_mean[p]) / s;
          z2 = z * z;
        }
        s_z[threadIdx.x] = z2;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        const int n = min(kPenaltyChunk, n_par - p0);
        for (int j = 0; j < n; ++j) pen += s_z[j];
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) penalty[k] = pen;
}

// Correlated penalty sets, one block per point, added to the prologue's
// diagonal term in set order (LikelihoodInterface.cpp:186-213).  Per set:
// d_i = theta[par_i] - prior_i; y_i = 0 + sum_j A[i][j] d_j with j ascending
// (one thread per row, ROOT's TMatrixD x TVectorD order); the set's value is
// sum_i d_i y_i from 0 in index order (one thread), or +inf when a bounded
// parameter is NaN or out of its range.  meta[s] = {first, n, has_matrix,
// has_bounds}; inv_t holds each matrix transposed so that row i is read down
// column i by consecutive threads.
__global__ void penalty_sets_kernel(const double* __restrict__ theta, int n_par,
                                    int n_sets, const int64_t* __restrict__ meta,
                                    const int32_t* __restrict__ par,
                                    const double* __restrict__ prior,
                                    const double* __restrict__ lo,
                                    const double* __restrict__ hi,
                                    const double* __restrict__ inv_t,
                                    const int64_t* __restrict__ mat_off,
                                    int64_t total_n, double* __restrict__ scratch,
                                    double* __restrict__ penalty) {
  const int k = blockIdx.x;
  const double* th = theta + (size_t)k * n_par;
  double* dv = scratch + (size_t)k * 2 * total_n;
  double* yv = dv + total_n;
  __shared__ int s_out;
  double pen = 0;
  if (threadIdx.x == 0) pen = penalty[k];
  for (int s = 0; s < n_sets; ++s) {
    const int64_t first = meta[4 * s], n = meta[4 * s + 1];
    const bool has_mat = meta[4 * s + 2] != 0, has_bounds = meta[4 * s + 3] != 0;
    if (threadIdx.x == 0) s_out = 0;
    __syncthreads();
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
      const double v = th[par[first + i]];
      dv[first + i] = v - prior[first + i];
      // NaN bounds compare false: no bound on that side
      if (has_bounds && (isnan(v) || v < lo[first + i] || v > hi[first + i])) s_out = 1;
    }
    __syncthreads();
    if (has_mat) {
      const double* A = inv_t + mat_off[s];
      for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        double sum = 0;
        for (int64_t j = 0; j < n; ++j) sum += A[j * n + i] * dv[first + j];
        yv[first + i] = 0.0 + sum;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      double buf = 0;
      if (has_mat)
        for (int64_t i = 0; i < n; ++i) buf += dv[first + i] * yv[first + i];
      if (s_out) buf = __builtin_huge_val();
      pen += buf;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) penalty[k] = pen;
}


// The fill kernel.  Wave roles in a block of C + NB waves:
//   waves [0, C)   evaluate every dial of the current tile (type-sorted
//                  records, one dial per lane) into the step's response buffer;
//   wave C         forms the event weights and the bin partials of the
//                  previous step while the evaluators work on the next one;
//   waves C+1 ..   NB-1 producers: producer p streams the tiles whose ordinal
//                  is p mod (NB-1) HBM -> LDS with DMA loads, NB-1 tiles
//                  ahead, into a ring of NB buffers.
// A step is one (tile, point); every wave meets one barrier per step.  Each
// role runs its own loop over the block's steps (the barriers pair up by
// count, not by place), so that a role keeps only its own state live across
// steps: with one shared loop the union of the three roles' state spilled
// scalar registers, reloaded by every wave at every step.

// Producer p: at the first step of tile i, producer i mod P loads tile i + P
// into the buffer of tile i - 1; at the last step of tile i, producer
// (i + 1) mod P waits for tile i + 1.  The next tile's header is held in
// registers one DMA period ahead, so that no DMA waits on a header load.
// With table waves the tile is copied in two parts and, as soon as the first
// (header, table dials' records and descriptors) has landed, the producer
// writes the tile's ordinal into ready[buffer] for the table waves.
template <typename R, int NB>
__device__ __forceinline__ void fill_producer(const FillArgs& a, unsigned char* ring,
                                              size_t buf_bytes, size_t hdr_off, int* ready,
                                              int me, int lane, int64_t t0, int64_t G,
                                              TileHdr hq) {
  constexpr int P = NB - 1;
  const unsigned char* rec = reinterpret_cast<const unsigned char*>(a.rec);
  const unsigned char* desc = reinterpret_cast<const unsigned char*>(a.desc);
  int64_t i = 0;
  for (int64_t t = t0; t < a.n_tiles; t += G, ++i) {
    for (int k = 0; k < a.k_points; ++k) {
      if (k == 0 && me == (int)(i % P)) {
        const int64_t tn = t + P * G;
        if (tn < a.n_tiles) {
          unsigned char* nbuf = ring + (size_t)((i + P) % NB) * buf_bytes;
          if (a.tab_waves > 0) {
            const int n2 = dma_tile_split<R>(hq, a.tiles + tn, rec, desc, a.ev, nbuf,
                                             a.rec_bytes, a.max_dials, hdr_off, lane);
            wait_vm_at_most(n2);
            if (lane == 0) ready[(i + P) % NB] = (int)(i + P);
          } else {
            dma_tile<R>(hq, rec, desc, a.ev, nbuf, a.rec_bytes, a.max_dials, lane);
            dma_header(a.tiles + tn, nbuf, hdr_off, lane);
          }
          if (tn + P * G < a.n_tiles) hq = uniform_header(a.tiles[tn + P * G]);
        }
      }
      if (k == a.k_points - 1 && me == (int)((i + 1) % P))
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      lds_barrier();
    }
  }
  // no LDS-DMA load of this block may land after its LDS is handed on
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// Table waves: the table dials (Tabulated, Kriged) of each step, with their
// global-memory loads issued one step ahead.  In step s a table wave finishes
// the dials whose table values it loaded in step s - 1, then reads the
// records of step s + 1's tile (for a new tile, once the producer has flagged
// its first part as landed) and starts their loads, which complete across
// the barrier.  Lane l of table wave w takes table dial 64 w + l; dials past
// the waves' lanes are done within the step.
template <typename R, bool kInputsLds, int NB>
__device__ __forceinline__ void fill_tables(const FillArgs& a, unsigned char* ring,
                                            size_t buf_bytes, size_t hdr_off,
                                            double* resp_base, const int* ready,
                                            const ClassRec* s_cls, int tw, int lane,
                                            int64_t t0, int64_t G) {
  const int q0 = tw * kWave + lane;
  const int stride = a.tab_waves * kWave;
  double tv[kTabBatch];
  uint64_t dt = 0;
  double tlo = 0.0, thi = 0.0;
  bool pend = false;
  // start the loads of this lane's round-0 table dial of (tile ordinal j, point k)
  auto issue = [&](int64_t j, int k) {
    const unsigned char* buf = ring + (size_t)(j % NB) * buf_bytes;
    const TileHdr* hp = reinterpret_cast<const TileHdr*>(buf + hdr_off);
    const int n_tab = __builtin_amdgcn_readfirstlane(hp->n_tab);
    pend = q0 < n_tab;
    if (pend) {
      dt = reinterpret_cast<const uint64_t*>(buf + a.rec_bytes)[q0];
      table_issue<R>((int)((dt >> 16) & 0xf), (int)((dt >> 20) & 0xff),
                     reinterpret_cast<const R*>(buf) + (int)(dt >> 48),
                     a.tab + (size_t)k * a.table_len, tv);
      const int cl = (int)((dt >> 28) & 0xf);
      if (kInputsLds && cl != kEscapeClass) {
        tlo = s_cls[cl].lo;
        thi = s_cls[cl].hi;
      } else {
        const InputRec& ir = a.xin[(size_t)k * a.n_inputs + (int)(dt & 0xffff)];
        tlo = ir.lo;
        thi = ir.hi;
      }
    }
  };
  auto await_tile = [&](int64_t j) {
    const volatile int* rd = ready;
    while (rd[j % NB] != (int)j) __builtin_amdgcn_s_sleep(1);
    asm volatile("" ::: "memory");
  };
  if (t0 < a.n_tiles) {
    await_tile(0);
    issue(0, 0);
  }
  int step = 0;
  int64_t i = 0;
  for (int64_t t = t0; t < a.n_tiles; t += G, ++i) {
    const unsigned char* cur = ring + (size_t)(i % NB) * buf_bytes;
    const TileHdr h = *reinterpret_cast<const TileHdr*>(cur + hdr_off);
    const int n_tab = __builtin_amdgcn_readfirstlane(h.n_tab);
    const R* s_rec = reinterpret_cast<const R*>(cur);
    const uint64_t* s_desc = reinterpret_cast<const uint64_t*>(cur + a.rec_bytes);
    for (int k = 0; k < a.k_points; ++k, ++step) {
      double* resp = resp_base + (size_t)(step & 1) * a.max_dials;
      const double* tk = a.tab + (size_t)k * a.table_len;
      if (pend) {
        double r = table_finish<R>((int)((dt >> 16) & 0xf), (int)((dt >> 20) & 0xff),
                                   s_rec + (int)(dt >> 48), tk, tv);
        // DialResponseSupervisor::process (DialResponseSupervisor.cpp:10-18)
        if (r < tlo) r = tlo;
        else if (r > thi) r = thi;
        resp[(int)((dt >> 32) & 0xffff)] = r;
      }
      for (int q = q0 + stride; q < n_tab; q += stride) {
        const uint64_t d = s_desc[q];
        const int type = (int)((d >> 16) & 0xf), n = (int)((d >> 20) & 0xff);
        double tw2[kTabBatch];
        table_issue<R>(type, n, s_rec + (int)(d >> 48), tk, tw2);
        double r = table_finish<R>(type, n, s_rec + (int)(d >> 48), tk, tw2);
        const int cl = (int)((d >> 28) & 0xf);
        double lo, hi;
        if (kInputsLds && cl != kEscapeClass) {
          lo = s_cls[cl].lo;
          hi = s_cls[cl].hi;
        } else {
          const InputRec& ir = a.xin[(size_t)k * a.n_inputs + (int)(d & 0xffff)];
          lo = ir.lo;
          hi = ir.hi;
        }
        if (r < lo) r = lo;
        else if (r > hi) r = hi;
        resp[(int)((d >> 32) & 0xffff)] = r;
      }
      // the next step's table loads
      if (k + 1 < a.k_points) {
        issue(i, k + 1);
      } else if (t + G < a.n_tiles) {
        await_tile(i + 1);
        issue(i + 1, 0);
      } else {
        pend = false;
      }
      lds_barrier();
    }
  }
}

// Event wave: the event weights and bin partials of step s, formed after the
// barrier that ends step s (while the evaluators run step s + 1).
template <bool kComp, int NB>
__device__ __forceinline__ void fill_events(const FillArgs& a, unsigned char* ring,
                                            size_t buf_bytes, size_t hdr_off,
                                            double* resp_base, int lane, int64_t t0,
                                            int64_t G) {
  int step = 0;
  int64_t i = 0;
  for (int64_t t = t0; t < a.n_tiles; t += G, ++i) {
    const unsigned char* cur = ring + (size_t)(i % NB) * buf_bytes;
    const TileHdr h = *reinterpret_cast<const TileHdr*>(cur + hdr_off);
    const int n_ev = __builtin_amdgcn_readfirstlane(h.n_events);
    const uint32_t first_partial = __builtin_amdgcn_readfirstlane(h.first_partial);
    const uint32_t first_event = __builtin_amdgcn_readfirstlane(h.first_event);
    for (int k = 0; k < a.k_points; ++k, ++step) {
      const double* resp = resp_base + (size_t)(step & 1) * a.max_dials;
      // The event record is read before the barrier: after it, a producer
      // may refill this buffer.
      EventRec er{0.0, INT_MAX, 0u};
      if (lane < n_ev)
        er = reinterpret_cast<const EventRec*>(cur + a.rec_bytes + (size_t)a.max_dials * 8)[lane];
      lds_barrier();
      double w = 0.0, w2 = 0.0, c = 0.0, c2 = 0.0;
      int bin = INT_MAX;
      if (lane < n_ev) {
        const int b0 = (int)(er.dials & 0xffff), cnt = (int)(er.dials >> 16);
        // EventDialCache.cpp:159-170: the product in dial-list order; the
        // loads are batched four at a time, the multiplies stay in order.
        double temp = 1;
        for (int j = 0; j < cnt; j += 4) {
          const double r0 = resp[b0 + j];
          const double r1 = (j + 1 < cnt) ? resp[b0 + j + 1] : 1.0;
          const double r2 = (j + 2 < cnt) ? resp[b0 + j + 2] : 1.0;
          const double r3 = (j + 3 < cnt) ? resp[b0 + j + 3] : 1.0;
          temp *= r0;
          if (j + 1 < cnt) temp *= r1;
          if (j + 2 < cnt) temp *= r2;
          if (j + 3 < cnt) temp *= r3;
        }
        if (temp > a.cap) temp = a.cap;
        w = er.base;
        w *= temp;
        bin = er.bin;
        if (k == 0 && a.weights_out)
          a.weights_out[a.ev_result[(int64_t)first_event + lane]] = w;
        if (isnan(w)) atomicOr(a.flags, 1);
        w2 = w * w;
      }
      // Inclusive segmented scan over the wave; bins are sorted, so equal
      // bins are one contiguous segment.
#pragma unroll
      for (int dlt = 1; dlt < kWave; dlt <<= 1) {
        const double ws = __shfl_up(w, dlt);
        const double w2s = __shfl_up(w2, dlt);
        const int bs = __shfl_up(bin, dlt);
        if (kComp) {
          const double cs = __shfl_up(c, dlt);
          const double c2s = __shfl_up(c2, dlt);
          if (lane >= dlt && bs == bin) {
            double sum, e;
            two_sum(ws, w, sum, e);
            w = sum;
            c = c + cs + e;
            two_sum(w2s, w2, sum, e);
            w2 = sum;
            c2 = c2 + c2s + e;
          }
        } else {
          if (lane >= dlt && bs == bin) {
            w = ws + w;
            w2 = w2s + w2;
          }
        }
      }
      const int prev = __shfl_up(bin, 1);
      const int next = __shfl_down(bin, 1);
      const bool valid = lane < n_ev;
      const bool head = valid && (lane == 0 || prev != bin);
      const bool tail = valid && (lane == n_ev - 1 || next != bin);
      const unsigned long long heads = __ballot(head);
      const int seg = __popcll(heads & ((2ull << lane) - 1ull)) - 1;
      if (tail) {
        const size_t slot = (size_t)k * a.n_partials + first_partial + seg;
        a.psum[slot] = w;
        a.psum2[slot] = w2;
        if (kComp) {
          a.pcomp[slot] = c;
          a.pcomp2[slot] = c2;
        }
      }
    }
  }
}

// Evaluator waves: every dial of the tile for point k into the step's
// response buffer.  Per round of eval_threads dials, the other 1-D dials go
// to the threads in order and the table dials (Tabulated, Kriged) to the
// threads in reverse order, so that while a tile's dials fit in one round
// the table dials, whose loads go to global memory, sit in the last waves by
// themselves (fewest waves and load instructions waiting on global memory;
// measured faster than spreading them over the waves).  A thread holding
// both kinds in a round starts the table loads, evaluates its other dial,
// then finishes the table dial.  The 2-D dials follow in their own loop, so
// that their register-hungry grid code stays out of this one.
template <typename R, bool kInputsLds, int NB, bool kGrid>
__device__ __forceinline__ void fill_evaluate(const FillArgs& a, unsigned char* ring,
                                              size_t buf_bytes, size_t hdr_off,
                                              double* resp_base, const double* s_x,
                                              const ClassRec* s_cls, int eval_threads,
                                              bool with_tables, int64_t t0, int64_t G) {
  const int rev = eval_threads - 1 - (int)threadIdx.x;
  int step = 0;
  int64_t i = 0;
  for (int64_t t = t0; t < a.n_tiles; t += G, ++i) {
    const unsigned char* cur = ring + (size_t)(i % NB) * buf_bytes;
    const TileHdr h = *reinterpret_cast<const TileHdr*>(cur + hdr_off);
    const int n_dials = __builtin_amdgcn_readfirstlane(h.n_dials);
    const int n_tab = with_tables ? __builtin_amdgcn_readfirstlane(h.n_tab) : 0;
    const int tab_begin = __builtin_amdgcn_readfirstlane(h.n_tab);  // first other 1-D dial
    const int grid_begin = kGrid ? __builtin_amdgcn_readfirstlane(h.grid_begin) : n_dials;
    const int n_other = grid_begin - tab_begin;
    const int span = n_other > n_tab ? n_other : n_tab;
    const R* s_rec = reinterpret_cast<const R*>(cur);
    const uint64_t* s_desc = reinterpret_cast<const uint64_t*>(cur + a.rec_bytes);
    for (int k = 0; k < a.k_points; ++k, ++step) {
      double* resp = resp_base + (size_t)(step & 1) * a.max_dials;
      const InputRec* inp = a.xin + (size_t)k * a.n_inputs;
      const double* xk = kInputsLds ? s_x + (size_t)k * a.n_inputs : nullptr;
      const double* tk = a.tab + (size_t)k * a.table_len;
      for (int r0 = 0; r0 < span; r0 += eval_threads) {
        // the table dial: start its loads and its supervisor bounds
        const int qt = r0 + rev;
        const bool has_t = qt < n_tab;
        uint64_t dt = 0;
        double tv[kTabBatch];
        double tlo = 0.0, thi = 0.0;
        if (has_t) {
          dt = s_desc[qt];
          const int cl = (int)((dt >> 28) & 0xf);
          table_issue<R>((int)((dt >> 16) & 0xf), (int)((dt >> 20) & 0xff),
                         s_rec + (int)(dt >> 48), tk, tv);
          if (kInputsLds && cl != kEscapeClass) {
            tlo = s_cls[cl].lo;
            thi = s_cls[cl].hi;
          } else {
            const InputRec& ir = inp[(int)(dt & 0xffff)];
            tlo = ir.lo;
            thi = ir.hi;
          }
        }
        // another 1-D dial meanwhile
        const int qn = r0 + (int)threadIdx.x;
        if (qn < n_other) {
          const uint64_t d = s_desc[tab_begin + qn];
          const int in = (int)(d & 0xffff);
          const int type = (int)((d >> 16) & 0xf);
          const int n = (int)((d >> 20) & 0xff);
          const int dest = (int)((d >> 32) & 0xffff);
          const int off = (int)(d >> 48);
          const int cl = (int)((d >> 28) & 0xf);
          double x, lo, hi;
          int flags;
          if (kInputsLds && cl != kEscapeClass) {
            x = xk[in];
            lo = s_cls[cl].lo;
            hi = s_cls[cl].hi;
            flags = s_cls[cl].flags;
          } else {
            const InputRec ir = inp[in];
            x = ir.x;
            lo = ir.lo;
            hi = ir.hi;
            flags = ir.flags;
          }
          double r = eval_dial<R>(type, n, s_rec + off, x, flags, tk);
          // DialResponseSupervisor::process (DialResponseSupervisor.cpp:10-18)
          if (r < lo) r = lo;
          else if (r > hi) r = hi;
          resp[dest] = r;
        }
        // the table dial's response
        if (has_t) {
          double r = table_finish<R>((int)((dt >> 16) & 0xf), (int)((dt >> 20) & 0xff),
                                     s_rec + (int)(dt >> 48), tk, tv);
          if (r < tlo) r = tlo;
          else if (r > thi) r = thi;
          resp[(int)((dt >> 32) & 0xffff)] = r;
        }
      }
      for (int d0 = grid_begin + (int)threadIdx.x; kGrid && d0 < n_dials; d0 += eval_threads) {
        const uint64_t d = s_desc[d0];
        const int type = (int)((d >> 16) & 0xf);
        const int dest = (int)((d >> 32) & 0xffff);
        const int off = (int)(d >> 48);
        const InputRec ir = inp[(int)(d & 0xffff)];
        // record word 0: the second input (y axis); the payload follows
        const InputRec ir2 = inp[(int)s_rec[off]];
        double r = eval_dial2d<R>(type, s_rec + off + 1, ir.x, ir2.x, ir.flags);
        if (r < ir.lo) r = ir.lo;
        else if (r > ir.hi) r = ir.hi;
        resp[dest] = r;
      }
      lds_barrier();
    }
  }
}

template <typename R, bool kComp, bool kInputsLds, int NB, bool kGrid>
__global__ void __launch_bounds__(1024, 6) fill_kernel(FillArgs a) {
  constexpr int P = NB - 1;  // producer waves
  extern __shared__ __align__(16) unsigned char lds[];
  const int nwaves = blockDim.x / kWave;
  // wave-uniform role index in a scalar register, so that the role branches
  // are scalar branches rather than exec-masked divergent code
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int lane = threadIdx.x % kWave;
  const int n_eval = nwaves - 1 - P;
  const int ev_wave = n_eval;
  const int prod0 = n_eval + 1;
  // LDS: [input x values if staged][class records][tile ring][responses]
  const size_t in_bytes = kInputsLds ? (size_t)a.k_points * a.n_inputs * sizeof(double) : 0;
  const size_t in_pad = (in_bytes + kClassBytes + 1023) / 1024 * 1024;
  double* s_x = reinterpret_cast<double*>(lds);
  ClassRec* s_cls = reinterpret_cast<ClassRec*>(lds + in_bytes);
  const size_t hdr_off = (size_t)a.rec_bytes + (size_t)a.max_dials * 8 +
                         kMaxTileEvents * sizeof(EventRec);
  const size_t buf_bytes = hdr_off + kHeaderBytes;
  unsigned char* const ring = lds + in_pad;  // NB tile buffers
  double* const resp_base = reinterpret_cast<double*>(ring + NB * buf_bytes);
  int* const ready = reinterpret_cast<int*>(resp_base + 2 * (size_t)a.max_dials);
  const int64_t G = gridDim.x;
  const int64_t t0 = blockIdx.x;
  const int tab0 = n_eval - a.tab_waves;  // first table wave

  if (kInputsLds) {
    const int n = a.k_points * a.n_inputs;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s_x[i] = a.xin[i].x;
  }
  if (threadIdx.x < kInputClasses) s_cls[threadIdx.x] = a.cls[threadIdx.x];
  // producer p loads tile ordinal p into buffer p; tile 0 is waited for now,
  // tile p > 0 by producer p at the end of tile p - 1
  // With table waves each producer clears the ready flag of its first
  // buffer (producer 0 also those of the buffers no producer fills now)
  // before it copies a tile there, then flags the tile.
  TileHdr hq{};
  if (wave >= prod0) {
    const int p = wave - prod0;
    const int64_t tp = t0 + p * G;
    if (a.tab_waves > 0 && lane == 0) {
      ready[p] = -1;
      if (p == 0)
        for (int b = P; b < NB; ++b) ready[b] = -1;
    }
    if (tp < a.n_tiles) {
      const TileHdr h0 = uniform_header(a.tiles[tp]);
      if (a.tab_waves > 0) {
        const int n2 = dma_tile_split<R>(h0, a.tiles + tp, reinterpret_cast<const unsigned char*>(a.rec),
                                         reinterpret_cast<const unsigned char*>(a.desc), a.ev,
                                         ring + (size_t)p * buf_bytes, a.rec_bytes, a.max_dials,
                                         hdr_off, lane);
        wait_vm_at_most(n2);
        if (lane == 0) ready[p] = p;
      } else {
        dma_tile<R>(h0, reinterpret_cast<const unsigned char*>(a.rec),
                    reinterpret_cast<const unsigned char*>(a.desc), a.ev,
                    ring + (size_t)p * buf_bytes, a.rec_bytes, a.max_dials, lane);
        dma_header(a.tiles + tp, ring + (size_t)p * buf_bytes, hdr_off, lane);
      }
    }
    if (tp + P * G < a.n_tiles) hq = uniform_header(a.tiles[tp + P * G]);
    if (p == 0) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  }
  lds_barrier();

  if (wave >= prod0)
    fill_producer<R, NB>(a, ring, buf_bytes, hdr_off, ready, wave - prod0, lane, t0, G, hq);
  else if (wave == ev_wave)
    fill_events<kComp, NB>(a, ring, buf_bytes, hdr_off, resp_base, lane, t0, G);
  else if (wave >= tab0)
    fill_tables<R, kInputsLds, NB>(a, ring, buf_bytes, hdr_off, resp_base, ready, s_cls,
                                   wave - tab0, lane, t0, G);
  else
    fill_evaluate<R, kInputsLds, NB, kGrid>(a, ring, buf_bytes, hdr_off, resp_base, s_x, s_cls,
                                            tab0 * kWave, a.tab_waves == 0, t0, G);
}

// The fill kernel instance for a precision, input staging, ring depth and
// whether the tables hold 2-D dials (kGrid: their loop is compiled in only
// where needed, so the other workloads keep the leaner kernel).
template <bool kGrid>
static const void* fill_fn(bool fp32, bool inputs_lds, int nb) {
  if (fp32)
    return nb == 3 ? (inputs_lds ? (const void*)fill_kernel<float, true, true, 3, kGrid>
                                 : (const void*)fill_kernel<float, true, false, 3, kGrid>)
                   : (inputs_lds ? (const void*)fill_kernel<float, true, true, 2, kGrid>
                                 : (const void*)fill_kernel<float, true, false, 2, kGrid>);
  return nb == 3 ? (inputs_lds ? (const void*)fill_kernel<double, false, true, 3, kGrid>
                               : (const void*)fill_kernel<double, false, false, 3, kGrid>)
                 : (inputs_lds ? (const void*)fill_kernel<double, false, true, 2, kGrid>
                               : (const void*)fill_kernel<double, false, false, 2, kGrid>);
}

// Wave-level sum of (s, c) pairs with a fixed butterfly shape: the result is
// the same on every run.  With kComp the pairs are combined by two-sum.
template <bool kComp>
__device__ __forceinline__ void wave_sum(double& s, double& c) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const double so = __shfl_xor(s, off);
    if (kComp) {
      const double co = __shfl_xor(c, off);
      double t, e;
      // combine in lane order so both partners compute the same value
      if ((threadIdx.x & off) == 0) two_sum(s, so, t, e);
      else two_sum(so, s, t, e);
      s = t;
      c = ((threadIdx.x & off) == 0 ? c + co : co + c) + e;
    } else {
      s = ((threadIdx.x & off) == 0) ? s + so : so + s;
    }
  }
}

// One wave per (point, bin), grid-strided over all K x n_bins of them: the
// lanes read the bin's partial slots coalesced, then a fixed-shape wave
// reduction; lane 0 writes the bin's sums and its LLH term.
template <bool kComp>
__global__ void __launch_bounds__(kFinalizeBlock)
    finalize_bins_kernel(const double* __restrict__ psum,
                         const double* __restrict__ psum2,
                         const double* __restrict__ pcomp,
                         const double* __restrict__ pcomp2, int64_t n_partials,
                         const uint32_t* __restrict__ bin_pbeg, int n_bins, int k_points,
                         const double* __restrict__ data, LlhParams cfg,
                         double* __restrict__ hist, double* __restrict__ hist2,
                         double* __restrict__ llh_terms) {
  const int lane = threadIdx.x % kWave;
  const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  const int64_t items = (int64_t)k_points * n_bins;
  for (int64_t it = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
       it < items; it += n_waves) {
    const int k = (int)(it / n_bins), b = (int)(it % n_bins);
    const size_t kp = (size_t)k * n_partials;
    double s = 0.0, s2 = 0.0, c = 0.0, c2 = 0.0;
    const uint32_t p0 = bin_pbeg[b], p1 = bin_pbeg[b + 1];
    for (uint32_t p = p0 + lane; p < p1; p += kWave) {
      if (kComp) {
        double t, e;
        two_sum(s, psum[kp + p], t, e);
        s = t;
        c += e + pcomp[kp + p];
        two_sum(s2, psum2[kp + p], t, e);
        s2 = t;
        c2 += e + pcomp2[kp + p];
      } else {
        s += psum[kp + p];
        s2 += psum2[kp + p];
      }
    }
    wave_sum<kComp>(s, c);
    wave_sum<kComp>(s2, c2);
    if (kComp) {
      s = s + c;
      s2 = s2 + c2;
    }
    if (lane == 0) {
      hist[(size_t)k * n_bins + b] = s;
      hist2[(size_t)k * n_bins + b] = s2;
      llh_terms[(size_t)k * n_bins + b] = data ? llh_bin(cfg, b, data[b], s, sqrt(s2)) : 0.0;
    }
  }
}

// One block per point: the per-bin LLH terms summed in a fixed order (each
// thread a strided run in bin order, then a fixed tree), plus the penalty.
// The results also go to the caller's pinned result block host_res
// ([total][stat][penalty], M entries each, then the NaN flag), written with
// system-scope stores, so that reading them after the stream's
// synchronisation needs no copy.
__global__ void __launch_bounds__(kTotalBlock)
    finalize_total_kernel(const double* __restrict__ llh_terms, int n_bins,
                          const double* __restrict__ penalty, double* stat_out,
                          double* pen_out, double* total_out, const int* flags,
                          double* host_res, int M) {
  __shared__ double s_red[kTotalBlock];
  const int k = blockIdx.x;
  const double* lb = llh_terms + (size_t)k * n_bins;
  double acc = 0.0;
  for (int b = threadIdx.x; b < n_bins; b += blockDim.x) acc += lb[b];
  s_red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) s_red[threadIdx.x] += s_red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double stat = s_red[0];
    const double pen = penalty[k];
    stat_out[k] = stat;
    pen_out[k] = pen;
    total_out[k] = stat + pen;
    __hip_atomic_store(host_res + k, stat + pen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(host_res + M + k, stat, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(host_res + 2 * M + k, pen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    if (k == 0)
      __hip_atomic_store(reinterpret_cast<int*>(host_res + 3 * M), *flags, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// First-match binning: LoaderUtils.cpp:32-40 / Bin.cpp:167-176.
constexpr int kMaxBinVars = 8;
__global__ void bin_kernel(int64_t n, int nv, const double* __restrict__ vals,
                           int nb, const double* __restrict__ edges,
                           int32_t* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += stride) {
    double v[kMaxBinVars];
    for (int j = 0; j < nv; ++j) v[j] = vals[e * nv + j];
    int found = -1;
    for (int b = 0; b < nb; ++b) {
      bool in = true;
      for (int j = 0; j < nv; ++j) {
        const double lo = edges[((int64_t)b * nv + j) * 2];
        const double hi = edges[((int64_t)b * nv + j) * 2 + 1];
        bool ok;
        if (lo == hi) ok = (lo == v[j]);
        else ok = !(lo > v[j]) && !(hi <= v[j]);
        in = in && ok;
      }
      if (in) {
        found = b;
        break;
      }
    }
    out[e] = found;
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// Words of device record per dial (per precision).
// Descriptor type code of a dial: the public type, or an internal code for a
// general spline / graph whose knot abscissae are not non-decreasing (NaN
// included), which keeps the reference's binary segment search on the device.
int device_type_code(int type, int n, const double* p) {
  if (type == GCM_DIAL_GENERAL) {
    for (int j = 0; j + 1 < n; ++j)
      if (!(p[2 + 3 * j + 2] <= p[2 + 3 * (j + 1) + 2])) return kTypeGeneralUnsorted;
  } else if (type == GCM_DIAL_GRAPH) {
    for (int j = 0; j + 1 < n; ++j)
      if (!(p[2 * j + 1] <= p[2 * (j + 1) + 1])) return kTypeGraphUnsorted;
  }
  return type;
}

int record_words(int type, int n, bool fp32, int64_t len) {
  switch (type) {
    case GCM_DIAL_BILINEAR:
    case GCM_DIAL_BICUBIC: return (int)(1 + len);  // second input index + payload
    case GCM_DIAL_NORM: return 0;
    case GCM_DIAL_SHIFT: return 1;
    case GCM_DIAL_MONOTONIC:
    case GCM_DIAL_COMPACT: return n + 2;
    case GCM_DIAL_UNIFORM: return 2 * n + 2;
    case GCM_DIAL_GENERAL: return 3 * n + 2;
    case GCM_DIAL_GRAPH: return 2 * n;
    case GCM_DIAL_TABULATED: return fp32 ? 2 : 1;
    case GCM_DIAL_KRIGED: return fp32 ? 2 * n : n;
    default: return -1;
  }
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
};

}  // namespace gcm

using namespace gcm;

struct gcm_context {
  int device = 0;
  hipStream_t stream = nullptr;
  int precision = GCM_PRECISION_FP64;
  gcm_launch_config launch{kDefaultBlock, kDefaultTileWords, kDefaultBlocksPerCU, 2};
  int n_cu = 256;
  bool built = false;
  bool weights_enabled = false;
  bool llh_set = false;
  int k_points = 0;
  int max_points = 1;
  int64_t n_events_all = 0, n_events = 0, n_dials = 0, n_tiles = 0;
  int64_t n_partials = 0;
  int n_inputs = 0, n_par = 0, n_bins = 0, n_tables = 0;
  int64_t table_len = 0;
  int max_dials = 1024;
  int tab_dials_mid = 0;  // table dials per tile, median over tiles
  int fin_blocks = 1;
  double cap = __builtin_huge_val();
  LlhParams llh{GCM_LLH_POISSON, 0, 1e150, -1.0, 0, 0, nullptr, nullptr, nullptr};
  bool has_grid = false;             // the tables hold 2-D dials
  double* d_llh_nominal = nullptr;   // [n_bins], BBNoUpdateWeights
  double* d_llh_extra = nullptr;     // [n_bins], SFGD/WAGASCI fractions
  uint8_t* d_llh_ignore = nullptr;   // [n_bins], zero-prediction-at-prior mask
  std::vector<void*> allocations;
  int64_t resident = 0;
  // device
  TileHdr* d_tiles = nullptr;
  uint64_t* d_desc = nullptr;
  void* d_rec = nullptr;
  EventRec* d_ev = nullptr;
  int32_t* d_ev_result = nullptr;
  uint32_t* d_bin_pbeg = nullptr;
  // K-point launches (K >= kPointsMin, fill_points_kernel): range partial
  // slots, point-interleaved inputs and tables
  int64_t n_pr = 0;
  uint32_t* d_tile_pr = nullptr;    // [n_tiles]
  double *d_prs = nullptr, *d_prs2 = nullptr, *d_prc = nullptr, *d_prc2 = nullptr;  // [n_pr][K]
  // the two-level per-bin pass (FinalizePlan): runs, long bins, run sums
  uint4* d_fin_runs = nullptr;
  uint4* d_fin_long = nullptr;
  int64_t n_fin_runs = 0, n_fin_long = 0, n_fin_items = 0;
  double *d_ips = nullptr, *d_ips2 = nullptr, *d_ipc = nullptr, *d_ipc2 = nullptr;  // [items][K]
  double* d_xT = nullptr;           // [n_inputs][K]
  double* d_tabT = nullptr;         // [table_len][K]
  bool points_ready = false;
  double *d_psum = nullptr, *d_psum2 = nullptr, *d_pcomp = nullptr,
         *d_pcomp2 = nullptr;
  int32_t *d_in_par = nullptr, *d_in_flags = nullptr;
  double *d_in_mlo = nullptr, *d_in_mrange = nullptr, *d_in_min = nullptr,
         *d_in_max = nullptr;
  double *d_theta = nullptr, *d_tab = nullptr;
  InputRec* d_xin = nullptr;
  ClassRec* d_cls = nullptr;  // [kInputClasses]
  double *d_data = nullptr, *d_prior_mean = nullptr, *d_prior_sigma = nullptr;
  double *d_hist = nullptr, *d_hist2 = nullptr;
  double *d_llh_part = nullptr, *d_stat = nullptr, *d_pen = nullptr,
         *d_total = nullptr, *d_penalty = nullptr;
  int* d_flags = nullptr;
  double* d_weights = nullptr;
  // pinned staging; the parameters and the results are coherent pinned
  // memory that the prologue reads and finalize_total_kernel writes (their
  // device addresses h_theta_dev, h_res_dev), so that a fill copies neither
  double* h_theta = nullptr;
  double* h_tab = nullptr;
  // per-point results {total[K], stat[K], penalty[K], NaN flag}: the device
  // block d_res and its pinned mirror h_res
  double* d_res = nullptr;
  double* h_res = nullptr;
  double* h_theta_dev = nullptr;
  double* h_res_dev = nullptr;
  bool tab_dirty = false, theta_dirty = false;
  hipEvent_t ev_staged = nullptr;  // staging buffers consumed by the device
  hipStream_t last_stream = nullptr;
  bool lds_attr_set = false;
  // hipGraph mode: one instantiated graph per point count K
  bool graph_mode = false;
  std::vector<std::pair<int, hipGraphExec_t>> graphs;
  void drop_graphs() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
  }
  // kernel timing: a ring of event triples {before fill, after fill, after
  // finalize}, harvested at every synchronisation
  static constexpr int kTimingRing = 64;
  bool timing = false;
  hipEvent_t tev[kTimingRing][3] = {};
  int64_t t_issued = 0, t_done = 0;
  double t_fill_ms = 0, t_fin_ms = 0;

  // correlated penalty sets (gcm_set_penalty_sets), concatenated
  int n_pen_sets = 0;
  int64_t* d_pen_meta = nullptr;  // [n_sets][4]: first, n, has_matrix, has_bounds
  int32_t* d_pen_par = nullptr;   // [sum n]
  double *d_pen_prior = nullptr, *d_pen_lo = nullptr, *d_pen_hi = nullptr;
  double* d_pen_inv_t = nullptr;  // per set [n][n], transposed (column i contiguous
                                  // over rows is row i of the inverse)
  int64_t* d_pen_mat_off = nullptr;  // [n_sets] offsets into d_pen_inv_t
  double* d_pen_scratch = nullptr;   // [max_points][2 * sum n]: deltas, rows
  int64_t pen_total_n = 0;
  // device table refresh (gcm_add_table_update)
  struct TableUpdate {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    void* d_jobs = nullptr;
    int n_jobs = 0;
  };
  std::vector<TableUpdate> tab_updates;
  std::vector<char> tab_covered;       // [n_tables]
  std::vector<int32_t> tab_in_idx;     // every job's input indices, in job order
  int32_t* d_tab_in_idx = nullptr;
  double* d_tab_in = nullptr;          // [max_points][tab_in_idx.size()]
  bool host_tab = false;               // gcm_set_table_values has been called
  std::vector<int64_t> tab_begin;      // [n_tables+1]

  template <typename T>
  int alloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess)
      return fail(GCM_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocations.push_back(*p);
    sizes.push_back((int64_t)(count * sizeof(T)));
    resident += (int64_t)(count * sizeof(T));
    return GCM_OK;
  }
  // Free one allocation made by alloc (no-op for NULL); the device must be
  // done with it.
  template <typename T>
  void release(T** p) {
    if (!*p) return;
    for (size_t i = 0; i < allocations.size(); ++i)
      if (allocations[i] == (void*)*p) {
        (void)hipFree(allocations[i]);
        resident -= sizes[i];
        allocations.erase(allocations.begin() + i);
        sizes.erase(sizes.begin() + i);
        break;
      }
    *p = nullptr;
  }
  void free_all() {
    for (void* p : allocations) (void)hipFree(p);
    allocations.clear();
    sizes.clear();
    resident = 0;
  }
  std::vector<int64_t> sizes;  // bytes of each entry of allocations
};

#define GCM_TRY(expr)         \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != GCM_OK) return rc_; \
  } while (0)

namespace {

template <typename T>
int upload(gcm_context* c, T** dst, const T* src, size_t count) {
  GCM_TRY(c->alloc(dst, count));
  if (count && src) GCM_HIP(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return GCM_OK;
}

int validate_dial(const gcm_tables* t, int64_t d, int* n_out) {
  const int type = t->dial_type[d];
  const int64_t len = t->data_begin[d + 1] - t->data_begin[d];
  if (len < 0) return fail(GCM_ERR_ARG, "dial data_begin not monotonic");
  int n = 0;
  switch (type) {
    case GCM_DIAL_NORM: if (len != 0) return fail(GCM_ERR_ARG, "norm dial with payload"); break;
    case GCM_DIAL_SHIFT: if (len != 1) return fail(GCM_ERR_ARG, "shift dial needs 1 value"); n = 1; break;
    case GCM_DIAL_MONOTONIC:
    case GCM_DIAL_COMPACT: n = (int)(len - 2); if (n < 2) return fail(GCM_ERR_ARG, "spline needs >= 2 knots"); break;
    case GCM_DIAL_UNIFORM: n = (int)((len - 2) / 2); if ((len - 2) % 2 || len < 8) return fail(GCM_ERR_ARG, "uniform spline needs 2n+2 >= 8 values"); break;
    case GCM_DIAL_GENERAL: n = (int)((len - 2) / 3); if ((len - 2) % 3 || n < 2) return fail(GCM_ERR_ARG, "general spline needs 3n+2 values, n >= 2"); break;
    case GCM_DIAL_GRAPH: n = (int)(len / 2); if (len % 2 || n < 1) return fail(GCM_ERR_ARG, "graph needs 2n values"); break;
    case GCM_DIAL_TABULATED: n = 1; if (len != 3) return fail(GCM_ERR_ARG, "tabulated dial needs table, index, fraction"); break;
    case GCM_DIAL_KRIGED: n = (int)((len - 1) / 2); if (len < 1 || (len - 1) % 2) return fail(GCM_ERR_ARG, "kriged dial needs table + pairs"); break;
    case GCM_DIAL_BILINEAR:
    case GCM_DIAL_BICUBIC: {
      const double* p = t->dial_data + t->data_begin[d];
      const int64_t nx = len >= 2 ? (int64_t)p[0] : 0, ny = len >= 2 ? (int64_t)p[1] : 0;
      if (nx < 2 || ny < 2 || len != 2 + nx + ny + nx * ny)
        return fail(GCM_ERR_ARG, "2-D dial needs nx, ny >= 2 and 2 + nx + ny + nx*ny values");
      if (!t->dial_input2) return fail(GCM_ERR_ARG, "2-D dial without dial_input2");
      const int in2 = t->dial_input2[d];
      if (in2 < 0 || in2 >= t->n_inputs) return fail(GCM_ERR_ARG, "2-D dial second input out of range");
      n = 0;  // the grid shape is read from the record
      break;
    }
    default: return fail(GCM_ERR_UNSUPPORTED, "unsupported dial type " + std::to_string(type));
  }
  if (n > 255) return fail(GCM_ERR_UNSUPPORTED, "more than 255 knots/entries in one dial");
  const int in = t->dial_input[d];
  if (in < 0 || in >= t->n_inputs) return fail(GCM_ERR_ARG, "dial input out of range");
  if (type == GCM_DIAL_TABULATED || type == GCM_DIAL_KRIGED) {
    const double* p = t->dial_data + t->data_begin[d];
    const int tb = (int)p[0];
    if (tb < 0 || tb >= t->n_tables) return fail(GCM_ERR_ARG, "dial table out of range");
    const int64_t tsize = t->table_begin[tb + 1] - t->table_begin[tb];
    if (type == GCM_DIAL_TABULATED) {
      const int64_t i = (int64_t)p[1];
      if (i < 0 || i + 1 >= tsize) return fail(GCM_ERR_ARG, "tabulated index out of table");
    } else {
      for (int j = 0; j < n; ++j) {
        const int64_t i = (int64_t)p[1 + 2 * j];
        if (i < 0 || i >= tsize) return fail(GCM_ERR_ARG, "kriged index out of table");
      }
    }
  }
  *n_out = n;
  return GCM_OK;
}

// Lay out the tables for the device.  Templated on the record word type.
template <typename R>
int build_layout(gcm_context* c, const gcm_tables* t) {
  const bool fp32 = sizeof(R) == 4;
  const int64_t N = t->n_events;
  // input classes: the distinct (lo, hi, flags) of the response supervisors
  // as the fill kernel applies them; inputs past the first kEscapeClass
  // distinct ones keep kEscapeClass (full record read per dial)
  std::vector<uint8_t> in_class((size_t)std::max(1, t->n_inputs), (uint8_t)kEscapeClass);
  {
    std::vector<ClassRec> classes(kInputClasses, ClassRec{0.0, 0.0, 0, 0, 0.0});
    int n_cls = 0;
    const double inf = __builtin_huge_val();
    for (int i = 0; i < t->n_inputs; ++i) {
      const gcm_input& g = t->inputs[i];
      ClassRec r{std::isnan(g.min_response) ? -inf : g.min_response,
                 std::isnan(g.max_response) ? inf : g.max_response, g.flags, 0, 0.0};
      int found = -1;
      for (int k = 0; k < n_cls && found < 0; ++k)
        if (std::memcmp(&classes[k].lo, &r.lo, 8) == 0 && std::memcmp(&classes[k].hi, &r.hi, 8) == 0 &&
            classes[k].flags == r.flags)
          found = k;
      if (found < 0 && n_cls < kEscapeClass) {
        classes[n_cls] = r;
        found = n_cls++;
      }
      if (found >= 0) in_class[i] = (uint8_t)found;
    }
    GCM_TRY(upload(c, &c->d_cls, classes.data(), classes.size()));
  }
  // validate dials, record knot counts
  std::vector<uint8_t> knots(t->n_dials);
  for (int64_t d = 0; d < t->n_dials; ++d) {
    int n;
    GCM_TRY(validate_dial(t, d, &n));
    knots[d] = (uint8_t)n;
    if (t->dial_type[d] == GCM_DIAL_BILINEAR || t->dial_type[d] == GCM_DIAL_BICUBIC)
      c->has_grid = true;
  }
  // dial slots per tile (the LDS descriptor and response capacity): 12% above
  // what a tile of mean-size dial records holds and at least the largest
  // binned event's dial count, a multiple of 128 (so the
  // descriptor buffer is whole KiB: the producer wave's DMA writes whole KiB
  // chunks) in [256, 1024]; a tile that reaches it is closed early
  {
    double bytes = 0;
    for (int64_t d = 0; d < t->n_dials; ++d)
      bytes += (double)record_words(t->dial_type[d], knots[d], fp32,
                                    t->data_begin[d + 1] - t->data_begin[d]) * sizeof(R);
    const double mean = t->n_dials ? std::max(bytes / (double)t->n_dials, 1.0) : 8.0;
    const double want = 1.12 * ((double)c->launch.tile_words * 8.0) / mean;
    // never below the largest event's dial count (an event is never split)
    int64_t most = 0;
    for (int64_t e = 0; e < N; ++e)
      if (t->bin[e] >= 0) most = std::max<int64_t>(most, t->dial_begin[e + 1] - t->dial_begin[e]);
    const int64_t slots = (std::max<int64_t>((int64_t)std::ceil(want), most) + 127) / 128 * 128;
    c->max_dials = (int)std::min<int64_t>(1024, std::max<int64_t>(256, slots));
  }
  // stable counting sort of binned events by bin
  std::vector<int64_t> count(t->n_bins + 1, 0);
  for (int64_t e = 0; e < N; ++e) {
    const int b = t->bin[e];
    if (b < -1 || b >= t->n_bins) return fail(GCM_ERR_ARG, "event bin out of range");
    if (t->dial_begin[e + 1] < t->dial_begin[e]) return fail(GCM_ERR_ARG, "dial_begin not monotonic");
    if (b >= 0) count[b + 1]++;
  }
  for (int b = 0; b < t->n_bins; ++b) count[b + 1] += count[b];
  const int64_t NB = count[t->n_bins];
  std::vector<int64_t> order(NB);
  {
    std::vector<int64_t> pos(count.begin(), count.end() - 1);
    for (int64_t e = 0; e < N; ++e)
      if (t->bin[e] >= 0) order[pos[t->bin[e]]++] = e;
  }
  // Inside each bin, events are ordered (stably) by the lookup table of their
  // first kriged dial, then of their first tabulated dial: neighbouring events
  // then read the same tables, so the table loads of a wave share cache lines.
  // The per-bin sums keep a fixed order; only which fixed order changes.
  {
    std::vector<uint32_t> key(NB);
    for (int64_t i = 0; i < NB; ++i) {
      const int64_t e = order[i];
      uint32_t kr = 0xffff, tb = 0xffff;
      for (int64_t d = t->dial_begin[e]; d < t->dial_begin[e + 1]; ++d) {
        const int ty = t->dial_type[d];
        if (ty == GCM_DIAL_KRIGED && kr == 0xffff)
          kr = (uint32_t)std::min<int64_t>((int64_t)t->dial_data[t->data_begin[d]], 0xfffe);
        else if (ty == GCM_DIAL_TABULATED && tb == 0xffff)
          tb = (uint32_t)std::min<int64_t>((int64_t)t->dial_data[t->data_begin[d]], 0xfffe);
      }
      key[i] = kr << 16 | tb;
    }
    std::vector<int64_t> idx;
    std::vector<int64_t> tmp;
    for (int b = 0; b < t->n_bins; ++b) {
      const int64_t b0 = count[b], b1 = count[b + 1];
      if (b1 - b0 < 2) continue;
      idx.resize(b1 - b0);
      for (int64_t i = b0; i < b1; ++i) idx[i - b0] = i;
      std::stable_sort(idx.begin(), idx.end(),
                       [&](int64_t x, int64_t y) { return key[x] < key[y]; });
      tmp.resize(b1 - b0);
      for (int64_t i = b0; i < b1; ++i) tmp[i - b0] = order[idx[i - b0]];
      std::copy(tmp.begin(), tmp.end(), order.begin() + b0);
    }
  }
  // tiles: greedy over sorted events
  const int words_per16 = 16 / (int)sizeof(R);
  const int64_t cap_words = (int64_t)c->launch.tile_words * 8 / (int64_t)sizeof(R);
  std::vector<TileHdr> tiles;
  {
    TileHdr cur{};
    int64_t cur_words = 0, cur_dials = 0, cur_events = 0;
    int64_t rec_pos = 0, dial_pos = 0;
    auto close = [&](int64_t ev_end) {
      (void)ev_end;
      cur.rec_words = (uint16_t)cur_words;
      cur.n_dials = (uint16_t)cur_dials;
      cur.n_events = (uint16_t)cur_events;
      tiles.push_back(cur);
      rec_pos += (cur_words + words_per16 - 1) / words_per16 * words_per16;
      dial_pos += (cur_dials + 1) / 2 * 2;
    };
    for (int64_t i = 0; i < NB; ++i) {
      const int64_t e = order[i];
      int64_t w = 0;
      const int64_t nd = t->dial_begin[e + 1] - t->dial_begin[e];
      for (int64_t d = t->dial_begin[e]; d < t->dial_begin[e + 1]; ++d)
        w += record_words(t->dial_type[d], knots[d], fp32, t->data_begin[d + 1] - t->data_begin[d]);
      if (w > cap_words || nd > c->max_dials)
        return fail(GCM_ERR_UNSUPPORTED, "one event's dials exceed the tile capacity; raise tile_words");
      if (cur_events > 0 && (cur_events + 1 > kMaxTileEvents ||
                             cur_dials + nd > c->max_dials ||
                             cur_words + w > cap_words)) {
        close(i);
        cur_words = cur_dials = cur_events = 0;
      }
      if (cur_events == 0) {
        cur = TileHdr{};
        cur.first_event = (uint32_t)i;
        cur.first_dial = (uint32_t)dial_pos;
        cur.rec_begin = (uint64_t)rec_pos;
      }
      cur_words += w;
      cur_dials += nd;
      cur_events += 1;
    }
    if (cur_events > 0) close(NB);
    if (dial_pos > UINT32_MAX || NB > UINT32_MAX)
      return fail(GCM_ERR_UNSUPPORTED, "more than 2^32 dials or events");
  }
  const int64_t T = (int64_t)tiles.size();
  const int64_t total_words = T ? (int64_t)tiles.back().rec_begin +
                                      (tiles.back().rec_words + words_per16 - 1) / words_per16 * words_per16
                                : 0;
  const int64_t total_desc = T ? (int64_t)tiles.back().first_dial + (tiles.back().n_dials + 1) / 2 * 2 : 0;
  // partial slots: one per distinct bin per tile
  std::vector<uint32_t> bin_np(t->n_bins + 1, 0);
  int64_t np = 0;
  for (int64_t ti = 0; ti < T; ++ti) {
    tiles[ti].first_partial = (uint32_t)np;
    int prev = -1;
    for (int j = 0; j < tiles[ti].n_events; ++j) {
      const int b = t->bin[order[tiles[ti].first_event + j]];
      if (j == 0 || b != prev) {
        np++;
        bin_np[b + 1]++;
      }
      prev = b;
    }
  }
  if (np > UINT32_MAX) return fail(GCM_ERR_UNSUPPORTED, "too many partial slots");
  for (int b = 0; b < t->n_bins; ++b) bin_np[b + 1] += bin_np[b];
  // K-point launches (fill_points_kernel): one partial slot per distinct bin
  // per range of a tile (kRanges contiguous event ranges, range r starting at
  // event r n / kRanges), numbered tile by tile and range by range
  std::vector<uint32_t> tile_pr((size_t)std::max<int64_t>(T, 1), 0);
  std::vector<uint32_t> bin_npr(t->n_bins + 1, 0);
  int64_t npr = 0;
  for (int64_t ti = 0; ti < T; ++ti) {
    tile_pr[ti] = (uint32_t)npr;
    const int n = tiles[ti].n_events;
    int prev = -1;
    for (int j = 0; j < n; ++j) {
      const int b = t->bin[order[tiles[ti].first_event + j]];
      bool start = false;
      for (int r = 0; r < kRanges; ++r) start = start || (r * n / kRanges == j);
      if (start || b != prev) {
        npr++;
        bin_npr[b + 1]++;
      }
      prev = b;
    }
  }
  if (npr > UINT32_MAX) return fail(GCM_ERR_UNSUPPORTED, "too many partial slots");
  for (int b = 0; b < t->n_bins; ++b) bin_npr[b + 1] += bin_npr[b];
  // The two-level plan of their per-bin pass: runs of at most `chunk` slots,
  // chunk about the square root of the longest bin's slot count (a multiple
  // of 8, at least 16), so that neither level adds a long chain of slots.
  uint32_t max_slots = 0;
  for (int b = 0; b < t->n_bins; ++b) max_slots = std::max(max_slots, bin_npr[b + 1] - bin_npr[b]);
  uint32_t chunk = 16;
  while ((uint64_t)chunk * chunk < max_slots) chunk += 8;
  std::vector<uint4> fin_runs, fin_long;
  fin_runs.reserve(t->n_bins);
  uint32_t fin_items = 0;
  for (int b = 0; b < t->n_bins; ++b) {
    const uint32_t p0 = bin_npr[b], n = bin_npr[b + 1] - p0;
    if (n <= chunk) {
      fin_runs.push_back(make_uint4(p0, n, (uint32_t)b, 0));
      continue;
    }
    const uint32_t m = (n + chunk - 1) / chunk;
    fin_long.push_back(make_uint4((uint32_t)b, fin_items, m, 0));
    for (uint32_t i = 0; i < m; ++i) {
      const uint32_t f = p0 + i * chunk;
      fin_runs.push_back(make_uint4(f, std::min(chunk, p0 + n - f), (uint32_t)b, fin_items + i + 1));
    }
    fin_items += m;
  }

  // fill records, descriptors and event arrays, tiles in parallel
  std::vector<R> rec((size_t)std::max<int64_t>(total_words, 1), R(0));
  std::vector<uint64_t> desc((size_t)std::max<int64_t>(total_desc, 1), 0);
  std::vector<EventRec> ev(NB);
  std::vector<int32_t> ev_result(NB);
  int nthreads = (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  std::vector<std::string> errs(nthreads);
  for (int th = 0; th < nthreads; ++th) {
    pool.emplace_back([&, th] {
      std::vector<int64_t> dial_ids;
      std::vector<int> slot;
      for (int64_t ti = th; ti < T; ti += nthreads) {
        TileHdr& h = tiles[ti];
        dial_ids.clear();
        slot.clear();
        int local = 0;
        for (int j = 0; j < h.n_events; ++j) {
          const int64_t se = h.first_event + j;
          const int64_t e = order[se];
          const int64_t d0 = t->dial_begin[e], d1 = t->dial_begin[e + 1];
          ev[se].base = t->base_weight[e];
          ev[se].bin = t->bin[e];
          ev_result[se] = (int32_t)e;
          ev[se].dials = (uint32_t)local | ((uint32_t)(d1 - d0) << 16);
          for (int64_t d = d0; d < d1; ++d) {
            dial_ids.push_back(d);
            slot.push_back(local++);
          }
        }
        // type-major order, stable otherwise: Tabulated and Kriged (also by
        // table), then the other 1-D types, then the 2-D types
        std::vector<int> idx(dial_ids.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
        auto sort_key = [&](int q) -> int64_t {
          const int64_t d = dial_ids[q];
          const int ty = t->dial_type[d];
          int64_t tb = 0;
          int rank = 16 + ty;
          if (ty == GCM_DIAL_TABULATED || ty == GCM_DIAL_KRIGED) {
            tb = (int64_t)t->dial_data[t->data_begin[d]];
            rank = ty;
          } else if (ty == GCM_DIAL_BILINEAR || ty == GCM_DIAL_BICUBIC) {
            rank = 32 + ty;
          }
          return (int64_t)rank << 40 | tb;
        };
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int x, int y) { return sort_key(x) < sort_key(y); });
        int64_t off = 0, tab_words = 0;
        uint32_t grid0 = (uint32_t)idx.size(), n_tab = 0;
        for (size_t q = 0; q < idx.size(); ++q) {
          const int64_t d = dial_ids[idx[q]];
          const int type = t->dial_type[d];
          const int n = knots[d];
          if ((type == GCM_DIAL_BILINEAR || type == GCM_DIAL_BICUBIC) && grid0 > q)
            grid0 = (uint32_t)q;
          const bool table_dial = type == GCM_DIAL_TABULATED || type == GCM_DIAL_KRIGED;
          const double* p = t->dial_data + t->data_begin[d];
          R* out = rec.data() + h.rec_begin + off;
          switch (type) {
            case GCM_DIAL_NORM: break;
            case GCM_DIAL_TABULATED:
            case GCM_DIAL_KRIGED: {
              const int tb = (int)p[0];
              const int64_t base = t->table_begin[tb];
              for (int j = 0; j < n; ++j) {
                const uint32_t sl = (uint32_t)(base + (int64_t)p[1 + 2 * j]);
                const float w = (float)p[2 + 2 * j];
                if (fp32) {
                  uint32_t wb;
                  std::memcpy(&wb, &w, 4);
                  std::memcpy(out + 2 * j, &sl, 4);
                  std::memcpy(out + 2 * j + 1, &wb, 4);
                } else {
                  uint32_t wb;
                  std::memcpy(&wb, &w, 4);
                  const uint64_t bits = (uint64_t)sl | ((uint64_t)wb << 32);
                  std::memcpy(out + j, &bits, 8);
                }
              }
              break;
            }
            case GCM_DIAL_BILINEAR:
            case GCM_DIAL_BICUBIC: {
              const int64_t len = t->data_begin[d + 1] - t->data_begin[d];
              out[0] = (R)t->dial_input2[d];
              for (int64_t j = 0; j < len; ++j) out[1 + j] = (R)p[j];
              break;
            }
            default: {
              const int64_t len = t->data_begin[d + 1] - t->data_begin[d];
              for (int64_t j = 0; j < len; ++j) out[j] = (R)p[j];
            }
          }
          const uint64_t in = (uint64_t)t->dial_input[d];
          const int code = device_type_code(type, n, p);
          desc[h.first_dial + q] = in | ((uint64_t)code << 16) | ((uint64_t)n << 20) |
                                   ((uint64_t)in_class[in] << 28) |
                                   ((uint64_t)slot[idx[q]] << 32) | ((uint64_t)off << 48);
          off += record_words(type, n, fp32, t->data_begin[d + 1] - t->data_begin[d]);
          if (table_dial) {
            ++n_tab;
            tab_words = off;
          }
        }
        h.grid_begin = (uint16_t)grid0;  // the 2-D types sort last
        h.n_tab = (uint16_t)n_tab;
        h.tab_words = (uint16_t)tab_words;
      }
    });
  }
  for (auto& th : pool) th.join();

  c->n_events = NB;
  c->n_dials = t->n_dials;
  c->n_tiles = T;
  c->n_partials = np;
  // table dials per tile that the table waves cover in one round: the
  // median over the tiles (a tile with more does the rest within its step;
  // covering the 99th percentile instead took an evaluator wave from the
  // other dials and was slower)
  {
    std::vector<int> nt(T);
    for (int64_t ti = 0; ti < T; ++ti) nt[ti] = tiles[ti].n_tab;
    int mid = 0;
    if (T > 0) {
      std::nth_element(nt.begin(), nt.begin() + T / 2, nt.end());
      mid = nt[T / 2];
    }
    c->tab_dials_mid = mid;
  }
  // the producer wave copies whole KiB chunks: pad both streams by 1 KiB
  desc.resize(desc.size() + 1024 / sizeof(uint64_t), 0);
  rec.resize(rec.size() + 1024 / sizeof(R), R(0));
  GCM_TRY(upload(c, &c->d_tiles, tiles.data(), tiles.size()));
  GCM_TRY(upload(c, &c->d_desc, desc.data(), desc.size()));
  {
    R* p = nullptr;
    GCM_TRY(upload(c, &p, rec.data(), rec.size()));
    c->d_rec = p;
  }
  GCM_TRY(upload(c, &c->d_ev, ev.data(), ev.size()));
  GCM_TRY(upload(c, &c->d_ev_result, ev_result.data(), ev_result.size()));
  GCM_TRY(upload(c, &c->d_bin_pbeg, bin_np.data(), bin_np.size()));
  c->n_pr = npr;
  GCM_TRY(upload(c, &c->d_tile_pr, tile_pr.data(), tile_pr.size()));
  c->n_fin_runs = (int64_t)fin_runs.size();
  c->n_fin_long = (int64_t)fin_long.size();
  c->n_fin_items = fin_items;
  GCM_TRY(upload(c, &c->d_fin_runs, fin_runs.data(), fin_runs.size()));
  GCM_TRY(upload(c, &c->d_fin_long, fin_long.data(), fin_long.size()));
  return GCM_OK;
}

}  // namespace

extern "C" {

const char* gcm_last_error(void) { return g_error.c_str(); }

int gcm_version(void) { return 10000; }

int gcm_device_count(int* count) {
  if (!count) return fail(GCM_ERR_ARG, "count is NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return GCM_OK;
}

int gcm_create(int device, gcm_context** out) {
  if (!out) return fail(GCM_ERR_ARG, "out is NULL");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(GCM_ERR_HIP, "no HIP device visible");
  if (device < 0 || device >= n) return fail(GCM_ERR_ARG, "device out of range");
  GCM_HIP(hipSetDevice(device));
  auto* c = new gcm_context();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->n_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_staged, hipEventDisableTiming) != hipSuccess) {
    delete c;
    return fail(GCM_ERR_HIP, "stream/event creation failed");
  }
  *out = c;
  return GCM_OK;
}

int gcm_destroy(gcm_context* c) {
  if (!c) return GCM_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->drop_graphs();
  for (auto& u : c->tab_updates) (void)hipModuleUnload(u.mod);
  c->tab_updates.clear();
  c->free_all();
  for (auto& tr : c->tev)
    for (auto& e : tr)
      if (e) (void)hipEventDestroy(e);
  if (c->h_theta) (void)hipHostFree(c->h_theta);
  if (c->h_tab) (void)hipHostFree(c->h_tab);
  if (c->h_res) (void)hipHostFree(c->h_res);
  if (c->ev_staged) (void)hipEventDestroy(c->ev_staged);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return GCM_OK;
}

int gcm_set_precision(gcm_context* c, int precision) {
  if (!c) return fail(GCM_ERR_ARG, "ctx is NULL");
  if (c->built) return fail(GCM_ERR_STATE, "precision must be set before gcm_build");
  if (precision != GCM_PRECISION_FP64 && precision != GCM_PRECISION_FP32)
    return fail(GCM_ERR_ARG, "unknown precision");
  c->precision = precision;
  return GCM_OK;
}

int gcm_set_launch_config(gcm_context* c, const gcm_launch_config* cfg) {
  if (!c || !cfg) return fail(GCM_ERR_ARG, "NULL argument");
  gcm_launch_config l = c->launch;
  if (cfg->block_threads) l.block_threads = cfg->block_threads;
  if (cfg->tile_words) {
    if (c->built && cfg->tile_words != c->launch.tile_words)
      return fail(GCM_ERR_STATE, "tile_words is fixed at gcm_build");
    l.tile_words = cfg->tile_words;
  }
  if (cfg->blocks_per_cu) l.blocks_per_cu = cfg->blocks_per_cu;
  if (cfg->ring_buffers) l.ring_buffers = cfg->ring_buffers;
  if (l.ring_buffers != 2 && l.ring_buffers != 3)
    return fail(GCM_ERR_ARG, "ring_buffers must be 2 or 3");
  if (l.block_threads < 64 || l.block_threads > 832 || l.block_threads % 64)
    return fail(GCM_ERR_ARG, "block_threads must be a multiple of 64 in [64, 832]");
  if (l.tile_words < 256 || l.tile_words > 8192 || l.tile_words % 128)
    return fail(GCM_ERR_ARG, "tile_words must be a multiple of 128 in [256, 8192]");
  if (l.blocks_per_cu < 1 || l.blocks_per_cu > 32)
    return fail(GCM_ERR_ARG, "blocks_per_cu must be in [1, 32]");
  c->launch = l;
  return GCM_OK;
}

int gcm_get_launch_config(gcm_context* c, gcm_launch_config* cfg) {
  if (!c || !cfg) return fail(GCM_ERR_ARG, "NULL argument");
  *cfg = c->launch;
  return GCM_OK;
}

int gcm_build(gcm_context* c, const gcm_tables* t, int32_t max_points) {
  if (!c || !t) return fail(GCM_ERR_ARG, "NULL argument");
  if (c->built) return fail(GCM_ERR_STATE, "context already built");
  if (max_points < 1) return fail(GCM_ERR_ARG, "max_points must be >= 1");
  if (t->n_events < 0 || t->n_dials < 0 || t->n_inputs < 0 || t->n_bins < 1 ||
      t->n_parameters < 1 || t->n_tables < 0)
    return fail(GCM_ERR_ARG, "invalid table sizes");
  if (t->n_inputs > 65536) return fail(GCM_ERR_UNSUPPORTED, "more than 65536 dial inputs");
  if (t->n_events && (!t->base_weight || !t->bin || !t->dial_begin))
    return fail(GCM_ERR_ARG, "event arrays missing");
  if (t->dial_begin && t->n_events && (t->dial_begin[0] != 0 || t->dial_begin[t->n_events] != t->n_dials))
    return fail(GCM_ERR_ARG, "dial_begin must span [0, n_dials]");
  if (t->n_dials && (!t->dial_type || !t->dial_input || !t->data_begin || !t->dial_data))
    return fail(GCM_ERR_ARG, "dial arrays missing");
  for (int i = 0; i < t->n_inputs; ++i)
    if (t->inputs[i].parameter < 0 || t->inputs[i].parameter >= t->n_parameters)
      return fail(GCM_ERR_ARG, "input parameter out of range");
  GCM_HIP(hipSetDevice(c->device));
  c->n_events_all = t->n_events;
  c->n_inputs = t->n_inputs;
  c->n_par = t->n_parameters;
  c->n_bins = t->n_bins;
  c->n_tables = t->n_tables;
  c->table_len = t->n_tables ? t->table_begin[t->n_tables] : 0;
  if (t->n_tables) c->tab_begin.assign(t->table_begin, t->table_begin + t->n_tables + 1);
  c->max_points = max_points;
  c->cap = std::isfinite(t->reweight_cap) ? t->reweight_cap : __builtin_huge_val();
  int rc = (c->precision == GCM_PRECISION_FP32) ? build_layout<float>(c, t)
                                                 : build_layout<double>(c, t);
  if (rc != GCM_OK) {
    c->free_all();
    return rc;
  }
  // inputs
  std::vector<int32_t> par(t->n_inputs), flg(t->n_inputs);
  std::vector<double> mlo(t->n_inputs), mr(t->n_inputs), mn(t->n_inputs), mx(t->n_inputs);
  const double inf = __builtin_huge_val();
  for (int i = 0; i < t->n_inputs; ++i) {
    const gcm_input& in = t->inputs[i];
    par[i] = in.parameter;
    flg[i] = in.flags;
    mlo[i] = in.mirror_lower;
    mr[i] = in.mirror_range;
    mn[i] = std::isnan(in.min_response) ? -inf : in.min_response;
    mx[i] = std::isnan(in.max_response) ? inf : in.max_response;
  }
  const size_t K = (size_t)max_points;
  // the per-bin pass: one wave per (point, bin), grid-strided over at most
  // kFinalizeMaxBlocks blocks (about what the device holds at once)
  {
    const int64_t items = (int64_t)max_points * c->n_bins;
    const int64_t per_block = kFinalizeBlock / kWave;
    c->fin_blocks = (int)std::max<int64_t>(1, std::min<int64_t>(kFinalizeMaxBlocks,
                                                                (items + per_block - 1) / per_block));
  }
  rc = GCM_OK;
  auto step = [&](int r) { if (rc == GCM_OK) rc = r; };
  step(upload(c, &c->d_in_par, par.data(), par.size()));
  step(upload(c, &c->d_in_flags, flg.data(), flg.size()));
  step(upload(c, &c->d_in_mlo, mlo.data(), mlo.size()));
  step(upload(c, &c->d_in_mrange, mr.data(), mr.size()));
  step(upload(c, &c->d_in_min, mn.data(), mn.size()));
  step(upload(c, &c->d_in_max, mx.data(), mx.size()));
  step(c->alloc(&c->d_theta, K * c->n_par));
  step(c->alloc(&c->d_xin, K * std::max(1, c->n_inputs)));
  step(c->alloc(&c->d_tab, K * std::max<int64_t>(1, c->table_len)));
  step(c->alloc(&c->d_psum, K * c->n_partials));
  step(c->alloc(&c->d_psum2, K * c->n_partials));
  if (c->precision == GCM_PRECISION_FP32) {
    step(c->alloc(&c->d_pcomp, K * c->n_partials));
    step(c->alloc(&c->d_pcomp2, K * c->n_partials));
  }
  if (max_points >= kPointsMin) {
    step(c->alloc(&c->d_prs, K * std::max<int64_t>(1, c->n_pr)));
    step(c->alloc(&c->d_prs2, K * std::max<int64_t>(1, c->n_pr)));
    if (c->precision == GCM_PRECISION_FP32) {
      step(c->alloc(&c->d_prc, K * std::max<int64_t>(1, c->n_pr)));
      step(c->alloc(&c->d_prc2, K * std::max<int64_t>(1, c->n_pr)));
    }
    step(c->alloc(&c->d_ips, K * std::max<int64_t>(1, c->n_fin_items)));
    step(c->alloc(&c->d_ips2, K * std::max<int64_t>(1, c->n_fin_items)));
    if (c->precision == GCM_PRECISION_FP32) {
      step(c->alloc(&c->d_ipc, K * std::max<int64_t>(1, c->n_fin_items)));
      step(c->alloc(&c->d_ipc2, K * std::max<int64_t>(1, c->n_fin_items)));
    }
    step(c->alloc(&c->d_xT, K * std::max(1, c->n_inputs)));
    step(c->alloc(&c->d_tabT, K * std::max<int64_t>(1, c->table_len)));
    c->points_ready = true;
  }
  step(c->alloc(&c->d_hist, K * c->n_bins));
  step(c->alloc(&c->d_hist2, K * c->n_bins));
  step(c->alloc(&c->d_llh_part, K * c->n_bins));  // per-bin LLH terms
  step(c->alloc(&c->d_res, 3 * K + 1));
  step(c->alloc(&c->d_penalty, K));
  step(c->alloc(&c->d_data, c->n_bins));
  step(c->alloc(&c->d_prior_mean, c->n_par));
  step(c->alloc(&c->d_prior_sigma, c->n_par));
  if (rc != GCM_OK) {
    c->free_all();
    return rc;
  }
  c->d_total = c->d_res;
  c->d_stat = c->d_res + K;
  c->d_pen = c->d_res + 2 * K;
  c->d_flags = reinterpret_cast<int*>(c->d_res + 3 * K);
  GCM_HIP(hipMemset(c->d_res, 0, (3 * K + 1) * sizeof(double)));
  GCM_HIP(hipMemset(c->d_flags, 0, sizeof(int)));
  GCM_HIP(hipMemset(c->d_tab, 0, K * std::max<int64_t>(1, c->table_len) * sizeof(double)));
  GCM_HIP(hipMemset(c->d_theta, 0, K * c->n_par * sizeof(double)));
  const unsigned coherent = hipHostMallocMapped | hipHostMallocCoherent;
  if (hipHostMalloc((void**)&c->h_theta, K * c->n_par * sizeof(double), coherent) != hipSuccess ||
      hipHostMalloc((void**)&c->h_tab, K * std::max<int64_t>(1, c->table_len) * sizeof(double)) != hipSuccess ||
      hipHostMalloc((void**)&c->h_res, (3 * K + 1) * sizeof(double), coherent) != hipSuccess)
    return fail(GCM_ERR_NOMEM, "pinned staging allocation failed");
  std::memset(c->h_theta, 0, K * c->n_par * sizeof(double));
  std::memset(c->h_res, 0, (3 * K + 1) * sizeof(double));
  GCM_HIP(hipHostGetDevicePointer((void**)&c->h_theta_dev, c->h_theta, 0));
  GCM_HIP(hipHostGetDevicePointer((void**)&c->h_res_dev, c->h_res, 0));
  // default likelihood: no data histogram -> stat term 0 until set
  c->built = true;
  return GCM_OK;
}

int gcm_resident_bytes(gcm_context* c, int64_t* bytes) {
  if (!c || !bytes) return fail(GCM_ERR_ARG, "NULL argument");
  *bytes = c->resident;
  return GCM_OK;
}

int gcm_counts(gcm_context* c, int64_t* events, int64_t* dials, int64_t* tiles) {
  if (!c) return fail(GCM_ERR_ARG, "ctx is NULL");
  if (events) *events = c->n_events;
  if (dials) *dials = c->n_dials;
  if (tiles) *tiles = c->n_tiles;
  return GCM_OK;
}

// Accumulate the elapsed times of every timed fill that has been issued.
static int harvest_timing(gcm_context* c) {
  while (c->t_done < c->t_issued) {
    hipEvent_t* e = c->tev[c->t_done % gcm_context::kTimingRing];
    GCM_HIP(hipEventSynchronize(e[2]));
    float a = 0, b = 0;
    GCM_HIP(hipEventElapsedTime(&a, e[0], e[1]));
    GCM_HIP(hipEventElapsedTime(&b, e[1], e[2]));
    c->t_fill_ms += a;
    c->t_fin_ms += b;
    c->t_done++;
  }
  return GCM_OK;
}

static int wait_staging(gcm_context* c) {
  GCM_HIP(hipEventSynchronize(c->ev_staged));
  return GCM_OK;
}

int gcm_set_parameters(gcm_context* c, const double* theta, int32_t n_parameters,
                       int32_t k_points) {
  if (!c || !theta) return fail(GCM_ERR_ARG, "NULL argument");
  if (!c->built) return fail(GCM_ERR_STATE, "gcm_build first");
  if (n_parameters != c->n_par) return fail(GCM_ERR_ARG, "parameter count mismatch");
  if (k_points < 1 || k_points > c->max_points) return fail(GCM_ERR_ARG, "k_points out of [1, max_points]");
  GCM_TRY(wait_staging(c));
  std::memcpy(c->h_theta, theta, (size_t)k_points * n_parameters * sizeof(double));
  c->k_points = k_points;
  c->theta_dirty = true;
  return GCM_OK;
}

int gcm_set_table_values(gcm_context* c, const double* values, int64_t table_len,
                         int32_t k_points) {
  if (!c || (!values && table_len)) return fail(GCM_ERR_ARG, "NULL argument");
  if (!c->built) return fail(GCM_ERR_STATE, "gcm_build first");
  if (table_len != c->table_len) return fail(GCM_ERR_ARG, "table length mismatch");
  if (k_points < 1 || k_points > c->max_points) return fail(GCM_ERR_ARG, "k_points out of [1, max_points]");
  if (table_len == 0) return GCM_OK;
  GCM_TRY(wait_staging(c));
  std::memcpy(c->h_tab, values, (size_t)k_points * table_len * sizeof(double));
  c->tab_dirty = true;
  if (!c->host_tab) {
    c->host_tab = true;
    c->drop_graphs();  // the captured uploads now include the tables
  }
  return GCM_OK;
}

int gcm_set_likelihood(gcm_context* c, const gcm_llh_config* cfg,
                       const double* data_hist, const double* prior_mean,
                       const double* prior_sigma) {
  if (!c || !cfg || !data_hist) return fail(GCM_ERR_ARG, "NULL argument");
  if (!c->built) return fail(GCM_ERR_STATE, "gcm_build first");
  if (cfg->type < 0 || cfg->type >= GCM_LLH_NTYPES)
    return fail(GCM_ERR_ARG, "unknown likelihood type");
  if (cfg->bb_no_update_weights && !cfg->nominal_err)
    return fail(GCM_ERR_ARG, "bb_no_update_weights needs nominal_err");
  GCM_HIP(hipSetDevice(c->device));
  // a fill enqueued on a caller stream may still read the buffers replaced here
  GCM_HIP(hipStreamSynchronize(c->last_stream ? c->last_stream : c->stream));
  c->llh.type = cfg->type;
  c->llh.use_poisson = cfg->use_poisson_likelihood;
  // BarlowBeestonBanff2022.h:85-88: hard limit on the fractional error
  c->llh.frac_limit = std::min(cfg->fractional_error_limit, 1.0E+152);
  c->llh.expected_min = cfg->expected_value_minimum;
  c->llh.lsq_poisson = cfg->lsq_poissonian_approximation;
  c->llh.no_update_weights = cfg->bb_no_update_weights;
  const size_t nb = (size_t)c->n_bins;
  c->llh.nominal_err = nullptr;
  c->llh.extra_fraction = nullptr;
  c->llh.ignore = nullptr;
  if (cfg->nominal_err) {
    if (!c->d_llh_nominal) GCM_TRY(c->alloc(&c->d_llh_nominal, nb));
    GCM_HIP(hipMemcpy(c->d_llh_nominal, cfg->nominal_err, nb * sizeof(double), hipMemcpyHostToDevice));
    c->llh.nominal_err = c->d_llh_nominal;
  }
  if (cfg->extra_fraction) {
    if (!c->d_llh_extra) GCM_TRY(c->alloc(&c->d_llh_extra, nb));
    GCM_HIP(hipMemcpy(c->d_llh_extra, cfg->extra_fraction, nb * sizeof(double), hipMemcpyHostToDevice));
    c->llh.extra_fraction = c->d_llh_extra;
  }
  if (cfg->ignore_bins) {
    if (!c->d_llh_ignore) GCM_TRY(c->alloc(&c->d_llh_ignore, nb));
    GCM_HIP(hipMemcpy(c->d_llh_ignore, cfg->ignore_bins, nb, hipMemcpyHostToDevice));
    c->llh.ignore = c->d_llh_ignore;
  }
  GCM_HIP(hipMemcpy(c->d_data, data_hist, c->n_bins * sizeof(double), hipMemcpyHostToDevice));
  std::vector<double> mean(c->n_par, 0.0), sig(c->n_par, 0.0);
  if (prior_mean && prior_sigma) {
    std::copy(prior_mean, prior_mean + c->n_par, mean.begin());
    std::copy(prior_sigma, prior_sigma + c->n_par, sig.begin());
  }
  GCM_HIP(hipMemcpy(c->d_prior_mean, mean.data(), c->n_par * sizeof(double), hipMemcpyHostToDevice));
  GCM_HIP(hipMemcpy(c->d_prior_sigma, sig.data(), c->n_par * sizeof(double), hipMemcpyHostToDevice));
  c->llh_set = true;
  c->drop_graphs();  // kernel arguments changed
  return GCM_OK;
}

int gcm_set_penalty_sets(gcm_context* c, int32_t n_sets, const gcm_penalty_set* sets) {
  if (!c || (n_sets > 0 && !sets)) return fail(GCM_ERR_ARG, "NULL argument");
  if (!c->built) return fail(GCM_ERR_STATE, "gcm_build first");
  if (n_sets < 0) return fail(GCM_ERR_ARG, "n_sets < 0");
  std::vector<int64_t> meta, mat_off;
  std::vector<int32_t> par;
  std::vector<double> prior, lo, hi, inv_t;
  const double nan = std::nan("");
  for (int s = 0; s < n_sets; ++s) {
    const gcm_penalty_set& ps = sets[s];
    if (ps.n < 0 || (ps.n > 0 && (!ps.parameters || !ps.prior)))
      return fail(GCM_ERR_ARG, "penalty set " + std::to_string(s) + ": missing arrays");
    if ((ps.lower == nullptr) != (ps.upper == nullptr))
      return fail(GCM_ERR_ARG, "penalty set " + std::to_string(s) + ": give both bounds or neither");
    const int64_t first = (int64_t)par.size(), n = ps.n;
    meta.insert(meta.end(), {first, n, ps.inv_cov ? 1 : 0, ps.lower ? 1 : 0});
    mat_off.push_back((int64_t)inv_t.size());
    for (int64_t i = 0; i < n; ++i) {
      if (ps.parameters[i] < 0 || ps.parameters[i] >= c->n_par)
        return fail(GCM_ERR_ARG, "penalty set " + std::to_string(s) + ": parameter index out of range");
      par.push_back(ps.parameters[i]);
      prior.push_back(ps.prior[i]);
      lo.push_back(ps.lower ? ps.lower[i] : nan);
      hi.push_back(ps.upper ? ps.upper[i] : nan);
    }
    if (ps.inv_cov)
      for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) inv_t.push_back(ps.inv_cov[i * n + j]);
  }
  GCM_HIP(hipSetDevice(c->device));
  GCM_HIP(hipStreamSynchronize(c->last_stream ? c->last_stream : c->stream));
  c->release(&c->d_pen_meta);
  c->release(&c->d_pen_par);
  c->release(&c->d_pen_prior);
  c->release(&c->d_pen_lo);
  c->release(&c->d_pen_hi);
  c->release(&c->d_pen_inv_t);
  c->release(&c->d_pen_mat_off);
  c->release(&c->d_pen_scratch);
  c->n_pen_sets = 0;
  c->pen_total_n = 0;
  c->drop_graphs();  // kernel arguments change
  if (n_sets == 0) return GCM_OK;
  GCM_TRY(upload(c, &c->d_pen_meta, meta.data(), meta.size()));
  GCM_TRY(upload(c, &c->d_pen_par, par.data(), par.size()));
  GCM_TRY(upload(c, &c->d_pen_prior, prior.data(), prior.size()));
  GCM_TRY(upload(c, &c->d_pen_lo, lo.data(), lo.size()));
  GCM_TRY(upload(c, &c->d_pen_hi, hi.data(), hi.size()));
  GCM_TRY(upload(c, &c->d_pen_inv_t, inv_t.data(), inv_t.size()));
  GCM_TRY(upload(c, &c->d_pen_mat_off, mat_off.data(), mat_off.size()));
  GCM_TRY(c->alloc(&c->d_pen_scratch, (size_t)c->max_points * 2 * std::max<size_t>(1, par.size())));
  c->n_pen_sets = n_sets;
  c->pen_total_n = (int64_t)par.size();
  return GCM_OK;
}

int gcm_add_table_update(gcm_context* c, const char* path, const char* kernel_name,
                         int32_t n_jobs, const gcm_table_job* jobs) {
  if (!c || !path || !kernel_name || (n_jobs > 0 && !jobs)) return fail(GCM_ERR_ARG, "NULL argument");
  if (!c->built) return fail(GCM_ERR_STATE, "gcm_build first");
  if (n_jobs < 1) return fail(GCM_ERR_ARG, "n_jobs < 1");
  if ((int64_t)n_jobs > 65535) return fail(GCM_ERR_ARG, "more than 65535 jobs in one update");
  c->tab_covered.resize((size_t)c->n_tables, 0);
  std::vector<char> covered = c->tab_covered;
  std::vector<int32_t> idx = c->tab_in_idx;
  std::vector<gcm_table_job_dev> dj((size_t)n_jobs);
  for (int j = 0; j < n_jobs; ++j) {
    const gcm_table_job& jb = jobs[j];
    const std::string where = "table job " + std::to_string(j) + ": ";
    if (jb.table < 0 || jb.table >= c->n_tables) return fail(GCM_ERR_ARG, where + "table out of range");
    if (covered[jb.table]) return fail(GCM_ERR_ARG, where + "table already has an update");
    if (jb.n_inputs < 0 || (jb.n_inputs > 0 && !jb.inputs)) return fail(GCM_ERR_ARG, where + "missing inputs");
    covered[jb.table] = 1;
    dj[j].table_offset = c->tab_begin[jb.table];
    dj[j].size = (int32_t)(c->tab_begin[jb.table + 1] - c->tab_begin[jb.table]);
    dj[j].input_offset = (int64_t)idx.size();
    dj[j].n_inputs = jb.n_inputs;
    dj[j].user_id = jb.user_id;
    dj[j].reserved = 0;
    for (int i = 0; i < jb.n_inputs; ++i) {
      if (jb.inputs[i] < 0 || jb.inputs[i] >= c->n_inputs)
        return fail(GCM_ERR_ARG, where + "input index out of range");
      idx.push_back(jb.inputs[i]);
    }
  }
  GCM_HIP(hipSetDevice(c->device));
  gcm_context::TableUpdate u;
  // a failed module call also leaves its error as the thread's last HIP
  // error, which a later hipGetLastError() after a kernel launch would
  // report: clear it on these paths
  hipError_t e = hipModuleLoad(&u.mod, path);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(GCM_ERR_ARG, std::string("cannot load code object ") + path + ": " + hipGetErrorString(e));
  }
  e = hipModuleGetFunction(&u.fn, u.mod, kernel_name);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    (void)hipModuleUnload(u.mod);
    return fail(GCM_ERR_ARG, std::string("no kernel ") + kernel_name + " in " + path);
  }
  GCM_HIP(hipStreamSynchronize(c->last_stream ? c->last_stream : c->stream));
  gcm_table_job_dev* d_jobs = nullptr;
  int rc = upload(c, &d_jobs, dj.data(), dj.size());
  if (rc == GCM_OK && idx.size() != c->tab_in_idx.size()) {
    c->release(&c->d_tab_in_idx);
    c->release(&c->d_tab_in);
    rc = upload(c, &c->d_tab_in_idx, idx.data(), idx.size());
    if (rc == GCM_OK) rc = c->alloc(&c->d_tab_in, (size_t)c->max_points * std::max<size_t>(1, idx.size()));
  }
  if (rc != GCM_OK) {
    c->release(&d_jobs);
    (void)hipModuleUnload(u.mod);
    return rc;
  }
  u.d_jobs = d_jobs;
  u.n_jobs = n_jobs;
  c->tab_updates.push_back(u);
  c->tab_covered = covered;
  c->tab_in_idx = idx;
  c->drop_graphs();
  return GCM_OK;
}

int gcm_clear_table_updates(gcm_context* c) {
  if (!c) return fail(GCM_ERR_ARG, "ctx is NULL");
  if (c->tab_updates.empty()) return GCM_OK;
  GCM_HIP(hipSetDevice(c->device));
  GCM_HIP(hipStreamSynchronize(c->last_stream ? c->last_stream : c->stream));
  c->drop_graphs();
  for (auto& u : c->tab_updates) {
    c->release(&u.d_jobs);
    (void)hipModuleUnload(u.mod);
  }
  c->tab_updates.clear();
  c->tab_covered.assign((size_t)c->n_tables, 0);
  c->tab_in_idx.clear();
  c->release(&c->d_tab_in_idx);
  c->release(&c->d_tab_in);
  return GCM_OK;
}

int gcm_get_table_values(gcm_context* c, int32_t k, double* values) {
  if (!c || !values) return fail(GCM_ERR_ARG, "NULL argument");
  if (!c->built) return fail(GCM_ERR_STATE, "gcm_build first");
  if (k < 0 || k >= c->max_points) return fail(GCM_ERR_ARG, "k out of [0, max_points)");
  if (c->table_len == 0) return GCM_OK;
  GCM_HIP(hipSetDevice(c->device));
  GCM_HIP(hipStreamSynchronize(c->last_stream ? c->last_stream : c->stream));
  GCM_HIP(hipMemcpy(values, c->d_tab + (size_t)k * c->table_len, (size_t)c->table_len * sizeof(double),
                    hipMemcpyDeviceToHost));
  return GCM_OK;
}

int gcm_enable_weights(gcm_context* c, int enable) {
  if (!c) return fail(GCM_ERR_ARG, "ctx is NULL");
  if (!c->built) return fail(GCM_ERR_STATE, "gcm_build first");
  if (enable && !c->d_weights) {
    GCM_TRY(c->alloc(&c->d_weights, (size_t)std::max<int64_t>(1, c->n_events_all)));
    GCM_HIP(hipMemset(c->d_weights, 0, (size_t)std::max<int64_t>(1, c->n_events_all) * sizeof(double)));
  }
  if (c->weights_enabled != (enable != 0)) {
    GCM_HIP(hipStreamSynchronize(c->last_stream ? c->last_stream : c->stream));
    c->drop_graphs();  // kernel arguments changed
  }
  c->weights_enabled = enable != 0;
  return GCM_OK;
}

// Per-kernel attributes that must be set outside stream capture.
static int prepare_launch(gcm_context* c) {
  if (c->lds_attr_set) return GCM_OK;
  for (int v = 0; v < 16; ++v) {
    const bool fp32 = v & 1, inl = v & 2, grid = v & 8;
    const int nb = (v & 4) ? 3 : 2;
    const void* f = grid ? fill_fn<true>(fp32, inl, nb) : fill_fn<false>(fp32, inl, nb);
    GCM_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxDynamicLds));
  }
  GCM_TRY(points_set_lds_attrs());
  c->lds_attr_set = true;
  return GCM_OK;
}

// Enqueue one evaluation of K points on stream s: with `copies` the staged
// parameters (read by the prologue from pinned memory) and the host-table
// upload, then the prologue, fill and finalize kernels (the last one writes
// the results to pinned memory), and the timing events (when tev is given).
// Used eagerly and under hipGraph stream capture.
static int enqueue_fill(gcm_context* c, hipStream_t s, int K, bool copies,
                        hipEvent_t* tev) {
  if (copies) {
    if (c->table_len && c->host_tab)
      GCM_HIP(hipMemcpyAsync(c->d_tab, c->h_tab, (size_t)K * c->table_len * sizeof(double),
                             hipMemcpyHostToDevice, s));
  }
  // K-point launches of at least kPointsMin points: one lane per point
  const bool points = K >= kPointsMin && c->points_ready;
  const int n_tab_idx = c->tab_updates.empty() ? 0 : (int)c->tab_in_idx.size();
  prologue_kernel<<<K, 256, 0, s>>>(copies ? c->h_theta_dev : nullptr, c->d_theta, c->n_par, c->d_in_par, c->d_in_mlo,
                                    c->d_in_mrange, c->d_in_min, c->d_in_max,
                                    c->d_in_flags, c->n_inputs, c->d_xin,
                                    c->d_tab_in_idx, n_tab_idx, c->d_tab_in,
                                    c->llh_set ? c->d_prior_mean : nullptr,
                                    c->llh_set ? c->d_prior_sigma : nullptr, c->d_penalty,
                                    c->d_flags, points ? c->d_xT : nullptr);
  GCM_HIP(hipGetLastError());
  if (c->n_pen_sets > 0) {
    penalty_sets_kernel<<<K, 256, 0, s>>>(c->d_theta, c->n_par, c->n_pen_sets, c->d_pen_meta,
                                          c->d_pen_par, c->d_pen_prior, c->d_pen_lo, c->d_pen_hi,
                                          c->d_pen_inv_t, c->d_pen_mat_off, c->pen_total_n,
                                          c->d_pen_scratch, c->d_penalty);
    GCM_HIP(hipGetLastError());
  }
  if (!c->tab_updates.empty()) {
    // device table refresh: the prologue gathered the jobs' mirrored inputs;
    // one launch of each update function over (job, point)
    const int n_idx = n_tab_idx;
    for (auto& u : c->tab_updates) {
      const void* jobs = u.d_jobs;
      double* tab = c->d_tab;
      int64_t tab_stride = c->table_len;
      const double* tin = c->d_tab_in;
      int64_t in_stride = n_idx;
      void* args[] = {&jobs, &tab, &tab_stride, &tin, &in_stride};
      GCM_HIP(hipModuleLaunchKernel(u.fn, (unsigned)u.n_jobs, (unsigned)K, 1, kTableUpdateBlock,
                                    1, 1, 0, s, args, nullptr));
    }
  }
  const bool fp32 = c->precision == GCM_PRECISION_FP32;
  FillArgs a;
  a.cls = c->d_cls;
  a.tiles = c->d_tiles;
  a.n_tiles = c->n_tiles;
  a.desc = c->d_desc;
  a.rec = c->d_rec;
  a.ev = c->d_ev;
  a.ev_result = c->d_ev_result;
  a.xin = c->d_xin;
  a.n_inputs = c->n_inputs;
  a.tab = c->d_tab;
  a.table_len = c->table_len;
  a.cap = c->cap;
  a.k_points = K;
  a.psum = c->d_psum;
  a.psum2 = c->d_psum2;
  a.pcomp = c->d_pcomp;
  a.pcomp2 = c->d_pcomp2;
  a.n_partials = c->n_partials;
  a.weights_out = c->weights_enabled ? c->d_weights : nullptr;
  a.flags = c->d_flags;
  a.rec_bytes = c->launch.tile_words * 8;
  a.max_dials = c->max_dials;
  // The inputs' x values are staged in LDS when they are small and the
  // staging still lets blocks_per_cu blocks share a CU; otherwise the
  // evaluators read them from global memory (L1/L2) and the LDS goes to
  // larger tiles.  The class records always sit in LDS.
  const size_t in_bytes = (size_t)K * c->n_inputs * sizeof(double);
  const size_t ring_lds =
      (size_t)c->launch.ring_buffers * ((size_t)a.rec_bytes + (size_t)a.max_dials * 8 +
                                        kMaxTileEvents * sizeof(EventRec) + kHeaderBytes) +
      2 * (size_t)a.max_dials * 8 + kReadyBytes;
  const size_t staged_pad = (in_bytes + kClassBytes + 1023) / 1024 * 1024;
  // Table waves: enough to take the median tile's table dials in one round,
  // while at least one evaluator wave is left for the other dials.
  {
    const int n_eval = c->launch.block_threads / kWave;
    const int want = (c->tab_dials_mid + kWave - 1) / kWave;
    a.tab_waves = std::max(0, std::min(want, n_eval - 1));
  }
  a.inputs_lds = in_bytes <= kInputsLdsMax &&
                 (staged_pad + ring_lds) * (size_t)c->launch.blocks_per_cu <= kLdsPerCU;
  const size_t in_pad = a.inputs_lds ? staged_pad : (kClassBytes + 1023) / 1024 * 1024;
  const size_t lds = in_pad + ring_lds;
  if (points && c->table_len) GCM_TRY(launch_transpose_tables(c->d_tab, c->table_len, K, c->d_tabT, s));
  if (tev) GCM_HIP(hipEventRecord(tev[0], s));
  if (points && c->n_tiles > 0) {
    const int nb = c->launch.ring_buffers;
    const size_t plds = points_lds_bytes(a, nb);
    PointsArgs pa{c->d_xT, c->d_tabT, c->d_tile_pr, c->d_prs, c->d_prs2, c->d_prc, c->d_prc2};
    GCM_TRY(launch_points(a, pa, fp32, c->has_grid, nb, c->n_cu, plds, s));
  } else if (c->n_tiles > 0) {
    const int64_t want = (int64_t)c->launch.blocks_per_cu * c->n_cu;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(c->n_tiles, want));
    const int nb = c->launch.ring_buffers;
    // evaluators + the event wave + nb - 1 producer waves
    const int threads = c->launch.block_threads + nb * kWave;
    const void* fn = c->has_grid ? fill_fn<true>(fp32, a.inputs_lds, nb)
                                 : fill_fn<false>(fp32, a.inputs_lds, nb);
    void* args[] = {&a};
    GCM_HIP(hipLaunchKernel(fn, dim3(grid), dim3(threads), args, lds, s));
    GCM_HIP(hipGetLastError());
  }
  if (tev) GCM_HIP(hipEventRecord(tev[1], s));
  // the per-bin pass over K x n_bins items, then one block per point
  const int fb = (int)std::max<int64_t>(1, std::min<int64_t>(
      c->fin_blocks, ((int64_t)K * c->n_bins + kFinalizeBlock / kWave - 1) / (kFinalizeBlock / kWave)));
  if (points) {
    PointsArgs pa{c->d_xT, c->d_tabT, c->d_tile_pr, c->d_prs, c->d_prs2, c->d_prc, c->d_prc2};
    FinalizePlan fp{c->d_fin_runs, c->n_fin_runs, c->d_fin_long, c->n_fin_long,
                    c->d_ips,      c->d_ips2,     c->d_ipc,      c->d_ipc2};
    GCM_TRY(launch_finalize_points(pa, fp, fp32, c->n_bins, K,
                                   c->llh_set ? c->d_data : nullptr, c->llh, c->d_hist,
                                   c->d_hist2, c->d_llh_part, s));
  } else if (fp32)
    finalize_bins_kernel<true><<<fb, kFinalizeBlock, 0, s>>>(
        c->d_psum, c->d_psum2, c->d_pcomp, c->d_pcomp2, c->n_partials, c->d_bin_pbeg,
        c->n_bins, K, c->llh_set ? c->d_data : nullptr, c->llh, c->d_hist, c->d_hist2,
        c->d_llh_part);
  else
    finalize_bins_kernel<false><<<fb, kFinalizeBlock, 0, s>>>(
        c->d_psum, c->d_psum2, c->d_pcomp, c->d_pcomp2, c->n_partials, c->d_bin_pbeg,
        c->n_bins, K, c->llh_set ? c->d_data : nullptr, c->llh, c->d_hist, c->d_hist2,
        c->d_llh_part);
  GCM_HIP(hipGetLastError());
  finalize_total_kernel<<<K, kTotalBlock, 0, s>>>(c->d_llh_part, c->n_bins, c->d_penalty,
                                                     c->d_stat, c->d_pen, c->d_total,
                                                     c->d_flags, c->h_res_dev, c->max_points);
  GCM_HIP(hipGetLastError());
  if (tev) GCM_HIP(hipEventRecord(tev[2], s));
  return GCM_OK;
}

int gcm_fill(gcm_context* c, void* stream_ptr) {
  if (!c) return fail(GCM_ERR_ARG, "ctx is NULL");
  if (!c->built) return fail(GCM_ERR_STATE, "gcm_build first");
  if (c->k_points < 1) return fail(GCM_ERR_STATE, "gcm_set_parameters first");
  hipStream_t s = stream_ptr ? (hipStream_t)stream_ptr : c->stream;
  GCM_HIP(hipSetDevice(c->device));
  const int K = c->k_points;
  if (c->graph_mode && !c->timing) {
    // one captured graph per point count: uploads from the pinned staging
    // buffers + the three kernels; replays pick up the staged values
    hipGraphExec_t exec = nullptr;
    for (auto& g : c->graphs)
      if (g.first == K) exec = g.second;
    if (!exec) {
      GCM_TRY(prepare_launch(c));
      GCM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_fill(c, c->stream, K, true, nullptr);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(c->stream, &graph);
      if (rc != GCM_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (e != hipSuccess) return fail(GCM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
      const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) return fail(GCM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
      c->graphs.emplace_back(K, exec);
    }
    GCM_HIP(hipGraphLaunch(exec, s));
    c->theta_dirty = c->tab_dirty = false;
  } else {
    hipEvent_t* tev = nullptr;
    if (c->timing) {
      if (c->t_issued - c->t_done >= gcm_context::kTimingRing) GCM_TRY(harvest_timing(c));
      tev = c->tev[c->t_issued % gcm_context::kTimingRing];
    }
    GCM_TRY(prepare_launch(c));
    GCM_TRY(enqueue_fill(c, s, K, c->theta_dirty || c->tab_dirty, tev));
    c->theta_dirty = c->tab_dirty = false;
    if (tev) c->t_issued++;
  }
  GCM_HIP(hipEventRecord(c->ev_staged, s));
  c->last_stream = s;
  return GCM_OK;
}

int gcm_enable_graph(gcm_context* c, int enable) {
  if (!c) return fail(GCM_ERR_ARG, "ctx is NULL");
  GCM_HIP(hipSetDevice(c->device));
  GCM_HIP(hipStreamSynchronize(c->last_stream ? c->last_stream : c->stream));
  c->drop_graphs();
  c->graph_mode = enable != 0;
  return GCM_OK;
}



int gcm_sync(gcm_context* c) {
  if (!c) return fail(GCM_ERR_ARG, "ctx is NULL");
  GCM_HIP(hipSetDevice(c->device));
  GCM_HIP(hipStreamSynchronize(c->last_stream ? c->last_stream : c->stream));
  if (c->timing) GCM_TRY(harvest_timing(c));
  if (c->built) {
    int* flag = reinterpret_cast<int*>(c->h_res + 3 * (size_t)c->max_points);
    if (*flag) {
      *flag = 0;  // the next fill's prologue clears the device flag
      return fail(GCM_ERR_NUMERIC, "NaN event weight");
    }
  }
  return GCM_OK;
}

int gcm_get_hist(gcm_context* c, int32_t k, double* sums, double* sums2) {
  if (!c) return fail(GCM_ERR_ARG, "ctx is NULL");
  if (!c->built) return fail(GCM_ERR_STATE, "gcm_build first");
  if (k < 0 || k >= std::max(1, c->k_points)) return fail(GCM_ERR_ARG, "point out of range");
  GCM_HIP(hipSetDevice(c->device));
  hipStream_t s = c->last_stream ? c->last_stream : c->stream;
  GCM_HIP(hipStreamSynchronize(s));
  if (sums) GCM_HIP(hipMemcpy(sums, c->d_hist + (size_t)k * c->n_bins, c->n_bins * sizeof(double), hipMemcpyDeviceToHost));
  if (sums2) GCM_HIP(hipMemcpy(sums2, c->d_hist2 + (size_t)k * c->n_bins, c->n_bins * sizeof(double), hipMemcpyDeviceToHost));
  return GCM_OK;
}

namespace {
// GundamUtils::almostEqual (GundamAlmostEqual.h:24-58) for two doubles.
bool almost_equal(double a, double b, double fraction_of_bits_lost) {
  if (!std::isfinite(a) || !std::isfinite(b)) return false;
  double eps = std::numeric_limits<double>::epsilon();
  double eps_a = eps, eps_b = eps;
  if (fraction_of_bits_lost > 0.0) {
    fraction_of_bits_lost = std::min(fraction_of_bits_lost, 0.99);
    eps = std::pow(eps, 1.0 - fraction_of_bits_lost);
  }
  if (std::isnormal(a * eps)) eps_a = std::abs(a * eps);
  if (std::isnormal(b * eps)) eps_b = std::abs(b * eps);
  return std::abs(a - b) < std::max(eps_a, eps_b);
}
}  // namespace

int gcm_validate_hist(gcm_context* c, int32_t k, const double* sums,
                      const double* sqrt_sum_w2, double fraction_of_bits_lost, int verbose,
                      int32_t* n_bad) {
  if (!c || !sums || !sqrt_sum_w2 || !n_bad) return fail(GCM_ERR_ARG, "NULL argument");
  if (!c->built) return fail(GCM_ERR_STATE, "gcm_build first");
  if (k < 0 || k >= std::max(1, c->k_points)) return fail(GCM_ERR_ARG, "point out of range");
  std::vector<double> s((size_t)c->n_bins), s2((size_t)c->n_bins);
  GCM_TRY(gcm_get_hist(c, k, s.data(), s2.data()));
  int32_t bad = 0;
  for (int b = 0; b < c->n_bins; ++b) {
    const double err = std::sqrt(s2[b]);
    if (almost_equal(sums[b], s[b], fraction_of_bits_lost) &&
        almost_equal(sqrt_sum_w2[b], err, fraction_of_bits_lost))
      continue;
    ++bad;
    if (verbose)
      std::fprintf(stderr, "gcm: invalid bin %d -- GPU: %.17g +/- %.17g CPU: %.17g +/- %.17g\n", b,
                   s[b], err, sums[b], sqrt_sum_w2[b]);
  }
  *n_bad = bad;
  return GCM_OK;
}

int gcm_get_weights(gcm_context* c, double* weights) {
  if (!c || !weights) return fail(GCM_ERR_ARG, "NULL argument");
  if (!c->weights_enabled || !c->d_weights) return fail(GCM_ERR_STATE, "gcm_enable_weights first");
  GCM_HIP(hipSetDevice(c->device));
  GCM_HIP(hipStreamSynchronize(c->last_stream ? c->last_stream : c->stream));
  // only binned events are written; keep the caller's values elsewhere
  std::vector<double> tmp((size_t)c->n_events_all);
  GCM_HIP(hipMemcpy(tmp.data(), c->d_weights, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<int32_t> res((size_t)c->n_events);
  GCM_HIP(hipMemcpy(res.data(), c->d_ev_result, res.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int32_t r : res) weights[r] = tmp[r];
  return GCM_OK;
}

int gcm_get_llh(gcm_context* c, double* total, double* stat, double* penalty) {
  if (!c) return fail(GCM_ERR_ARG, "ctx is NULL");
  if (!c->built || c->k_points < 1) return fail(GCM_ERR_STATE, "nothing evaluated yet");
  GCM_HIP(hipSetDevice(c->device));
  GCM_HIP(hipStreamSynchronize(c->last_stream ? c->last_stream : c->stream));
  const size_t K = (size_t)c->k_points;
  const size_t M = (size_t)c->max_points;
  if (total) std::memcpy(total, c->h_res, K * sizeof(double));
  if (stat) std::memcpy(stat, c->h_res + M, K * sizeof(double));
  if (penalty) std::memcpy(penalty, c->h_res + 2 * M, K * sizeof(double));
  return GCM_OK;
}

int gcm_enable_timing(gcm_context* c, int enable) {
  if (!c) return fail(GCM_ERR_ARG, "ctx is NULL");
  GCM_HIP(hipSetDevice(c->device));
  if (enable && !c->tev[0][0]) {
    for (auto& tr : c->tev)
      for (auto& e : tr) GCM_HIP(hipEventCreate(&e));
  }
  if (c->timing) GCM_TRY(harvest_timing(c));
  c->timing = enable != 0;
  c->t_issued = c->t_done = 0;
  c->t_fill_ms = c->t_fin_ms = 0;
  return GCM_OK;
}

int gcm_get_timing(gcm_context* c, double* fill_ms, double* finalize_ms, int64_t* count) {
  if (!c) return fail(GCM_ERR_ARG, "ctx is NULL");
  GCM_HIP(hipSetDevice(c->device));
  GCM_TRY(harvest_timing(c));
  if (fill_ms) *fill_ms = c->t_fill_ms;
  if (finalize_ms) *finalize_ms = c->t_fin_ms;
  if (count) *count = c->t_done;
  return GCM_OK;
}

int gcm_bin_events(int device, int64_t n_events, int32_t n_vars, const double* values,
                   int32_t n_bins, const double* edges, int32_t* out_bin) {
  if (n_events < 0 || n_vars < 1 || n_vars > kMaxBinVars || n_bins < 0 ||
      (n_events && (!values || !out_bin)) || (n_bins && !edges))
    return fail(GCM_ERR_ARG, "invalid binning arguments");
  if (n_events == 0) return GCM_OK;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    return fail(GCM_ERR_HIP, "no such HIP device");
  GCM_HIP(hipSetDevice(device));
  double *dv = nullptr, *de = nullptr;
  int32_t* db = nullptr;
  const size_t nv = (size_t)n_events * n_vars, ne = (size_t)std::max(1, n_bins) * n_vars * 2;
  GCM_HIP(hipMalloc((void**)&dv, nv * sizeof(double)));
  hipError_t e1 = hipMalloc((void**)&de, ne * sizeof(double));
  hipError_t e2 = hipMalloc((void**)&db, (size_t)n_events * sizeof(int32_t));
  int rc = GCM_OK;
  if (e1 != hipSuccess || e2 != hipSuccess) rc = fail(GCM_ERR_NOMEM, "hipMalloc failed");
  if (rc == GCM_OK && hipMemcpy(dv, values, nv * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = fail(GCM_ERR_HIP, "copy failed");
  if (rc == GCM_OK && n_bins && hipMemcpy(de, edges, (size_t)n_bins * n_vars * 2 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = fail(GCM_ERR_HIP, "copy failed");
  if (rc == GCM_OK) {
    const int grid = (int)std::min<int64_t>((n_events + 255) / 256, 65536);
    bin_kernel<<<grid, 256>>>(n_events, n_vars, dv, n_bins, de, db);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = fail(GCM_ERR_HIP, "binning kernel failed");
  }
  if (rc == GCM_OK && hipMemcpy(out_bin, db, (size_t)n_events * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess) rc = fail(GCM_ERR_HIP, "copy failed");
  (void)hipFree(dv);
  (void)hipFree(de);
  (void)hipFree(db);
  return rc;
}

}  // extern "C"
