// gcm_points.hip -- the K-point evaluation path (K >= kPointsMin points per
// launch): one lane per point.  Kept in its own translation unit so that the
// one-point fill kernel of gcm.hip compiles alone (its register allocation
// is at the edge of the 80 VGPRs its occupancy allows).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dial_eval.hpp"
#include "fill_common.hpp"

namespace gcm {

// Lane j's value, broadcast to the wave (j uniform).
__device__ __forceinline__ uint32_t lane_u32(uint32_t v, int j) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, j);
}

// K points per launch with one lane per point (K >= kPointsMin).  The tiles
// stream into the same LDS ring as in fill_kernel; each tile's events are cut
// into kRanges contiguous ranges, one per evaluator wave.  A wave walks its
// events in order and each event's dials in dial-list order: the descriptor
// and the record are read once per dial (LDS broadcast, type and sizes in
// scalar registers, so the wave never diverges on the dial type) and the
// dial is evaluated for 64 points at once, x from the input-major copy xT
// [input][K] and table values from the point-interleaved tables tabT
// [slot][K].  The response product, the cap and the base weight follow
// EventDialCache.cpp:156-171 per point; the bin sums of each (range, bin) run
// are added in event order and stored point-interleaved, [slot][K], one
// partial slot per (tile, range, bin).  Each producer, after its tile has
// landed, writes the tile's event-major index (dial-list position -> the
// type-sorted descriptor) into the buffer, so no evaluator searches.

// Launch bounds: the block's kRanges + NB - 1 waves, and registers for six
// waves per SIMD, so that two blocks share a CU.
template <typename R, bool kComp, int NB, bool kGrid>
__global__ void __launch_bounds__((kRanges + NB - 1) * kWave, 6)
    fill_points_kernel(FillArgs a, PointsArgs pa) {
  constexpr int P = NB - 1;  // producer waves
  constexpr int W = kRanges;  // evaluator waves
  extern __shared__ __align__(16) unsigned char lds[];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int lane = threadIdx.x % kWave;
  const int K = a.k_points;
  const int n_chunks = (K + kWave - 1) / kWave;
  // LDS: [class records][NB tile buffers: records, descriptors, events,
  // header, event-major index]
  ClassRec* s_cls = reinterpret_cast<ClassRec*>(lds);
  const size_t cls_pad = (kClassBytes + 1023) / 1024 * 1024;
  const size_t ev_off = (size_t)a.rec_bytes + (size_t)a.max_dials * 8;
  const size_t hdr_off = ev_off + kMaxTileEvents * sizeof(EventRec);
  const size_t inv_off = hdr_off + kHeaderBytes;
  const size_t buf_bytes = inv_off + ((size_t)a.max_dials * 2 + 15) / 16 * 16;
  unsigned char* const ring = lds + cls_pad;
  const unsigned char* rec = reinterpret_cast<const unsigned char*>(a.rec);
  const unsigned char* desc = reinterpret_cast<const unsigned char*>(a.desc);
  const int64_t G = gridDim.x;
  const int64_t t0 = blockIdx.x;
  if (threadIdx.x < kInputClasses) s_cls[threadIdx.x] = a.cls[threadIdx.x];

  // the producer that waited for a tile writes its event-major index
  auto index_tile = [&](unsigned char* buf) {
    const TileHdr& th = *reinterpret_cast<const TileHdr*>(buf + hdr_off);
    const uint64_t* sd = reinterpret_cast<const uint64_t*>(buf + a.rec_bytes);
    uint16_t* inv = reinterpret_cast<uint16_t*>(buf + inv_off);
    for (int q = lane; q < th.n_dials; q += kWave)
      inv[(int)((sd[q] >> 32) & 0xffff)] = (uint16_t)q;
  };
  TileHdr hq{};
  if (wave >= W) {
    const int p = wave - W;
    const int64_t tp = t0 + p * G;
    if (tp < a.n_tiles) {
      dma_tile<R>(a.tiles[tp], rec, desc, a.ev, ring + (size_t)p * buf_bytes,
                  a.rec_bytes, a.max_dials, lane);
      dma_header(a.tiles + tp, ring + (size_t)p * buf_bytes, hdr_off, lane);
    }
    if (tp + P * G < a.n_tiles) hq = a.tiles[tp + P * G];
    if (p == 0) {
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      if (tp < a.n_tiles) index_tile(ring);
    }
  }
  lds_barrier();

  int64_t i = 0;  // ordinal of the tile within this block
  for (int64_t t = t0; t < a.n_tiles; t += G, ++i) {
    unsigned char* cur = ring + (size_t)(i % NB) * buf_bytes;
    for (int c = 0; c < n_chunks; ++c) {
      if (wave >= W) {
        const int me = wave - W;
        if (c == 0 && me == (int)(i % P)) {
          const int64_t tn = t + P * G;
          if (tn < a.n_tiles) {
            unsigned char* nbuf = ring + (size_t)((i + P) % NB) * buf_bytes;
            dma_tile<R>(hq, rec, desc, a.ev, nbuf, a.rec_bytes, a.max_dials, lane);
            dma_header(a.tiles + tn, nbuf, hdr_off, lane);
            if (tn + P * G < a.n_tiles) hq = a.tiles[tn + P * G];
          }
        }
        if (c == n_chunks - 1 && me == (int)((i + 1) % P)) {
          asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
          if (t + G < a.n_tiles) index_tile(ring + (size_t)((i + 1) % NB) * buf_bytes);
        }
        lds_barrier();
        continue;
      }
      // ---- evaluator wave `wave`: range `wave` of the tile, points of chunk c
      const TileHdr h = *reinterpret_cast<const TileHdr*>(cur + hdr_off);
      const R* s_rec = reinterpret_cast<const R*>(cur);
      const uint64_t* s_desc = reinterpret_cast<const uint64_t*>(cur + a.rec_bytes);
      const EventRec* s_ev = reinterpret_cast<const EventRec*>(cur + ev_off);
      const uint16_t* s_inv = reinterpret_cast<const uint16_t*>(cur + inv_off);
      const int n = h.n_events;
      const int e0 = wave * n / W, e1 = (wave + 1) * n / W;
      // segment heads over the tile: range starts and bin changes
      bool head = false;
      if (lane < n) {
        bool start = false;
#pragma unroll
        for (int r = 0; r < W; ++r) start = start || (r * n / W == lane);
        head = start || s_ev[lane].bin != s_ev[lane - 1].bin;
      }
      const unsigned long long heads = __ballot(head);
      const uint32_t slot0 = pa.tile_pr[t];
      const int k = c * kWave + lane;
      const int kk = k < K ? k : K - 1;  // lanes past K compute on point K-1, store nothing
      double s = 0.0, s2 = 0.0, cs = 0.0, cs2 = 0.0;
      int prev_bin = INT_MAX;
      int64_t slot = -1;
      // One event done: cap, base weight, the point-0 weight copy, the NaN
      // flag, then the (range, bin) run sums in event order.
      auto finish_event = [&](int e, const EventRec& er, double temp) {
        if (temp > a.cap) temp = a.cap;
        double w = er.base;
        w *= temp;
        if (c == 0 && lane == 0 && a.weights_out)
          a.weights_out[a.ev_result[(int64_t)h.first_event + e]] = w;
        if (k < K && isnan(w)) atomicOr(a.flags, 1);
        const double w2 = w * w;
        if (e == e0 || er.bin != prev_bin) {
          if (slot >= 0 && k < K) {
            pa.prs[slot * K + k] = s;
            pa.prs2[slot * K + k] = s2;
            if (kComp) {
              pa.prc[slot * K + k] = cs;
              pa.prc2[slot * K + k] = cs2;
            }
          }
          slot = (int64_t)slot0 + __popcll(heads & ((2ull << e) - 1ull)) - 1;
          s = w;
          s2 = w2;
          cs = cs2 = 0.0;
          prev_bin = er.bin;
        } else if (kComp) {
          double tt, ee;
          two_sum(s, w, tt, ee);
          s = tt;
          cs += ee;
          two_sum(s2, w2, tt, ee);
          s2 = tt;
          cs2 += ee;
        } else {
          s = s + w;
          s2 = s2 + w2;
        }
      };
      if (e0 < e1) {
        // the range's dials are the event-major slots [S0, S1), taken 64 at
        // a time: lane j reads the descriptor of slot sb + j, and the walk
        // over the batch broadcasts it with readlane; the points' x of the
        // next dial is loaded while the current one is evaluated
        const int S0 = (int)(s_ev[e0].dials & 0xffff);
        const EventRec last = s_ev[e1 - 1];
        const int S1 = (int)(last.dials & 0xffff) + (int)(last.dials >> 16);
        int e = e0;
        EventRec er = s_ev[e];
        int ev_end = (int)(er.dials & 0xffff) + (int)(er.dials >> 16);
        double temp = 1;
        for (int sb = S0; sb < S1; sb += kWave) {
          const int nbat = S1 - sb < kWave ? S1 - sb : kWave;
          const uint64_t dv = s_desc[s_inv[sb + (lane < nbat ? lane : 0)]];
          const uint32_t vlo = (uint32_t)dv, vhi = (uint32_t)(dv >> 32);
          double xn = pa.xT[(int64_t)(lane_u32(vlo, 0) & 0xffff) * K + kk];
          for (int j = 0; j < nbat; ++j) {
            const int sl = sb + j;
            while (sl >= ev_end) {  // events that end here (dialless ones too)
              finish_event(e, er, temp);
              temp = 1;
              er = s_ev[++e];
              ev_end = (int)(er.dials & 0xffff) + (int)(er.dials >> 16);
            }
            const uint32_t dlo = lane_u32(vlo, j), dhi = lane_u32(vhi, j);
            const double x = xn;
            if (j + 1 < nbat) xn = pa.xT[(int64_t)(lane_u32(vlo, j + 1) & 0xffff) * K + kk];
            const int type = (int)((dlo >> 16) & 0xf);
            const int nk = (int)((dlo >> 20) & 0xff);
            const int off = (int)(dhi >> 16);
            // the supervisor does not depend on the point: the class record
            // (a wave-uniform LDS read), or for the escape class the input's
            // own record
            const int cl = (int)((dlo >> 28) & 0xf);
            double lo, hi;
            int flags;
            if (cl != kEscapeClass) {
              lo = s_cls[cl].lo;
              hi = s_cls[cl].hi;
              flags = __builtin_amdgcn_readfirstlane(s_cls[cl].flags);
            } else {
              const InputRec* ir = a.xin + (dlo & 0xffff);
              lo = ir->lo;
              hi = ir->hi;
              flags = __builtin_amdgcn_readfirstlane(ir->flags);
            }
            double r;
            if (kGrid && (type == GCM_DIAL_BILINEAR || type == GCM_DIAL_BICUBIC)) {
              const int in2 = __builtin_amdgcn_readfirstlane((int)s_rec[off]);
              const double x2 = pa.xT[(int64_t)in2 * K + kk];
              r = eval_dial2d<R>(type, s_rec + off + 1, x, x2, flags);
            } else {
              r = eval_dial<R, true>(type, nk, s_rec + off, x, flags, pa.tabT + kk, K);
            }
            // DialResponseSupervisor::process (DialResponseSupervisor.cpp:10-18)
            if (r < lo) r = lo;
            else if (r > hi) r = hi;
            temp *= r;
          }
        }
        for (; e < e1; ++e) {  // the last event with dials, and dialless ones after it
          finish_event(e, er, temp);
          temp = 1;
          if (e + 1 < e1) er = s_ev[e + 1];
        }
      }
      if (slot >= 0 && k < K) {
        pa.prs[slot * K + k] = s;
        pa.prs2[slot * K + k] = s2;
        if (kComp) {
          pa.prc[slot * K + k] = cs;
          pa.prc2[slot * K + k] = cs2;
        }
      }
      lds_barrier();
    }
  }
}

// Rows [p0, p1) of the [row][K] arrays v, v2 (and their compensations vc,
// vc2 with kComp) added for point k in row order onto (s, s2, c, c2): eight
// rows' loads in flight, coalesced over the points, Knuth two-sums with
// kComp.
template <bool kComp>
__device__ __forceinline__ void add_rows(const double* __restrict__ v,
                                         const double* __restrict__ v2,
                                         const double* __restrict__ vc,
                                         const double* __restrict__ vc2, uint32_t p0,
                                         uint32_t p1, int K, int k, double& s, double& s2,
                                         double& c, double& c2) {
  for (uint32_t pb = p0; pb < p1; pb += 8) {
    double x[8], x2[8], xc[8], xc2[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const bool ok = pb + u < p1;
      const size_t q = (size_t)(ok ? pb + u : p0) * K + k;
      x[u] = ok ? v[q] : 0.0;
      x2[u] = ok ? v2[q] : 0.0;
      if (kComp) {
        xc[u] = ok ? vc[q] : 0.0;
        xc2[u] = ok ? vc2[q] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (pb + u >= p1) break;
      if (kComp) {
        double t, e;
        two_sum(s, x[u], t, e);
        s = t;
        c += e + xc[u];
        two_sum(s2, x2[u], t, e);
        s2 = t;
        c2 += e + xc2[u];
      } else {
        s += x[u];
        s2 += x2[u];
      }
    }
  }
}

// A bin's sums for point k, and its LLH term.
template <bool kComp>
__device__ __forceinline__ void finish_bin(int b, int k, int n_bins, double s, double s2,
                                           double c, double c2, const double* data,
                                           const LlhParams& cfg, double* hist, double* hist2,
                                           double* llh_terms) {
  if (kComp) {
    s = s + c;
    s2 = s2 + c2;
  }
  hist[(size_t)k * n_bins + b] = s;
  hist2[(size_t)k * n_bins + b] = s2;
  llh_terms[(size_t)k * n_bins + b] = data ? llh_bin(cfg, b, data[b], s, sqrt(s2)) : 0.0;
}

// Level 1 of the per-bin pass of a K-point launch: one wave per (run, chunk
// of 64 points), one lane per point; the run's range slots are added in slot
// order; a bin of one run is finished here, a longer bin's run sum is stored
// for level 2.
template <bool kComp>
__global__ void __launch_bounds__(kFinalizeBlock)
    finalize_runs_kernel(PointsArgs pa, FinalizePlan fp, int n_bins, int k_points,
                         const double* __restrict__ data, LlhParams cfg,
                         double* __restrict__ hist, double* __restrict__ hist2,
                         double* __restrict__ llh_terms) {
  const int lane = threadIdx.x % kWave;
  const int K = k_points;
  const int n_chunks = (K + kWave - 1) / kWave;
  const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  const int64_t items = (int64_t)n_chunks * fp.n_runs;
  for (int64_t it = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
       it < items; it += n_waves) {
    const int64_t r = it % fp.n_runs;
    const int k = (int)(it / fp.n_runs) * kWave + lane;
    if (k >= K) continue;
    const uint4 run = fp.runs[r];
    double s = 0.0, s2 = 0.0, c = 0.0, c2 = 0.0;
    add_rows<kComp>(pa.prs, pa.prs2, pa.prc, pa.prc2, run.x, run.x + run.y, K, k, s, s2, c, c2);
    if (run.w == 0) {
      finish_bin<kComp>((int)run.z, k, n_bins, s, s2, c, c2, data, cfg, hist, hist2, llh_terms);
    } else {
      const size_t q = (size_t)(run.w - 1) * K + k;
      fp.ips[q] = s;
      fp.ips2[q] = s2;
      if (kComp) {
        fp.ipc[q] = c;
        fp.ipc2[q] = c2;
      }
    }
  }
}

// Level 2: one wave per (long bin, chunk of 64 points); the bin's run sums
// are added in run order, then the bin is finished.
template <bool kComp>
__global__ void __launch_bounds__(kFinalizeBlock)
    finalize_long_bins_kernel(FinalizePlan fp, int n_bins, int k_points,
                              const double* __restrict__ data, LlhParams cfg,
                              double* __restrict__ hist, double* __restrict__ hist2,
                              double* __restrict__ llh_terms) {
  const int lane = threadIdx.x % kWave;
  const int K = k_points;
  const int n_chunks = (K + kWave - 1) / kWave;
  const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  const int64_t items = (int64_t)n_chunks * fp.n_long;
  for (int64_t it = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
       it < items; it += n_waves) {
    const int64_t j = it % fp.n_long;
    const int k = (int)(it / fp.n_long) * kWave + lane;
    if (k >= K) continue;
    const uint4 lb = fp.long_bins[j];
    double s = 0.0, s2 = 0.0, c = 0.0, c2 = 0.0;
    add_rows<kComp>(fp.ips, fp.ips2, fp.ipc, fp.ipc2, lb.y, lb.y + lb.z, K, k, s, s2, c, c2);
    finish_bin<kComp>((int)lb.x, k, n_bins, s, s2, c, c2, data, cfg, hist, hist2, llh_terms);
  }
}

// Point-interleaved copy of the lookup tables for fill_points_kernel:
// tabT[s][k] = tab[k][s].
__global__ void transpose_tables_kernel(const double* __restrict__ tab, int64_t len, int K,
                                        double* __restrict__ tabT) {
  const int64_t n = len * K;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t sl = i / K, k = i % K;
    tabT[i] = tab[k * len + sl];
  }
}

template <bool kGrid>
static const void* points_fn(bool fp32, int nb) {
  if (fp32)
    return nb == 3 ? (const void*)fill_points_kernel<float, true, 3, kGrid>
                   : (const void*)fill_points_kernel<float, true, 2, kGrid>;
  return nb == 3 ? (const void*)fill_points_kernel<double, false, 3, kGrid>
                 : (const void*)fill_points_kernel<double, false, 2, kGrid>;
}

size_t points_lds_bytes(const FillArgs& a, int nb) {
  const size_t buf = (size_t)a.rec_bytes + (size_t)a.max_dials * 8 +
                     kMaxTileEvents * sizeof(EventRec) + kHeaderBytes +
                     ((size_t)a.max_dials * 2 + 15) / 16 * 16;
  return (kClassBytes + 1023) / 1024 * 1024 + (size_t)nb * buf;
}

int points_set_lds_attrs() {
  for (int v = 0; v < 8; ++v) {
    const bool fp32 = v & 1, grid = v & 4;
    const int nb = (v & 2) ? 3 : 2;
    const void* f = grid ? points_fn<true>(fp32, nb) : points_fn<false>(fp32, nb);
    GCM_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxDynamicLds));
  }
  return GCM_OK;
}

int launch_points(const FillArgs& a, const PointsArgs& pa, bool fp32, bool grid, int nb,
                  int n_cu, size_t lds, hipStream_t s) {
  const int threads = (kRanges + nb - 1) * kWave;
  const void* fn = grid ? points_fn<true>(fp32, nb) : points_fn<false>(fp32, nb);
  // a persistent grid: as many blocks as are resident at once (registers
  // and LDS), never more than there are tiles
  int per_cu = 0;
  GCM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds));
  const int blocks = (int)std::max<int64_t>(
      1, std::min<int64_t>(a.n_tiles, (int64_t)std::max(1, per_cu) * n_cu));
  FillArgs aa = a;
  PointsArgs pp = pa;
  void* args[] = {&aa, &pp};
  GCM_HIP(hipLaunchKernel(fn, dim3(blocks), dim3(threads), args, lds, s));
  GCM_HIP(hipGetLastError());
  return GCM_OK;
}

int launch_finalize_points(const PointsArgs& pa, const FinalizePlan& fp, bool fp32, int n_bins,
                           int k_points, const double* data, const LlhParams& cfg,
                           double* hist, double* hist2, double* llh_terms, hipStream_t s) {
  const int64_t n_chunks = (k_points + kWave - 1) / kWave;
  auto blocks = [](int64_t waves) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(
        kFinalizeMaxBlocks, (waves + kFinalizeBlock / kWave - 1) / (kFinalizeBlock / kWave)));
  };
  const int b1 = blocks(n_chunks * fp.n_runs);
  if (fp32)
    finalize_runs_kernel<true><<<b1, kFinalizeBlock, 0, s>>>(pa, fp, n_bins, k_points, data, cfg,
                                                             hist, hist2, llh_terms);
  else
    finalize_runs_kernel<false><<<b1, kFinalizeBlock, 0, s>>>(pa, fp, n_bins, k_points, data,
                                                              cfg, hist, hist2, llh_terms);
  GCM_HIP(hipGetLastError());
  if (fp.n_long > 0) {
    const int b2 = blocks(n_chunks * fp.n_long);
    if (fp32)
      finalize_long_bins_kernel<true><<<b2, kFinalizeBlock, 0, s>>>(fp, n_bins, k_points, data,
                                                                    cfg, hist, hist2, llh_terms);
    else
      finalize_long_bins_kernel<false><<<b2, kFinalizeBlock, 0, s>>>(fp, n_bins, k_points, data,
                                                                     cfg, hist, hist2, llh_terms);
    GCM_HIP(hipGetLastError());
  }
  return GCM_OK;
}

int launch_transpose_tables(const double* tab, int64_t len, int k_points, double* tabT,
                            hipStream_t s) {
  const int64_t n = len * k_points;
  transpose_tables_kernel<<<(int)std::min<int64_t>((n + 255) / 256, 4096), 256, 0, s>>>(
      tab, len, k_points, tabT);
  GCM_HIP(hipGetLastError());
  return GCM_OK;
}

}  // namespace gcm
