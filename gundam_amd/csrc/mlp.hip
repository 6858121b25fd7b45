// mlp.hip -- the learned surrogate: an MLP n_in -> H -> H -> n_out (GELU)
// that maps a parameter point to all predicted log bin contents, trained on
// the GPU with hand-written BF16 MFMA kernels (FP32 accumulation) and Adam.
//
// GEMM: C[M x N] = A[M x K] . B[K x N] (+ bias, optional GELU), each operand
// stored either with K contiguous or with M (N) contiguous, read with 16-byte
// loads along its contiguous dimension, converted to BF16 and staged in LDS as
// [row][k] tiles, and multiplied with v_mfma_f32_32x32x16_bf16 (gfx950).
// Lane layout of that instruction (checked on the device): lane l holds
// A[l%32][8(l/32)+j] and B[8(l/32)+j][l%32] (j < 8); accumulator r of lane l
// is C[8(r/4) + 4(l/32) + r%4][l%32].
//
// Parameters, flat FP32 master copy: W1[H][n_in] b1[H] W2[H][H] b2[H]
// W3[n_out][H] b3[n_out]; the GEMMs read BF16 copies of the weights (W1 rows
// padded to a multiple of 8 elements).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/gcm.h"

namespace gcm {
int fail(int code, const std::string& msg);  // gcm.hip
}

namespace {

using gcm::fail;

#define MLP_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(GCM_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Block tile BM x BN (64 x 64, or 128 x 128 for products with plenty of
// work) with a K step of 64; the waves form a 2 x WN grid (WN = 2 for the
// 64 x 64 tile, 4 for the 128 x 128 one), each owning a (BM/2) x (BN/WN)
// sub-tile of 32 x 32 accumulators.  Operand tiles are
// double-buffered in LDS as [row][k] BF16 with a padded row of 72 elements
// (144 B, 16-B aligned for the 128-bit operand reads); the K step after next
// is loaded into registers while the current one is multiplied, one barrier
// per step.  A 128 x 128 tile reads each operand byte once per 128 outputs
// instead of 64, halving the operand traffic from L2 per FLOP.
constexpr int TM = 64, TN = 64, TK = 64;
constexpr int kWaveMlp = 64;
constexpr int kBig = 128;  // the large tile's side
// The 128 x 128 tile runs in blocks of eight waves (2 x 4, each 64 x 32):
// half the registers per wave of a 2 x 2 layout, so four waves per SIMD
// instead of two hide each other's waits (measured 10% per training step at
// 4096 rows); the 64 x 64 tile runs in four (2 x 2, each 32 x 32).
constexpr int kBigWN = 4;
constexpr int LDK = TK + 8;
constexpr int kMaxSplit = 16;
constexpr int kMinSplitSteps = 16;  // split only long reductions (K >= 1024)

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

struct GemmArgs {
  int M, N, K;
  // element (m, k) of A: A[m * lda + k] when A is K-contiguous, else A[k * lda + m]
  const void* A;
  int64_t lda;
  // element (k, n) of B: B[n * ldb + k] when B is K-contiguous, else B[k * ldb + n]
  const void* B;
  int64_t ldb;
  int vec;    // both operands meet the alignment of the unchecked vector loads
  int yfast;  // tile order: M tiles fastest (B the larger operand), else N tiles
  float* C;   // C[m * ldc + n] (FP32), or
  __bf16* Cb;  // the same in BF16, when set (an operand of later GEMMs only)
  int64_t ldc;
  // split-K launches (gridDim.z > 1): split z writes its partial product to
  // ws + z * split_stride (dense, ld N); a reduction launch adds them in
  // split order and applies the epilogue
  float* ws;
  int64_t split_stride;
  int k_chunk;        // K range of one split, a multiple of TK
  // epilogue: + bias[n]; GELU with the pre-activation kept in Z; or, with gz,
  // the backward pass through a GELU (C *= gelu'(gz), C dense)
  const float* bias;  // [N] or null
  float* Z;           // pre-activation copy (GELU only) or null
  int gelu;
  const float* gz;
  // The loss head fused into the output layer (unsplit launches, head set):
  // for y = C + bias, dY = 2 (y - T) scale is stored in BF16 to dyb (ld N)
  // instead of C, with T's row tidx[row] (or row) at tgt + row * ldt; each
  // block writes its column sums of dY to hcol[tile row][n] and its sum of
  // squared errors to hsq[tile row][tile column], each in a fixed order.
  int head;
  const float* tgt;
  const int32_t* tidx;
  int64_t ldt;
  __bf16* dyb;
  float scale;
  float* hcol;
  float* hsq;
  // Adam fused into the epilogue (adam set; no bias, GELU or gz; C dense,
  // ldc == N): output element o = row N + col is the gradient of parameter
  // ap[o], whose moments am1[o], am2[o] and BF16 copy apb[row ldpb + col]
  // are updated in place with it (adam_one), and nothing is stored to C.
  int adam;
  float* ap;
  float* am1;
  float* am2;
  __bf16* apb;
  int64_t ldpb;
  float lr, b1, b2, eps, c1, c2;
};

// One ROWS x 64 (k) operand tile (ROWS = 64 or 128) held in registers
// between its global load and its LDS store, for an operand that is
// K-contiguous (KC) or row-contiguous, FP32 or BF16 (BF).  load() is
// unchecked along K (full K steps of an aligned operand); load_checked()
// reads rows >= R and k >= kend as zero.
//  KC: thread t holds rows t/8 + 32 i (i < ROWS/32), k = 8 (t%8) .. + 7, read
//      as 16-byte vectors along k, stored as one 16-byte LDS write per row.
//  row-contiguous: thread t holds rows 2 (t%32) + e + 64 j (e < 2,
//      j < ROWS/64), k = 8 (t/32) .. + 7, read as 2-element vectors along
//      the rows (one per k), stored transposed as one 16-byte LDS write per
//      row.
//  With NT = 512 threads the rows are spread over twice the threads: KC rows
//  t/8 + 64 i, row-contiguous rows 2 (t%32) + e + 64 (2 j + t/256) with
//  k = 8 ((t/32) % 8).
template <bool KC, bool BF, int ROWS, int NT = 256>
struct Stage {
  using E = typename std::conditional<BF, __bf16, float>::type;
  static constexpr int RPP = NT / 8;            // KC: rows per pass of the block
  static constexpr int RB = NT / 256;           // row-contiguous: 64-row blocks per pass
  static constexpr int NI = ROWS / RPP;         // KC: rows per thread
  static constexpr int NJ = ROWS / (64 * RB);   // row-contiguous: row pairs per thread
  static constexpr int NS = KC ? NI : 2 * NJ;   // rows per thread
  static_assert(NI >= 1 && NJ >= 1, "tile too small for the block");
  __device__ __forceinline__ static int kc_row(int tid, int i) { return tid / 8 + RPP * i; }
  __device__ __forceinline__ static int rc_row(int tid, int j) {
    return 2 * (tid % 32) + 64 * (j * RB + tid / 256);
  }
  __device__ __forceinline__ static int rc_k(int tid) { return 8 * ((tid / 32) % 8); }
  // raw loaded values: converted to BF16 (FP32 operands) only when stored,
  // so that the loads stay in flight across the multiply of the previous step
  float f[NS][8];      // FP32: [row slot][k]
  bf16x8 hk[NI];       // BF16, K-contiguous: [row slot] -> 8 k
  bf16x2 hr[NJ][8];    // BF16, row-contiguous: [pair][k] -> 2 rows

  // Rows past R are read as row R - 1 (K-contiguous) or as rows R - 2, R - 1
  // (row-contiguous, R even): they only feed output rows that are not stored.
  __device__ __forceinline__ void load(const E* base, int64_t ld, int R, int r0, int k0,
                                       int tid) {
    if constexpr (KC) {
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const int r = min(r0 + kc_row(tid, i), R - 1);
        const E* p = base + (int64_t)r * ld + k0 + (tid % 8) * 8;
        if constexpr (BF) {
          hk[i] = *reinterpret_cast<const bf16x8*>(p);
        } else {
          const float4 u = *reinterpret_cast<const float4*>(p);
          const float4 w = *reinterpret_cast<const float4*>(p + 4);
          f[i][0] = u.x; f[i][1] = u.y; f[i][2] = u.z; f[i][3] = u.w;
          f[i][4] = w.x; f[i][5] = w.y; f[i][6] = w.z; f[i][7] = w.w;
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int r = min(r0 + rc_row(tid, j), R - 2);
        const E* p = base + (int64_t)(k0 + rc_k(tid)) * ld + r;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          if constexpr (BF) {
            hr[j][q] = *reinterpret_cast<const bf16x2*>(p + q * ld);
          } else {
            const float2 v = *reinterpret_cast<const float2*>(p + q * ld);
            f[2 * j][q] = v.x;
            f[2 * j + 1][q] = v.y;
          }
        }
      }
    }
  }

  // Branch-free: every element is loaded (an out-of-range one from element 0
  // of the operand, which exists) and replaced by zero afterwards, so the
  // loads are all in flight together.
  __device__ __forceinline__ void load_checked(const E* base, int64_t ld, int R, int r0, int kend,
                                               int k0, int tid) {
    const E zero = (E)0.0f;
    if constexpr (KC) {
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const int r = r0 + kc_row(tid, i), k = k0 + (tid % 8) * 8;
        E v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const bool ok = r < R && k + j < kend;
          v[j] = base[ok ? (int64_t)r * ld + k + j : 0];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const bool ok = r < R && k + j < kend;
          if constexpr (BF) hk[i][j] = ok ? v[j] : zero;
          else f[i][j] = ok ? v[j] : zero;
        }
      }
    } else {
#pragma unroll
      for (int jj = 0; jj < NJ; ++jj) {
        const int r = r0 + rc_row(tid, jj), k = k0 + rc_k(tid);
        E v[8][2];
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int e = 0; e < 2; ++e) {
            const bool ok = r + e < R && k + q < kend;
            v[q][e] = base[ok ? (int64_t)(k + q) * ld + r + e : 0];
          }
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int e = 0; e < 2; ++e) {
            const bool ok = r + e < R && k + q < kend;
            if constexpr (BF) hr[jj][q][e] = ok ? v[q][e] : zero;
            else f[2 * jj + e][q] = ok ? v[q][e] : zero;
          }
      }
    }
  }

  __device__ __forceinline__ void store(__bf16* S, int tid) const {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      bf16x8 v;
      if constexpr (BF && KC) {
        v = hk[i];
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          if constexpr (BF) v[j] = hr[i / 2][j][i % 2];
          else v[j] = (__bf16)f[i][j];
        }
      }
      const int r = KC ? kc_row(tid, i) : rc_row(tid, i / 2) + (i % 2);
      const int k = KC ? (tid % 8) * 8 : rc_k(tid);
      *reinterpret_cast<bf16x8*>(S + r * LDK + k) = v;
    }
  }
};

// One Adam step (Kingma & Ba 2015) of parameter p with gradient gi and
// moments m1, m2 (updated in place); c1, c2 are the bias corrections.
__device__ __forceinline__ float adam_one(float p, float gi, float& m1, float& m2, float lr,
                                          float b1, float b2, float eps, float c1, float c2) {
  const float a = b1 * m1 + (1.0f - b1) * gi;
  const float b = b2 * m2 + (1.0f - b2) * gi * gi;
  m1 = a;
  m2 = b;
  return p - lr * (a / c1) / (sqrtf(b / c2) + eps);
}

__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// The GEMM epilogue of element (row, col) of C: bias (bcol = bias[col],
// loaded by the caller), GELU (the pre-activation kept in Z) or the GELU
// derivative, then the store.
// With gz, gzv is gz's element (row, col), loaded by the caller.
__device__ __forceinline__ void finish_store(const GemmArgs& g, int row, int col, float v,
                                             float bcol, float gzv) {
  const int64_t o = (int64_t)row * g.ldc + col;
  if (g.bias) v += bcol;
  if (g.gelu) {
    g.Z[o] = v;
    v = gelu(v);
  }
  if (g.gz) v *= gelu_grad(gzv);
  if (g.Cb) g.Cb[o] = (__bf16)v;
  else g.C[o] = v;
}

// One K step (64) of the wave's (BM/2) x (BN/WN) sub-tile from the LDS tiles
// (the waves form a 2 x WN grid).
template <int BM, int BN, int WN>
__device__ __forceinline__ void mfma_step(const __bf16* As, const __bf16* Bs,
                                          f32x16 (&acc)[BM / 64][BN / WN / 32], int wm, int wn,
                                          int lane) {
  constexpr int MI = BM / 64, NJ = BN / WN / 32;
  const __bf16* as = As + (wm * (BM / 2) + lane % 32) * LDK + (lane / 32) * 8;
  const __bf16* bs = Bs + (wn * (BN / WN) + lane % 32) * LDK + (lane / 32) * 8;
#pragma unroll
  for (int kk = 0; kk < TK; kk += 16) {
    bf16x8 a[MI], b[NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i) a[i] = *reinterpret_cast<const bf16x8*>(as + i * 32 * LDK + kk);
#pragma unroll
    for (int j = 0; j < NJ; ++j) b[j] = *reinterpret_cast<const bf16x8*>(bs + j * 32 * LDK + kk);
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
      for (int j = 0; j < NJ; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
}

// LDS bytes of a BM x BN tile's double-buffered operands.
constexpr int gemm_lds_bytes(int BM, int BN) { return 2 * (BM + BN) * LDK * 2; }

// Grid (ceil(N / BN), ceil(M / BM), splits); split z covers
// k in [z k_chunk, min(K, (z + 1) k_chunk)).  Aligned operands run the full
// K steps pipelined with unchecked loads (edge tiles by clamping the rows);
// the K tail and unaligned operands take checked, unpipelined steps.
template <int BM, int BN, bool AKC, bool ABF, bool BKC, bool BBF, int WN = 2>
__global__ void __launch_bounds__(128 * WN) gemm_bf16_kernel(GemmArgs g) {
  constexpr int NT = 128 * WN;  // threads: 2 x WN waves
  extern __shared__ __align__(16) unsigned char gemm_lds[];
  __bf16* const As[2] = {reinterpret_cast<__bf16*>(gemm_lds),
                         reinterpret_cast<__bf16*>(gemm_lds) + BM * LDK};
  __bf16* const Bs[2] = {reinterpret_cast<__bf16*>(gemm_lds) + 2 * BM * LDK,
                         reinterpret_cast<__bf16*>(gemm_lds) + 2 * BM * LDK + BN * LDK};
  constexpr int MI = BM / 64, NJ = BN / WN / 32;
  using SA = Stage<AKC, ABF, BM, NT>;
  using SB = Stage<BKC, BBF, BN, NT>;
  const auto* A = reinterpret_cast<const typename SA::E*>(g.A);
  const auto* B = reinterpret_cast<const typename SB::E*>(g.B);
  const int tid = threadIdx.x;
  const int lane = tid % 64, wave = tid / 64;
  const int wm = wave / WN, wn = wave % WN;
  // Workgroups are dispatched to the 8 XCDs round-robin in linear order; each
  // XCD gets a contiguous range of tiles, ordered so that the tiles sharing
  // the larger operand's tiles share that XCD's L2.
  const int gx = gridDim.x, gy = gridDim.y;
  const int nblk = gx * gy * gridDim.z;
  const int lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const int q = nblk / 8, rem = nblk % 8, xcd = lin % 8;
  const int t = xcd * q + min(xcd, rem) + lin / 8;
  int bx, by;
  if (g.yfast) {
    by = t % gy;
    bx = (t / gy) % gx;
  } else {
    bx = t % gx;
    by = (t / gx) % gy;
  }
  const int bz = t / (gx * gy);
  const int m0 = by * BM, n0 = bx * BN;
  const int kb = bz * g.k_chunk;
  const int ke = min(g.K, kb + g.k_chunk);
  f32x16 acc[MI][NJ];
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[i][j] = f32x16{0};
  SA sa, sa2;
  SB sb, sb2;
  const int nfull = (g.vec && ke > kb) ? (ke - kb) / TK : 0;
  if (nfull > 0) {
    // Two register sets: step s + 2 is loaded while step s is multiplied and
    // step s + 1 (loaded one step earlier) is written to LDS.  Loads past the
    // last full step re-read that step (in bounds, unused), so no load is
    // conditional.
    auto kstep = [&](int st) { return kb + min(st, nfull - 1) * TK; };
    sa.load(A, g.lda, g.M, m0, kstep(0), tid);
    sb.load(B, g.ldb, g.N, n0, kstep(0), tid);
    sa2.load(A, g.lda, g.M, m0, kstep(1), tid);
    sb2.load(B, g.ldb, g.N, n0, kstep(1), tid);
    sa.store(As[0], tid);
    sb.store(Bs[0], tid);
    __syncthreads();
    int st = 0;
    for (; st + 2 < nfull; st += 2) {
      // LDS 0: step st; set 2: step st + 1
      sa.load(A, g.lda, g.M, m0, kstep(st + 2), tid);
      sb.load(B, g.ldb, g.N, n0, kstep(st + 2), tid);
      __builtin_amdgcn_sched_barrier(0);  // issue the loads before the multiply
      mfma_step<BM, BN, WN>(As[0], Bs[0], acc, wm, wn, lane);
      sa2.store(As[1], tid);
      sb2.store(Bs[1], tid);
      __syncthreads();
      // LDS 1: step st + 1; set 1: step st + 2
      sa2.load(A, g.lda, g.M, m0, kstep(st + 3), tid);
      sb2.load(B, g.ldb, g.N, n0, kstep(st + 3), tid);
      __builtin_amdgcn_sched_barrier(0);
      mfma_step<BM, BN, WN>(As[1], Bs[1], acc, wm, wn, lane);
      sa.store(As[0], tid);
      sb.store(Bs[0], tid);
      __syncthreads();
    }
    // one or two full steps left: LDS 0 holds step st, set 2 step st + 1
    mfma_step<BM, BN, WN>(As[0], Bs[0], acc, wm, wn, lane);
    if (st + 1 < nfull) {
      sa2.store(As[1], tid);
      sb2.store(Bs[1], tid);
      __syncthreads();
      mfma_step<BM, BN, WN>(As[1], Bs[1], acc, wm, wn, lane);
    }
    __syncthreads();
  }
  int k0 = kb + nfull * TK;
  for (; k0 < ke; k0 += TK) {
    sa.load_checked(A, g.lda, g.M, m0, ke, k0, tid);
    sb.load_checked(B, g.ldb, g.N, n0, ke, k0, tid);
    sa.store(As[0], tid);
    sb.store(Bs[0], tid);
    __syncthreads();
    mfma_step<BM, BN, WN>(As[0], Bs[0], acc, wm, wn, lane);
    __syncthreads();
  }
  const int splits = gridDim.z;
  if (g.head) {
    // the loss head: per lane, its rows of each of its columns in a fixed
    // order; lane l then adds lane l + 32 (the other rows of the same
    // columns), the two waves of a column half add in wave order through LDS
    float* hs = reinterpret_cast<float*>(gemm_lds);  // [2][BN] column sums, [2 WN] squared errors
    float csum[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) csum[j] = 0.0f;
    float sq = 0.0f;
#pragma unroll
    for (int i = 0; i < MI; ++i) {
      // the target rows of this lane's 16 rows, then each column's 16 target
      // values, all loaded unconditionally (indices clamped) so that they
      // are in flight together
      int64_t src[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = min(m0 + wm * (BM / 2) + i * 32 + (r / 4) * 8 + (lane / 32) * 4 + r % 4,
                            g.M - 1);
        src[r] = (g.tidx ? (int64_t)g.tidx[row] : (int64_t)row) * g.ldt;
      }
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int col = n0 + wn * (BN / WN) + j * 32 + lane % 32;
        const int cc = min(col, g.N - 1);
        const float bcol = g.bias[cc];
        float tv[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) tv[r] = g.tgt[src[r] + cc];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int row = m0 + wm * (BM / 2) + i * 32 + (r / 4) * 8 + (lane / 32) * 4 + r % 4;
          if (row < g.M && col < g.N) {
            const float e = (acc[i][j][r] + bcol) - tv[r];
            const float gr = 2.0f * e * g.scale;
            g.dyb[(int64_t)row * g.N + col] = (__bf16)gr;
            csum[j] += gr;
            sq += e * e;
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NJ; ++j) csum[j] += __shfl_down(csum[j], 32);
#pragma unroll
    for (int off = kWaveMlp / 2; off > 0; off >>= 1) {
      const float o = __shfl_xor(sq, off);
      sq = (lane & off) == 0 ? sq + o : o + sq;
    }
    if (lane < 32) {
#pragma unroll
      for (int j = 0; j < NJ; ++j) hs[wm * BN + wn * (BN / WN) + j * 32 + lane] = csum[j];
    }
    if (lane == 0) hs[2 * BN + wave] = sq;
    __syncthreads();
    for (int c = tid; c < BN; c += blockDim.x)
      if (n0 + c < g.N) g.hcol[(int64_t)by * g.N + n0 + c] = hs[c] + hs[BN + c];
    if (tid == 0) {
      float q = hs[2 * BN];
      for (int w = 1; w < 2 * WN; ++w) q += hs[2 * BN + w];
      g.hsq[(int64_t)by * gx + bx] = q;
    }
    return;
  }
  // Adam on the parameters whose gradients the tile holds, in the operand
  // layouts that produce them (weight gradients dY^T h: both operands
  // row-contiguous): values and moments of eight rows loaded
  // unconditionally (indices clamped), in flight together, ahead of the
  // bounds-checked updates
  if constexpr (!AKC && !BKC) {
    if (g.adam && splits == 1) {
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const int col = n0 + wn * (BN / WN) + j * 32 + lane % 32;
          const int cc = min(col, g.N - 1);
#pragma unroll
          for (int h = 0; h < 16; h += 8) {
            float pv[8], av[8], bv[8];
#pragma unroll
            for (int r = h; r < h + 8; ++r) {
              const int row = min(m0 + wm * (BM / 2) + i * 32 + (r / 4) * 8 + (lane / 32) * 4 +
                                      r % 4,
                                  g.M - 1);
              const int64_t o = (int64_t)row * g.N + cc;
              pv[r - h] = g.ap[o];
              av[r - h] = g.am1[o];
              bv[r - h] = g.am2[o];
            }
#pragma unroll
            for (int r = h; r < h + 8; ++r) {
              const int row = m0 + wm * (BM / 2) + i * 32 + (r / 4) * 8 + (lane / 32) * 4 + r % 4;
              if (row < g.M && col < g.N) {
                const int64_t o = (int64_t)row * g.N + col;
                const float w = adam_one(pv[r - h], acc[i][j][r], av[r - h], bv[r - h], g.lr,
                                         g.b1, g.b2, g.eps, g.c1, g.c2);
                g.ap[o] = w;
                g.am1[o] = av[r - h];
                g.am2[o] = bv[r - h];
                g.apb[(int64_t)row * g.ldpb + col] = (__bf16)w;
              }
            }
          }
        }
      return;
    }
  }
  float* P = g.ws + (int64_t)bz * g.split_stride;
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int col = n0 + wn * (BN / WN) + j * 32 + lane % 32;
      const float bcol = (g.bias && col < g.N) ? g.bias[col] : 0.0f;  // one load per lane
      // the GELU derivative's pre-activations, loaded unconditionally
      // (indices clamped) so that all 16 are in flight together
      float gzv[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = min(m0 + wm * (BM / 2) + i * 32 + (r / 4) * 8 + (lane / 32) * 4 + r % 4,
                            g.M - 1);
        gzv[r] = g.gz ? g.gz[(int64_t)row * g.ldc + min(col, g.N - 1)] : 0.0f;
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm * (BM / 2) + i * 32 + (r / 4) * 8 + (lane / 32) * 4 + r % 4;
        if (row < g.M && col < g.N) {
          if (splits == 1) finish_store(g, row, col, acc[i][j][r], bcol, gzv[r]);
          else P[(int64_t)row * g.N + col] = acc[i][j][r];
        }
      }
    }
}

// Split-K epilogue: C = sum over s (in order) of ws[s][m][n], then the
// GEMM's epilogue (bias, GELU keeping the pre-activation, or the GELU
// derivative of the backward pass), for NS splits (2 .. kMaxSplit, one
// instance each, so that every split's load is unconditional and all are in
// flight together).  The workspace is dense (ld N), so the elements are taken
// flat, grid-stride: four adjacent columns per thread with 16-byte loads when
// N is a multiple of 4 (the workspace and, with gz, the dense pre-activations
// then 16-byte aligned), else one.  Every split launch has fewer than 384
// tiles of 128 x 128 (M N < 2^23): 32-bit indices.
template <int NS>
__global__ void splitk_reduce_kernel(GemmArgs g) {
  const int n = g.M * g.N;
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x, ts = gridDim.x * blockDim.x;
  if ((g.N & 3) == 0) {
    // the pre-activations are loaded with the partials (ldc == N with gz;
    // without gz a workspace vector, unused)
    const float* zsrc = g.gz ? g.gz : g.ws;
    // with Adam, the parameters and moments likewise (ldc == N)
    const float* psrc = g.adam ? g.ap : g.ws;
    const float* asrc = g.adam ? g.am1 : g.ws;
    const float* bsrc = g.adam ? g.am2 : g.ws;
    for (int i = 4 * t0; i < n; i += 4 * ts) {
      const float4 z = *reinterpret_cast<const float4*>(zsrc + i);
      const float4 pq = *reinterpret_cast<const float4*>(psrc + i);
      float4 aq = *reinterpret_cast<const float4*>(asrc + i);
      float4 bq = *reinterpret_cast<const float4*>(bsrc + i);
      float4 w[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) w[s] = *reinterpret_cast<const float4*>(g.ws + s * n + i);
      float4 v = w[0];
#pragma unroll
      for (int s = 1; s < NS; ++s) {
        v.x += w[s].x;
        v.y += w[s].y;
        v.z += w[s].z;
        v.w += w[s].w;
      }
      if (g.adam) {
        float4 u;
        u.x = adam_one(pq.x, v.x, aq.x, bq.x, g.lr, g.b1, g.b2, g.eps, g.c1, g.c2);
        u.y = adam_one(pq.y, v.y, aq.y, bq.y, g.lr, g.b1, g.b2, g.eps, g.c1, g.c2);
        u.z = adam_one(pq.z, v.z, aq.z, bq.z, g.lr, g.b1, g.b2, g.eps, g.c1, g.c2);
        u.w = adam_one(pq.w, v.w, aq.w, bq.w, g.lr, g.b1, g.b2, g.eps, g.c1, g.c2);
        *reinterpret_cast<float4*>(g.ap + i) = u;
        *reinterpret_cast<float4*>(g.am1 + i) = aq;
        *reinterpret_cast<float4*>(g.am2 + i) = bq;
        const int row = i / g.N;
        __bf16* pb = g.apb + (int64_t)row * g.ldpb + (i - row * g.N);
        pb[0] = (__bf16)u.x;
        pb[1] = (__bf16)u.y;
        pb[2] = (__bf16)u.z;
        pb[3] = (__bf16)u.w;
        continue;
      }
      const int row = i / g.N, c = i - row * g.N;
      const float e[4] = {v.x, v.y, v.z, v.w}, ze[4] = {z.x, z.y, z.z, z.w};
#pragma unroll
      for (int q = 0; q < 4; ++q)
        finish_store(g, row, c + q, e[q], g.bias ? g.bias[c + q] : 0.0f, ze[q]);
    }
    return;
  }
  for (int i = t0; i < n; i += ts) {
    float w[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) w[s] = g.ws[s * n + i];
    float v = w[0];
#pragma unroll
    for (int s = 1; s < NS; ++s) v += w[s];
    if (g.adam) {
      float a = g.am1[i], b = g.am2[i];
      const float u = adam_one(g.ap[i], v, a, b, g.lr, g.b1, g.b2, g.eps, g.c1, g.c2);
      g.ap[i] = u;
      g.am1[i] = a;
      g.am2[i] = b;
      const int row = i / g.N;
      g.apb[(int64_t)row * g.ldpb + (i - row * g.N)] = (__bf16)u;
      continue;
    }
    const int row = i / g.N, c = i - row * g.N;
    finish_store(g, row, c, v, g.bias ? g.bias[c] : 0.0f, g.gz ? g.gz[i] : 0.0f);
  }
}

const void* splitk_reduce_fn(int splits) {
  switch (splits) {
#define GCM_REDUCE_CASE(i) \
  case i:                  \
    return (const void*)splitk_reduce_kernel<i>;
    GCM_REDUCE_CASE(2) GCM_REDUCE_CASE(3) GCM_REDUCE_CASE(4) GCM_REDUCE_CASE(5)
    GCM_REDUCE_CASE(6) GCM_REDUCE_CASE(7) GCM_REDUCE_CASE(8) GCM_REDUCE_CASE(9)
    GCM_REDUCE_CASE(10) GCM_REDUCE_CASE(11) GCM_REDUCE_CASE(12) GCM_REDUCE_CASE(13)
    GCM_REDUCE_CASE(14) GCM_REDUCE_CASE(15) GCM_REDUCE_CASE(16)
#undef GCM_REDUCE_CASE
  }
  return nullptr;
}
static_assert(kMaxSplit == 16, "one splitk_reduce_kernel instance per split count");

// d *= gelu'(z) elementwise (the backward pass through a GELU layer after an
// unsplit GEMM with a small output tiling, whose epilogue would run this on
// few threads).
__global__ void gelu_backward_kernel(float* __restrict__ d, const float* __restrict__ z,
                                     int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] *= gelu_grad(z[i]);
}

// Column sums over M rows in a fixed order: the rows are cut into S slices
// (S = col_slices(M)), one block per (64 columns, slice); a block's 16 row
// sub-slices sum their rows in order and are added in order; with S > 1 the
// slice sums go to part[s][n] and colsum_final_kernel adds them in slice
// order.  One launch sums the columns of up to two M x N matrices: d[z] into
// out[z] for blockIdx.z = z.
constexpr int kColsumSlices = 16;
constexpr int kMaxColSlices = 32;
int col_slices(int M) { return std::max(1, std::min(kMaxColSlices, M / 256)); }

// Rows [r0, r1) of a slice.
__device__ __forceinline__ void slice_rows(int M, int S, int s, int* r0, int* r1) {
  *r0 = (int)((int64_t)M * s / S);
  *r1 = (int)((int64_t)M * (s + 1) / S);
}

// With ap[z] set (a single slice, so that the sum is the gradient), the
// sums are the gradients of parameters ap[z][n] (biases), which take their
// Adam step in place (moments am1, am2; the step's scalars st) instead of
// being stored.
struct AdamScalars {
  float lr, b1, b2, eps, c1, c2;
};
struct ColsumPair {
  const float* d[2];
  float* out[2];
  float* ap[2];
  float* am1[2];
  float* am2[2];
  AdamScalars st;
};

// Parameter i's Adam step with gradient gi, or out[i] = gi without ap.
__device__ __forceinline__ void store_or_step(float* out, float* ap, float* am1, float* am2,
                                              const AdamScalars& st, int64_t i, float gi) {
  if (ap) {
    float a = am1[i], b = am2[i];
    ap[i] = adam_one(ap[i], gi, a, b, st.lr, st.b1, st.b2, st.eps, st.c1, st.c2);
    am1[i] = a;
    am2[i] = b;
  } else {
    out[i] = gi;
  }
}

__global__ void __launch_bounds__(64 * kColsumSlices)
colsum_kernel(ColsumPair pr, int M, int N) {
  const float* __restrict__ d = pr.d[blockIdx.z];
  float* __restrict__ out = pr.out[blockIdx.z];
  __shared__ float part[kColsumSlices][65];
  const int c = blockIdx.x * 64 + threadIdx.x;
  int r0, r1;
  slice_rows(M, gridDim.y, blockIdx.y, &r0, &r1);
  float s = 0.0f;
  if (c < N) {
#pragma unroll 4
    for (int m = r0 + threadIdx.y; m < r1; m += kColsumSlices) s += d[(int64_t)m * N + c];
  }
  part[threadIdx.y][threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.y == 0 && c < N) {
    float t = 0.0f;
    for (int i = 0; i < kColsumSlices; ++i) t += part[i][threadIdx.x];
    const int z = blockIdx.z;
    store_or_step(out, pr.ap[z], pr.am1[z], pr.am2[z], pr.st, (int64_t)blockIdx.y * N + c, t);
  }
}

// The final passes of up to three column sums in one launch: segment
// blockIdx.y adds its S slice sums part[s][n] in slice order into db[n], or
// with ap set takes the Adam step of parameters ap[n] with them (as
// ColsumPair).
struct ColsumFinal {
  const float* part;
  int S, N;
  float* db;
  float *ap = nullptr, *am1 = nullptr, *am2 = nullptr;
};
struct ColsumFinals {
  ColsumFinal f[3];
  AdamScalars st;
};

__global__ void colsum_final_kernel(ColsumFinals fs) {
  const ColsumFinal sg = fs.f[blockIdx.y];
  const float* __restrict__ part = sg.part;
  float* __restrict__ db = sg.db;
  const int S = sg.S, N = sg.N;
  constexpr int U = 16;  // slices loaded together (indices clamped), added in order
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
    float t = 0.0f;
    for (int s0 = 0; s0 < S; s0 += U) {
      float v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = part[(size_t)min(s0 + u, S - 1) * N + c];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (s0 + u < S) t = s0 + u == 0 ? v[u] : t + v[u];
    }
    store_or_step(db, sg.ap, sg.am1, sg.am2, fs.st, c, t);
  }
}

// The loss head of the backward pass, laid out like colsum_kernel:
// dY = 2 (Y - T) / (M N), stored in BF16 (the GEMMs that read it round their
// operands to BF16 anyway), the (slice) column sums of dY in FP32, and the
// block's sum of squared errors in partial[slice][blockIdx.x] (each in a
// fixed order).
__global__ void __launch_bounds__(64 * kColsumSlices)
mse_colsum_kernel(const float* __restrict__ y, const float* __restrict__ t,
                  __bf16* __restrict__ dy, int M, int N, float scale, float* __restrict__ out,
                  float* __restrict__ partial) {
  __shared__ float part[kColsumSlices][65];
  __shared__ float sq[kColsumSlices][65];
  const int c = blockIdx.x * 64 + threadIdx.x;
  int r0, r1;
  slice_rows(M, gridDim.y, blockIdx.y, &r0, &r1);
  float s = 0.0f, e2 = 0.0f;
  if (c < N) {
#pragma unroll 4
    for (int m = r0 + threadIdx.y; m < r1; m += kColsumSlices) {
      const int64_t i = (int64_t)m * N + c;
      const float e = y[i] - t[i];
      const float g = 2.0f * e * scale;
      dy[i] = (__bf16)g;
      s += g;
      e2 += e * e;
    }
  }
  part[threadIdx.y][threadIdx.x] = s;
  sq[threadIdx.y][threadIdx.x] = e2;
  __syncthreads();
  if (threadIdx.y == 0) {
    float a = 0.0f, q = 0.0f;
    for (int i = 0; i < kColsumSlices; ++i) {
      a += part[i][threadIdx.x];
      q += sq[i][threadIdx.x];
    }
    if (c < N) out[(size_t)blockIdx.y * N + c] = a;
    part[0][threadIdx.x] = q;
  }
  __syncthreads();
  if (threadIdx.y == 0 && threadIdx.x == 0) {
    float q = 0.0f;
    for (int i = 0; i < 64; ++i) q += part[0][i];
    partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = q;
  }
}

// acc[0] += the sum of part[0 .. n) in FP64, in a fixed order (each thread
// its elements i = t mod 256 in order, then a fixed tree over the threads):
// one block.
__global__ void __launch_bounds__(256) loss_sum_kernel(const float* __restrict__ part, int n,
                                                       double* __restrict__ acc) {
  __shared__ double sm[256];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) v += part[i];
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[0] += sm[0];
}

// dst row r = src row idx[r], n floats per row, over the rows flattened:
// one 16-byte vector per thread and step (one float when n % 4 != 0), grid
// stride, so that every thread copies (rows * n < 2^30, checked by the host,
// keeps the 32-bit index and its last stride in range).
__device__ __forceinline__ void gather_flat(const float* __restrict__ src,
                                            float* __restrict__ dst,
                                            const int32_t* __restrict__ idx, int rows, int n) {
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x, ts = gridDim.x * blockDim.x;
  if ((n & 3) == 0) {
    const int n4 = n / 4, total = rows * n4;
    const float4* s4 = reinterpret_cast<const float4*>(src);
    float4* d4 = reinterpret_cast<float4*>(dst);
    for (int i = t0; i < total; i += ts) {
      const int r = i / n4;
      d4[i] = s4[(int64_t)idx[r] * n4 + (i - r * n4)];
    }
  } else {
    const int total = rows * n;
    for (int i = t0; i < total; i += ts) {
      const int r = i / n;
      dst[i] = src[(int64_t)idx[r] * n + (i - r * n)];
    }
  }
}

// The minibatch's rows of the training set: inputs (cx columns) and targets
// (ct columns; none with ct = 0).
__global__ void gather_rows_kernel(const float* __restrict__ sx, const float* __restrict__ st,
                                   const int32_t* __restrict__ idx, int rows, int cx, int ct,
                                   float* __restrict__ dx, float* __restrict__ dt) {
  gather_flat(sx, dx, idx, rows, cx);
  if (ct) gather_flat(st, dt, idx, rows, ct);
}

// Where flat parameter i lives in the BF16 weight copy: W1 with its rows
// padded to a multiple of 8 elements (16-byte aligned rows for the GEMM's
// vector loads), then W2 and W3; biases (read in FP32) have no copy (-1).
struct Bf16Map {
  int64_t off_b1, off_w2, off_b2, off_w3, off_b3, bw2, bw3;
  int n_in, n_in_pad;
};

__device__ __forceinline__ int64_t bf16_index(const Bf16Map& q, int64_t i) {
  if (i < q.off_b1) return (i / q.n_in) * q.n_in_pad + i % q.n_in;
  if (i >= q.off_w2 && i < q.off_b2) return q.bw2 + (i - q.off_w2);
  if (i >= q.off_w3 && i < q.off_b3) return q.bw3 + (i - q.off_w3);
  return -1;
}

// Adam on the FP32 master copy; refreshes the BF16 copy.  Four parameters
// per thread with 16-byte loads and stores (the arrays are 16-byte aligned);
// their BF16 copies as one 8-byte store when they are consecutive and aligned
// there, else one by one.  Parameters [gap0, gap0 + gap) are skipped (updated
// by the epilogue that produced their gradient); gap0 and gap are multiples
// of 4.

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m1, float* __restrict__ m2,
                            __bf16* __restrict__ pb, Bf16Map q, int64_t n, int64_t gap0,
                            int64_t gap, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
  typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
  const int64_t n4 = (n - gap) / 4;  // over the parameters outside the gap
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n4;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = 4 * v < gap0 ? 4 * v : 4 * v + gap;
    const float4 gv = *reinterpret_cast<const float4*>(g + i);
    float4 a = *reinterpret_cast<const float4*>(m1 + i);
    float4 b = *reinterpret_cast<const float4*>(m2 + i);
    float4 w = *reinterpret_cast<const float4*>(p + i);
    w.x = adam_one(w.x, gv.x, a.x, b.x, lr, b1, b2, eps, c1, c2);
    w.y = adam_one(w.y, gv.y, a.y, b.y, lr, b1, b2, eps, c1, c2);
    w.z = adam_one(w.z, gv.z, a.z, b.z, lr, b1, b2, eps, c1, c2);
    w.w = adam_one(w.w, gv.w, a.w, b.w, lr, b1, b2, eps, c1, c2);
    *reinterpret_cast<float4*>(m1 + i) = a;
    *reinterpret_cast<float4*>(m2 + i) = b;
    *reinterpret_cast<float4*>(p + i) = w;
    const int64_t j0 = bf16_index(q, i), j3 = bf16_index(q, i + 3);
    if (j0 >= 0 && j3 == j0 + 3 && (j0 & 3) == 0) {
      bf16x4 h;
      h[0] = (__bf16)w.x;
      h[1] = (__bf16)w.y;
      h[2] = (__bf16)w.z;
      h[3] = (__bf16)w.w;
      *reinterpret_cast<bf16x4*>(pb + j0) = h;
    } else {
      const float e[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int64_t j = bf16_index(q, i + k);
        if (j >= 0) pb[j] = (__bf16)e[k];
      }
    }
  }
  // the last (n - gap) % 4 parameters
  for (int64_t j = 4 * n4 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n - gap;
       j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = j < gap0 ? j : j + gap;
    float a = m1[i], b = m2[i];
    const float w = adam_one(p[i], g[i], a, b, lr, b1, b2, eps, c1, c2);
    m1[i] = a;
    m2[i] = b;
    p[i] = w;
    const int64_t k = bf16_index(q, i);
    if (k >= 0) pb[k] = (__bf16)w;
  }
}

__global__ void to_bf16_kernel(const float* __restrict__ p, __bf16* __restrict__ pb, Bf16Map q,
                               int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = bf16_index(q, i);
    if (j >= 0) pb[j] = (__bf16)p[i];
  }
}

int grid_for(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 4096); }

// Surrogate likelihood head: one block per row.  Bin b's predicted content is
// pred = exp(mean[b] + scale[b] * y[b]) (mean = -inf marks a bin with no
// events: pred = 0), its MC variance pred * ratio[b], and the row's chi2 is
// the Barlow-Beeston OA2021 sum over bins (BarlowBeestonBanff2022.h:150-294
// with the default options: frac limit 1e150, expected minimum -1, penalty
// included), reduced in a fixed tree order.
constexpr int kHeadBlock = 256;

__device__ double bb_chi2(double data, double pred, double mcvar) {
  const double inf = __builtin_huge_val();
  if (!isfinite(mcvar) || mcvar < 0.0) return inf;
  if (pred <= -1.0) return (pred <= 0.0 && data > 0.0) ? inf : 0.0;
  const double dmin = 2.2250738585072014e-308;
  const double deps = 2.220446049250313e-16;
  double newmc = pred, penalty = 0.0;
  if (mcvar > 0.0) {
    const double f = fmin(sqrt(mcvar) / pred, 1e150);
    const double temp = pred * f * f - 1;
    const double temp2 = temp * temp + 4 * data * f * f;
    if (temp2 < 0) return inf;
    const double beta = (-temp + sqrt(temp2)) / 2.;
    newmc = fmax(pred * beta, dmin);
    penalty = (beta - 1) * (beta - 1) / (2 * f * f);
  }
  double stat;
  if (data <= deps) stat = newmc;
  else if (newmc > dmin) stat = newmc - data + data * (log(data) - log(newmc));
  else stat = dmin - data + data * (log(data) - log(dmin));
  return 2.0 * stat + 2.0 * penalty;
}

__global__ void __launch_bounds__(kHeadBlock)
head_llh_kernel(const float* __restrict__ y, int n_out, const double* __restrict__ mean,
                const double* __restrict__ scale, const double* __restrict__ data,
                const double* __restrict__ ratio, double* __restrict__ out) {
  __shared__ double s[kHeadBlock];
  const int row = blockIdx.x;
  const float* yr = y + (int64_t)row * n_out;
  double acc = 0.0;
  for (int b = threadIdx.x; b < n_out; b += kHeadBlock) {
    const double pred = exp(mean[b] + scale[b] * (double)yr[b]);
    acc += bb_chi2(data[b], pred, pred * ratio[b]);
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int st = kHeadBlock / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[row] = s[0];
}

}  // namespace

// Scratch of gcm_mlp_predict_llh_noalloc: the activations and split-K
// partials of one batch live in this fixed array (one copy per GPU, held
// until the process ends) instead of in memory the library allocates.  In
// host code the name is not a device address: kernels get the address that
// hipGetSymbolAddress returns for the current device (noalloc_scratch below).
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, resolved once per device with
// that device current.
int noalloc_scratch(int device, unsigned char** out) {
  if (device < 0 || device >= kMaxDevices) return fail(GCM_ERR_ARG, "device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(GCM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch_ptr[device];
  return GCM_OK;
}
}  // namespace

struct gcm_mlp {
  int device = 0;
  int n_in = 0, H = 0, n_out = 0;
  hipStream_t stream = nullptr;
  unsigned char* scratch = nullptr;  // this device's noalloc scratch array
  int64_t n_params = 0;
  int64_t off_w1 = 0, off_b1 = 0, off_w2 = 0, off_b2 = 0, off_w3 = 0, off_b3 = 0;
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  __bf16* pb = nullptr;  // BF16 weights, layout of Bf16Map
  Bf16Map bmap{};
  int64_t n_bf16 = 0;
  int64_t adam_t = 0;
  // activations for a batch of up to cap rows
  int cap = 0;
  // (the GELU outputs h1, h2 and the loss gradient dy are read only by
  // GEMMs, which take BF16 operands: they are stored in BF16)
  float *x = nullptr, *z1 = nullptr, *z2 = nullptr, *y = nullptr, *t = nullptr, *d2 = nullptr,
        *d1 = nullptr, *partial = nullptr, *colpart = nullptr;
  // the fused loss head's per-tile column sums and squared errors
  float *hcol = nullptr, *hsq = nullptr;
  __bf16 *h1 = nullptr, *h2 = nullptr, *dy = nullptr;
  // split-K partial sums
  float* ws = nullptr;
  size_t ws_cap = 0;
  // resident training set; the shuffled row order of two consecutive epochs
  // (order: 2 n_data on the device; order_host: the same in pinned host
  // memory, each half's copy marked done by order_ev[half]); the sampled
  // minibatch losses summed on the device, one double per epoch
  float *dx = nullptr, *dt = nullptr;
  int32_t* order = nullptr;
  int32_t* order_host = nullptr;
  hipEvent_t order_ev[2] = {nullptr, nullptr};
  int64_t n_data = 0;
  double* loss_dev = nullptr;
  int loss_cap = 0;
  // likelihood head: mean, scale, data, ratio (n_out each) and per-row chi2
  double* head = nullptr;
  double* chi2 = nullptr;
  int chi2_cap = 0;
  std::vector<void*> allocs;

  template <typename T>
  int alloc(T** ptr, size_t count) {
    *ptr = nullptr;
    if (hipMalloc((void**)ptr, std::max<size_t>(1, count) * sizeof(T)) != hipSuccess)
      return fail(GCM_ERR_NOMEM, "hipMalloc failed (mlp)");
    allocs.push_back(*ptr);
    return GCM_OK;
  }
  int ensure_batch(int rows) {
    if (rows <= cap) return GCM_OK;
    for (void* q : {(void*)x, (void*)z1, (void*)h1, (void*)z2, (void*)h2, (void*)y,
                    (void*)t, (void*)dy, (void*)d2, (void*)d1, (void*)hcol, (void*)hsq}) {
      if (!q) continue;
      (void)hipFree(q);
      allocs.erase(std::remove(allocs.begin(), allocs.end(), q), allocs.end());
    }
    cap = rows;
    int rc = GCM_OK;
    auto step = [&](int r) { if (rc == GCM_OK) rc = r; };
    step(alloc(&x, (size_t)rows * n_in));
    step(alloc(&z1, (size_t)rows * H));
    step(alloc(&h1, (size_t)rows * H));
    step(alloc(&z2, (size_t)rows * H));
    step(alloc(&h2, (size_t)rows * H));
    step(alloc(&y, (size_t)rows * n_out));
    step(alloc(&t, (size_t)rows * n_out));
    step(alloc(&dy, (size_t)rows * n_out));
    step(alloc(&d2, (size_t)rows * H));
    step(alloc(&d1, (size_t)rows * H));
    const size_t row_tiles = ((size_t)rows + 63) / 64;  // at most, whatever the tile
    step(alloc(&hcol, row_tiles * n_out));
    step(alloc(&hsq, row_tiles * (((size_t)n_out + 63) / 64)));
    return rc;
  }
  int ensure_ws(size_t floats) {
    if (floats <= ws_cap) return GCM_OK;
    if (ws) {
      MLP_HIP(hipStreamSynchronize(stream));  // queued launches may still read it
      (void)hipFree(ws);
      allocs.erase(std::remove(allocs.begin(), allocs.end(), (void*)ws), allocs.end());
      ws = nullptr;
    }
    ws_cap = floats;
    return alloc(&ws, floats);
  }
};

namespace {

// Whether an operand meets the alignment of the kernel's unchecked loads:
// 16-byte vectors along K (row pitch a multiple of the vector), 2-element
// vectors along the rows (even row pitch and an even row count R).
bool vec_ok(const void* p, int64_t ld, int kc, int bf16, int R) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  const int es = bf16 ? 2 : 4;
  if (kc) return a % 16 == 0 && (ld * es) % 16 == 0;
  return a % (2 * es) == 0 && ld % 2 == 0 && R % 2 == 0;
}

// The kernel instance for a tile side (64 or 128), its waves along N (2 or
// 4) and operand layout selector sel = A K-contiguous | A BF16 | B
// K-contiguous | B BF16 (bits 3..0).
template <int T, int WN>
const void* gemm_fn(int sel) {
  switch (sel) {
#define GCM_GEMM_CASE(i) \
  case i:                \
    return (const void*)gemm_bf16_kernel<T, T, ((i)&8) != 0, ((i)&4) != 0, ((i)&2) != 0, ((i)&1) != 0, WN>;
    GCM_GEMM_CASE(0) GCM_GEMM_CASE(1) GCM_GEMM_CASE(2) GCM_GEMM_CASE(3)
    GCM_GEMM_CASE(4) GCM_GEMM_CASE(5) GCM_GEMM_CASE(6) GCM_GEMM_CASE(7)
    GCM_GEMM_CASE(8) GCM_GEMM_CASE(9) GCM_GEMM_CASE(10) GCM_GEMM_CASE(11)
    GCM_GEMM_CASE(12) GCM_GEMM_CASE(13) GCM_GEMM_CASE(14) GCM_GEMM_CASE(15)
#undef GCM_GEMM_CASE
  }
  return nullptr;
}

// The large tile's LDS is past the default 64 KB: allow it (once per process).
hipError_t gemm_set_lds_attrs() {
  for (int sel = 0; sel < 16; ++sel) {
    const hipError_t e = hipFuncSetAttribute(gemm_fn<kBig, kBigWN>(sel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             gemm_lds_bytes(kBig, kBig));
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_gemm(bool big, int sel, dim3 grid, hipStream_t stream, const GemmArgs& g) {
  const void* fn = big ? gemm_fn<kBig, kBigWN>(sel) : gemm_fn<TM, 2>(sel);
  GemmArgs a = g;
  void* args[] = {&a};
  const hipError_t e = hipLaunchKernel(fn, grid, dim3(128 * (big ? kBigWN : 2)), args,
                                       big ? gemm_lds_bytes(kBig, kBig) : gemm_lds_bytes(TM, TN),
                                       stream);
  return e != hipSuccess ? e : hipGetLastError();
}

// C[M x N] = A . B (+ bias, GELU with the pre-activation in Z), operand
// layouts as in GemmArgs.  With gz, C is dense (ldc == N) and is multiplied by
// gelu'(gz) elementwise (the backward pass through a GELU layer).
//
// Tile and split-K plan of an M x N x K product (the same on every path, so
// that every entry point computes the same sums in the same order):
//  - the 128 x 128 tile when M, N >= 128, K has at least four steps and the
//    product has at least 4096 tile-K-steps of work; else 64 x 64;
//  - split along K (a block's K steps are bound by load latency, so more
//    blocks in flight finish sooner; shorter reductions gain less than the
//    extra reduction launch costs):
//      64 x 64: outputs of fewer than 128 tiles with at least 16 K steps,
//        towards 1024 blocks, into up to 16 chunks of at least two K steps;
//      128 x 128: outputs of fewer than 384 tiles with at least 16 K steps,
//        towards 512 blocks, into up to 8 chunks of at least eight K steps.
//    The partials go to a workspace and a reduction launch sums them in
//    split order, so the result does not depend on scheduling.
struct GemmPlan {
  bool big;
  int T, splits, chunk_steps;
};

GemmPlan gemm_plan(int M, int N, int K) {
  const int steps = std::max(1, (K + TK - 1) / TK);
  const int64_t big_tiles = (int64_t)((M + kBig - 1) / kBig) * ((N + kBig - 1) / kBig);
  GemmPlan p{};
  p.big = M >= kBig && N >= kBig && steps >= 4 && big_tiles * steps >= 4096;
  p.T = p.big ? kBig : TM;
  const int tiles = ((M + p.T - 1) / p.T) * ((N + p.T - 1) / p.T);
  int splits = 1;
  if (p.big) {
    if (tiles < 384 && steps >= kMinSplitSteps)
      splits = std::max(1, std::min({steps / 8, (512 + tiles - 1) / tiles, 8}));
  } else if (tiles < 128 && steps >= kMinSplitSteps) {
    splits = std::max(1, std::min({steps / 2, (1024 + tiles - 1) / tiles, kMaxSplit}));
  }
  p.chunk_steps = (steps + splits - 1) / splits;
  p.splits = (steps + p.chunk_steps - 1) / p.chunk_steps;
  return p;
}

// Split-K partial floats an M x N x K product needs (0 when not split).
size_t gemm_ws_floats(int M, int N, int K) {
  const GemmPlan p = gemm_plan(M, N, K);
  return p.splits > 1 ? (size_t)p.splits * M * N : 0;
}

// Where a GEMM runs: a stream and a fixed split-K workspace that is never
// grown (the allocation-free path).  Without one, the handle's stream and its
// workspace, grown on demand.
struct Exec {
  hipStream_t stream;
  float* ws;
  size_t ws_cap;  // floats
};

// One Adam step applied to a dense M x N block of parameters by the GEMM
// that produces their gradient (GemmArgs::adam): the block's FP32 values,
// moments and BF16 copy, the step's scalars.
struct AdamStep {
  float *p, *m1, *m2;
  __bf16* pb;  // BF16 copy, row pitch ldpb (null for biases: none)
  int64_t ldpb;
  float lr, b1, b2, eps, c1, c2;
};

// C is FP32 (C) or BF16 (Cb, with C null); with adam the product is the
// gradient of adam's parameters, which are updated instead of storing C.
int gemm(gcm_mlp* m, int M, int N, int K, const void* A, int64_t lda, int a_kc, int a_bf16,
         const void* B, int64_t ldb, int b_kc, int b_bf16, float* C, int64_t ldc,
         const float* bias, float* Z, const float* gz = nullptr, const Exec* ex = nullptr,
         __bf16* Cb = nullptr, const GemmArgs* head = nullptr, const AdamStep* adam = nullptr) {
  if (M <= 0 || N <= 0) return GCM_OK;
  if (adam && (bias || Z || gz || Cb || head || ex || ldc != N || a_kc || b_kc))
    return fail(GCM_ERR_ARG,
                "gemm: the Adam epilogue needs a plain dense output of row-contiguous operands");
  if (gz && (ldc != N || Cb)) return fail(GCM_ERR_ARG, "gemm: GELU backward needs a dense FP32 output");
  const hipStream_t stream = ex ? ex->stream : m->stream;
  const GemmPlan p = gemm_plan(M, N, K);
  GemmArgs g{M, N, K, A, lda, B, ldb, vec_ok(A, lda, a_kc, a_bf16, M) && vec_ok(B, ldb, b_kc, b_bf16, N),
             (int64_t)N * K * (b_bf16 ? 2 : 4) > (int64_t)M * K * (a_bf16 ? 2 : 4),
             C, Cb, ldc, nullptr, 0, p.chunk_steps * TK, bias, Z, Z != nullptr, gz};
  const int sel = (a_kc ? 8 : 0) | (a_bf16 ? 4 : 0) | (b_kc ? 2 : 0) | (b_bf16 ? 1 : 0);
  const dim3 grid((N + p.T - 1) / p.T, (M + p.T - 1) / p.T, p.splits);
  if (head) {  // the fused loss head (the caller checked that the launch is unsplit)
    if (p.splits != 1 || !bias) return fail(GCM_ERR_ARG, "gemm: loss head needs an unsplit launch with a bias");
    g.head = 1;
    g.tgt = head->tgt;
    g.tidx = head->tidx;
    g.ldt = head->ldt;
    g.dyb = head->dyb;
    g.scale = head->scale;
    g.hcol = head->hcol;
    g.hsq = head->hsq;
  }
  if (adam) {
    g.adam = 1;
    g.ap = adam->p;
    g.am1 = adam->m1;
    g.am2 = adam->m2;
    g.apb = adam->pb;
    g.ldpb = adam->ldpb;
    g.lr = adam->lr;
    g.b1 = adam->b1;
    g.b2 = adam->b2;
    g.eps = adam->eps;
    g.c1 = adam->c1;
    g.c2 = adam->c2;
  }
  if (p.splits == 1) {
    // the GELU derivative in the epilogue when the output tiling is wide
    // (>= 256 blocks); a small tiling leaves it to one wide elementwise pass
    const bool gz_in_epilogue = gz && (int64_t)grid.x * grid.y >= 256;
    if (!gz_in_epilogue) g.gz = nullptr;
    MLP_HIP(launch_gemm(p.big, sel, grid, stream, g));
    if (gz && !gz_in_epilogue) {
      gelu_backward_kernel<<<grid_for((int64_t)M * N), 256, 0, stream>>>(C, gz, (int64_t)M * N);
      MLP_HIP(hipGetLastError());
    }
    return GCM_OK;
  }
  const int64_t mn = (int64_t)M * N;
  g.split_stride = mn;
  if (ex) {
    if ((size_t)p.splits * mn > ex->ws_cap) return fail(GCM_ERR_ARG, "gemm: workspace too small");
    g.ws = ex->ws;
  } else {
    int rc = m->ensure_ws((size_t)p.splits * mn);
    if (rc) return rc;
    g.ws = m->ws;
  }
  MLP_HIP(launch_gemm(p.big, sel, grid, stream, g));
  const void* red = splitk_reduce_fn(p.splits);
  if (!red) return fail(GCM_ERR_ARG, "gemm: split count out of range");
  void* args[] = {&g};
  MLP_HIP(hipLaunchKernel(red, dim3(grid_for(N % 4 == 0 ? mn / 4 : mn)), dim3(256), args, 0,
                          stream));
  MLP_HIP(hipGetLastError());
  return GCM_OK;
}

// Forward: h1 = gelu(x W1^T + b1), h2 = gelu(h1 W2^T + b2), y = h2 W3^T +
// b3, with the pre-activations in z1, z2 (FP32) and h1, h2 in BF16.
int forward(gcm_mlp* m, int rows, const float* x, float* z1, __bf16* h1, float* z2,
            __bf16* h2, float* y, const Exec* ex) {
  const __bf16* w = m->pb;
  int rc = gemm(m, rows, m->H, m->n_in, x, m->n_in, 1, 0, w, m->bmap.n_in_pad, 1, 1, nullptr,
                m->H, m->p + m->off_b1, z1, nullptr, ex, h1);
  if (rc) return rc;
  rc = gemm(m, rows, m->H, m->H, h1, m->H, 1, 1, w + m->bmap.bw2, m->H, 1, 1, nullptr, m->H,
            m->p + m->off_b2, z2, nullptr, ex, h2);
  if (rc) return rc;
  return gemm(m, rows, m->n_out, m->H, h2, m->H, 1, 1, w + m->bmap.bw3, m->H, 1, 1, y,
              m->n_out, m->p + m->off_b3, nullptr, nullptr, ex);
}

// Forward on rows already in m->x, into the handle's buffers.
int forward(gcm_mlp* m, int rows) {
  return forward(m, rows, m->x, m->z1, m->h1, m->z2, m->h2, m->y, nullptr);
}

// Whether a training step of `rows` fuses the loss head into the output
// layer: when that launch is not split along K.
bool head_fused(const gcm_mlp* m, int rows) {
  return gemm_plan(rows, m->n_out, m->H).splits == 1;
}

// The training forward pass on the rows in m->x.  With the loss head fused
// (head_fused), the output layer writes dY (BF16) to m->dy from the targets'
// rows (row r at tgt + (tidx ? tidx[r] : r) * n_out) and its per-tile column
// sums and squared errors, and no output y; otherwise y goes to m->y.
int forward_train(gcm_mlp* m, int rows, const float* tgt, const int32_t* tidx) {
  if (!head_fused(m, rows)) return forward(m, rows);
  const __bf16* w = m->pb;
  int rc = gemm(m, rows, m->H, m->n_in, m->x, m->n_in, 1, 0, w, m->bmap.n_in_pad, 1, 1,
                nullptr, m->H, m->p + m->off_b1, m->z1, nullptr, nullptr, m->h1);
  if (rc) return rc;
  rc = gemm(m, rows, m->H, m->H, m->h1, m->H, 1, 1, w + m->bmap.bw2, m->H, 1, 1, nullptr, m->H,
            m->p + m->off_b2, m->z2, nullptr, nullptr, m->h2);
  if (rc) return rc;
  GemmArgs hd{};
  hd.tgt = tgt;
  hd.tidx = tidx;
  hd.ldt = m->n_out;
  hd.dyb = m->dy;
  hd.scale = 1.0f / (float)((int64_t)rows * m->n_out);
  hd.hcol = m->hcol;
  hd.hsq = m->hsq;
  return gemm(m, rows, m->n_out, m->H, m->h2, m->H, 1, 1, w + m->bmap.bw3, m->H, 1, 1, nullptr,
              m->n_out, m->p + m->off_b3, nullptr, nullptr, nullptr, nullptr, &hd);
}

// Loss and gradients into m->g for the rows in m->x / m->t, after
// forward_train (with the loss head fused, dY and its partial sums are
// already there).  The loss (the mean squared error) is returned in *loss
// (a wait for the stream), or its sum of squared errors is added to
// *loss_acc on the device.  With w3_adam the output layer's weights W3 take
// their Adam step in the dW3 product (their gradient is not stored in g).
// One training step's Adam update fused into the backward pass: W3 always;
// with all also W2, W1 and the biases, so that no parameter is left for the
// Adam launch (the loss head must be fused).
struct FusedAdam {
  AdamScalars st;
  bool all;
};

int backward(gcm_mlp* m, int rows, double* loss, double* loss_acc = nullptr,
             const FusedAdam* fa = nullptr) {
  const bool fused = head_fused(m, rows);
  if (fa && fa->all && !fused) return fail(GCM_ERR_ARG, "backward: fused Adam needs the loss head");
  auto step_of = [&](int64_t off, int64_t bf_off, int64_t ldpb) {
    return AdamStep{m->p + off, m->m1 + off, m->m2 + off, m->pb + bf_off, ldpb, fa->st.lr,
                    fa->st.b1, fa->st.b2, fa->st.eps, fa->st.c1, fa->st.c2};
  };
  AdamStep s3{}, s2{}, s1{};
  if (fa) s3 = step_of(m->off_w3, m->bmap.bw3, m->H);
  if (fa && fa->all) {
    s2 = step_of(m->off_w2, m->bmap.bw2, m->H);
    s1 = step_of(m->off_w1, 0, m->bmap.n_in_pad);
  }
  const AdamStep* w3_adam = fa ? &s3 : nullptr;
  const AdamStep* w2_adam = fa && fa->all ? &s2 : nullptr;
  const AdamStep* w1_adam = fa && fa->all ? &s1 : nullptr;
  const int64_t ny = (int64_t)rows * m->n_out;
  const int nb = (m->n_out + 63) / 64;  // mse_colsum column blocks
  const int S = col_slices(rows);        // row slices: one partial per (slice, block)
  const __bf16* w = m->pb;
  float* g = m->g;
  const dim3 cs_block(64, kColsumSlices);
  // The bias gradients are column sums of dY, dH2 and dH1, read only by the
  // optimizer: dY's partial sums come with dY (the loss head's per-tile sums,
  // or mse_colsum_kernel's per-slice sums), dH2's and dH1's are taken in one
  // launch after the last GEMM, and one launch adds all partial sums into g.
  // colpart holds dY's slice sums (unfused), then dH2's, then dH1's.
  float* const part_y = m->colpart;
  float* const part_2 = part_y + (size_t)S * m->n_out;
  float* const part_1 = part_2 + (size_t)S * m->H;
  ColsumFinals fin{};
  int nfin = 0;
  // the loss partials: per (row tile, column tile) of the output layer when
  // fused, else per (row slice, column block) of mse_colsum_kernel
  const int head_tile = gemm_plan(rows, m->n_out, m->H).T;
  const int head_rows = (rows + head_tile - 1) / head_tile;
  const size_t n_loss = fused ? (size_t)head_rows * ((m->n_out + head_tile - 1) / head_tile)
                              : (size_t)nb * S;
  if (fa) fin.st = fa->st;
  // with every step fused, each bias takes its step where its sum completes
  auto bias_final = [&](const float* part, int S_, int N, int64_t off) {
    ColsumFinal f{part, S_, N, g + off};
    if (fa && fa->all) {
      f.ap = m->p + off;
      f.am1 = m->m1 + off;
      f.am2 = m->m2 + off;
    }
    return f;
  };
  if (fused) {
    fin.f[nfin++] = bias_final(m->hcol, head_rows, m->n_out, m->off_b3);
  } else {
    mse_colsum_kernel<<<dim3(nb, S), cs_block, 0, m->stream>>>(
        m->y, m->t, m->dy, rows, m->n_out, 1.0f / (float)ny, S > 1 ? part_y : g + m->off_b3,
        m->partial);
    if (S > 1) fin.f[nfin++] = ColsumFinal{part_y, S, m->n_out, g + m->off_b3};
  }
  MLP_HIP(hipGetLastError());
  // dW3 = dY^T h2 : (n_out x rows) . (rows x H); with w3_adam W3 takes its
  // Adam step in this product's epilogue instead of storing dW3, after dH2
  // has read W3
  auto dw3 = [&] {
    return gemm(m, m->n_out, m->H, rows, m->dy, m->n_out, 0, 1, m->h2, m->H, 0, 1,
                g + m->off_w3, m->H, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                w3_adam);
  };
  // dH2 = (dY W3) * gelu'(z2) : (rows x n_out) . (n_out x H)
  auto dh2 = [&] {
    return gemm(m, rows, m->H, m->n_out, m->dy, m->n_out, 1, 1, w + m->bmap.bw3, m->H, 0, 1,
                m->d2, m->H, nullptr, nullptr, m->z2);
  };
  int rc = w3_adam ? dh2() : dw3();
  if (rc) return rc;
  rc = w3_adam ? dw3() : dh2();
  if (rc) return rc;
  // dW2 = dH2^T h1 (with w2_adam W2 takes its Adam step here, after dH1 has
  // read W2)
  auto dw2 = [&] {
    return gemm(m, m->H, m->H, rows, m->d2, m->H, 0, 0, m->h1, m->H, 0, 1, g + m->off_w2, m->H,
                nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, w2_adam);
  };
  // dH1 = (dH2 W2) * gelu'(z1)
  auto dh1 = [&] {
    return gemm(m, rows, m->H, m->H, m->d2, m->H, 1, 0, w + m->bmap.bw2, m->H, 0, 1, m->d1, m->H,
                nullptr, nullptr, m->z1);
  };
  rc = w2_adam ? dh1() : dw2();
  if (rc) return rc;
  rc = w2_adam ? dw2() : dh1();
  if (rc) return rc;
  // dW1 = dH1^T x
  rc = gemm(m, m->H, m->n_in, rows, m->d1, m->H, 0, 0, m->x, m->n_in, 0, 0, g + m->off_w1,
            m->n_in, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, w1_adam);
  if (rc) return rc;
  const bool bias_step = fa && fa->all && S == 1;  // else in the final pass
  const ColsumPair pr{{m->d2, m->d1},
                      {S > 1 ? part_2 : g + m->off_b2, S > 1 ? part_1 : g + m->off_b1},
                      {bias_step ? m->p + m->off_b2 : nullptr, bias_step ? m->p + m->off_b1 : nullptr},
                      {bias_step ? m->m1 + m->off_b2 : nullptr, bias_step ? m->m1 + m->off_b1 : nullptr},
                      {bias_step ? m->m2 + m->off_b2 : nullptr, bias_step ? m->m2 + m->off_b1 : nullptr},
                      fa ? fa->st : AdamScalars{}};
  colsum_kernel<<<dim3((m->H + 63) / 64, S, 2), cs_block, 0, m->stream>>>(pr, rows, m->H);
  if (S > 1) {
    fin.f[nfin++] = bias_final(part_2, S, m->H, m->off_b2);
    fin.f[nfin++] = bias_final(part_1, S, m->H, m->off_b1);
  }
  if (nfin) {
    int n_max = 0;
    for (int i = 0; i < nfin; ++i) n_max = std::max(n_max, fin.f[i].N);
    colsum_final_kernel<<<dim3((n_max + 255) / 256, nfin), 256, 0, m->stream>>>(fin);
  }
  MLP_HIP(hipGetLastError());
  if (loss_acc) {  // the sum of squared errors added on the device, no wait
    loss_sum_kernel<<<1, 256, 0, m->stream>>>(fused ? m->hsq : m->partial, (int)n_loss, loss_acc);
    MLP_HIP(hipGetLastError());
  }
  if (loss) {
    std::vector<float> part(n_loss);
    MLP_HIP(hipMemcpyAsync(part.data(), fused ? m->hsq : m->partial, part.size() * sizeof(float),
                           hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    double s = 0;
    for (float v : part) s += v;
    *loss = s / (double)ny;
  }
  return GCM_OK;
}

}  // namespace

extern "C" {

int gcm_mlp_create(int device, int n_in, int n_hidden, int n_out, uint64_t seed,
                   gcm_mlp** out) {
  if (!out || n_in < 1 || n_hidden < 1 || n_out < 1) return fail(GCM_ERR_ARG, "invalid MLP shape");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    return fail(GCM_ERR_HIP, "no such HIP device");
  MLP_HIP(hipSetDevice(device));
  auto* m = new gcm_mlp();
  m->device = device;
  m->n_in = n_in;
  m->H = n_hidden;
  m->n_out = n_out;
  m->off_w1 = 0;
  m->off_b1 = m->off_w1 + (int64_t)n_hidden * n_in;
  m->off_w2 = m->off_b1 + n_hidden;
  m->off_b2 = m->off_w2 + (int64_t)n_hidden * n_hidden;
  m->off_w3 = m->off_b2 + n_hidden;
  m->off_b3 = m->off_w3 + (int64_t)n_out * n_hidden;
  m->n_params = m->off_b3 + n_out;
  const int n_in_pad = (n_in + 7) / 8 * 8;
  const int64_t bw2 = (int64_t)n_hidden * n_in_pad, bw3 = bw2 + (int64_t)n_hidden * n_hidden;
  m->bmap = Bf16Map{m->off_b1, m->off_w2, m->off_b2, m->off_w3, m->off_b3, bw2, bw3, n_in, n_in_pad};
  m->n_bf16 = bw3 + (int64_t)n_out * n_hidden;
  int rc = GCM_OK;
  auto step = [&](int r) { if (rc == GCM_OK) rc = r; };
  step(m->alloc(&m->p, m->n_params));
  step(m->alloc(&m->g, m->n_params));
  step(m->alloc(&m->m1, m->n_params));
  step(m->alloc(&m->m2, m->n_params));
  step(m->alloc(&m->pb, m->n_bf16));
  step(m->alloc(&m->partial, (size_t)(n_out + 63) / 64 * kMaxColSlices));  // mse_colsum blocks
  step(m->alloc(&m->colpart, (size_t)kMaxColSlices * (n_out + 2 * (size_t)n_hidden)));
  if (rc == GCM_OK && hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
    rc = fail(GCM_ERR_HIP, "stream creation failed");
  step(noalloc_scratch(device, &m->scratch));
  if (rc == GCM_OK && gemm_set_lds_attrs() != hipSuccess)
    rc = fail(GCM_ERR_HIP, "GEMM LDS attribute");
  if (rc != GCM_OK) {
    for (void* q : m->allocs) (void)hipFree(q);
    delete m;
    return rc;
  }
  // He-style initialisation, zero biases.
  std::vector<float> host(m->n_params, 0.0f);
  std::mt19937_64 rng(seed);
  auto init = [&](int64_t off, int64_t count, int fan_in) {
    std::normal_distribution<float> d(0.0f, std::sqrt(2.0f / fan_in));
    for (int64_t i = 0; i < count; ++i) host[off + i] = d(rng);
  };
  init(m->off_w1, (int64_t)n_hidden * n_in, n_in);
  init(m->off_w2, (int64_t)n_hidden * n_hidden, n_hidden);
  init(m->off_w3, (int64_t)n_out * n_hidden, n_hidden);
  MLP_HIP(hipMemcpy(m->p, host.data(), m->n_params * 4, hipMemcpyHostToDevice));
  // on the library's (non-blocking) stream, ordered before the BF16 refresh
  MLP_HIP(hipMemsetAsync(m->m1, 0, m->n_params * 4, m->stream));
  MLP_HIP(hipMemsetAsync(m->m2, 0, m->n_params * 4, m->stream));
  MLP_HIP(hipMemsetAsync(m->pb, 0, m->n_bf16 * sizeof(__bf16), m->stream));  // W1 row padding
  to_bf16_kernel<<<grid_for(m->n_params), 256, 0, m->stream>>>(m->p, m->pb, m->bmap, m->n_params);
  MLP_HIP(hipStreamSynchronize(m->stream));
  *out = m;
  return GCM_OK;
}

int gcm_mlp_destroy(gcm_mlp* m) {
  if (!m) return GCM_OK;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void* q : m->allocs) (void)hipFree(q);
  if (m->order_host) (void)hipHostFree(m->order_host);
  for (hipEvent_t e : m->order_ev)
    if (e) (void)hipEventDestroy(e);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return GCM_OK;
}

int gcm_mlp_param_count(gcm_mlp* m, int64_t* count) {
  if (!m || !count) return fail(GCM_ERR_ARG, "NULL argument");
  *count = m->n_params;
  return GCM_OK;
}

int gcm_mlp_get_params(gcm_mlp* m, float* out) {
  if (!m || !out) return fail(GCM_ERR_ARG, "NULL argument");
  MLP_HIP(hipSetDevice(m->device));
  MLP_HIP(hipStreamSynchronize(m->stream));
  MLP_HIP(hipMemcpy(out, m->p, m->n_params * 4, hipMemcpyDeviceToHost));
  return GCM_OK;
}

int gcm_mlp_set_params(gcm_mlp* m, const float* in) {
  if (!m || !in) return fail(GCM_ERR_ARG, "NULL argument");
  MLP_HIP(hipSetDevice(m->device));
  MLP_HIP(hipStreamSynchronize(m->stream));
  MLP_HIP(hipMemcpy(m->p, in, m->n_params * 4, hipMemcpyHostToDevice));
  to_bf16_kernel<<<grid_for(m->n_params), 256, 0, m->stream>>>(m->p, m->pb, m->bmap, m->n_params);
  MLP_HIP(hipStreamSynchronize(m->stream));
  return GCM_OK;
}

int gcm_mlp_predict(gcm_mlp* m, const float* x, int64_t rows, float* y) {
  if (!m || (rows && (!x || !y)) || rows < 0) return fail(GCM_ERR_ARG, "invalid arguments");
  MLP_HIP(hipSetDevice(m->device));
  const int chunk = 8192;
  int rc = m->ensure_batch((int)std::min<int64_t>(rows, chunk));
  if (rc) return rc;
  for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
    const int n = (int)std::min<int64_t>(chunk, rows - r0);
    MLP_HIP(hipMemcpyAsync(m->x, x + r0 * m->n_in, (size_t)n * m->n_in * 4, hipMemcpyHostToDevice, m->stream));
    if ((rc = forward(m, n))) return rc;
    MLP_HIP(hipMemcpyAsync(y + r0 * m->n_out, m->y, (size_t)n * m->n_out * 4, hipMemcpyDeviceToHost, m->stream));
  }
  MLP_HIP(hipStreamSynchronize(m->stream));
  return GCM_OK;
}

// Likelihood head over the network outputs (see head_llh_kernel).
int gcm_mlp_set_head(gcm_mlp* m, const double* mean, const double* scale,
                     const double* data, const double* var_ratio) {
  if (!m || !mean || !scale || !data || !var_ratio) return fail(GCM_ERR_ARG, "NULL argument");
  MLP_HIP(hipSetDevice(m->device));
  MLP_HIP(hipStreamSynchronize(m->stream));
  if (!m->head) {
    int rc = m->alloc(&m->head, (size_t)4 * m->n_out);
    if (rc) return rc;
  }
  const size_t n = (size_t)m->n_out * sizeof(double);
  MLP_HIP(hipMemcpy(m->head, mean, n, hipMemcpyHostToDevice));
  MLP_HIP(hipMemcpy(m->head + m->n_out, scale, n, hipMemcpyHostToDevice));
  MLP_HIP(hipMemcpy(m->head + 2 * (size_t)m->n_out, data, n, hipMemcpyHostToDevice));
  MLP_HIP(hipMemcpy(m->head + 3 * (size_t)m->n_out, var_ratio, n, hipMemcpyHostToDevice));
  return GCM_OK;
}

int gcm_mlp_predict_llh(gcm_mlp* m, const float* x, int64_t rows, double* chi2) {
  if (!m || (rows && (!x || !chi2)) || rows < 0) return fail(GCM_ERR_ARG, "invalid arguments");
  if (!m->head) return fail(GCM_ERR_STATE, "gcm_mlp_set_head first");
  MLP_HIP(hipSetDevice(m->device));
  const int chunk = 8192;
  int rc = m->ensure_batch((int)std::min<int64_t>(rows, chunk));
  if (rc) return rc;
  if (m->chi2_cap < std::min<int64_t>(rows, chunk)) {
    if (m->chi2) {
      MLP_HIP(hipStreamSynchronize(m->stream));
      (void)hipFree(m->chi2);
      m->allocs.erase(std::remove(m->allocs.begin(), m->allocs.end(), (void*)m->chi2), m->allocs.end());
      m->chi2 = nullptr;
    }
    m->chi2_cap = (int)std::min<int64_t>(rows, chunk);
    if ((rc = m->alloc(&m->chi2, m->chi2_cap))) return rc;
  }
  const double* h = m->head;
  const int64_t no = m->n_out;
  for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
    const int n = (int)std::min<int64_t>(chunk, rows - r0);
    MLP_HIP(hipMemcpyAsync(m->x, x + r0 * m->n_in, (size_t)n * m->n_in * 4, hipMemcpyHostToDevice, m->stream));
    if ((rc = forward(m, n))) return rc;
    head_llh_kernel<<<n, kHeadBlock, 0, m->stream>>>(m->y, m->n_out, h, h + no, h + 2 * no,
                                                     h + 3 * no, m->chi2);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipMemcpyAsync(chi2 + r0, m->chi2, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  }
  MLP_HIP(hipStreamSynchronize(m->stream));
  return GCM_OK;
}

// gcm_mlp_predict_llh without allocation: device input and output, every
// intermediate in this GPU's `scratch` array (through the address resolved
// at gcm_mlp_create), enqueued on `stream` without synchronising.  Layout of
// the scratch, each part 256-byte aligned: z1, h1, z2, h2 [rows][H], y
// [rows][n_out], then the split-K partials of the largest split GEMM.  The
// same kernels with the same split decisions as gcm_mlp_predict_llh, so the
// results are the same for the same rows (h1, h2 in BF16 in their FP32-sized
// slots).  Nothing in the scratch needs to
// start at zero (every kernel writes what it later reads).
int gcm_mlp_predict_llh_noalloc(gcm_mlp* m, const float* x_device, int64_t rows,
                                double* chi2_device, void* stream_ptr) {
  if (!m || rows < 0 || (rows && (!x_device || !chi2_device)))
    return fail(GCM_ERR_ARG, "invalid arguments");
  if (!m->head) return fail(GCM_ERR_STATE, "gcm_mlp_set_head first");
  if (!m->scratch) return fail(GCM_ERR_STATE, "no scratch address for this device");
  if (rows == 0) return GCM_OK;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const int r = (int)std::min<int64_t>(rows, INT32_MAX);
  const size_t act = up((size_t)r * m->H * 4), out = up((size_t)r * m->n_out * 4);
  const size_t ws = std::max({gemm_ws_floats(r, m->H, m->n_in), gemm_ws_floats(r, m->H, m->H),
                              gemm_ws_floats(r, m->n_out, m->H)});
  const size_t need = 4 * act + out + up(ws * 4);
  if (rows > INT32_MAX || need > kNoAllocScratchBytes)
    return fail(GCM_ERR_ARG, "gcm_mlp_predict_llh_noalloc: " + std::to_string(rows) +
                                 " rows need " + std::to_string(need) + " bytes of scratch, more than " +
                                 std::to_string(kNoAllocScratchBytes));
  MLP_HIP(hipSetDevice(m->device));
  unsigned char* base = m->scratch;
  float* z1 = reinterpret_cast<float*>(base);
  __bf16* h1 = reinterpret_cast<__bf16*>(base + act);
  float* z2 = reinterpret_cast<float*>(base + 2 * act);
  __bf16* h2 = reinterpret_cast<__bf16*>(base + 3 * act);
  float* y = reinterpret_cast<float*>(base + 4 * act);
  const Exec ex{stream_ptr ? (hipStream_t)stream_ptr : m->stream,
                reinterpret_cast<float*>(base + 4 * act + out), ws};
  int rc = forward(m, r, x_device, z1, h1, z2, h2, y, &ex);
  if (rc) return rc;
  const double* h = m->head;
  const int64_t no = m->n_out;
  head_llh_kernel<<<r, kHeadBlock, 0, ex.stream>>>(y, m->n_out, h, h + no, h + 2 * no, h + 3 * no,
                                                   chi2_device);
  MLP_HIP(hipGetLastError());
  return GCM_OK;
}

// Loss and gradient of the mean squared error on the given rows (no update).
int gcm_mlp_gradients(gcm_mlp* m, const float* x, const float* t, int rows, float* grad,
                      double* loss) {
  if (!m || !x || !t || rows < 1) return fail(GCM_ERR_ARG, "invalid arguments");
  MLP_HIP(hipSetDevice(m->device));
  int rc = m->ensure_batch(rows);
  if (rc) return rc;
  MLP_HIP(hipMemcpyAsync(m->x, x, (size_t)rows * m->n_in * 4, hipMemcpyHostToDevice, m->stream));
  MLP_HIP(hipMemcpyAsync(m->t, t, (size_t)rows * m->n_out * 4, hipMemcpyHostToDevice, m->stream));
  if ((rc = forward_train(m, rows, m->t, nullptr))) return rc;
  if ((rc = backward(m, rows, loss))) return rc;
  if (grad) {
    MLP_HIP(hipMemcpyAsync(grad, m->g, m->n_params * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
  }
  return GCM_OK;
}

// Upload a training set (rows of normalised inputs and targets).
int gcm_mlp_set_data(gcm_mlp* m, const float* x, const float* t, int64_t rows) {
  if (!m || !x || !t || rows < 1) return fail(GCM_ERR_ARG, "invalid arguments");
  MLP_HIP(hipSetDevice(m->device));
  MLP_HIP(hipStreamSynchronize(m->stream));
  if (m->dx) {
    for (void* q : {(void*)m->dx, (void*)m->dt, (void*)m->order}) {
      (void)hipFree(q);
      m->allocs.erase(std::remove(m->allocs.begin(), m->allocs.end(), q), m->allocs.end());
    }
    m->dx = m->dt = nullptr;
    m->order = nullptr;
  }
  if (m->order_host) {
    (void)hipHostFree(m->order_host);
    m->order_host = nullptr;
  }
  int rc = m->alloc(&m->dx, (size_t)rows * m->n_in);
  if (rc == GCM_OK) rc = m->alloc(&m->dt, (size_t)rows * m->n_out);
  if (rc == GCM_OK) rc = m->alloc(&m->order, 2 * (size_t)rows);
  if (rc == GCM_OK && hipHostMalloc((void**)&m->order_host, 2 * (size_t)rows * sizeof(int32_t),
                                    hipHostMallocDefault) != hipSuccess) {
    m->order_host = nullptr;
    rc = fail(GCM_ERR_NOMEM, "hipHostMalloc failed (mlp)");
  }
  for (int h = 0; h < 2 && rc == GCM_OK; ++h)
    if (!m->order_ev[h] &&
        hipEventCreateWithFlags(&m->order_ev[h], hipEventDisableTiming) != hipSuccess)
      rc = fail(GCM_ERR_HIP, "event creation failed (mlp)");
  if (rc) return rc;
  MLP_HIP(hipMemcpy(m->dx, x, (size_t)rows * m->n_in * 4, hipMemcpyHostToDevice));
  MLP_HIP(hipMemcpy(m->dt, t, (size_t)rows * m->n_out * 4, hipMemcpyHostToDevice));
  m->n_data = rows;
  return GCM_OK;
}

// Adam epochs over the resident training set in shuffled minibatches; the
// mean minibatch loss of each epoch goes to loss_out[epoch] (optional).
int gcm_mlp_train(gcm_mlp* m, int epochs, int batch, float lr, uint64_t seed,
                  double* loss_out) {
  if (!m || epochs < 0 || batch < 1) return fail(GCM_ERR_ARG, "invalid arguments");
  if (!m->dx) return fail(GCM_ERR_STATE, "gcm_mlp_set_data first");
  if ((int64_t)batch * std::max(m->n_in, m->n_out) >= (int64_t(1) << 30))
    return fail(GCM_ERR_ARG, "minibatch too large (rows x columns >= 2^30)");
  MLP_HIP(hipSetDevice(m->device));
  int rc = m->ensure_batch(batch);
  if (rc) return rc;
  std::vector<int32_t> order(m->n_data);
  for (int64_t i = 0; i < m->n_data; ++i) order[i] = (int32_t)i;
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const int64_t n = m->n_data;
  const int nb = (int)(n / batch);     // minibatches per epoch
  // W3 takes its Adam step in the epilogue of its gradient's product when its
  // range is aligned for the 16-byte paths (adam_kernel skips it),
  const bool w3_fused = m->off_w3 % 4 == 0 && (m->off_b3 - m->off_w3) % 4 == 0;
  // and every parameter takes its step in the backward pass (no Adam launch)
  // when W2's range is aligned too and the loss head is fused
  const bool all_fused = w3_fused && m->off_w2 % 4 == 0 && head_fused(m, batch);
  const int n_sampled = (nb + 15) / 16;  // of which the loss is taken: every 16th
  if (loss_out && epochs > m->loss_cap) {
    MLP_HIP(hipStreamSynchronize(m->stream));
    if (m->loss_dev) {
      (void)hipFree(m->loss_dev);
      m->allocs.erase(std::remove(m->allocs.begin(), m->allocs.end(), (void*)m->loss_dev),
                      m->allocs.end());
      m->loss_dev = nullptr;
    }
    m->loss_cap = 0;
    if ((rc = m->alloc(&m->loss_dev, (size_t)epochs))) return rc;
    m->loss_cap = epochs;
  }
  if (loss_out && epochs > 0)
    MLP_HIP(hipMemsetAsync(m->loss_dev, 0, (size_t)epochs * sizeof(double), m->stream));
  // Epoch ep's order: the next shuffle of the running permutation, in half
  // ep % 2 of the device and pinned host buffers.  It is shuffled and queued
  // while the device still runs the previous epoch; the half's host copy is
  // rewritten only once its copy two epochs back has completed.
  auto stage = [&](int ep) -> int {
    const int h = ep & 1;
    MLP_HIP(hipEventSynchronize(m->order_ev[h]));
    std::shuffle(order.begin(), order.end(), rng);
    std::copy(order.begin(), order.end(), m->order_host + h * n);
    MLP_HIP(hipMemcpyAsync(m->order + h * n, m->order_host + h * n, (size_t)n * 4,
                           hipMemcpyHostToDevice, m->stream));
    MLP_HIP(hipEventRecord(m->order_ev[h], m->stream));
    return GCM_OK;
  };
  if (epochs > 0 && (rc = stage(0))) return rc;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int nbatch = 0; nbatch < nb; ++nbatch) {
      const int32_t* idx = m->order + (ep & 1) * n + (int64_t)nbatch * batch;
      // the minibatch's input rows (and target rows, unless the loss head
      // reads them from the training set itself)
      const bool fused = head_fused(m, batch);
      const int ct = fused ? 0 : m->n_out;
      gather_rows_kernel<<<grid_for((int64_t)batch * ((std::max(m->n_in, ct) + 3) / 4)), 256, 0,
                           m->stream>>>(m->dx, m->dt, idx, batch, m->n_in, ct, m->x, m->t);
      if ((rc = forward_train(m, batch, m->dt, idx))) return rc;
      // the loss of every 16th minibatch, summed on the device
      const bool want = loss_out != nullptr && (nbatch % 16) == 0;
      m->adam_t++;
      const float c1 = 1.0f - std::pow(b1, (float)m->adam_t);
      const float c2 = 1.0f - std::pow(b2, (float)m->adam_t);
      const FusedAdam fa{{lr, b1, b2, eps, c1, c2}, all_fused};
      if ((rc = backward(m, batch, nullptr, want ? m->loss_dev + ep : nullptr,
                         w3_fused ? &fa : nullptr)))
        return rc;
      if (!all_fused) {
        // Adam on the parameters outside W3 (all of them without W3's step)
        const int64_t gap = w3_fused ? m->off_b3 - m->off_w3 : 0;
        adam_kernel<<<grid_for((m->n_params - gap + 3) / 4), 256, 0, m->stream>>>(
            m->p, m->g, m->m1, m->m2, m->pb, m->bmap, m->n_params, w3_fused ? m->off_w3 : 0,
            gap, lr, b1, b2, eps, c1, c2);
        MLP_HIP(hipGetLastError());
      }
    }
    if (ep + 1 < epochs && (rc = stage(ep + 1))) return rc;
  }
  if (loss_out && epochs > 0) {
    std::vector<double> acc(epochs);
    MLP_HIP(hipMemcpyAsync(acc.data(), m->loss_dev, (size_t)epochs * sizeof(double),
                           hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    const double per = (double)batch * m->n_out * std::max(1, n_sampled);
    for (int ep = 0; ep < epochs; ++ep) loss_out[ep] = nb > 0 ? acc[ep] / per : 0.0;
  }
  MLP_HIP(hipStreamSynchronize(m->stream));
  return GCM_OK;
}

}  // extern "C"
