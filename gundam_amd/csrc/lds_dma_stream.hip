// lds_dma_stream.hip -- calibration program for the fill kernel's tile
// stream: how fast can the LDS-DMA load (global_load_lds_dwordx4, 16 bytes per
// lane, the instruction the fill kernel's producer waves use) stream HBM into
// LDS, as a function of the bytes each CU keeps in flight?
//
// Each block has one DMA wave that copies tiles of T bytes into a ring of D
// LDS buffers, keeping D tiles in flight (s_waitcnt vmcnt with the loads of
// D - 1 tiles still outstanding), grid-strided over a 4 GiB buffer; blocks per
// CU are swept.  Cache policy 0 (default) and 2 (non-temporal, as the fill
// kernel streams) are both reported.  A plain VGPR read stream is the
// reference.
//   usage: lds_dma_stream
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef __attribute__((address_space(3))) void* lds_ptr_t;
typedef __attribute__((address_space(1))) void* glb_ptr_t;

template <int T, int D, int POL>
__global__ void dma_ring(const unsigned char* __restrict__ src, size_t ntiles,
                         unsigned long long* sink) {
  extern __shared__ __align__(16) unsigned char lds[];
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  int it = 0;
  for (size_t c = blockIdx.x; c < ntiles; c += gridDim.x, ++it) {
    unsigned char* dst = lds + (size_t)(it % D) * T;
    const unsigned char* s = src + c * T;
#pragma unroll
    for (int j = 0; j < T; j += 1024)
      __builtin_amdgcn_global_load_lds((glb_ptr_t)(s + j + lane * 16), (lds_ptr_t)(dst + j), 16,
                                       0, POL);
    if constexpr (D == 1) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    else if constexpr (D == 2) asm volatile("s_waitcnt vmcnt(%0)" ::"n"(T / 1024) : "memory");
    else asm volatile("s_waitcnt vmcnt(%0)" ::"n"(2 * T / 1024) : "memory");
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  if (lds[lane] == 0x5a && lane == 999) sink[0] = 1;  // keeps the loads live
}

__global__ void read_vgpr(const uint4* __restrict__ src, size_t nchunks,
                          unsigned long long* sink) {
  unsigned acc = 0;
  for (size_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const uint4* p = src + c * 1024;
#pragma unroll 4
    for (int i = threadIdx.x; i < 1024; i += blockDim.x) {
      const uint4 v = p[i];
      acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
  }
  if (acc == 0x12345678u) sink[0] = acc;
}

static const size_t kBytes = 4ull << 30;
static unsigned char* g_src = nullptr;
static unsigned long long* g_sink = nullptr;
static hipEvent_t g_e0, g_e1;

template <typename F>
static double best_gbs(F launch) {
  launch();
  if (hipDeviceSynchronize() != hipSuccess) exit(1);
  float best = 1e30f;
  for (int r = 0; r < 3; ++r) {
    if (hipEventRecord(g_e0) != hipSuccess) exit(1);
    launch();
    if (hipEventRecord(g_e1) != hipSuccess || hipEventSynchronize(g_e1) != hipSuccess) exit(1);
    float ms = 0;
    if (hipEventElapsedTime(&ms, g_e0, g_e1) != hipSuccess) exit(1);
    if (ms < best) best = ms;
  }
  return kBytes / (best * 1e-3) / 1e9;
}

template <int T, int D, int POL>
static void sweep(int cus) {
  const size_t lds = (size_t)T * D;
  if (hipFuncSetAttribute((const void*)dma_ring<T, D, POL>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
    exit(1);
  for (int bpc : {1, 2, 3, 4, 6}) {
    if (lds * bpc > 160 * 1024) continue;
    const double g = best_gbs([&] {
      dma_ring<T, D, POL><<<cus * bpc, 64, lds>>>(g_src, kBytes / T, g_sink);
    });
    printf("dma tile=%5d in-flight tiles/block=%d policy=%d blocks/CU=%d in-flight/CU=%3zu KiB : %5.0f GB/s\n",
           T, D, POL, bpc, (size_t)bpc * T * D / 1024, g);
    fflush(stdout);
  }
}

int main() {
  if (hipMalloc((void**)&g_src, kBytes) != hipSuccess ||
      hipMalloc((void**)&g_sink, 8) != hipSuccess || hipMemset(g_src, 1, kBytes) != hipSuccess ||
      hipEventCreate(&g_e0) != hipSuccess || hipEventCreate(&g_e1) != hipSuccess) {
    fprintf(stderr, "setup failed\n");
    return 1;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess) return 1;
  for (int bpc : {4, 8}) {
    const double g = best_gbs([&] {
      read_vgpr<<<cus * bpc, 512>>>((const uint4*)g_src, kBytes / 16384, g_sink);
    });
    printf("vgpr read stream, blocks/CU=%d x 512 threads : %5.0f GB/s\n", bpc, g);
  }
  sweep<16384, 1, 0>(cus);
  sweep<16384, 2, 0>(cus);
  sweep<24576, 1, 0>(cus);
  sweep<24576, 2, 0>(cus);
  sweep<16384, 2, 2>(cus);
  sweep<24576, 1, 2>(cus);
  sweep<24576, 2, 2>(cus);
  return 0;
}
