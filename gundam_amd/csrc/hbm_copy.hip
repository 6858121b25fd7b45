// hbm_copy.hip -- calibration program for the HBM roofline: streams N bytes
// from one device buffer to another with 16-byte vector loads/stores and
// reports the achieved bandwidth.  Under `rocprofv3 --pmc FETCH_SIZE` (or
// WRITE_SIZE) its known traffic (N bytes read, N bytes written per launch)
// calibrates the counters used to measure the fill kernel's HBM traffic.
//   usage: hbm_copy [bytes=1073741824] [reps=5]
// The launch is swept over grid sizes (blocks per CU) and the best time is
// reported as the achievable copy bandwidth.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                            size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

int main(int argc, char** argv) {
  const size_t bytes = argc > 1 ? strtoull(argv[1], nullptr, 10) : (1ull << 30);
  const int reps = argc > 2 ? atoi(argv[2]) : 5;
  const size_t n = bytes / 16;
  uint4 *a = nullptr, *b = nullptr;
  if (hipMalloc((void**)&a, n * 16) != hipSuccess || hipMalloc((void**)&b, n * 16) != hipSuccess) {
    fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  if (hipMemset(a, 1, n * 16) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1;
  float best = 1e30f;
  int best_mult = 0, launches = 0;
  for (int mult : {2, 4, 8, 16, 32}) {
    const int grid = cus * mult;
    copy_kernel<<<grid, 256>>>(a, b, n);  // warm-up
    ++launches;
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    for (int r = 0; r < reps; ++r) {
      if (hipEventRecord(e0) != hipSuccess) return 1;
      copy_kernel<<<grid, 256>>>(a, b, n);
      ++launches;
      if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
      float ms = 0;
      if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
      if (ms < best) {
        best = ms;
        best_mult = mult;
      }
    }
  }
  printf("hbm_copy bytes=%zu launches=%d best_ms=%.4f blocks_per_cu=%d read+write_GBps=%.1f\n",
         n * 16, launches, best, best_mult, 2.0 * n * 16 / (best * 1e-3) / 1e9);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
