// probe.hip -- streaming-read probe (rts_probe.h): how fast the device reads a buffer once
// with the scan kernels' load pattern (non-temporal uint4 per lane), for grids and unroll
// depths swept by the benchmark.
#include "common.h"
#include "rts_probe.h"

namespace rts {

template <int U>
__global__ __launch_bounds__(kBlock) void k_probe_read(const uint4 *p, uint64_t n16, uint4 *sink) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  uint4 acc = make_uint4(0, 0, 0, 0);
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  for (; i + (U - 1) * stride < n16; i += U * stride) {
    uint4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = load_stream(p + i + u * stride);
#pragma unroll
    for (int u = 0; u < U; u++) {
      acc.x ^= v[u].x;
      acc.y ^= v[u].y;
      acc.z ^= v[u].z;
      acc.w ^= v[u].w;
    }
  }
  for (; i < n16; i += stride) {
    const uint4 v = load_stream(p + i);
    acc.x ^= v.x;
    acc.y ^= v.y;
    acc.z ^= v.z;
    acc.w ^= v.w;
  }
  // a lane whose XOR is zero writes nothing; otherwise the store keeps the loads alive
  if ((acc.x | acc.y | acc.z | acc.w) == 0xFFFFFFFFu) *sink = acc;
}

}  // namespace rts

extern "C" hipError_t rts_probe_read(const void *buf, size_t bytes, int blocks_per_cu, int unroll,
                                     unsigned long long *sink, hipStream_t stream) {
  using namespace rts;
  if (!buf || !sink || bytes % 16 || ((uintptr_t)buf & 15) || blocks_per_cu < 1 || blocks_per_cu > 64)
    return set_error(hipErrorInvalidValue, "rts_probe_read");
  const unsigned grid = (unsigned)(device_cu_count() * blocks_per_cu);
  const uint4 *p = static_cast<const uint4 *>(buf);
  uint4 *s = reinterpret_cast<uint4 *>(sink);
  switch (unroll) {
    case 1: hipLaunchKernelGGL(k_probe_read<1>, dim3(grid), dim3(kBlock), 0, stream, p, bytes / 16, s); break;
    case 2: hipLaunchKernelGGL(k_probe_read<2>, dim3(grid), dim3(kBlock), 0, stream, p, bytes / 16, s); break;
    case 4: hipLaunchKernelGGL(k_probe_read<4>, dim3(grid), dim3(kBlock), 0, stream, p, bytes / 16, s); break;
    case 8: hipLaunchKernelGGL(k_probe_read<8>, dim3(grid), dim3(kBlock), 0, stream, p, bytes / 16, s); break;
    default: return set_error(hipErrorInvalidValue, "rts_probe_read: unroll");
  }
  return set_error(hipGetLastError(), "rts_probe_read");
}
