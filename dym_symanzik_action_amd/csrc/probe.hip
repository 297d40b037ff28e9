// probe.hip -- measurement kernels for the roofline of the sweep.
//
// k_gather_probe: random uint16 gathers from a P x P table, GPL independent
// gathers in flight per lane, each in a different (random) row, so that each
// one is a separate line as the sweep's staple products are: the measured
// random-gather ceiling (L2 requests) the sweep is bound by.
// k_copy_probe: a plain device copy of known size, the scale for the
// FETCH_SIZE / WRITE_SIZE hardware counters.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "probe.hpp"

namespace dsa {

namespace {

template <int GPL>
__global__ __launch_bounds__(256) void k_gather_probe(const uint16_t *__restrict__ table, uint32_t P, uint32_t rounds,
                                                      uint32_t *sink) {
  uint32_t st[GPL];
  const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
  for (int j = 0; j < GPL; ++j) st[j] = (lane * 2654435761u + (uint32_t)j * 40503u) % P;
  uint32_t col = (lane * 7919u) % P;
  uint32_t acc = 0;
  for (uint32_t r = 0; r < rounds; ++r) {
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < GPL; ++j) s += table[st[j] * P + col];
    acc ^= s;
    col = s % P;  // the next column depends on this round, like the next proposal
  }
  sink[lane] = acc;  // keeps every gather live
}

__global__ void k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

hipError_t launch_gather_probe(const uint16_t *table, uint32_t P, uint32_t blocks, uint32_t rounds, uint32_t *sink,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_gather_probe<24>, dim3(blocks), dim3(256), 0, s, table, P, rounds, sink);
  return hipGetLastError();
}

hipError_t launch_copy_probe(const void *src, void *dst, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 16;
  uint32_t blocks = (uint32_t)((n + 255) / 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, s, (const uint4 *)src, (uint4 *)dst, n);
  return hipGetLastError();
}

}  // namespace dsa
