// probe.hpp -- measurement kernels (probe.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dsa {

// blocks x 256 lanes, each doing `rounds` x 24 random gathers from table[P*P].
hipError_t launch_gather_probe(const uint16_t *table, uint32_t P, uint32_t blocks, uint32_t rounds, uint32_t *sink,
                               hipStream_t s);
// dst <- src, bytes a multiple of 16.
hipError_t launch_copy_probe(const void *src, void *dst, size_t bytes, hipStream_t s);

}  // namespace dsa
