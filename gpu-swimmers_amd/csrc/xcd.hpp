// XCD-aware workgroup numbering.
//
// MI355X hands workgroups to its 8 XCDs (each with its own L2) round-robin by
// workgroup id.  Kernels whose neighbouring work items share cache lines give
// each XCD one contiguous range of the `nb` logical blocks instead: launched
// with round_xcd(nb) workgroups, workgroup b takes block
// (b % 8) * round_xcd(nb)/8 + b / 8 (a block >= nb means nothing to do).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lbs {

constexpr uint32_t XCDS = 8;
__host__ __device__ inline uint32_t round_xcd(uint32_t nb) { return (nb + XCDS - 1) / XCDS * XCDS; }
__device__ __forceinline__ uint32_t xcd_block(uint32_t nb8) {
  return (blockIdx.x % XCDS) * (nb8 / XCDS) + blockIdx.x / XCDS;
}

}  // namespace lbs
