// Spatial ordering of particle storage.
//
// Swimmer and tracer fields live on the device in a private "slot" order that
// keeps particles that are close in space close in memory, so that the 64
// particles of a wavefront gather their 4x4x4 velocity stencils from the same
// L2 lines.  perm[slot] is the user's index of the particle in each slot; the
// host mirrors stay in user order and H2D/D2H permute on the device.
// Re-ordering is a stable radix sort of Morton keys of the containing site,
// so it is deterministic.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "d3q15.hpp"
#include "internal.hpp"

namespace lbs {
namespace {

constexpr int BLOCK = 256;

__device__ __forceinline__ uint32_t spread_bits10(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__device__ __forceinline__ uint32_t site_coord(double x, int n) {
  if (!(x > -2147483648.0 && x < 2147483647.0)) return 0;
  int v = int(floor(x)) % n;
  return uint32_t(v < 0 ? v + n : v);
}

__global__ __launch_bounds__(BLOCK) void k_morton_keys(int np, const double* __restrict__ r, Geom g,
                                                       uint32_t* __restrict__ keys, int* __restrict__ idx) {
  const int s = int(blockIdx.x * BLOCK + threadIdx.x);
  if (s >= np) return;
  const uint32_t x = site_coord(r[s], g.nx), y = site_coord(r[size_t(np) + s], g.ny),
                 z = site_coord(r[2 * size_t(np) + s], g.nz);
  keys[s] = (spread_bits10(x) << 2) | (spread_bits10(y) << 1) | spread_bits10(z);
  idx[s] = s;
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_gather(int np, int ncomp, const int* __restrict__ order,
                                                  const T* __restrict__ src, T* __restrict__ dst) {
  const int t = int(blockIdx.x * BLOCK + threadIdx.x);
  if (t >= np) return;
  const int from = order[t];
  for (int c = 0; c < ncomp; ++c) dst[size_t(c) * np + t] = src[size_t(c) * np + from];
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_scatter(int np, int ncomp, const int* __restrict__ order,
                                                   const T* __restrict__ src, T* __restrict__ dst) {
  const int t = int(blockIdx.x * BLOCK + threadIdx.x);
  if (t >= np) return;
  const int to = order[t];
  for (int c = 0; c < ncomp; ++c) dst[size_t(c) * np + to] = src[size_t(c) * np + t];
}

__global__ __launch_bounds__(BLOCK) void k_iota(int np, int* __restrict__ p) {
  const int t = int(blockIdx.x * BLOCK + threadIdx.x);
  if (t < np) p[t] = t;
}

inline unsigned blocks(int n) { return unsigned((n + BLOCK - 1) / BLOCK); }

}  // namespace

hipError_t launch_iota(int np, int* p, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3(blocks(np)), dim3(BLOCK), 0, s, np, p);
  return hipGetLastError();
}

hipError_t sort_scratch_bytes(int np, size_t* bytes) {
  *bytes = 0;
  if (np <= 0) return hipSuccess;
  return rocprim::radix_sort_pairs(nullptr, *bytes, static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                   static_cast<int*>(nullptr), static_cast<int*>(nullptr), size_t(np), 0, 30);
}

hipError_t launch_spatial_order(int np, const double* r, const Geom& g, uint32_t* keys, uint32_t* keys_alt,
                                int* idx, int* order_out, void* scratch, size_t scratch_bytes, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton_keys, dim3(blocks(np)), dim3(BLOCK), 0, s, np, r, g, keys, idx);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t bytes = scratch_bytes;
  return rocprim::radix_sort_pairs(scratch, bytes, keys, keys_alt, idx, order_out, size_t(np), 0, 30, s);
}

template <typename T>
hipError_t launch_gather(int np, int ncomp, const int* order, const T* src, T* dst, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather<T>, dim3(blocks(np)), dim3(BLOCK), 0, s, np, ncomp, order, src, dst);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scatter(int np, int ncomp, const int* order, const T* src, T* dst, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter<T>, dim3(blocks(np)), dim3(BLOCK), 0, s, np, ncomp, order, src, dst);
  return hipGetLastError();
}

template hipError_t launch_gather<uint64_t>(int, int, const int*, const uint64_t*, uint64_t*, hipStream_t);
template hipError_t launch_gather<int>(int, int, const int*, const int*, int*, hipStream_t);
template hipError_t launch_scatter<uint64_t>(int, int, const int*, const uint64_t*, uint64_t*, hipStream_t);

}  // namespace lbs
