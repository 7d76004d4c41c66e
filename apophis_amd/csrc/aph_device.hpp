// aph_device.hpp -- device-side helpers of the hydro flux path (gfx950, FP64):
// pack-layout arithmetic and reductions, on top of the point functions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "aph_point.hpp"

namespace aph {

// ------------------------------------------------------------ layout
__host__ __device__ __forceinline__ int ext(const aph_block &b, int d, int ndim) {
  return d < ndim ? b.nx[d] + 2 * b.ng : b.nx[d];
}
__host__ __device__ __forceinline__ int gst(const aph_block &b, int d, int ndim) {
  return d < ndim ? b.ng : 0;
}
__host__ __device__ __forceinline__ int64_t ncell(const aph_block &b, int ndim) {
  return (int64_t)ext(b, 0, ndim) * ext(b, 1, ndim) * ext(b, 2, ndim);
}

// *addr = min(*addr, v) for positive doubles, through the IEEE bit pattern (monotone for
// x >= 0); the atomic is skipped when *addr already holds a value no larger (a relaxed
// read first).
__device__ __forceinline__ void fold_min_pos(double *addr, double v) {
  unsigned long long *a = reinterpret_cast<unsigned long long *>(addr);
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  if (bits < __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMin(a, bits);
}

// Min of v over a workgroup of NT threads (every thread calls it), folded into *addr
// (positive doubles): one atomic per workgroup, and none when *addr already holds a value
// no larger -- thousands of workgroups folding into one address otherwise serialise on it.
// The result is the exact minimum whatever the order.
template <int NT>
__device__ __forceinline__ void workgroup_min_pos(double *addr, double v) {
  static_assert(NT % 64 == 0 && NT <= 1024, "whole waves");
  __shared__ double part[NT / 64];
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 1; k < NT / 64; ++k) v = fmin(v, part[k]);
    fold_min_pos(addr, v);
  }
}

}  // namespace aph
