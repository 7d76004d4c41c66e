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

// Positive-double atomic min through the IEEE bit pattern (monotone for x >= 0).
__device__ __forceinline__ void atomic_min_pos(double *addr, double v) {
  atomicMin(reinterpret_cast<unsigned long long *>(addr),
            (unsigned long long)__double_as_longlong(v));
}

}  // namespace aph
