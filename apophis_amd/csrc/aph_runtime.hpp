// aph_runtime.hpp -- host-side objects behind the C ABI: context, block set, plan.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/apophis_hydro.h"

struct aph_context {
  hipStream_t stream = nullptr;
  int device = 0;
  std::string err;
  double *d_scalar = nullptr;  // small device scratch for reductions (aph_global_sum)
  int64_t d_scalar_n = 0;
};

struct aph_blockset {
  std::vector<aph_block> h;  // host copy of the table
  aph_block *d = nullptr;    // device copy
  int32_t nblocks = 0, ndim = 0, nscalars = 0, nvar = 0;
  int64_t ncells = 0;
  int64_t max_ncell = 0;       // max entire cells per block
  int64_t max_interior = 0;    // max interior cells per block
  int64_t max_pencils[3] = {0, 0, 0};  // reference loop pencils per block per axis
  int32_t max_ext[3] = {0, 0, 0};
};

namespace aph {

int set_error(aph_context *ctx, int code, const char *fmt, ...);

// aph_global_sum_noalloc's fixed scratch: partial sums for at most this many blocks
// (512 KiB per device), and its per-device set-up (aph_plain.hip).
constexpr int kNoAllocMaxBlocks = 65536;
int noalloc_scratch_init(int device);

#define APH_HIP_CHECK(ctx, call)                                                      \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return aph::set_error((ctx), APH_E_HIP, "%s failed: %s", #call,                 \
                            hipGetErrorString(e_));                                   \
  } while (0)

#define APH_LAUNCH_CHECK(ctx)                                                         \
  do {                                                                                \
    hipError_t e_ = hipGetLastError();                                                \
    if (e_ != hipSuccess)                                                             \
      return aph::set_error((ctx), APH_E_HIP, "kernel launch failed: %s",             \
                            hipGetErrorString(e_));                                   \
  } while (0)

// Monopole gravity's whole pass over every block of bs (aph_gravity.hip): bins from the
// density of `field` floored at dfloor, tables, and the field (3 components) into grav.
int monopole_pass(aph_context *ctx, aph_monopole *m, const aph_blockset *bs,
                  const double *field, double dfloor, double *grav);

// Poisson gravity's whole pass (aph_poisson.hip) on the density of `field` (floored at
// dfloor): grav (3 components) and, if non-null, phi.
int poisson_pass(aph_context *ctx, aph_poisson *ps, const double *field, double dfloor,
                 double *grav, double *phi);
const aph_blockset *poisson_blockset(const aph_poisson *ps);

// Resolve the (first_block, nblocks) sub-range convention of the plain entry points.
int block_range(aph_context *ctx, const aph_blockset *bs, int32_t first, int32_t n,
                int32_t *lo, int32_t *cnt);

}  // namespace aph
