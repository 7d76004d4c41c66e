// aph_runtime.hip -- context, block set and error handling of the C ABI.
#include "aph_runtime.hpp"

#include <algorithm>
#include <cstring>
#include <new>

#include "aph_device.hpp"

namespace aph {

int set_error(aph_context *ctx, int code, const char *fmt, ...) {
  if (ctx) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

int block_range(aph_context *ctx, const aph_blockset *bs, int32_t first, int32_t n,
                int32_t *lo, int32_t *cnt) {
  if (!bs) return set_error(ctx, APH_E_INVALID, "null block set");
  if (first < 0 || first > bs->nblocks)
    return set_error(ctx, APH_E_INVALID, "first_block %d out of range [0,%d]", first,
                     bs->nblocks);
  int32_t c = (n < 0) ? bs->nblocks - first : n;
  if (c < 0 || first + c > bs->nblocks)
    return set_error(ctx, APH_E_INVALID, "block range [%d,+%d) exceeds %d blocks", first, n,
                     bs->nblocks);
  *lo = first;
  *cnt = c;
  return APH_OK;
}

}  // namespace aph

extern "C" {

int aph_abi_version(void) { return APH_ABI_VERSION; }

int aph_context_create(void *stream, aph_context **out) {
  if (!out) return APH_E_INVALID;
  aph_context *c = new (std::nothrow) aph_context();
  if (!c) return APH_E_NOMEM;
  c->stream = static_cast<hipStream_t>(stream);
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return APH_E_HIP;
  }
  // aph_global_sum_noalloc's scratch address on this device; a failed lookup is kept
  // and reported by that entry point alone, so every other entry point still works
  (void)aph::noalloc_scratch_init(c->device);
  *out = c;
  return APH_OK;
}

int aph_context_destroy(aph_context *ctx) {
  if (!ctx) return APH_OK;
  if (ctx->d_scalar) (void)hipFree(ctx->d_scalar);
  delete ctx;
  return APH_OK;
}

const char *aph_last_error(const aph_context *ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

int aph_context_synchronize(aph_context *ctx) {
  if (!ctx) return APH_E_INVALID;
  APH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return APH_OK;
}

int aph_blockset_create(aph_context *ctx, const aph_block *blocks, int32_t nblocks,
                        int32_t ndim, int32_t nscalars, aph_blockset **out) {
  using namespace aph;
  if (!ctx || !out || (!blocks && nblocks > 0) || nblocks < 0)
    return set_error(ctx, APH_E_INVALID, "aph_blockset_create: bad arguments");
  if (ndim < 1 || ndim > 3) return set_error(ctx, APH_E_INVALID, "ndim must be 1..3");
  if (nscalars != 0 && nscalars != 1 && nscalars != 7 && nscalars != 9)
    return set_error(ctx, APH_E_INVALID,
                     "nscalars=%d unsupported (0, 1, 7 = ONe composition, 9 = ONe + one level set)",
                     nscalars);
  aph_blockset *bs = new (std::nothrow) aph_blockset();
  if (!bs) return set_error(ctx, APH_E_NOMEM, "host allocation failed");
  bs->h.assign(blocks, blocks + nblocks);
  bs->nblocks = nblocks;
  bs->ndim = ndim;
  bs->nscalars = nscalars;
  bs->nvar = 5 + nscalars;
  int64_t base = 0;
  for (int b = 0; b < nblocks; ++b) {
    const aph_block &k = bs->h[b];
    for (int d = 0; d < 3; ++d) {
      if (k.nx[d] < 1 || (d >= ndim && k.nx[d] != 1)) {
        delete bs;
        return set_error(ctx, APH_E_INVALID, "block %d: bad nx[%d]=%d", b, d, k.nx[d]);
      }
      bs->max_ext[d] = std::max(bs->max_ext[d], ext(k, d, ndim));
    }
    if (k.ng < 1 || k.cell_base != base) {
      delete bs;
      return set_error(ctx, APH_E_INVALID, "block %d: bad ng or cell_base", b);
    }
    for (int d = 0; d < ndim; ++d)
      if (k.nx[d] < k.ng) {  // ghost exchange and reflection read ng interior layers
        delete bs;
        return set_error(ctx, APH_E_INVALID, "block %d: nx[%d]=%d smaller than ng=%d", b, d,
                         k.nx[d], k.ng);
      }
    for (int o = 0; o < 27; ++o) {
      int nb = k.nbr[o];
      if (nb < -1 || nb >= nblocks) {
        delete bs;
        return set_error(ctx, APH_E_INVALID, "block %d: neighbour %d out of range", b, nb);
      }
      if (nb >= 0) {
        const aph_block &q = blocks[nb];
        if (q.nx[0] != k.nx[0] || q.nx[1] != k.nx[1] || q.nx[2] != k.nx[2] || q.ng != k.ng) {
          delete bs;
          return set_error(ctx, APH_E_INVALID,
                           "block %d: neighbour %d has a different shape (mixed-level "
                           "neighbours are not supported)", b, nb);
        }
      }
    }
    int64_t nc = ncell(k, ndim);
    base += nc;
    bs->max_ncell = std::max(bs->max_ncell, nc);
    bs->max_interior = std::max(bs->max_interior, (int64_t)k.nx[0] * k.nx[1] * k.nx[2]);
    // pencils visited by the reference loops (hydro.cpp:541-549, 623-627, 688)
    int64_t e1 = k.nx[1] + (ndim >= 2 ? 2 : 0), e2 = k.nx[2] + (ndim >= 3 ? 2 : 0);
    int64_t e0 = k.nx[0] + 2;
    bs->max_pencils[0] = std::max(bs->max_pencils[0], e1 * e2);
    bs->max_pencils[1] = std::max(bs->max_pencils[1], e0 * e2);
    bs->max_pencils[2] = std::max(bs->max_pencils[2], e0 * e1);
  }
  bs->ncells = base;
  if (nblocks > 0) {
    if (hipMalloc(&bs->d, sizeof(aph_block) * nblocks) != hipSuccess) {
      delete bs;
      return set_error(ctx, APH_E_NOMEM, "device allocation of the block table failed");
    }
    hipError_t e = hipMemcpy(bs->d, bs->h.data(), sizeof(aph_block) * nblocks,
                             hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(bs->d);
      delete bs;
      return set_error(ctx, APH_E_HIP, "block table upload failed: %s", hipGetErrorString(e));
    }
  }
  *out = bs;
  return APH_OK;
}

int aph_blockset_destroy(aph_context *ctx, aph_blockset *bs) {
  (void)ctx;
  if (!bs) return APH_OK;
  if (bs->d) (void)hipFree(bs->d);
  delete bs;
  return APH_OK;
}

int64_t aph_blockset_ncells(const aph_blockset *bs) { return bs ? bs->ncells : -1; }

}  // extern "C"
