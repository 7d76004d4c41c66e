// aph_amr.hip -- multilevel meshes in the plain form: cross-level ghost values and flux
// correction, executed from host-built plans.
//
// The reference leaves both to Parthenon (apophis_driver.cpp:68,74-76,94): the boundary
// exchange of the cons field restricts fine data into coarse ghosts (RestrictAverage) and
// prolongates coarse data into fine ghosts (ProlongateSharedMinMod), and the flux
// correction replaces a coarse face flux by the area-weighted mean of the fine fluxes
// across it.  Here the host resolves the mesh tree once into flat op lists; each op is one
// ghost cell (or one coarse face) whose sources are a handful of interior cells (fine
// faces), so both passes are a single gather kernel over the ops: every op reads cells
// no op writes, and the result does not depend on the order the ops run in.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <set>
#include <utility>
#include <vector>

#include "aph_adjoint_host.hpp"
#include "aph_device.hpp"
#include "aph_runtime.hpp"

struct aph_amr_plan {
  aph_context *ctx = nullptr;
  const aph_blockset *bs = nullptr;  // for the boundary conditions after the fill
  bool physical = false;             // some block has a face on a physical domain face
  int32_t nvar = 0, ndim = 0;
  int64_t nops = 0, nfixes = 0, nrefs = 0;
  aph_amr_op *ops = nullptr;
  aph_flux_fix *fixes = nullptr;
  aph_amr_ref *refs = nullptr;
  // for the adjoint: the block of each op's ghost cell, and the cells the ops read as a
  // transposed list (rows of ref indices per cell, ascending)
  int32_t *op_block = nullptr;
  int64_t *tgt = nullptr, *tgt_stride = nullptr, *row = nullptr;
  int32_t *col = nullptr;
  int64_t ntgt = 0;
};

namespace aph {
namespace {

constexpr int kThreads = 256;

template <int NV>
__device__ __forceinline__ double amr_value(const aph_amr_ref *__restrict__ refs,
                                            aph_amr_val v, int comp,
                                            const double *__restrict__ u) {
  aph_amr_ref r = refs[v.first];
  double s = u[r.offset + comp * r.stride];
  for (int k = 1; k < v.count; ++k) {
    r = refs[v.first + k];
    s += u[r.offset + comp * r.stride];
  }
  return v.count == 1 ? s : s * (1.0 / v.count);
}

__device__ __forceinline__ double sgn1(double x) { return x < 0.0 ? -1.0 : 1.0; }

template <int NV>
__global__ void __launch_bounds__(kThreads)
    k_amr_fill(const aph_amr_op *__restrict__ ops, int64_t nops,
               const aph_amr_ref *__restrict__ refs, int ndim, double *u, unsigned *epoch) {
  // the fused form's launch epoch rides on this launch (one launch fewer per stage)
  if (epoch && blockIdx.x == 0 && threadIdx.x == 0) *epoch = *epoch + 1u;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nops;
       t += (int64_t)gridDim.x * blockDim.x) {
    const aph_amr_op op = ops[t];
    const int mirror = op.kind >> APH_AMR_MIRROR_BIT;  // slots 1..6 beyond a reflecting face
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double c = amr_value<NV>(refs, op.val[0], v, u);
      double out = c;
      if ((op.kind & 0xff) == APH_AMR_PROLONG) {
        for (int d = 0; d < ndim; ++d) {
          double lo = amr_value<NV>(refs, op.val[1 + 2 * d], v, u);
          double hi = amr_value<NV>(refs, op.val[2 + 2 * d], v, u);
          if (v == 1 + d) {  // the mirror image negates the momentum normal to the face
            if (mirror & (1 << (2 * d))) lo = -lo;
            if (mirror & (2 << (2 * d))) hi = -hi;
          }
          const double a = c - lo;
          const double b = hi - c;
          const double g = (0.5 * (sgn1(a) + sgn1(b))) * fmin(fabs(a), fabs(b));
          out = out + (op.sign[d] * 0.25) * g;
        }
      }
      u[op.dst + v * op.dst_stride] = out;
    }
  }
}

template <int NV>
__global__ void __launch_bounds__(kThreads)
    k_flux_fix(const aph_flux_fix *__restrict__ fixes, int64_t nfix,
               const aph_amr_ref *__restrict__ refs, double *f1, double *f2, double *f3) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nfix;
       t += (int64_t)gridDim.x * blockDim.x) {
    const aph_flux_fix fx = fixes[t];
    double *f = fx.dir == 0 ? f1 : (fx.dir == 1 ? f2 : f3);
#pragma unroll
    for (int v = 0; v < NV; ++v) f[fx.dst + v * fx.dst_stride] = amr_value<NV>(refs, fx.val, v, f);
  }
}

// Host-side checks that keep every device access inside the fields.
struct Locator {
  const aph_blockset *bs;
  std::vector<int64_t> base;  // nvar * cell_base per block
  explicit Locator(const aph_blockset *b) : bs(b) {
    for (const auto &k : b->h) base.push_back((int64_t)b->nvar * k.cell_base);
  }
  // block holding component-0 offset `off` with component stride `stride`, or -1
  int find(int64_t off, int64_t stride) const {
    auto it = std::upper_bound(base.begin(), base.end(), off);
    if (it == base.begin()) return -1;
    const int b = (int)(it - base.begin()) - 1;
    const aph_block &k = bs->h[b];
    int64_t nc = 1;
    for (int d = 0; d < 3; ++d) nc *= (d < bs->ndim) ? k.nx[d] + 2 * k.ng : k.nx[d];
    if (stride != nc || off - base[b] >= nc) return -1;
    return b;
  }
  // whether the cell at `off` of block b is an interior cell
  bool interior(int b, int64_t off) const {
    const aph_block &k = bs->h[b];
    int N[3], g[3];
    for (int d = 0; d < 3; ++d) {
      g[d] = (d < bs->ndim) ? k.ng : 0;
      N[d] = k.nx[d] + 2 * g[d];
    }
    int64_t c = off - base[b];
    const int i = (int)(c % N[0]);
    c /= N[0];
    const int j = (int)(c % N[1]);
    const int kk = (int)(c / N[1]);
    const int id[3] = {i, j, kk};
    for (int d = 0; d < 3; ++d)
      if (id[d] < g[d] || id[d] >= g[d] + k.nx[d]) return false;
    return true;
  }
};

// Refinement criteria (src/refinement/gradient.cpp, other.cpp): one workgroup per block
// reduces the criterion's quantity over the reference's index ranges with max, which is
// independent of the order, so the result is the same bits as any serial loop.
__global__ void __launch_bounds__(kThreads)
    k_refine_tag(const aph_block *__restrict__ blocks, int ndim, int nvar, int crit, double p0,
                 double p1, const double *__restrict__ prim, int32_t *tags, double *values) {
  const aph_block &bk = blocks[blockIdx.x];
  const int N0 = ext(bk, 0, ndim), N1 = ext(bk, 1, ndim), N2 = ext(bk, 2, ndim);
  const int g[3] = {gst(bk, 0, ndim), gst(bk, 1, ndim), gst(bk, 2, ndim)};
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const double *w = prim + (int64_t)nvar * bk.cell_base;
  auto at = [&](int v, int i, int j, int k) -> double {
    return w[v * nc + ((int64_t)k * N1 + j) * N0 + i];
  };
  // inclusive ghost-inclusive ranges
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = g[d];
    hi[d] = g[d] + bk.nx[d] - 1;
  }
  if (crit == APH_REFINE_PRESSURE_GRADIENT) {
    for (int d = 0; d < ndim; ++d) {
      --lo[d];
      ++hi[d];
    }
  } else if (crit == APH_REFINE_VELOCITY_GRADIENT) {
    --lo[0];
    ++hi[0];
    --lo[1];
    ++hi[1];
  } else {
    ++hi[0];  // the reference's i range runs to ib.e + 1
  }
  const int n0 = hi[0] - lo[0] + 1, n1 = hi[1] - lo[1] + 1, n2 = hi[2] - lo[2] + 1;
  const int64_t n = (int64_t)n0 * n1 * n2;
  double m = 0.0;
  if (!(crit == APH_REFINE_PRESSURE_GRADIENT && ndim == 1)) {
    for (int64_t t = threadIdx.x; t < n; t += blockDim.x) {
      const int i = lo[0] + (int)(t % n0);
      const int j = lo[1] + (int)((t / n0) % n1);
      const int k = lo[2] + (int)(t / ((int64_t)n0 * n1));
      double q;
      if (crit == APH_REFINE_PRESSURE_GRADIENT) {
        const double a = 0.5 * (at(4, i + 1, j, k) - at(4, i - 1, j, k));
        const double b = 0.5 * (at(4, i, j + 1, k) - at(4, i, j - 1, k));
        double s2 = a * a + b * b;
        if (ndim == 3) {
          const double c = 0.5 * (at(4, i, j, k + 1) - at(4, i, j, k - 1));
          s2 = s2 + c * c;
        }
        q = sqrt(s2) / at(4, i, j, k);
      } else if (crit == APH_REFINE_VELOCITY_GRADIENT) {
        const double vgy = fabs(at(2, i + 1, j, k) - at(2, i - 1, j, k)) * 0.5;
        const double vgx = fabs(at(1, i, j + 1, k) - at(1, i, j - 1, k)) * 0.5;
        q = sqrt(vgx * vgx + vgy * vgy);
      } else {
        q = at(0, i, j, k);
      }
      m = (m < q) ? q : m;  // std::max(m, q)
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double r = __shfl_xor(m, o);
    m = (m < r) ? r : m;
  }
  __shared__ double part[kThreads / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = (m < part[k]) ? part[k] : m;
    int tag = 0;
    if (crit == APH_REFINE_PRESSURE_GRADIENT) {
      if (ndim > 1) tag = (m > p0) ? 1 : ((m < 0.25 * p0) ? -1 : 0);
    } else if (crit == APH_REFINE_VELOCITY_GRADIENT) {
      tag = (m > p0) ? 1 : ((m < 0.5 * p0) ? -1 : 0);
    } else {
      tag = (m > p1) ? 1 : ((m < p0) ? -1 : 0);
    }
    tags[blockIdx.x] = tag;
    if (values) values[blockIdx.x] = m;
  }
}

}  // namespace

// The cross-level fill of cons; when epoch is non-null the same launch bumps the fused
// form's launch epoch (it is launched even without ops then).
int amr_fill_epoch(aph_context *ctx, const aph_amr_plan *p, double *cons, unsigned *epoch);
// The physical boundary conditions of aph_fill_ghosts alone (aph_plain.hip).
int physical_bc(aph_context *ctx, const aph_blockset *bs, double *cons);

}  // namespace aph

using namespace aph;

extern "C" {

int aph_refinement_tag(aph_context *ctx, const aph_blockset *bs, int32_t criterion,
                       const double params[2], const double *prim, int32_t *tags,
                       double *values) {
  if (!ctx) return APH_E_INVALID;
  if (!bs || !params || !prim || !tags)
    return set_error(ctx, APH_E_INVALID, "aph_refinement_tag: null");
  if (criterion < APH_REFINE_PRESSURE_GRADIENT || criterion > APH_REFINE_MAX_DENSITY)
    return set_error(ctx, APH_E_INVALID, "aph_refinement_tag: unknown criterion %d", criterion);
  if (criterion == APH_REFINE_VELOCITY_GRADIENT && bs->ndim < 2)
    return set_error(ctx, APH_E_INVALID, "aph_refinement_tag: the velocity gradient needs ndim >= 2");
  const int need = criterion == APH_REFINE_MAX_DENSITY ? 1 : 2;
  for (int b = 0; b < bs->nblocks; ++b)
    if (bs->ndim > 1 && bs->h[b].ng < need)
      return set_error(ctx, APH_E_INVALID, "aph_refinement_tag: block %d has %d ghost layers, "
                       "the criterion reads %d", b, bs->h[b].ng, need);
  if (!bs->nblocks) return APH_OK;
  k_refine_tag<<<bs->nblocks, kThreads, 0, ctx->stream>>>(bs->d, bs->ndim, bs->nvar, criterion,
                                                          params[0], params[1], prim, tags,
                                                          values);
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

int aph_amr_plan_create(aph_context *ctx, const aph_blockset *bs, const aph_amr_op *ops,
                        int64_t nops, const aph_flux_fix *fixes, int64_t nfixes,
                        const aph_amr_ref *refs, int64_t nrefs, aph_amr_plan **out) {
  if (!ctx) return APH_E_INVALID;
  if (!bs || !out || nops < 0 || nfixes < 0 || nrefs < 0 || (nops && !ops) ||
      (nfixes && !fixes) || (nrefs && !refs))
    return set_error(ctx, APH_E_INVALID, "aph_amr_plan_create: bad arguments");
  if (bs->ndim < 2) return set_error(ctx, APH_E_INVALID, "aph_amr_plan_create: ndim must be 2 or 3");
  const Locator loc(bs);
  auto check_val = [&](const aph_amr_val &v, bool want_interior, int *blk) -> bool {
    if (v.count < 1 || v.count > 8 || v.first < 0 || (int64_t)v.first + v.count > nrefs)
      return false;
    for (int k = 0; k < v.count; ++k) {
      const aph_amr_ref &r = refs[v.first + k];
      const int b = loc.find(r.offset, r.stride);
      if (b < 0) return false;
      if (want_interior && !loc.interior(b, r.offset)) return false;
      if (blk) *blk = b;
    }
    return true;
  };
  for (int64_t t = 0; t < nops; ++t) {
    const aph_amr_op &op = ops[t];
    const int b = loc.find(op.dst, op.dst_stride);
    if (b < 0 || loc.interior(b, op.dst))
      return set_error(ctx, APH_E_INVALID, "amr op %lld: destination is not a ghost cell",
                       (long long)t);
    const int kind = op.kind & 0xff, mirror = op.kind >> APH_AMR_MIRROR_BIT;
    if ((kind != APH_AMR_MEAN && kind != APH_AMR_PROLONG) || mirror < 0 ||
        mirror >= (1 << (2 * bs->ndim)) || (kind == APH_AMR_MEAN && mirror))
      return set_error(ctx, APH_E_INVALID, "amr op %lld: unknown kind %d", (long long)t, op.kind);
    const int nval = kind == APH_AMR_MEAN ? 1 : 1 + 2 * bs->ndim;
    for (int k = 0; k < nval; ++k)
      if (!check_val(op.val[k], true, nullptr))
        return set_error(ctx, APH_E_INVALID, "amr op %lld: value %d does not reference "
                         "interior cells", (long long)t, k);
    if (kind == APH_AMR_PROLONG)
      for (int d = 0; d < bs->ndim; ++d)
        if (op.sign[d] != -1 && op.sign[d] != 1)
          return set_error(ctx, APH_E_INVALID, "amr op %lld: sign must be +-1", (long long)t);
  }
  // faces (flux array, component-0 offset) written and read: a face written twice, or
  // both written and read, would race inside the one fix launch (on a mesh of three or
  // more levels a block can be corrected on one face and a source on another)
  std::set<std::pair<int, int64_t>> dst_faces, src_faces;
  for (int64_t t = 0; t < nfixes; ++t) {
    const aph_flux_fix &fx = fixes[t];
    const int b = loc.find(fx.dst, fx.dst_stride);
    int sb = -1;
    if (b < 0 || fx.dir < 0 || fx.dir >= bs->ndim || !check_val(fx.val, false, &sb))
      return set_error(ctx, APH_E_INVALID, "flux fix %lld: bad face references", (long long)t);
    if (!dst_faces.emplace(fx.dir, fx.dst).second)
      return set_error(ctx, APH_E_INVALID, "flux fix %lld: face corrected twice", (long long)t);
    for (int k = 0; k < fx.val.count; ++k)
      src_faces.emplace(fx.dir, refs[fx.val.first + k].offset);
  }
  for (const auto &f : dst_faces)
    if (src_faces.count(f))
      return set_error(ctx, APH_E_INVALID,
                       "flux fix: face %lld of flux array %d is both corrected and a source",
                       (long long)f.second, f.first);
  aph_amr_plan *p = new (std::nothrow) aph_amr_plan();
  if (!p) return set_error(ctx, APH_E_NOMEM, "host allocation failed");
  p->ctx = ctx;
  p->bs = bs;
  for (const aph_block &k : bs->h)
    for (int f = 0; f < 2 * bs->ndim; ++f) {
      const int sdn = (f >> 1) == 0 ? 1 : ((f >> 1) == 1 ? 3 : 9);
      if (k.nbr[13 + ((f & 1) ? sdn : -sdn)] < 0 && k.bc[f] != APH_BC_AMR) p->physical = true;
    }
  p->nvar = bs->nvar;
  p->ndim = bs->ndim;
  p->nops = nops;
  p->nfixes = nfixes;
  p->nrefs = nrefs;
  auto up = [&](void **dst, const void *src, size_t bytes) -> bool {
    if (!bytes) return true;
    if (hipMalloc(dst, bytes) != hipSuccess) return false;
    return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  // the adjoint's view: each op's block, and the refs of the ops grouped by the cell they
  // read (a cell's contributions are then summed in one fixed order, with no atomics)
  std::vector<int32_t> op_block((size_t)nops);
  std::vector<std::pair<int64_t, int32_t>> pairs;  // (cell offset, ref index)
  for (int64_t t = 0; t < nops; ++t) {
    const aph_amr_op &op = ops[t];
    op_block[t] = loc.find(op.dst, op.dst_stride);
    const int nval = (op.kind & 0xff) == APH_AMR_MEAN ? 1 : 1 + 2 * bs->ndim;
    for (int k = 0; k < nval; ++k)
      for (int r = op.val[k].first; r < op.val[k].first + op.val[k].count; ++r)
        pairs.emplace_back(refs[r].offset, r);
  }
  std::sort(pairs.begin(), pairs.end());
  std::vector<int64_t> tgt, tstride, row;
  std::vector<int32_t> col;
  for (size_t q = 0; q < pairs.size(); ++q) {
    if (q == 0 || pairs[q].first != pairs[q - 1].first) {
      tgt.push_back(pairs[q].first);
      tstride.push_back(refs[pairs[q].second].stride);
      row.push_back((int64_t)q);
    }
    col.push_back(pairs[q].second);
  }
  row.push_back((int64_t)pairs.size());
  p->ntgt = (int64_t)tgt.size();
  if (!up((void **)&p->ops, ops, sizeof(aph_amr_op) * nops) ||
      !up((void **)&p->fixes, fixes, sizeof(aph_flux_fix) * nfixes) ||
      !up((void **)&p->refs, refs, sizeof(aph_amr_ref) * nrefs) ||
      !up((void **)&p->op_block, op_block.data(), sizeof(int32_t) * op_block.size()) ||
      !up((void **)&p->tgt, tgt.data(), sizeof(int64_t) * tgt.size()) ||
      !up((void **)&p->tgt_stride, tstride.data(), sizeof(int64_t) * tstride.size()) ||
      !up((void **)&p->row, row.data(), sizeof(int64_t) * row.size()) ||
      !up((void **)&p->col, col.data(), sizeof(int32_t) * col.size())) {
    aph_amr_plan_destroy(ctx, p);
    return set_error(ctx, APH_E_NOMEM, "amr plan device allocation failed");
  }
  *out = p;
  return APH_OK;
}

int aph_amr_plan_destroy(aph_context *ctx, aph_amr_plan *p) {
  if (!ctx) return APH_E_INVALID;
  if (!p) return APH_OK;
  for (void *q : {(void *)p->ops, (void *)p->fixes, (void *)p->refs, (void *)p->op_block,
                  (void *)p->tgt, (void *)p->tgt_stride, (void *)p->row, (void *)p->col})
    if (q) (void)hipFree(q);
  delete p;
  return APH_OK;
}

#define APH_AMR_NV(nv, CALL)                                                          \
  switch (nv) {                                                                       \
    case 5: { constexpr int NV = 5; CALL; } break;                                    \
    case 6: { constexpr int NV = 6; CALL; } break;                                    \
    case 12: { constexpr int NV = 12; CALL; } break;                                  \
    case 14: { constexpr int NV = 14; CALL; } break;                                  \
    default: return set_error(ctx, APH_E_INVALID, "unsupported nvar %d", nv);         \
  }

int aph_amr_fill(aph_context *ctx, const aph_amr_plan *p, double *cons) {
  if (!ctx) return APH_E_INVALID;
  if (!p || !cons) return set_error(ctx, APH_E_INVALID, "aph_amr_fill: null");
  const int rc = aph::amr_fill_epoch(ctx, p, cons, nullptr);
  if (rc != APH_OK || !p->physical) return rc;
  // the edge and corner ghost cells beyond the domain copy the cross-level ghosts beside them
  return aph::physical_bc(ctx, p->bs, cons);
}

int aph_amr_flux_correct(aph_context *ctx, const aph_amr_plan *p, double *f1, double *f2,
                         double *f3) {
  if (!ctx) return APH_E_INVALID;
  if (!p || !f1 || (p->ndim >= 2 && !f2) || (p->ndim >= 3 && !f3))
    return set_error(ctx, APH_E_INVALID, "aph_amr_flux_correct: null");
  if (!p->nfixes) return APH_OK;
  const unsigned grid = (unsigned)std::min<int64_t>((p->nfixes + kThreads - 1) / kThreads, 4096);
  APH_AMR_NV(p->nvar, (k_flux_fix<NV><<<grid, kThreads, 0, ctx->stream>>>(
                          p->fixes, p->nfixes, p->refs, f1, f2, f3)));
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

}  // extern "C"

namespace aph {

void amr_adj_view(const aph_amr_plan *p, AdjAmr &v) {
  v.ops = p->ops;
  v.nops = p->nops;
  v.refs = p->refs;
  v.nrefs = p->nrefs;
  v.op_block = p->op_block;
  v.tgt = p->tgt;
  v.tgt_stride = p->tgt_stride;
  v.row = p->row;
  v.col = p->col;
  v.ntgt = p->ntgt;
}

int amr_fill_epoch(aph_context *ctx, const aph_amr_plan *p, double *cons, unsigned *epoch) {
  if (!p->nops && !epoch) return APH_OK;
  const unsigned grid =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>((p->nops + kThreads - 1) / kThreads, 4096));
  APH_AMR_NV(p->nvar, (k_amr_fill<NV><<<grid, kThreads, 0, ctx->stream>>>(
                          p->ops, p->nops, p->refs, p->ndim, cons, epoch)));
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

}  // namespace aph
