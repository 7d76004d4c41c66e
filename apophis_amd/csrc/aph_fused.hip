// aph_fused.hip -- the batched form: one persistent launch per RK2 stage over every
// block of a block set (uneven sizes allowed), fusing ConsToPrim, reconstruction,
// the Riemann solve, the flux-divergence update and (optionally) the gravity source.
//
// Work item = one tile of one block.  A 64-lane workgroup takes items in a static
// persistent schedule; for each item it sweeps x, then y, then z:
//   * a lane owns one (pencil, segment) of the tile and marches it (aph_march.hpp),
//     loading conserved variables straight from HBM -- from the neighbouring block
//     or through the physical-boundary rule when the stencil leaves the block, so no
//     ghost exchange pass and no ghost storage reads are needed;
//   * the x and y sweeps accumulate du = A*F(+) - A*F(-) per cell in LDS in the
//     reference's order (x first, then y, then z, FluxDivHelper upstream);
//   * the last sweep finishes du, applies gam0*u + gam1*u_n + beta*dt*(-du/V), the
//     gravity source (gravity.cpp:386-407) and, in stage 2, the time-step estimate of
//     the new state (hydro.cpp:369-419) min-reduced on the device.
// Floors are applied on the fly every time a conserved state is read, once per
// stage, which is where the reference's ConsToPrim applies them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <map>
#include <vector>

#include "aph_adjoint_host.hpp"
#include "aph_fused_kernel.hpp"

namespace aph {

constexpr int kDefaultWaitLimit = 1 << 20;

// aph_amr.hip: the cross-level fill, bumping the multilevel launch epoch in the same launch
int amr_fill_epoch(aph_context *ctx, const aph_amr_plan *p, double *cons, unsigned *epoch);

namespace {

__global__ void k_fill(double *p, double v) { *p = v; }

// A multilevel stage whose cross-level wait ran out (its fine tiles were not all
// resident): the result is wrong, so make it visibly so.
__global__ void k_amr_check(const int *err, double *u_out) {
  if (*err) *u_out = __builtin_nan("");
}

// dt policy of the upstream EvolutionDriver: time += dt; dt <- min(2 dt, estimate),
// clipped so that the step does not pass tlim.  A multilevel wait that ran out (err set)
// makes dt and time NaN, so every later step and any read of them shows it.
__global__ void k_dt_finalize(double *dt, double *time, const double *est, double tlim,
                              int fixed, const int *err) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (err && *err) {
    *dt = __builtin_nan("");
    *time = __builtin_nan("");
    return;
  }
  const double d = *dt;
  const double t = *time + d;
  *time = t;
  if (!fixed) {
    double nd = (d < 0.1 * 1.79769313486231570815e308) ? 2.0 * d : d;
    const double e = *est;
    nd = (e < nd) ? e : nd;
    if (t < tlim && (tlim - t) < nd) nd = tlim - t;
    if (!(t < tlim)) nd = 0.0;
    *dt = nd;
  }
}

}  // namespace
}  // namespace aph

using namespace aph;

struct aph_plan {
  const aph_blockset *bs = nullptr;
  int recon = 0, rsolver = 0;
  aph_eos eos{};
  double g[3] = {0, 0, 0};
  int grav = 0;
  aph_monopole *mono = nullptr;  // monopole gravity before every stage (aph_plan_set_monopole)
  aph_poisson *poisson = nullptr;  // Poisson gravity before every stage (aph_plan_set_poisson)
  double *d_gfield = nullptr;    // their field: 3 components, pack layout
  std::vector<aph_tile> user_tiles;  // aph_plan_set_tile overrides, keyed by nx[0]
  // autotuner choices, keyed by block shape: the tile and the segment scale (1: the
  // shortest segments that keep every lane busy, 2: twice as long where they divide)
  struct Tuned {
    int4 shape;
    aph_tile tile;
    int seg_scale;
  };
  std::vector<Tuned> tuned;
  int only_class = -1;               // autotuner: item list restricted to one class
  int wg = 64;                       // lanes per workgroup (autotuner: 64, 128 or 256)
  std::vector<aph_tile> tiles;       // resolved tile per class
  std::vector<int> seg_scale;        // resolved segment scale per class
  std::vector<int4> class_shape;  // class index -> block shape (nx0, nx1, nx2)
  int4 *d_items = nullptr;
  int nitems = 0;
  int grid = 0;
  size_t lds = 0;
  size_t lds_reserve = 0;  // aph_plan_set_lds_reserve: LDS per workgroup at least this
  double *d_est = nullptr;  // device dt estimate accumulator
  int integrator = APH_INTEGRATOR_RK2;
  double *d_u2 = nullptr;   // third register (rk3)
  bool dirty = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  aph::AdjScratch adj;
  // multilevel mesh (aph_plan_set_amr): one launch per stage whose items run the finest
  // level first; a coarse tile waits for the count of fine tiles that wrote its interface
  // region (per block face: region index; per region: fine tiles per launch from the item
  // list, a running count, and the launch epoch)
  const aph_amr_plan *amr = nullptr;
  aph_amr_face *d_amr_face = nullptr;
  int32_t *d_amr_cblk = nullptr;  // per block face: the coarse block of a role-1 face
  double *d_amr_buf = nullptr;
  std::vector<int32_t> level;
  std::vector<aph_amr_face> faces_h;
  std::vector<int32_t> face_region;
  int nregions = 0;
  int *d_region = nullptr;
  unsigned long long *d_count = nullptr;
  unsigned *d_expect = nullptr;
  unsigned *d_epoch = nullptr;
  int *d_err = nullptr;
  int wait_limit = kDefaultWaitLimit;  // polls of a multilevel wait before it gives up
  bool amr_faces = false;  // the block set has faces to another level: needs set_amr
};

namespace {


template <int RECON, int RS, int NV>
KernelFn pick_dim(int ndim) {
  switch (ndim) {
    case 1: return k_fused<RECON, RS, NV, 1>;
    case 2: return k_fused<RECON, RS, NV, 2>;
    default: return k_fused<RECON, RS, NV, 3>;
  }
}

template <int RECON, int RS>
KernelFn pick_nv(int nv, int ndim) {
  switch (nv) {
    case 5: return pick_dim<RECON, RS, 5>(ndim);
    case 6: return pick_dim<RECON, RS, 6>(ndim);
    default: return nullptr;
  }
}

KernelFn pick_kernel(int recon, int rs, int nv, int ndim) {
  if (nv != 5 && nv != 6) return pick_kernel_wide(recon, rs, nv, ndim);
#define APH_PICK(R, S) \
  if (recon == R && rs == S) return pick_nv<R, S>(nv, ndim);
  APH_PICK(RECON_PPM, RS_HLLC)
  APH_PICK(RECON_PPM, RS_HLLE)
  APH_PICK(RECON_PPM, RS_LHLLC)
  APH_PICK(RECON_PLM, RS_HLLC)
  APH_PICK(RECON_PLM, RS_HLLE)
  APH_PICK(RECON_PLM, RS_LHLLC)
  APH_PICK(RECON_DC, RS_HLLC)
  APH_PICK(RECON_DC, RS_HLLE)
  APH_PICK(RECON_DC, RS_LHLLC)
  APH_PICK(RECON_DC, RS_NONE)  // hydro switched off (hydro_none.hpp:17-35)
#undef APH_PICK
  return nullptr;
}

// Low-storage coefficients of stage s (1-based) of an integrator (upstream Parthenon:
// rk1, rk2, rk3, vl2) and the stage's reconstruction (vl2 predicts with DC,
// hydro.cpp:114-139).  Returns the number of stages (0: unknown integrator).
int stage_coef(int integrator, int s, int recon, double &gam0, double &gam1, double &beta,
               int &srecon) {
  static const double kC[4][3][3] = {
      {{0.0, 1.0, 1.0}, {0, 0, 0}, {0, 0, 0}},                           // rk1
      {{0.0, 1.0, 1.0}, {0.5, 0.5, 0.5}, {0, 0, 0}},                     // rk2
      {{0.0, 1.0, 1.0}, {0.25, 0.75, 0.25}, {2.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0}},  // rk3
      {{0.0, 1.0, 0.5}, {0.0, 1.0, 1.0}, {0, 0, 0}}};                    // vl2
  static const int kN[4] = {1, 2, 3, 2};
  if (integrator < 1 || integrator > 4) return 0;
  const int n = kN[integrator - 1];
  if (s < 1 || s > n) return 0;
  gam0 = kC[integrator - 1][s - 1][0];
  gam1 = kC[integrator - 1][s - 1][1];
  beta = kC[integrator - 1][s - 1][2];
  srecon = (integrator == APH_INTEGRATOR_VL2 && s == 1) ? RECON_DC : recon;
  return n;
}

// Largest divisor of n that does not exceed cap.
int divisor_at_most(int n, int cap) {
  for (int d = std::min(n, cap); d > 1; --d)
    if (n % d == 0) return d;
  return 1;
}

aph_tile default_tile(const aph_block &k, int ndim) {
  const int cap = (ndim == 3) ? 8 : (ndim == 2 ? 16 : 256);
  aph_tile t{k.nx[0], 1, 1, 1};
  t.tx = divisor_at_most(k.nx[0], cap);
  if (ndim >= 2) t.ty = divisor_at_most(k.nx[1], cap);
  if (ndim >= 3) t.tz = divisor_at_most(k.nx[2], cap);
  return t;
}

// Segment length along an axis: the largest divisor of the tile extent that keeps at
// least `lanes` (pencil, segment) units per sweep, so every lane of the workgroup works;
// scale 2 doubles it where that still divides the extent (fewer recomputed halo cells,
// some lanes idle).
int seg_len(int t_d, int pencils, int lanes, int scale) {
  int best = 1;
  for (int s = 1; s <= t_d; ++s)
    if (t_d % s == 0 && (int64_t)pencils * (t_d / s) >= lanes) best = s;
  if (scale == 2 && t_d % (2 * best) == 0) best *= 2;
  return best;
}

int build_items(aph_context *ctx, aph_plan *p) {
  const aph_blockset *bs = p->bs;
  const int ndim = bs->ndim;
  p->class_shape.clear();
  std::vector<int> cls_of(bs->nblocks);
  for (int b = 0; b < bs->nblocks; ++b) {
    const aph_block &k = bs->h[b];
    int c = -1;
    for (size_t q = 0; q < p->class_shape.size(); ++q) {
      const int4 &sh = p->class_shape[q];
      if (sh.x == k.nx[0] && sh.y == k.nx[1] && sh.z == k.nx[2]) c = (int)q;
    }
    if (c < 0) {
      if ((int)p->class_shape.size() >= kMaxClasses)
        return set_error(ctx, APH_E_INVALID, "more than %d block shapes", kMaxClasses);
      p->class_shape.push_back(make_int4(k.nx[0], k.nx[1], k.nx[2], b));
      c = (int)p->class_shape.size() - 1;
    }
    cls_of[b] = c;
  }
  // tile per class: a user tile for edge n = nx0 (aph_plan_set_tile) if one is set,
  // else the autotuner's choice for the shape, else the default
  std::vector<aph_tile> tl(p->class_shape.size());
  std::vector<int> sc(p->class_shape.size(), 1);
  for (size_t c = 0; c < p->class_shape.size(); ++c) {
    const aph_block &k = bs->h[p->class_shape[c].w];
    aph_tile t = default_tile(k, ndim);
    for (const auto &u : p->tuned)
      if (u.shape.x == k.nx[0] && u.shape.y == k.nx[1] && u.shape.z == k.nx[2]) {
        t = u.tile;
        sc[c] = u.seg_scale;
      }
    for (const auto &u : p->user_tiles)
      if (u.n == k.nx[0]) {
        t = u;
        sc[c] = 1;
      }
    tl[c] = t;
  }
  // per XCD lists, greedy by cost (cells) so every L2 sees whole blocks
  constexpr int kXcd = 8;
  size_t maxcells = 0;
  std::vector<int4> items;
  // one phase without a multilevel mesh; with one, a phase per level, the finest first
  const bool multi = !p->level.empty();
  const int top = multi ? *std::max_element(p->level.begin(), p->level.end()) : 0;
  for (int phase = 0; phase <= top; ++phase) {
    std::vector<std::vector<int4>> lists(kXcd);
    std::vector<int64_t> load(kXcd, 0);
    for (int b = 0; b < bs->nblocks; ++b) {
      if (p->only_class >= 0 && cls_of[b] != p->only_class) continue;
      if (multi && p->level[b] != top - phase) continue;
      const aph_block &k = bs->h[b];
      const aph_tile &t = tl[cls_of[b]];
      const int tt[3] = {t.tx, t.ty, t.tz};
      for (int d = 0; d < 3; ++d)
        if (tt[d] < 1 || k.nx[d] % tt[d] != 0)
          return set_error(ctx, APH_E_INVALID,
                           "tile %dx%dx%d does not divide block %d (%dx%dx%d)", t.tx, t.ty,
                           t.tz, b, k.nx[0], k.nx[1], k.nx[2]);
      maxcells = std::max(maxcells, (size_t)t.tx * t.ty * t.tz);
      int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      for (int z = 0; z < k.nx[2]; z += t.tz)
        for (int y = 0; y < k.nx[1]; y += t.ty)
          for (int xx = 0; xx < k.nx[0]; xx += t.tx)
            lists[x].push_back(make_int4(b | (cls_of[b] << 24), xx, y, z));
      load[x] += (int64_t)k.nx[0] * k.nx[1] * k.nx[2];
    }
    size_t maxlen = 0;
    for (auto &l : lists) maxlen = std::max(maxlen, l.size());
    const size_t at = items.size();
    items.resize(at + maxlen * kXcd, make_int4(-1, 0, 0, 0));
    for (int x = 0; x < kXcd; ++x)
      for (size_t r = 0; r < lists[x].size(); ++r) items[at + r * kXcd + x] = lists[x][r];
  }
  if (p->amr) {
    // fine tiles per interface region and launch: tiles on a face to a coarser leaf
    std::vector<unsigned> expect(std::max(p->nregions, 1), 0u);
    for (const int4 &it : items) {
      if (it.x < 0) continue;
      const int b = it.x & 0xFFFFFF;
      const aph_tile &t = tl[it.x >> 24];
      const int tt[3] = {t.tx, t.ty, t.tz}, t0[3] = {it.y, it.z, it.w};
      for (int d = 0; d < ndim; ++d)
        for (int side = 0; side < 2; ++side) {
          const int fi = b * 6 + 2 * d + side;
          const bool on = side ? (t0[d] + tt[d] == bs->h[b].nx[d]) : (t0[d] == 0);
          if (on && p->faces_h[fi].role == 1) ++expect[p->face_region[fi]];
        }
    }
    APH_HIP_CHECK(ctx, hipMemcpy(p->d_expect, expect.data(), sizeof(unsigned) * expect.size(),
                                 hipMemcpyHostToDevice));
    APH_HIP_CHECK(ctx, hipMemset(p->d_count, 0, sizeof(unsigned long long) * expect.size()));
    APH_HIP_CHECK(ctx, hipMemset(p->d_epoch, 0, sizeof(unsigned)));
    APH_HIP_CHECK(ctx, hipMemset(p->d_err, 0, sizeof(int)));
  }
  if (p->d_items) (void)hipFree(p->d_items);
  p->d_items = nullptr;
  p->nitems = (int)items.size();
  if (!items.empty()) {
    if (hipMalloc(&p->d_items, sizeof(int4) * items.size()) != hipSuccess)
      return set_error(ctx, APH_E_NOMEM, "plan item list allocation failed");
    APH_HIP_CHECK(ctx, hipMemcpy(p->d_items, items.data(), sizeof(int4) * items.size(),
                                 hipMemcpyHostToDevice));
  }
  p->lds = (ndim >= 2) ? maxcells * bs->nvar * sizeof(double) : 0;
  if (p->lds > 64 * 1024)
    return set_error(ctx, APH_E_INVALID, "tile too large for LDS (%zu bytes)", p->lds);
  p->lds = std::max(p->lds, p->lds_reserve);
  p->tiles.assign(tl.begin(), tl.end());
  p->seg_scale.assign(sc.begin(), sc.end());
  // persistent grid: resident workgroups, a multiple of the XCD count
  // every kernel variant the plan launches (vl2: the first-order predictor too); the
  // grid takes the smallest residency so that a multilevel launch keeps every
  // workgroup resident whichever variant runs
  std::vector<int> recons{p->recon};
  if (p->integrator == APH_INTEGRATOR_VL2) recons.push_back(RECON_DC);
  int dev = ctx->device, ncu = 0, occ = 1 << 30;
  APH_HIP_CHECK(ctx, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  for (int r : recons) {
    KernelFn fn = p->amr ? pick_kernel_amr(r, p->rsolver, bs->nvar, ndim)
                         : pick_kernel(r, p->rsolver, bs->nvar, ndim);
    if (!fn) return set_error(ctx, APH_E_INVALID, "no fused kernel for this plan");
    if (p->lds > 64 * 1024)  // beyond the default dynamic LDS limit (a reservation)
      APH_HIP_CHECK(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)p->lds));
    int o = 0;
    APH_HIP_CHECK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                           &o, reinterpret_cast<const void *>(fn), p->wg, p->lds));
    occ = std::min(occ, o);
  }
  occ = std::max(occ, 1);
  int grid = ncu * occ;
  grid = std::max(kXcd, (grid / kXcd) * kXcd);
  p->grid = std::min<int>(grid, std::max(kXcd, p->nitems));
  p->dirty = false;
  return APH_OK;
}

FusedArgs make_args(aph_plan *p, int stage, int nstages, double gam0, double gam1,
                    double beta, double *u_in, const double *u_n, double *u_out,
                    const double *dt_dev, double cfl) {
  FusedArgs a{};
  a.blocks = p->bs->d;
  a.items = p->d_items;
  a.nitems = p->nitems;
  a.ndim = p->bs->ndim;
  for (size_t c = 0; c < p->tiles.size(); ++c) {
    const aph_tile &t = p->tiles[c];
    TileCfg &k = a.cfg[c];
    k.t[0] = t.tx;
    k.t[1] = t.ty;
    k.t[2] = t.tz;
    for (int d = 0; d < 3; ++d) {
      const int A = (d == 0) ? 1 : 0, B = (d == 2) ? 1 : 2;
      k.s[d] = (d < a.ndim) ? seg_len(k.t[d], k.t[A] * k.t[B], p->wg, p->seg_scale[c]) : 1;
    }
  }
  a.eos = p->eos;
  a.g[0] = p->g[0];
  a.g[1] = p->g[1];
  a.g[2] = p->g[2];
  a.grav = p->grav || p->mono || p->poisson;
  a.gfield = (p->mono || p->poisson) ? p->d_gfield : nullptr;
  a.u_in = u_in;
  a.u_n = u_n;
  a.u_out = u_out;
  a.dt = dt_dev;
  a.gam0 = gam0;
  a.gam1 = gam1;
  a.beta = beta;
  a.dt_min = p->d_est;
  a.cfl = cfl;
  a.has_un = stage > 1;
  a.last = stage == nstages;
  a.amr_face = p->d_amr_face;
  a.amr_buf = p->d_amr_buf;
  a.amr_region = p->d_region;
  a.amr_count = p->d_count;
  a.amr_expect = p->d_expect;
  a.amr_epoch = p->d_epoch;
  a.amr_err = p->d_err;
  a.amr_wait_limit = p->wait_limit;
  return a;
}

int launch_kernel(aph_context *ctx, aph_plan *p, int recon, int stage, int nst, double gam0,
                  double gam1, double beta, double *u_in, const double *u_n, double *u_out,
                  const double *dt_dev, double cfl) {
  if (p->dirty) {
    int rc = build_items(ctx, p);
    if (rc) return rc;
  }
  if (p->nitems == 0) return APH_OK;
  if (p->amr_faces && !p->amr)
    return set_error(ctx, APH_E_INVALID,
                     "fused form: the block set has faces to another refinement level; "
                     "call aph_plan_set_amr first");
  KernelFn fn = p->amr ? pick_kernel_amr(recon, p->rsolver, p->bs->nvar, p->bs->ndim)
                       : pick_kernel(recon, p->rsolver, p->bs->nvar, p->bs->ndim);
  if (!fn) return set_error(ctx, APH_E_INVALID, "no fused kernel for this stage");
  if (p->mono) {  // the solver's tasks on the stage input, then the stage reads the field
    int rc = monopole_pass(ctx, p->mono, p->bs, u_in, p->eos.dfloor, p->d_gfield);
    if (rc) return rc;
  }
  if (p->poisson) {
    int rc = poisson_pass(ctx, p->poisson, u_in, p->eos.dfloor, p->d_gfield, nullptr);
    if (rc) return rc;
  }
  FusedArgs a = make_args(p, stage, nst, gam0, gam1, beta, u_in, u_n, u_out, dt_dev, cfl);
  if (!p->amr) {
    hipLaunchKernelGGL(fn, dim3(p->grid), dim3(p->wg), p->lds, ctx->stream, a);
    APH_LAUNCH_CHECK(ctx);
    return APH_OK;
  }
  // multilevel mesh: cross-level ghosts of the stage input, then one launch whose items
  // run the finest level first (each tile records the fluxes on its faces to a coarser
  // leaf; a tile with a face to a finer region waits for them, then takes their mean)
  int rc = amr_fill_epoch(ctx, p->amr, u_in, p->d_epoch);  // also bumps the epoch
  if (rc) return rc;
  hipLaunchKernelGGL(fn, dim3(p->grid), dim3(p->wg), p->lds, ctx->stream, a);
  APH_LAUNCH_CHECK(ctx);
  k_amr_check<<<1, 1, 0, ctx->stream>>>(p->d_err, u_out);
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

// Stage `stage` (1-based) of the plan's integrator: u_out = gam0 u_in + gam1 u_n + beta dt
// (-div F(u_in)) [+ gravity]; u_n is ignored in stage 1 (it is u_in there).
int launch_stage(aph_context *ctx, aph_plan *p, int stage, double *u_in, const double *u_n,
                 double *u_out, const double *dt_dev, double cfl) {
  if (p->dirty) {
    int rc = build_items(ctx, p);
    if (rc) return rc;
  }
  double gam0, gam1, beta;
  int srecon;
  const int nst = stage_coef(p->integrator, stage, p->recon, gam0, gam1, beta, srecon);
  if (!nst) return set_error(ctx, APH_E_INVALID, "stage %d out of range", stage);
  return launch_kernel(ctx, p, srecon, stage, nst, gam0, gam1, beta, u_in, u_n, u_out, dt_dev,
                       cfl);
}

int plan_nstages(const aph_plan *p) {
  double g0, g1, b;
  int r;
  return stage_coef(p->integrator, 1, p->recon, g0, g1, b, r);
}

// The adjoint's view of stage s of the plan's integrator (its coefficients and
// reconstruction); 0 when s is out of range.
int adj_coef(const aph_plan *p, int s, AdjCoef &c, int &srecon) {
  double g0, g1, b;
  const int n = stage_coef(p->integrator, s, p->recon, g0, g1, b, srecon);
  if (!n) return 0;
  c = AdjCoef{g0, g1, b, s, p->recon};
  return n;
}

// The adjoint entry points' checks: a plan whose reconstruction / Riemann pair (and vl2's
// first-order predictor) the adjoint kernels cover, without self-gravity.
int adj_plan_ok(aph_context *ctx, const aph_plan *p) {
  if (p->mono || p->poisson)
    return set_error(ctx, APH_E_INVALID, "adjoint: monopole / Poisson gravity is not supported");
  if (!adj_supported(p->recon, p->rsolver, p->bs->nvar, p->bs->ndim) ||
      (p->integrator == APH_INTEGRATOR_VL2 &&
       !adj_supported(RECON_DC, p->rsolver, p->bs->nvar, p->bs->ndim)))
    return set_error(ctx, APH_E_INVALID, "adjoint: unsupported (recon, riemann, nvar)");
  if (p->integrator == APH_INTEGRATOR_RK3 && !p->d_u2)
    return set_error(ctx, APH_E_STATE, "adjoint: rk3 plan without its third register");
  return APH_OK;
}

// The multilevel view of the adjoint (empty on a single-level plan).
AdjAmr adj_amr(const aph_plan *p) {
  AdjAmr v;
  if (!p->amr) return v;
  amr_adj_view(p->amr, v);
  v.faces = p->d_amr_face;
  v.cblk = p->d_amr_cblk;
  return v;
}

// One adjoint stage of the plan; on a multilevel mesh the stage input's cross-level ghost
// cells are filled first (the forward stage reads them; they hold no state).
int plan_adj_stage(aph_context *ctx, aph_plan *p, int sr, const AdjCoef &c, const double *x,
                   const double *dt_dev, const double *lo, const double *acc, double *li,
                   double *ln, int add) {
  if (p->amr) {
    int rc = amr_fill_epoch(ctx, p->amr, const_cast<double *>(x), nullptr);
    if (rc) return rc;
    const AdjAmr v = adj_amr(p);
    return adj_stage(ctx, p->bs, sr, p->rsolver, p->eos, p->g, p->grav, p->adj, c, x, dt_dev,
                     lo, acc, li, ln, add, &v);
  }
  return adj_stage(ctx, p->bs, sr, p->rsolver, p->eos, p->g, p->grav, p->adj, c, x, dt_dev,
                   lo, acc, li, ln, add);
}

// Candidate tiles for one block shape and workgroup size: divisors of the block edges,
// x extent at least 4 where the block allows (coalesced rows), one to sixteen cells per
// lane (a whole small block is a candidate too), du tile within the LDS limit.
std::vector<aph_tile> tile_candidates(const int4 &sh, int ndim, int nvar, int wg) {
  auto divs = [](int n) {
    std::vector<int> d;
    for (int i = 1; i <= n; ++i)
      if (n % i == 0) d.push_back(i);
    return d;
  };
  const int64_t lds_cells = 64 * 1024 / (8 * nvar);
  const int64_t whole = (int64_t)sh.x * sh.y * sh.z;
  std::vector<int> dy = ndim >= 2 ? divs(sh.y) : std::vector<int>{1};
  std::vector<int> dz = ndim >= 3 ? divs(sh.z) : std::vector<int>{1};
  std::vector<aph_tile> out;
  for (int tx : divs(sh.x)) {
    if (tx < std::min(4, sh.x)) continue;
    for (int ty : dy)
      for (int tz : dz) {
        const int64_t cells = (int64_t)tx * ty * tz;
        if (ndim >= 2 && cells > lds_cells) continue;
        if (cells > 16 * wg && cells != whole) continue;
        if (cells < wg && cells != whole) continue;
        out.push_back(aph_tile{sh.x, tx, ty, tz});
      }
  }
  return out;
}

void drop_graph(aph_plan *p) {
  if (p->exec) {
    (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
  }
  if (p->graph) {
    (void)hipGraphDestroy(p->graph);
    p->graph = nullptr;
  }
}

}  // namespace

extern "C" {

int aph_plan_create(aph_context *ctx, const aph_blockset *bs, int32_t recon, int32_t rsolver,
                    const aph_eos *eos, const double g[3], aph_plan **out) {
  if (!ctx) return APH_E_INVALID;
  if (!bs || !eos || !out) return set_error(ctx, APH_E_INVALID, "aph_plan_create: null");
  if (!pick_kernel(recon, rsolver, bs->nvar, bs->ndim))
    return set_error(ctx, APH_E_INVALID,
                     "fused form: unsupported (reconstruction=%d, riemann=%d, nvar=%d)", recon,
                     rsolver, bs->nvar);
  const int need = recon == APH_RECON_PPM ? 3 : 2;
  bool amr_faces = false;
  for (int b = 0; b < bs->nblocks; ++b)
    for (int f = 0; f < 2 * bs->ndim; ++f)
      if (bs->h[b].bc[f] == APH_BC_AMR) amr_faces = true;
  for (int b = 0; b < bs->nblocks; ++b)
    for (int d = 0; d < bs->ndim; ++d)
      if (bs->h[b].nx[d] < need)
        return set_error(ctx, APH_E_INVALID,
                         "fused form reads ghosts from neighbours: block %d axis %d has %d "
                         "cells, stencil needs %d", b, d, bs->h[b].nx[d], need);
  aph_plan *p = new (std::nothrow) aph_plan();
  if (!p) return set_error(ctx, APH_E_NOMEM, "host allocation failed");
  p->bs = bs;
  p->amr_faces = amr_faces;
  p->recon = recon;
  p->rsolver = rsolver;
  p->eos = *eos;
  if (g) {
    p->g[0] = g[0];
    p->g[1] = g[1];
    p->g[2] = g[2];
    p->grav = (g[0] != 0.0 || g[1] != 0.0 || g[2] != 0.0);
  }
  if (hipMalloc(&p->d_est, sizeof(double)) != hipSuccess) {
    delete p;
    return set_error(ctx, APH_E_NOMEM, "plan scratch allocation failed");
  }
  int rc = build_items(ctx, p);
  if (rc) {
    aph_plan_destroy(ctx, p);
    return rc;
  }
  *out = p;
  return APH_OK;
}

int aph_plan_destroy(aph_context *ctx, aph_plan *p) {
  (void)ctx;
  if (!p) return APH_OK;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->d_items) (void)hipFree(p->d_items);
  if (p->d_est) (void)hipFree(p->d_est);
  if (p->d_u2) (void)hipFree(p->d_u2);
  if (p->d_gfield) (void)hipFree(p->d_gfield);
  if (p->d_amr_face) (void)hipFree(p->d_amr_face);
  if (p->d_amr_cblk) (void)hipFree(p->d_amr_cblk);
  if (p->d_amr_buf) (void)hipFree(p->d_amr_buf);
  for (void *q : {(void *)p->d_region, (void *)p->d_count, (void *)p->d_expect,
                  (void *)p->d_epoch, (void *)p->d_err})
    if (q) (void)hipFree(q);
  adj_release(p->adj);
  delete p;
  return APH_OK;
}

int aph_plan_set_amr(aph_context *ctx, aph_plan *p, const aph_amr_plan *amr,
                     const aph_amr_face *faces, const int32_t *level,
                     int64_t buffer_doubles) {
  if (!ctx) return APH_E_INVALID;
  if (!p || !amr || !faces || !level || buffer_doubles < 0)
    return set_error(ctx, APH_E_INVALID, "aph_plan_set_amr: bad arguments");
  const aph_blockset *bs = p->bs;
  const int nd = bs->ndim, nv = bs->nvar;
  if (nd < 2) return set_error(ctx, APH_E_INVALID, "aph_plan_set_amr: ndim must be 2 or 3");
  if (!pick_kernel_amr(p->recon, p->rsolver, nv, nd))
    return set_error(ctx, APH_E_INVALID, "aph_plan_set_amr: the fused form runs multilevel "
                     "meshes with nvar 5 or 6 (this block set has %d)", nv);
  const int64_t nch = (nd == 3) ? 4 : 2;
  auto extents = [&](int b, int d, int64_t &n_a, int64_t &n_b) {
    const aph_block &k = bs->h[b];
    const int A = (d == 0) ? 1 : 0, B = (d == 2) ? 1 : 2;
    n_a = k.nx[A];
    n_b = (B < nd) ? k.nx[B] : 1;
  };
  // the regions: a face to a finer region (role 2) owns [base, base + extent)
  struct Region { int block, face; int64_t end; };
  std::map<int64_t, Region> regions;
  for (int b = 0; b < bs->nblocks; ++b) {
    if (level[b] < 0 || level[b] > 30)
      return set_error(ctx, APH_E_INVALID, "aph_plan_set_amr: block %d level %d", b, level[b]);
    const aph_block &k = bs->h[b];
    for (int f = 0; f < 6; ++f) {
      const aph_amr_face &af = faces[b * 6 + f];
      const int d = f / 2;
      const bool amr_face = d < nd && k.bc[f] == APH_BC_AMR;
      if (!amr_face) {
        if (af.role != 0)
          return set_error(ctx, APH_E_INVALID, "block %d face %d: role on a non-AMR face", b, f);
        continue;
      }
      if (af.role != 1 && af.role != 2)
        return set_error(ctx, APH_E_INVALID, "block %d face %d: role %d on an AMR face", b, f,
                         af.role);
      if (af.role != 2) continue;
      int64_t n_a, n_b;
      extents(b, d, n_a, n_b);
      if (af.na != n_a || af.qa != 0 || af.qb != 0)
        return set_error(ctx, APH_E_INVALID, "block %d face %d: bad coarse region", b, f);
      const int64_t end = af.base + n_a * n_b * nch * nv;
      if (af.base < 0 || end > buffer_doubles)
        return set_error(ctx, APH_E_INVALID, "block %d face %d: outside the interface buffer",
                         b, f);
      if (!regions.emplace(af.base, Region{b, f, end}).second)
        return set_error(ctx, APH_E_INVALID, "block %d face %d: region shared with another face",
                         b, f);
    }
  }
  int64_t prev_end = 0;
  for (const auto &r : regions) {
    if (r.first < prev_end)
      return set_error(ctx, APH_E_INVALID, "block %d face %d: overlapping interface regions",
                       r.second.block, r.second.face);
    prev_end = r.second.end;
  }
  // a face to a coarser leaf (role 1) writes into the region of the opposite face of a
  // block one level coarser, which runs in the next phase
  std::vector<int32_t> cblk((size_t)6 * bs->nblocks, -1);
  for (int b = 0; b < bs->nblocks; ++b) {
    for (int f = 0; f < 2 * nd; ++f) {
      const aph_amr_face &af = faces[b * 6 + f];
      if (af.role != 1) continue;
      const auto it = regions.find(af.base);
      if (it == regions.end())
        return set_error(ctx, APH_E_INVALID, "block %d face %d: no coarse region at %lld", b, f,
                         (long long)af.base);
      const Region &r = it->second;
      if (r.face != (f ^ 1) || level[r.block] != level[b] - 1)
        return set_error(ctx, APH_E_INVALID, "block %d face %d (level %d): region of block %d "
                         "face %d (level %d) is not the opposite face one level coarser", b, f,
                         level[b], r.block, r.face, level[r.block]);
      int64_t n_a, n_b;
      extents(b, f / 2, n_a, n_b);
      if (af.qa < 0 || af.qb < 0 || af.na != faces[r.block * 6 + r.face].na ||
          (af.qa + n_a + 1) / 2 > af.na)
        return set_error(ctx, APH_E_INVALID, "block %d face %d: bad fine offsets", b, f);
      const int64_t last = af.base +
                           (((af.qb + n_b - 1) / 2) * af.na + (af.qa + n_a - 1) / 2) * nch * nv +
                           nch * nv - 1;  // last element the face touches
      if (last >= r.end)
        return set_error(ctx, APH_E_INVALID, "block %d face %d: outside its coarse region", b,
                         f);
      cblk[b * 6 + f] = r.block;
    }
  }
  // region index per block face: a role-2 face's own region, a role-1 face's target
  std::map<int64_t, int> region_index;
  for (const auto &r : regions) region_index.emplace(r.first, (int)region_index.size());
  p->face_region.assign((size_t)6 * bs->nblocks, -1);
  for (int b = 0; b < bs->nblocks; ++b)
    for (int f = 0; f < 6; ++f)
      if (faces[b * 6 + f].role) p->face_region[b * 6 + f] = region_index.at(faces[b * 6 + f].base);
  p->faces_h.assign(faces, faces + (size_t)6 * bs->nblocks);
  p->nregions = (int)region_index.size();
  if (p->d_amr_face) (void)hipFree(p->d_amr_face);
  if (p->d_amr_cblk) (void)hipFree(p->d_amr_cblk);
  if (p->d_amr_buf) (void)hipFree(p->d_amr_buf);
  for (void *q : {(void *)p->d_region, (void *)p->d_count, (void *)p->d_expect,
                  (void *)p->d_epoch, (void *)p->d_err})
    if (q) (void)hipFree(q);
  p->d_amr_cblk = nullptr;
  p->d_amr_face = nullptr;
  p->d_amr_buf = nullptr;
  p->d_region = nullptr;
  p->d_count = nullptr;
  p->d_expect = nullptr;
  p->d_epoch = nullptr;
  p->d_err = nullptr;
  const size_t nr = (size_t)std::max(p->nregions, 1);
  if (hipMalloc(&p->d_region, sizeof(int32_t) * 6 * bs->nblocks) != hipSuccess ||
      hipMalloc(&p->d_count, sizeof(unsigned long long) * nr) != hipSuccess ||
      hipMalloc(&p->d_expect, sizeof(unsigned) * nr) != hipSuccess ||
      hipMalloc(&p->d_epoch, sizeof(unsigned)) != hipSuccess ||
      hipMalloc(&p->d_err, sizeof(int)) != hipSuccess)
    return set_error(ctx, APH_E_NOMEM, "aph_plan_set_amr: device allocation failed");
  APH_HIP_CHECK(ctx, hipMemcpy(p->d_region, p->face_region.data(),
                               sizeof(int32_t) * 6 * bs->nblocks, hipMemcpyHostToDevice));
  APH_HIP_CHECK(ctx, hipMemset(p->d_err, 0, sizeof(int)));  // read by aph_plan_status
  if (hipMalloc(&p->d_amr_face, sizeof(aph_amr_face) * 6 * bs->nblocks) != hipSuccess ||
      hipMalloc(&p->d_amr_cblk, sizeof(int32_t) * 6 * bs->nblocks) != hipSuccess ||
      hipMalloc(&p->d_amr_buf, sizeof(double) * std::max<int64_t>(buffer_doubles, 1)) !=
          hipSuccess)
    return set_error(ctx, APH_E_NOMEM, "aph_plan_set_amr: device allocation failed");
  APH_HIP_CHECK(ctx, hipMemcpy(p->d_amr_face, faces, sizeof(aph_amr_face) * 6 * bs->nblocks,
                               hipMemcpyHostToDevice));
  APH_HIP_CHECK(ctx, hipMemcpy(p->d_amr_cblk, cblk.data(), sizeof(int32_t) * 6 * bs->nblocks,
                               hipMemcpyHostToDevice));
  APH_HIP_CHECK(ctx, hipMemset(p->d_amr_buf, 0, sizeof(double) * std::max<int64_t>(buffer_doubles, 1)));
  p->amr = amr;
  p->level.assign(level, level + bs->nblocks);
  p->dirty = true;
  drop_graph(p);
  return APH_OK;
}

int aph_plan_set_tile(aph_context *ctx, aph_plan *p, const aph_tile *t) {
  if (!ctx) return APH_E_INVALID;
  if (!p || !t) return set_error(ctx, APH_E_INVALID, "aph_plan_set_tile: null");
  if (t->tx < 1 || t->ty < 1 || t->tz < 1)
    return set_error(ctx, APH_E_INVALID, "tile extents must be positive");
  std::vector<aph_tile> keep = p->user_tiles;
  bool found = false;
  for (auto &u : p->user_tiles)
    if (u.n == t->n) {
      u = *t;
      found = true;
    }
  if (!found) p->user_tiles.push_back(*t);
  int rc = build_items(ctx, p);
  if (rc) {
    p->user_tiles = keep;
    build_items(ctx, p);
    return rc;
  }
  drop_graph(p);
  return APH_OK;
}

int aph_plan_set_poisson(aph_context *ctx, aph_plan *p, aph_poisson *ps) {
  if (!ctx) return APH_E_INVALID;
  if (!p) return set_error(ctx, APH_E_INVALID, "aph_plan_set_poisson: null plan");
  if (ps && poisson_blockset(ps) != p->bs)
    return set_error(ctx, APH_E_INVALID, "aph_plan_set_poisson: solver of another block set");
  if (ps && !p->d_gfield) {
    APH_HIP_CHECK(ctx, hipMalloc(&p->d_gfield, sizeof(double) * 3 *
                                                   std::max<int64_t>(p->bs->ncells, 1)));
  }
  p->poisson = ps;
  drop_graph(p);
  return APH_OK;
}

int aph_plan_set_lds_reserve(aph_context *ctx, aph_plan *p, int64_t bytes) {
  if (!ctx) return APH_E_INVALID;
  if (!p || bytes < 0 || bytes > 160 * 1024)
    return set_error(ctx, APH_E_INVALID, "aph_plan_set_lds_reserve: 0..160 KiB");
  const size_t keep = p->lds_reserve;
  p->lds_reserve = (size_t)bytes;
  p->dirty = true;
  int rc = build_items(ctx, p);
  if (rc) {
    p->lds_reserve = keep;
    p->dirty = true;
    build_items(ctx, p);
    return rc;
  }
  drop_graph(p);
  return APH_OK;
}

int aph_plan_set_monopole(aph_context *ctx, aph_plan *p, aph_monopole *m) {
  if (!ctx) return APH_E_INVALID;
  if (!p) return set_error(ctx, APH_E_INVALID, "aph_plan_set_monopole: null plan");
  if (m && !p->d_gfield) {
    APH_HIP_CHECK(ctx, hipMalloc(&p->d_gfield, sizeof(double) * 3 *
                                                   std::max<int64_t>(p->bs->ncells, 1)));
  }
  p->mono = m;
  drop_graph(p);
  return APH_OK;
}

int aph_plan_set_workgroup(aph_context *ctx, aph_plan *p, int32_t lanes) {
  if (!ctx) return APH_E_INVALID;
  if (!p || (lanes != 64 && lanes != 128 && lanes != 256))
    return set_error(ctx, APH_E_INVALID, "aph_plan_set_workgroup: lanes must be 64, 128 or 256");
  const int keep = p->wg;
  p->wg = lanes;
  p->dirty = true;
  int rc = build_items(ctx, p);
  if (rc) {
    p->wg = keep;
    p->dirty = true;
    build_items(ctx, p);
    return rc;
  }
  drop_graph(p);
  return APH_OK;
}

int aph_plan_autotune(aph_context *ctx, aph_plan *p, const double *u_in, double *u_scratch,
                      const double *dt_dev, int32_t reps, char *report, size_t report_len) {
  if (!ctx) return APH_E_INVALID;
  if (!p || !u_in || !u_scratch || !dt_dev || reps < 1 || (const double *)u_scratch == u_in)
    return set_error(ctx, APH_E_INVALID, "aph_plan_autotune: bad arguments");
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  APH_HIP_CHECK(ctx, hipStreamIsCapturing(ctx->stream, &cs));
  if (cs != hipStreamCaptureStatusNone)
    return set_error(ctx, APH_E_STATE, "aph_plan_autotune: stream is capturing a graph");
  if (p->dirty) {
    int rc = build_items(ctx, p);
    if (rc) return rc;
  }
  const std::vector<int4> shapes = p->class_shape;
  const int ndim = p->bs->ndim;
  std::string rep;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  APH_HIP_CHECK(ctx, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return set_error(ctx, APH_E_HIP, "hipEventCreate failed");
  }
  int rc = APH_OK;
  char line[200];
  // average stage-1 launch time (ms) of class c with the plan's current settings; false
  // when that configuration cannot be built (e.g. the tile exceeds LDS)
  auto time_class = [&](int c, double &avg) -> bool {
    p->only_class = c;
    if (build_items(ctx, p) != APH_OK) return false;
    rc = launch_kernel(ctx, p, p->recon, 1, 2, 0.0, 1.0, 1.0, const_cast<double *>(u_in),
                       nullptr, u_scratch, dt_dev, 0.0);
    if (rc) return false;
    if (hipEventRecord(e0, ctx->stream) != hipSuccess) {
      rc = set_error(ctx, APH_E_HIP, "autotune: hipEventRecord failed");
      return false;
    }
    for (int r = 0; r < reps && rc == APH_OK; ++r)
      rc = launch_kernel(ctx, p, p->recon, 1, 2, 0.0, 1.0, 1.0, const_cast<double *>(u_in),
                         nullptr, u_scratch, dt_dev, 0.0);
    if (rc) return false;
    float ms = 0.f;
    if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = set_error(ctx, APH_E_HIP, "autotune: event timing failed");
      return false;
    }
    avg = ms / reps;
    return true;
  };
  auto user_tile = [&](const int4 &sh) {
    for (const auto &u : p->user_tiles)
      if (u.n == sh.x) return true;
    return false;
  };
  auto set_tuned = [&](const int4 &sh, const aph_tile &t, int scale) {
    auto &tv = p->tuned;
    tv.erase(std::remove_if(tv.begin(), tv.end(),
                            [&](const aph_plan::Tuned &e) {
                              return e.shape.x == sh.x && e.shape.y == sh.y && e.shape.z == sh.z;
                            }),
             tv.end());
    tv.push_back({sh, t, scale});
  };
  auto segs = [&](const aph_tile &t, int wg, int scale, int *sg) {
    const int tt[3] = {t.tx, t.ty, t.tz};
    for (int d = 0; d < 3; ++d) {
      const int A = (d == 0) ? 1 : 0, B = (d == 2) ? 1 : 2;
      sg[d] = d < ndim ? seg_len(tt[d], tt[A] * tt[B], wg, scale) : 1;
    }
  };
  // the workgroup size is one per launch, so it is tuned for the whole plan: for each
  // size, the best (tile, segment scale) of every block shape, then the size whose summed
  // best times are smallest
  struct Best {
    aph_tile tile;
    int scale;
    double ms;
  };
  const int wgs[3] = {64, 128, 256};
  double best_total = 1e300;
  int best_wg = p->wg;
  std::vector<Best> best_cfg;
  for (int wg : wgs) {
    if (rc) break;
    p->wg = wg;
    std::vector<Best> cfg(shapes.size());
    double total = 0.0;
    bool ok = true;
    for (size_t c = 0; c < shapes.size() && rc == APH_OK && ok; ++c) {
      const int4 sh = shapes[c];
      snprintf(line, sizeof line, "workgroup %d, block %dx%dx%d:%s\n", wg, sh.x, sh.y, sh.z,
               user_tile(sh) ? " tile set by aph_plan_set_tile, not tuned" : "");
      rep += line;
      std::vector<aph_tile> cand;
      if (user_tile(sh)) {
        for (const auto &u : p->user_tiles)
          if (u.n == sh.x) cand.push_back(u);
      } else {
        cand = tile_candidates(sh, ndim, p->bs->nvar, wg);
      }
      Best b{default_tile(p->bs->h[sh.w], ndim), 1, 1e300};
      for (const aph_tile &t : cand) {
        for (int scale = 1; scale <= (user_tile(sh) ? 1 : 2); ++scale) {
          int s1[3], s2[3];
          segs(t, wg, 1, s1);
          segs(t, wg, scale, s2);
          if (scale == 2 && s1[0] == s2[0] && s1[1] == s2[1] && s1[2] == s2[2]) continue;
          if (!user_tile(sh)) set_tuned(sh, t, scale);
          double avg = 0.0;
          if (!time_class((int)c, avg)) {
            if (rc) break;
            continue;
          }
          snprintf(line, sizeof line, "  %3dx%3dx%3d  segments %d/%d/%d  %.4f ms\n", t.tx, t.ty,
                   t.tz, s2[0], s2[1], s2[2], avg);
          rep += line;
          if (avg < b.ms) b = Best{t, scale, avg};
        }
        if (rc) break;
      }
      if (b.ms >= 1e300) ok = false;
      cfg[c] = b;
      total += b.ms;
    }
    if (rc) break;
    // a larger workgroup has to win by more than 2%: the stage-1 launches timed here on
    // the initial state rank the sizes within about that much of each other, and on the
    // 256^3 blast 128 lanes timed 0.4% faster here but ran 4% slower over 100 steps
    if (ok && total < (wg == wgs[0] ? 1.0 : 0.98) * best_total) {
      best_total = total;
      best_wg = wg;
      best_cfg = cfg;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  p->only_class = -1;
  p->wg = best_wg;
  for (size_t c = 0; c < best_cfg.size(); ++c)
    if (!user_tile(shapes[c])) set_tuned(shapes[c], best_cfg[c].tile, best_cfg[c].scale);
  snprintf(line, sizeof line, "chosen: workgroup %d\n", best_wg);
  rep += line;
  for (size_t c = 0; c < best_cfg.size(); ++c) {
    int sg[3];
    segs(best_cfg[c].tile, best_wg, best_cfg[c].scale, sg);
    snprintf(line, sizeof line, "  block %dx%dx%d: tile %dx%dx%d, segments %d/%d/%d\n",
             shapes[c].x, shapes[c].y, shapes[c].z, best_cfg[c].tile.tx, best_cfg[c].tile.ty,
             best_cfg[c].tile.tz, sg[0], sg[1], sg[2]);
    rep += line;
  }
  p->dirty = true;
  const int rb = build_items(ctx, p);
  drop_graph(p);
  if (report && report_len > 0) {
    const size_t n = std::min(report_len - 1, rep.size());
    memcpy(report, rep.data(), n);
    report[n] = '\0';
  }
  if (rc) return rc;
  return rb;
}

int aph_plan_get_launch(aph_context *ctx, aph_plan *p, int32_t block, int32_t out[4]) {
  if (!ctx) return APH_E_INVALID;
  if (!p || !out || block < 0 || block >= p->bs->nblocks)
    return set_error(ctx, APH_E_INVALID, "aph_plan_get_launch: bad arguments");
  if (p->dirty) {
    int rc = build_items(ctx, p);
    if (rc) return rc;
  }
  const aph_block &k = p->bs->h[block];
  for (size_t c = 0; c < p->class_shape.size(); ++c) {
    const int4 &sh = p->class_shape[c];
    if (sh.x == k.nx[0] && sh.y == k.nx[1] && sh.z == k.nx[2]) {
      const aph_tile &t = p->tiles[c];
      const int tt[3] = {t.tx, t.ty, t.tz};
      out[0] = p->wg;
      for (int d = 0; d < 3; ++d) {
        const int A = (d == 0) ? 1 : 0, B = (d == 2) ? 1 : 2;
        out[1 + d] = d < p->bs->ndim ? seg_len(tt[d], tt[A] * tt[B], p->wg, p->seg_scale[c]) : 1;
      }
      return APH_OK;
    }
  }
  return set_error(ctx, APH_E_STATE, "aph_plan_get_launch: block shape not in plan");
}

int aph_plan_get_tile(aph_context *ctx, aph_plan *p, int32_t block, aph_tile *out) {
  if (!ctx) return APH_E_INVALID;
  if (!p || !out || block < 0 || block >= p->bs->nblocks)
    return set_error(ctx, APH_E_INVALID, "aph_plan_get_tile: bad arguments");
  if (p->dirty) {
    int rc = build_items(ctx, p);
    if (rc) return rc;
  }
  const aph_block &k = p->bs->h[block];
  for (size_t c = 0; c < p->class_shape.size(); ++c) {
    const int4 &sh = p->class_shape[c];
    if (sh.x == k.nx[0] && sh.y == k.nx[1] && sh.z == k.nx[2]) {
      *out = p->tiles[c];
      out->n = k.nx[0];
      return APH_OK;
    }
  }
  return set_error(ctx, APH_E_STATE, "aph_plan_get_tile: block shape not in plan");
}

int aph_plan_set_integrator(aph_context *ctx, aph_plan *p, int32_t integrator) {
  if (!ctx) return APH_E_INVALID;
  if (!p) return set_error(ctx, APH_E_INVALID, "aph_plan_set_integrator: null plan");
  if (integrator < APH_INTEGRATOR_RK1 || integrator > APH_INTEGRATOR_VL2)
    return set_error(ctx, APH_E_INVALID, "[Apophis]: Integrator not recognized. Exiting.");
  if (integrator == APH_INTEGRATOR_VL2 &&
      !pick_kernel(RECON_DC, p->rsolver, p->bs->nvar, p->bs->ndim))
    return set_error(ctx, APH_E_INVALID, "vl2: no first-order kernel for this Riemann solver");
  if (integrator == APH_INTEGRATOR_RK3 && !p->d_u2) {
    if (hipMalloc(&p->d_u2, sizeof(double) * p->bs->nvar * p->bs->ncells) != hipSuccess) {
      p->d_u2 = nullptr;
      return set_error(ctx, APH_E_NOMEM, "rk3 register allocation failed");
    }
  }
  p->integrator = integrator;
  p->dirty = true;  // the grid is sized over the kernel variants the integrator launches
  drop_graph(p);
  return APH_OK;
}

int aph_plan_status(aph_context *ctx, aph_plan *p) {
  if (!ctx) return APH_E_INVALID;
  if (!p) return set_error(ctx, APH_E_INVALID, "aph_plan_status: null plan");
  if (!p->d_err) return APH_OK;
  APH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  int err = 0;
  APH_HIP_CHECK(ctx, hipMemcpy(&err, p->d_err, sizeof(int), hipMemcpyDeviceToHost));
  if (!err) return APH_OK;
  APH_HIP_CHECK(ctx, hipMemset(p->d_err, 0, sizeof(int)));  // reported: cleared
  return set_error(ctx, APH_E_STATE,
                   "multilevel stage: a coarse tile's wait for the finer level's interface "
                   "fluxes ran out; the states, dt and time written since the last "
                   "aph_plan_status are invalid (NaN)");
}

int aph_plan_set_wait_limit(aph_context *ctx, aph_plan *p, int32_t polls) {
  if (!ctx) return APH_E_INVALID;
  if (!p) return set_error(ctx, APH_E_INVALID, "aph_plan_set_wait_limit: null plan");
  p->wait_limit = polls == 0 ? kDefaultWaitLimit : polls;
  drop_graph(p);  // the limit is a launch argument
  return APH_OK;
}

int aph_fused_stage(aph_context *ctx, aph_plan *p, int32_t stage, double *u_in,
                    const double *u_n, double *u_out, const double *dt_dev) {
  if (!ctx) return APH_E_INVALID;
  if (!p) return set_error(ctx, APH_E_INVALID, "aph_fused_stage: null plan");
  const int nst = plan_nstages(p);
  const bool empty = p->bs->nblocks == 0;  // an empty block set: state pointers may be null
  if ((!empty && (!u_in || !u_out || (stage > 1 && !u_n))) || !dt_dev || stage < 1 ||
      stage > nst)
    return set_error(ctx, APH_E_INVALID, "aph_fused_stage: bad arguments");
  if (u_in == u_out) return set_error(ctx, APH_E_INVALID, "u_out must not alias u_in");
  if (stage == 1) {  // the step's dt estimate starts here
    k_fill<<<1, 1, 0, ctx->stream>>>(p->d_est, 1.79769313486231570815e308);
    APH_LAUNCH_CHECK(ctx);
  }
  return launch_stage(ctx, p, stage, u_in, u_n, u_out, dt_dev, 0.0);
}

int aph_plan_step(aph_context *ctx, aph_plan *p, double *u0, double *u1, double *dt_dev,
                  double *time_dev, double cfl, double tlim, int32_t fixed_dt) {
  if (!ctx) return APH_E_INVALID;
  if (!p || !dt_dev || !time_dev ||
      (p->bs->nblocks > 0 && (!u0 || !u1 || u0 == u1)))  // empty set: states may be null
    return set_error(ctx, APH_E_INVALID, "aph_plan_step: bad arguments");
  k_fill<<<1, 1, 0, ctx->stream>>>(p->d_est, 1.79769313486231570815e308);
  APH_LAUNCH_CHECK(ctx);
  // registers: U0 holds the step's start state until the last stage overwrites it
  // (a lane reads its cell's U0 before writing it); stage states ping-pong in U1 / U2
  int rc = launch_stage(ctx, p, 1, u0, nullptr, u1, dt_dev, cfl);
  if (rc) return rc;
  switch (p->integrator) {
    case APH_INTEGRATOR_RK1:
      APH_HIP_CHECK(ctx, hipMemcpyAsync(u0, u1, sizeof(double) * p->bs->nvar * p->bs->ncells,
                                        hipMemcpyDeviceToDevice, ctx->stream));
      break;
    case APH_INTEGRATOR_RK3:
      rc = launch_stage(ctx, p, 2, u1, u0, p->d_u2, dt_dev, cfl);
      if (rc) return rc;
      rc = launch_stage(ctx, p, 3, p->d_u2, u0, u0, dt_dev, cfl);
      break;
    default:  // rk2, vl2
      rc = launch_stage(ctx, p, 2, u1, u0, u0, dt_dev, cfl);
  }
  if (rc) return rc;
  k_dt_finalize<<<1, 1, 0, ctx->stream>>>(dt_dev, time_dev, p->d_est, tlim, fixed_dt,
                                          p->amr ? p->d_err : nullptr);
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

int aph_plan_graph_capture(aph_context *ctx, aph_plan *p, int32_t nsteps, double *u0,
                           double *u1, double *dt_dev, double *time_dev, double cfl,
                           double tlim, int32_t fixed_dt) {
  if (!ctx) return APH_E_INVALID;
  if (!p || nsteps < 1) return set_error(ctx, APH_E_INVALID, "aph_plan_graph_capture: bad args");
  if (!ctx->stream)
    return set_error(ctx, APH_E_STATE, "graph capture needs a non-default stream");
  if (p->poisson)
    return set_error(ctx, APH_E_STATE,
                     "graph capture: the Poisson solve tests its convergence on the host");
  if (p->dirty) {
    int rc = build_items(ctx, p);
    if (rc) return rc;
  }
  if (p->exec) {
    (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
  }
  if (p->graph) {
    (void)hipGraphDestroy(p->graph);
    p->graph = nullptr;
  }
  APH_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = APH_OK;
  for (int s = 0; s < nsteps && rc == APH_OK; ++s)
    rc = aph_plan_step(ctx, p, u0, u1, dt_dev, time_dev, cfl, tlim, fixed_dt);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
  if (rc != APH_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess)
    return set_error(ctx, APH_E_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return set_error(ctx, APH_E_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
  }
  p->graph = graph;
  return APH_OK;
}

int aph_plan_graph_launch(aph_context *ctx, aph_plan *p) {
  if (!ctx) return APH_E_INVALID;
  if (!p) return set_error(ctx, APH_E_INVALID, "null plan");
  if (!p->exec) return set_error(ctx, APH_E_STATE, "no captured graph in this plan");
  APH_HIP_CHECK(ctx, hipGraphLaunch(p->exec, ctx->stream));
  return APH_OK;
}

int aph_plan_estimate_timestep(aph_context *ctx, aph_plan *p, const double *cons, double cfl,
                               double *dt_out) {
  if (!ctx) return APH_E_INVALID;
  if (!p || !dt_out || (!cons && p->bs->nblocks > 0))
    return set_error(ctx, APH_E_INVALID, "aph_plan_estimate_timestep: null");
  const aph_blockset *bs = p->bs;
  k_fill<<<1, 1, 0, ctx->stream>>>(dt_out, 1.79769313486231570815e308);
  if (bs->nblocks > 0) {
    int64_t x = std::min<int64_t>(64, (bs->max_interior + 255) / 256);
    dim3 gr((unsigned)std::max<int64_t>(1, x), (unsigned)bs->nblocks);
#define APH_PDT(NVv, ND) k_plan_dt<NVv, ND><<<gr, 256, 0, ctx->stream>>>(bs->d, p->eos, cons, cfl, dt_out)
    if (bs->nvar == 5) {
      if (bs->ndim == 1) APH_PDT(5, 1); else if (bs->ndim == 2) APH_PDT(5, 2); else APH_PDT(5, 3);
    } else if (bs->nvar == 6) {
      if (bs->ndim == 1) APH_PDT(6, 1); else if (bs->ndim == 2) APH_PDT(6, 2); else APH_PDT(6, 3);
    } else {
      PlanDtFn f = pick_plan_dt_wide(bs->nvar, bs->ndim);
      if (!f) return set_error(ctx, APH_E_INVALID, "aph_plan_estimate_timestep: unsupported nvar");
      hipLaunchKernelGGL(f, gr, dim3(256), 0, ctx->stream, bs->d, p->eos, cons, cfl, dt_out);
    }
#undef APH_PDT
  }
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

}  // extern "C"

extern "C" {

int aph_adjoint_stage(aph_context *ctx, aph_plan *p, int32_t stage, const double *u_in,
                      const double *dt_dev, const double *lam_out, const double *lam_acc,
                      double *lam_in, double *lam_n) {
  if (!ctx) return APH_E_INVALID;
  if (!p || !u_in || !dt_dev || !lam_out || !lam_in)
    return set_error(ctx, APH_E_INVALID, "aph_adjoint_stage: bad arguments");
  int rc = adj_plan_ok(ctx, p);
  if (rc) return rc;
  if (lam_in == lam_out || (lam_acc && lam_acc == lam_in))
    return set_error(ctx, APH_E_INVALID, "aph_adjoint_stage: lam_in must not alias inputs");
  AdjCoef c;
  int sr;
  if (!adj_coef(p, stage, c, sr))
    return set_error(ctx, APH_E_INVALID, "aph_adjoint_stage: stage %d out of range", stage);
  return plan_adj_stage(ctx, p, sr, c, u_in, dt_dev, lam_out, lam_acc, lam_in, lam_n, 0);
}

int aph_adjoint_step(aph_context *ctx, aph_plan *p, double *u0, double *u1,
                     const double *dt_dev, double *lam) {
  if (!ctx) return APH_E_INVALID;
  if (!p || !u0 || !u1 || !dt_dev || !lam || u0 == u1)
    return set_error(ctx, APH_E_INVALID, "aph_adjoint_step: bad arguments");
  int rc = adj_plan_ok(ctx, p);
  if (rc) return rc;
  rc = adj_prepare(ctx, p->bs, p->recon, p->adj);
  if (rc) return rc;
  // recompute the stage-1 state from the step's checkpoint, then reverse the stages
  if (plan_nstages(p) > 1) {
    rc = launch_stage(ctx, p, 1, u0, nullptr, u1, dt_dev, 0.0);
    if (rc) return rc;
  }
  return aph_adjoint_step_stored(ctx, p, u0, u1, dt_dev, lam);
}

int aph_adjoint_step_stored(aph_context *ctx, aph_plan *p, const double *u0, const double *u1,
                            const double *dt_dev, double *lam) {
  if (!ctx) return APH_E_INVALID;
  if (!p || !u0 || !u1 || !dt_dev || !lam || u0 == u1 || lam == u0 || lam == u1)
    return set_error(ctx, APH_E_INVALID, "aph_adjoint_step_stored: bad arguments");
  int rc = adj_plan_ok(ctx, p);
  if (rc) return rc;
  rc = adj_prepare(ctx, p->bs, p->recon, p->adj);
  if (rc) return rc;
  const aph_blockset *bs = p->bs;
  AdjScratch &A = p->adj;
  const int n = plan_nstages(p);
  AdjCoef c[3];
  int sr[3];
  for (int s = 1; s <= n; ++s) adj_coef(p, s, c[s - 1], sr[s - 1]);
  auto stage = [&](int s, const double *x, const double *lo, const double *acc, double *li,
                   double *ln, int add) {
    return plan_adj_stage(ctx, p, sr[s - 1], c[s - 1], x, dt_dev, lo, acc, li, ln, add);
  };
  if (n == 1) {  // rk1: U0' = U0 + dt L(U0); lam must not be both input and output
    APH_HIP_CHECK(ctx, hipMemcpyAsync(A.lam1, lam, sizeof(double) * bs->nvar * bs->ncells,
                                      hipMemcpyDeviceToDevice, ctx->stream));
    return stage(1, u0, A.lam1, nullptr, lam, nullptr, 0);
  }
  if (n == 2) {  // rk2, vl2: stage 2 (input u1), then stage 1 (input u0)
    rc = stage(2, u1, lam, nullptr, A.lam1, A.lam0p, 0);
    if (rc) return rc;
    return stage(1, u0, A.lam1, A.lam0p, lam, nullptr, 0);
  }
  // rk3: the stage-2 state X3 = stage2(u1, u0) is recomputed into the plan's third
  // register; then stage 3 (input X3), stage 2 (input u1), stage 1 (input u0), the
  // gam1 parts of stages 3 and 2 accumulated into the cotangent of u0
  if (!A.lam2 && hipMalloc(&A.lam2, sizeof(double) * bs->nvar * bs->ncells) != hipSuccess) {
    A.lam2 = nullptr;
    return set_error(ctx, APH_E_NOMEM, "adjoint: rk3 cotangent register allocation failed");
  }
  rc = launch_stage(ctx, p, 2, const_cast<double *>(u1), u0, p->d_u2, dt_dev, 0.0);
  if (rc) return rc;
  rc = stage(3, p->d_u2, lam, nullptr, A.lam1, A.lam0p, 0);
  if (rc) return rc;
  rc = stage(2, u1, A.lam1, nullptr, A.lam2, A.lam0p, 1);
  if (rc) return rc;
  return stage(1, u0, A.lam2, A.lam0p, lam, nullptr, 0);
}

}  // extern "C"
