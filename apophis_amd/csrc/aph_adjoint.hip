// aph_adjoint.hip -- adjoint (reverse-mode) kernels of one fused RK2 stage, with frozen dt.
//
// Forward stage (aph_fused.hip):  U_out = gam0 U + gam1 U_n + beta dt (-(1/V) sum_d A_d
// [F(c+1/2) - F(c-1/2)]) + gravity(prim(U)), with F = Riemann(recon_L, recon_R) of prim(U)
// on the stencil (neighbour blocks / boundary rule beyond the block).  Given the cotangent
// L_out of U_out this computes
//   L_in = (gam0 [+ gam1 in stage 1]) L_out + (stencil part) [+ L_acc],
//   L_n  = gam1 L_out                                   (stage 2).
// Structure (deterministic, no atomics), per stage:
//   PPM, split form (the PPM+HLLC adjoint does not fit one kernel's registers at two waves
//   per SIMD):
//     k_adj_prim       P = ConsToPrim(U) of every interior cell, in the pack layout and in
//                      the x-transposed layout, and L_out x-transposed;
//     k_adj_face       (every axis in one launch, blockIdx.z = axis) one lane per (block,
//                      pencil, segment) re-marches the forward reconstruction over its
//                      faces from P and back-propagates the face cotangents
//                      mu_f = beta dt A/V (L_out(f) - L_out(f-1)) through the Riemann
//                      solver: QL_bar / QR_bar (per axis), the cotangents of every cell's
//                      two reconstructed face states;
//     k_adj_recon      (every axis in one launch) one lane per (block, pencil, segment,
//                      field) applies the adjoint of that field's reconstruction of each
//                      cell to them.
//   PLM, one march per axis:
//     k_adj_axes       (every axis in one launch) one lane per (block, pencil, segment)
//                      re-marches the forward (ConsToPrim, reconstruction, Riemann) and
//                      back-propagates mu_f through the Riemann solver and both adjacent
//                      reconstructions.
//   Both forms write the primitive-space cotangent of a segment's own cells to the axis'
//   own Pbar array (the axes are independent) and that of the <= 3 cells it reads beyond
//   the segment on each side to a 6-slot spill record (primitive space; at a
//   reflecting-boundary image the normal velocity's cotangent is flipped back to the
//   interior cell's sign).
//   k_adj_gather     per interior cell: the axes' Pbar summed x, y, z, plus the spills of
//                    the segments / blocks / boundary images that read this cell (fixed
//                    order), the gravity adjoint, J_prim^T, and the gam terms.
// The x sweeps of the split form keep P, L_out and QL_bar / QR_bar in the x-transposed
// layout (y fastest, then z, then x): the lanes of a wave take adjacent pencils, so every
// load of the march reads adjacent addresses.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "aph_adjoint_kernels.hpp"

namespace aph {

// the composition counts NV 12 and 14 (aph_adjoint_wide.hip)
AxisKernels pick_adj_axis_wide(int recon, int rs, int nv, int nd, int d);
KFn pick_gather_wide(int nv, int nd, bool xt);
KFn pick_prim_wide(int nv, int nd);

AxisKernels pick_adj_axis(int recon, int rs, int nv, int nd, int d) {
  if (nv == 12 || nv == 14) return pick_adj_axis_wide(recon, rs, nv, nd, d);
  AxisKernels k;
#define APH_AX(R, S)                                                             \
  if (recon == R && rs == S) {                                                    \
    if (nv == 5) axis_nd<R, S, 5>(nd, d, k);                                      \
    if (nv == 6) axis_nd<R, S, 6>(nd, d, k);                                      \
  }
  APH_AX(RECON_PPM, RS_HLLC)
  APH_AX(RECON_PPM, RS_HLLE)
  APH_AX(RECON_PPM, RS_LHLLC)
  APH_AX(RECON_PLM, RS_HLLC)
  APH_AX(RECON_PLM, RS_HLLE)
  APH_AX(RECON_PLM, RS_LHLLC)
  APH_AX(RECON_DC, RS_HLLC)   // vl2's first-order predictor stage
  APH_AX(RECON_DC, RS_HLLE)
  APH_AX(RECON_DC, RS_LHLLC)
#undef APH_AX
  return k;
}

bool adj_supported(int recon, int rs, int nv, int nd) {
  const AxisKernels k = pick_adj_axis(recon, rs, nv, nd, 0);
  return k.face != nullptr || k.axis != nullptr;
}

KFn pick_nd(int nd, KFn k1, KFn k2, KFn k3) {
  return nd == 1 ? k1 : (nd == 2 ? k2 : k3);
}

KFn pick_gather(int nv, int nd, bool xt) {
  if (nv == 12 || nv == 14) return pick_gather_wide(nv, nd, xt);
  if (nv == 5)
    return xt ? pick_nd(nd, k_adj_gather<5, 1, true>, k_adj_gather<5, 2, true>,
                        k_adj_gather<5, 3, true>)
              : pick_nd(nd, k_adj_gather<5, 1, false>, k_adj_gather<5, 2, false>,
                        k_adj_gather<5, 3, false>);
  if (nv == 6)
    return xt ? pick_nd(nd, k_adj_gather<6, 1, true>, k_adj_gather<6, 2, true>,
                        k_adj_gather<6, 3, true>)
              : pick_nd(nd, k_adj_gather<6, 1, false>, k_adj_gather<6, 2, false>,
                        k_adj_gather<6, 3, false>);
  return nullptr;
}

// the multilevel kernels: which 0 = k_adj_amr_prim, 1 = k_adj_amr_ghost, 2 =
// k_adj_amr_scatter (multilevel meshes are 2D or 3D)
template <int NV>
KFn amr_kernel(int nd, int which) {
  if (which == 0) return nd == 3 ? k_adj_amr_prim<NV, 3> : k_adj_amr_prim<NV, 2>;
  if (which == 1) return nd == 3 ? k_adj_amr_ghost<NV, 3> : k_adj_amr_ghost<NV, 2>;
  return k_adj_amr_scatter<NV>;
}

KFn pick_amr(int nv, int nd, int which) {
  if (nv == 5) return amr_kernel<5>(nd, which);
  if (nv == 6) return amr_kernel<6>(nd, which);
  return nullptr;
}

KFn pick_prim(int nv, int nd) {
  if (nv == 12 || nv == 14) return pick_prim_wide(nv, nd);
  if (nv == 5) return pick_nd(nd, k_adj_prim<5, 1>, k_adj_prim<5, 2>, k_adj_prim<5, 3>);
  if (nv == 6) return pick_nd(nd, k_adj_prim<6, 1>, k_adj_prim<6, 2>, k_adj_prim<6, 3>);
  return nullptr;
}

int adj_prepare(aph_context *ctx, const aph_blockset *bs, int recon, AdjScratch &s) {
  const int nv = bs->nvar, nf = nv + 2, nd = bs->ndim, nb = bs->nblocks;
  if (recon == RECON_PPM && !s.prim) {  // buffers of the split form
    if (hipMalloc(&s.prim, sizeof(double) * nf * bs->ncells) != hipSuccess ||
        hipMalloc(&s.prim_t, sizeof(double) * nf * bs->ncells) != hipSuccess ||
        hipMalloc(&s.lam_t, sizeof(double) * nv * bs->ncells) != hipSuccess ||
        hipMalloc(&s.qlb, sizeof(double) * nf * bs->ncells * nd) != hipSuccess ||
        hipMalloc(&s.qrb, sizeof(double) * nf * bs->ncells * nd) != hipSuccess)
      return set_error(ctx, APH_E_NOMEM, "adjoint scratch allocation failed");
  }
  if (s.pbar) return APH_OK;
  // segment cap per axis: the longest segments (fewest recomputed halo cells) that still
  // give >= 2 waves of 64 lanes per SIMD (MI355X: 256 CUs x 4 SIMDs); up to 16 cells for
  // the split (PPM) form, whose passes move each segment's halo through HBM, 8 for the
  // one-march (PLM) form
  constexpr int64_t kTargetLanes = 2LL * 64 * 1024;
  const int cap0 = (recon == RECON_PPM) ? kSplitSegMax : 8;
  for (int d = 0; d < nd; ++d) {
    const int A = (d == 0) ? 1 : 0, B = (d == 2) ? 1 : 2;
    int best = cap0;
    int64_t best_lanes = -1;
    for (int cap = cap0; cap >= 3; --cap) {
      int64_t lanes = 0;
      for (int b = 0; b < nb; ++b) {
        const aph_block &k = bs->h[b];
        lanes += (int64_t)k.nx[A] * k.nx[B] * (k.nx[d] / adj_seg(k.nx[d], cap));
      }
      if (lanes > best_lanes) {
        best_lanes = lanes;
        best = cap;
      }
      if (lanes >= kTargetLanes) {
        best = cap;
        break;
      }
    }
    s.segcap[d] = best;
  }
  std::vector<int64_t> base(3 * (size_t)nb, 0);
  int64_t tot[3] = {0, 0, 0};
  for (int d = 0; d < nd; ++d)
    for (int b = 0; b < nb; ++b) {
      const aph_block &k = bs->h[b];
      const int A = (d == 0) ? 1 : 0, B = (d == 2) ? 1 : 2;
      const int nseg = k.nx[d] / adj_seg(k.nx[d], s.segcap[d]);
      base[(size_t)d * nb + b] = tot[d];
      tot[d] += (int64_t)k.nx[A] * k.nx[B] * nseg * 6 * nf;
    }
  if (hipMalloc(&s.pbar, sizeof(double) * nf * bs->ncells * bs->ndim) != hipSuccess ||
      hipMalloc(&s.lam1, sizeof(double) * nv * bs->ncells) != hipSuccess ||
      hipMalloc(&s.lam0p, sizeof(double) * nv * bs->ncells) != hipSuccess ||
      hipMalloc(&s.spill_base, sizeof(int64_t) * base.size()) != hipSuccess)
    return set_error(ctx, APH_E_NOMEM, "adjoint scratch allocation failed");
  for (int d = 0; d < nd; ++d)
    if (hipMalloc(&s.spill[d], sizeof(double) * std::max<int64_t>(tot[d], 1)) != hipSuccess)
      return set_error(ctx, APH_E_NOMEM, "adjoint spill allocation failed");
  APH_HIP_CHECK(ctx, hipMemcpy(s.spill_base, base.data(), sizeof(int64_t) * base.size(),
                               hipMemcpyHostToDevice));
  APH_HIP_CHECK(ctx, hipMemset(s.pbar, 0, sizeof(double) * nf * bs->ncells * bs->ndim));
  APH_HIP_CHECK(ctx, hipMemset(s.lam1, 0, sizeof(double) * nv * bs->ncells));
  APH_HIP_CHECK(ctx, hipMemset(s.lam0p, 0, sizeof(double) * nv * bs->ncells));
  s.ncells = bs->ncells;
  return APH_OK;
}

void adj_release(AdjScratch &s) {
  if (s.contrib) (void)hipFree(s.contrib);
  s.contrib = nullptr;
  s.ncontrib = 0;
  for (double *p : {s.pbar, s.lam1, s.lam2, s.lam0p, s.prim, s.prim_t, s.lam_t, s.qlb, s.qrb})
    if (p) (void)hipFree(p);
  for (auto &p : s.spill)
    if (p) (void)hipFree(p);
  if (s.spill_base) (void)hipFree(s.spill_base);
  s = AdjScratch{};
}

int adj_stage(aph_context *ctx, const aph_blockset *bs, int recon, int rs, const aph_eos &eos,
              const double *g, int grav, AdjScratch &s, const AdjCoef &c, const double *u_in,
              const double *dt_dev, const double *lam_out, const double *lam_acc,
              double *lam_in, double *lam_n, int lam_n_add, const AdjAmr *amr) {
  if (bs->nblocks == 0) return APH_OK;
  int rc = adj_prepare(ctx, bs, c.plan_recon, s);
  if (rc) return rc;
  if (amr && amr->nrefs > s.ncontrib) {  // per ref of the ghost ops, NV cotangents
    if (s.contrib) (void)hipFree(s.contrib);
    s.contrib = nullptr;
    s.ncontrib = 0;
    if (hipMalloc(&s.contrib, sizeof(double) * bs->nvar * amr->nrefs) != hipSuccess)
      return set_error(ctx, APH_E_NOMEM, "adjoint scratch allocation failed");
    s.ncontrib = amr->nrefs;
  }
  const int nd = bs->ndim;
  AxisKernels ks[3];
  for (int d = 0; d < nd; ++d) {
    ks[d] = pick_adj_axis(recon, rs, bs->nvar, nd, d);  // the stage's reconstruction
    if (!ks[d].face && !ks[d].axis)
      return set_error(ctx, APH_E_INVALID, "adjoint: unsupported combination");
  }
  AdjArgs a{};
  a.blocks = bs->d;
  a.nblocks = bs->nblocks;
  a.eos = eos;
  a.g[0] = g[0];
  a.g[1] = g[1];
  a.g[2] = g[2];
  a.grav = grav;
  a.u_in = u_in;
  a.lam_out = lam_out;
  a.lam_acc = lam_acc;
  a.lam_in = lam_in;
  a.lam_n = lam_n;
  a.dt = dt_dev;
  a.gam0 = c.gam0;
  a.gam1 = c.gam1;
  a.beta = c.beta;
  a.stage = c.stage;
  a.lam_n_add = lam_n_add;
  a.pbar = s.pbar;
  a.pbar_axis = (int64_t)(bs->nvar + 2) * bs->ncells;
  a.qb_axis = (int64_t)(bs->nvar + 2) * bs->ncells;
  a.prim = s.prim;
  a.prim_t = s.prim_t;
  a.lam_t = s.lam_t;
  a.qlb = s.qlb;
  a.qrb = s.qrb;
  for (int d = 0; d < 3; ++d) a.spill[d] = s.spill[d];
  a.spill_base = s.spill_base;
  a.pbar_xt = ks[0].face != nullptr;  // the split form writes the x axis' Pbar transposed
  for (int d = 0; d < 3; ++d) a.segcap[d] = s.segcap[d];
  if (amr) a.amr = *amr;
  a.contrib = s.contrib;
  const int nf = bs->nvar + 2;
  // prim / gather: one 32 x TY (x, y) tile (32 x 1 in 1D) per workgroup pass
  auto tile_launch = [&](int ty, dim3 &grid, dim3 &block) {
    const int tyd = nd == 1 ? 1 : ty;
    int64_t tiles = 1;
    for (const auto &k : bs->h)
      tiles = std::max<int64_t>(tiles, (int64_t)((k.nx[0] + 31) / 32) *
                                           ((k.nx[1] + tyd - 1) / tyd) * k.nx[2]);
    grid = dim3((unsigned)std::min<int64_t>(64, tiles), (unsigned)bs->nblocks);
    block = dim3((unsigned)(32 * tyd));
  };
  dim3 gp, bp, gg, bg;
  tile_launch(prim_tile_y(nf), gp, bp);
  tile_launch(a.pbar_xt ? gather_tile_y(nf) : kGatherTileYMarch, gg, bg);
  if (ks[0].face) hipLaunchKernelGGL(pick_prim(bs->nvar, nd), gp, bp, 0, ctx->stream, a);
  const unsigned gops = amr ? (unsigned)std::max<int64_t>(
                                  1, std::min<int64_t>((amr->nops + 255) / 256, 4096))
                            : 1u;
  if (ks[0].face && amr && amr->nops)  // P of the cross-level ghost cells the sweeps read
    hipLaunchKernelGGL(pick_amr(bs->nvar, nd, 0), dim3(gops), dim3(256), 0, ctx->stream, a);
  const unsigned nbk = (unsigned)bs->nblocks;
  int64_t maxu_all = 0;
  for (int d = 0; d < nd; ++d) {
    int64_t maxu = 0;
    for (const auto &k : bs->h) {
      const int A = (d == 0) ? 1 : 0, B = (d == 2) ? 1 : 2;
      maxu = std::max<int64_t>(maxu, (int64_t)k.nx[A] * k.nx[B] *
                                         (k.nx[d] / adj_seg(k.nx[d], s.segcap[d])));
    }
    maxu_all = std::max(maxu_all, maxu);
  }
  if (ks[0].face) {  // split form: the faces of every axis, then their reconstructions
    hipLaunchKernelGGL(ks[0].face, dim3((unsigned)((maxu_all + 63) / 64), nbk, (unsigned)nd),
                       dim3(64), 0, ctx->stream, a);
    hipLaunchKernelGGL(ks[0].recon,
                       dim3((unsigned)((maxu_all * nf + 63) / 64), nbk, (unsigned)nd), dim3(64),
                       0, ctx->stream, a);
  }
  if (ks[0].axis)  // one-march form: all axes in one launch, z = axis
    hipLaunchKernelGGL(ks[0].axis, dim3((unsigned)((maxu_all + 63) / 64), nbk, (unsigned)nd),
                       dim3(64), 0, ctx->stream, a);
  hipLaunchKernelGGL(pick_gather(bs->nvar, nd, a.pbar_xt), gg, bg, 0, ctx->stream, a);
  if (amr && amr->nops) {  // the cross-level ghosts' cotangents back to the cells they read
    hipLaunchKernelGGL(pick_amr(bs->nvar, nd, 1), dim3(gops), dim3(256), 0, ctx->stream, a);
    const unsigned gt = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>((amr->ntgt + 255) / 256, 4096));
    hipLaunchKernelGGL(pick_amr(bs->nvar, nd, 2), dim3(gt), dim3(256), 0, ctx->stream, a);
  }
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

}  // namespace aph
