// aph_adjoint_host.hpp -- host interface of the adjoint stage kernels (aph_adjoint.hip).
#pragma once

#include <cstdint>

#include "aph_runtime.hpp"

namespace aph {

// Device scratch of the adjoint: primitive-space cotangents (NV+2 fields, pack layout),
// per-axis spill records and their per-block offsets, two cotangent registers used by a
// whole RK2 step, and for the split (PPM) form the primitive state in two layouts, the
// x-transposed L_out and the face-state cotangents.  Allocated on first use, released
// with the plan.
struct AdjScratch {
  double *contrib = nullptr;  // multilevel: per referenced cell of the ghost ops, NV values
  int64_t ncontrib = 0;
  double *pbar = nullptr;
  double *prim = nullptr, *prim_t = nullptr, *lam_t = nullptr, *qlb = nullptr, *qrb = nullptr;
  double *spill[3] = {nullptr, nullptr, nullptr};
  int64_t *spill_base = nullptr;
  double *lam1 = nullptr, *lam0p = nullptr;
  double *lam2 = nullptr;  // third cotangent register of a three-stage step (rk3)
  int64_t ncells = 0;
  int segcap[3] = {8, 8, 8};  // longest segment per axis (shorter when blocks are few)
};

// One stage of the plan's integrator: u_out = gam0 u_in + gam1 u_n + beta dt L(u_in), with
// stage == 1 when u_n is u_in (the step's first stage); plan_recon sizes the scratch (the
// split form's buffers when it is PPM), the recon argument of adj_stage is the stage's.
struct AdjCoef {
  double gam0, gam1, beta;
  int stage;
  int plan_recon;
};
// A multilevel mesh seen by the adjoint (all null / 0 on a single-level plan): the fused
// form's interface table with the coarse block of each face to a coarser leaf, the
// cross-level ghost ops with the block of each op's ghost cell, and the cells the ops read
// as rows of a transposed list (row t: cell tgt[t] with component stride tgt_stride[t],
// read through refs col[row[t] .. row[t+1]) in ascending ref order).
struct AdjAmr {
  const aph_amr_face *faces = nullptr;
  const int32_t *cblk = nullptr;
  const aph_amr_op *ops = nullptr;
  int64_t nops = 0;
  const aph_amr_ref *refs = nullptr;
  int64_t nrefs = 0;
  const int32_t *op_block = nullptr;
  const int64_t *tgt = nullptr, *tgt_stride = nullptr, *row = nullptr;
  const int32_t *col = nullptr;
  int64_t ntgt = 0;
};
// The ghost-op part of the view, from an amr plan (aph_amr.hip).
void amr_adj_view(const aph_amr_plan *p, AdjAmr &v);

int adj_stage(aph_context *ctx, const aph_blockset *bs, int recon, int rs, const aph_eos &eos,
              const double *g, int grav, AdjScratch &s, const AdjCoef &c, const double *u_in,
              const double *dt_dev, const double *lam_out, const double *lam_acc,
              double *lam_in, double *lam_n, int lam_n_add, const AdjAmr *amr = nullptr);
int adj_prepare(aph_context *ctx, const aph_blockset *bs, int recon, AdjScratch &s);
void adj_release(AdjScratch &s);
bool adj_supported(int recon, int rs, int nv, int nd);

}  // namespace aph
