// aph_adjoint_wide.hip -- the adjoint kernels (aph_adjoint_kernels.hpp) instantiated for the
// composition counts: NV = 12 (5 hydro + 7 composition fields, composition = ONe) and
// NV = 14 (with one level set of two fields, nlset = 1), hydro.cpp:141-202.  A translation
// unit of its own so that these large instantiations compile beside the NV 5 / 6 ones.
#include "aph_adjoint_kernels.hpp"

namespace aph {

AxisKernels pick_adj_axis_wide(int recon, int rs, int nv, int nd, int d) {
  AxisKernels k;
#define APH_AX(R, S)                                                             \
  if (recon == R && rs == S) {                                                    \
    if (nv == 12) axis_nd<R, S, 12>(nd, d, k);                                    \
    if (nv == 14) axis_nd<R, S, 14>(nd, d, k);                                    \
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

namespace {
KFn nd_pick(int nd, KFn k1, KFn k2, KFn k3) { return nd == 1 ? k1 : (nd == 2 ? k2 : k3); }
}  // namespace

KFn pick_gather_wide(int nv, int nd, bool xt) {
  if (nv == 12)
    return xt ? nd_pick(nd, k_adj_gather<12, 1, true>, k_adj_gather<12, 2, true>,
                        k_adj_gather<12, 3, true>)
              : nd_pick(nd, k_adj_gather<12, 1, false>, k_adj_gather<12, 2, false>,
                        k_adj_gather<12, 3, false>);
  if (nv == 14)
    return xt ? nd_pick(nd, k_adj_gather<14, 1, true>, k_adj_gather<14, 2, true>,
                        k_adj_gather<14, 3, true>)
              : nd_pick(nd, k_adj_gather<14, 1, false>, k_adj_gather<14, 2, false>,
                        k_adj_gather<14, 3, false>);
  return nullptr;
}

KFn pick_prim_wide(int nv, int nd) {
  if (nv == 12) return nd_pick(nd, k_adj_prim<12, 1>, k_adj_prim<12, 2>, k_adj_prim<12, 3>);
  if (nv == 14) return nd_pick(nd, k_adj_prim<14, 1>, k_adj_prim<14, 2>, k_adj_prim<14, 3>);
  return nullptr;
}

}  // namespace aph
