// aph_fused_wide.hip -- k_fused / k_plan_dt for the composition scalar counts of the
// reference's hydro package (hydro.cpp:143-177): composition = ONe adds 6 species
// densities and Ye (NV = 12), each level set two more fields (nlset = 1: NV = 14).
// Compiled apart from aph_fused.hip so the two instantiation sets build in parallel.
#include "aph_fused_kernel.hpp"

namespace aph {
namespace {

template <int RECON, int RS, int NV>
KernelFn dim_wide(int ndim) {
  switch (ndim) {
    case 1: return k_fused<RECON, RS, NV, 1>;
    case 2: return k_fused<RECON, RS, NV, 2>;
    default: return k_fused<RECON, RS, NV, 3>;
  }
}

template <int RECON, int RS>
KernelFn nv_wide(int nv, int ndim) {
  switch (nv) {
    case 12: return dim_wide<RECON, RS, 12>(ndim);
    case 14: return dim_wide<RECON, RS, 14>(ndim);
    default: return nullptr;
  }
}

template <int NV>
PlanDtFn plan_dt_dim(int ndim) {
  switch (ndim) {
    case 1: return k_plan_dt<NV, 1>;
    case 2: return k_plan_dt<NV, 2>;
    default: return k_plan_dt<NV, 3>;
  }
}

}  // namespace

KernelFn pick_kernel_wide(int recon, int rs, int nv, int ndim) {
#define APH_PICK(R, S) \
  if (recon == R && rs == S) return nv_wide<R, S>(nv, ndim);
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

PlanDtFn pick_plan_dt_wide(int nv, int ndim) {
  if (nv == 12) return plan_dt_dim<12>(ndim);
  if (nv == 14) return plan_dt_dim<14>(ndim);
  return nullptr;
}

}  // namespace aph
