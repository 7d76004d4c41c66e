// aph_fused_amr.hip -- the multilevel-mesh variant of k_fused (block faces to another
// level record or correct their interface fluxes), NV = 5 and 6, ndim 2 and 3.  Compiled
// apart from aph_fused.hip so the single-level kernels carry none of that code and the
// instantiation sets build in parallel.
#include "aph_fused_kernel.hpp"

namespace aph {
namespace {

template <int RECON, int RS, int NV>
KernelFn dim_amr(int ndim) {
  return ndim == 2 ? k_fused<RECON, RS, NV, 2, true> : k_fused<RECON, RS, NV, 3, true>;
}

template <int RECON, int RS>
KernelFn nv_amr(int nv, int ndim) {
  if (ndim < 2) return nullptr;
  switch (nv) {
    case 5: return dim_amr<RECON, RS, 5>(ndim);
    case 6: return dim_amr<RECON, RS, 6>(ndim);
    default: return nullptr;
  }
}

}  // namespace

KernelFn pick_kernel_amr(int recon, int rs, int nv, int ndim) {
#define APH_PICK(R, S) \
  if (recon == R && rs == S) return nv_amr<R, S>(nv, ndim);
  APH_PICK(RECON_PPM, RS_HLLC)
  APH_PICK(RECON_PPM, RS_HLLE)
  APH_PICK(RECON_PPM, RS_LHLLC)
  APH_PICK(RECON_PLM, RS_HLLC)
  APH_PICK(RECON_PLM, RS_HLLE)
  APH_PICK(RECON_PLM, RS_LHLLC)
  APH_PICK(RECON_DC, RS_HLLC)
  APH_PICK(RECON_DC, RS_HLLE)
  APH_PICK(RECON_DC, RS_LHLLC)
#undef APH_PICK
  return nullptr;
}

}  // namespace aph
