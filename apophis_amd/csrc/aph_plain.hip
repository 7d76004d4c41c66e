// aph_plain.hip -- the plain form: one kernel per reference task function, fluxes
// and primitives round-trip through HBM exactly as in the reference task graph.
//
//   aph_cons_to_prim                 <- ConsToPrim<EOS_t>         hydro.cpp:422-533
//   aph_calculate_fluxes             <- CalculateFluxes<F,R,S>    hydro.cpp:535-744
//   aph_update_with_flux_divergence  <- Update::UpdateWithFluxDivergence (upstream)
//   aph_apply_gravity                <- ApplyGravity              gravity.cpp:374-410
//   aph_fill_ghosts                  <- boundary exchange + ApplyBoundaryConditions
//   aph_estimate_timestep            <- EstimateTimestep<euler>   hydro.cpp:351-420
//   aph_global_sum                   <- GlobalQuantHst            hydro.cpp:746-771
#include <algorithm>
#include <mutex>

#include "aph_device.hpp"
#include "aph_march.hpp"
#include "aph_runtime.hpp"

// Per-block partial sums of aph_global_sum_noalloc: a fixed device array instead of an
// allocation, one copy per GPU, held until the process ends.  Host code never passes
// `scratch` itself (in host code the name stands in for the symbol, not a device
// address); it passes the address hipGetSymbolAddress returned for the current device.
static __device__ __align__(256) double scratch[aph::kNoAllocMaxBlocks];

namespace aph {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};     // device address of `scratch`, per device
hipError_t g_scratch_err[kMaxDevices] = {};  // a failed lookup, per device

constexpr int kThreads = 256;

__device__ __forceinline__ void decompose(int64_t t, int nx0, int nx1, int &i, int &j,
                                          int &k) {
  i = (int)(t % nx0);
  int64_t r = t / nx0;
  j = (int)(r % nx1);
  k = (int)(r / nx1);
}

// ---------------------------------------------------------------- ConsToPrim
template <int NV>
__global__ void __launch_bounds__(kThreads)
    k_cons_to_prim(const aph_block *__restrict__ blocks, int first, int ndim, aph_eos eos,
                   double *__restrict__ cons, double *__restrict__ prim,
                   double *__restrict__ gamma) {
  const aph_block &bk = blocks[first + blockIdx.y];
  const int64_t nc = ncell(bk, ndim);
  double *u = cons + (int64_t)NV * bk.cell_base;
  double *w = prim + (int64_t)NV * bk.cell_base;
  double *gm = gamma + (int64_t)2 * bk.cell_base;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
       c += (int64_t)gridDim.x * blockDim.x) {
    double uu[NV], ww[NV], gg[2];
#pragma unroll
    for (int v = 0; v < NV; ++v) uu[v] = u[v * nc + c];
    cons_to_prim_pt<NV>(uu, eos, ww, gg);
    u[c] = uu[0];
    u[4 * nc + c] = uu[4];
#pragma unroll
    for (int v = 0; v < NV; ++v) w[v * nc + c] = ww[v];
    gm[c] = gg[0];
    gm[nc + c] = gg[1];
  }
}

// ---------------------------------------------------------------- fluxes
// One lane per pencil of the reference's loop nest (including its one transverse
// ghost row); the lane marches the pencil (aph_march.hpp) and writes cons.flux.
template <int RECON, int RS, int NV, int DIR>
__global__ void __launch_bounds__(64)
    k_flux(const aph_block *__restrict__ blocks, int first, int ndim,
           const double *__restrict__ prim, const double *__restrict__ gamma,
           double *__restrict__ flux) {
  const aph_block &bk = blocks[first + blockIdx.y];
  const int N0 = ext(bk, 0, ndim), N1 = ext(bk, 1, ndim), N2 = ext(bk, 2, ndim);
  const int g[3] = {gst(bk, 0, ndim), gst(bk, 1, ndim), gst(bk, 2, ndim)};
  const int64_t st[3] = {1, N0, (int64_t)N0 * N1};
  const int64_t nc = (int64_t)N0 * N1 * N2;
  constexpr int T1 = (DIR == 0) ? 1 : 0;
  constexpr int T2 = (DIR == 2) ? 1 : 2;
  int lo1 = g[T1], hi1 = g[T1] + bk.nx[T1] - 1;
  int lo2 = g[T2], hi2 = g[T2] + bk.nx[T2] - 1;
  if (DIR == 0) {  // hydro.cpp:544-549
    if (ndim >= 2) { --lo1; ++hi1; }
    if (ndim >= 3) { --lo2; ++hi2; }
  } else if (DIR == 1) {  // :623-627
    --lo1; ++hi1;
    if (ndim >= 3) { --lo2; ++hi2; }
  } else {  // :688
    --lo1; ++hi1; --lo2; ++hi2;
  }
  const int n1 = hi1 - lo1 + 1, n2 = hi2 - lo2 + 1;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n1 * n2) return;
  const int64_t base = (int64_t)(lo1 + p % n1) * st[T1] + (int64_t)(lo2 + p / n1) * st[T2];
  const double *w = prim + (int64_t)NV * bk.cell_base;
  const double *gm = gamma + (int64_t)2 * bk.cell_base;
  double *f = flux + (int64_t)NV * bk.cell_base;
  const int64_t sd = st[DIR];
  auto fetch = [&](int c, double *dst) {
    const int64_t o = base + c * sd;
#pragma unroll
    for (int v = 0; v < NV; ++v) dst[v] = w[v * nc + o];
    dst[NV] = gm[o];
    dst[NV + 1] = gm[nc + o];
  };
  auto convert = [](const double *src, double *dst) {
#pragma unroll
    for (int v = 0; v < NV + 2; ++v) dst[v] = src[v];
  };
  auto emit = [&](int fc, const double *fo, const double *, const double *) {
    const int64_t o = base + fc * sd;
#pragma unroll
    for (int v = 0; v < NV; ++v) f[v * nc + o] = fo[v];
  };
  march<RECON, RS, NV, 1 + DIR, NV + 2>(g[DIR], g[DIR] + bk.nx[DIR], fetch, convert, emit);
}

// ---------------------------------------------------------------- update
template <int NV>
__global__ void __launch_bounds__(kThreads)
    k_update(const aph_block *__restrict__ blocks, int first, int ndim,
             double *__restrict__ u0, const double *__restrict__ u1,
             const double *__restrict__ f1, const double *__restrict__ f2,
             const double *__restrict__ f3, double gam0, double gam1, double beta_dt) {
  const aph_block &bk = blocks[first + blockIdx.y];
  const int N0 = ext(bk, 0, ndim), N1 = ext(bk, 1, ndim), N2 = ext(bk, 2, ndim);
  const int64_t nc = (int64_t)N0 * N1 * N2, sy = N0, sz = (int64_t)N0 * N1;
  const int gx = gst(bk, 0, ndim), gy = gst(bk, 1, ndim), gz = gst(bk, 2, ndim);
  const int64_t nint = (int64_t)bk.nx[0] * bk.nx[1] * bk.nx[2];
  const int64_t off = (int64_t)NV * bk.cell_base;
  const double a0 = bk.area[0], a1 = bk.area[1], a2 = bk.area[2], vol = bk.vol;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nint;
       t += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    decompose(t, bk.nx[0], bk.nx[1], i, j, k);
    const int64_t c = (int64_t)(k + gz) * sz + (int64_t)(j + gy) * sy + (i + gx);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int64_t o = off + v * nc + c;
      double du = (a0 * f1[o + 1] - a0 * f1[o]);
      if (ndim >= 2) du += a1 * f2[o + sy] - a1 * f2[o];
      if (ndim == 3) du += a2 * f3[o + sz] - a2 * f3[o];
      double div = -du / vol;
      u0[o] = gam0 * u0[o] + gam1 * u1[o] + beta_dt * div;
    }
  }
}

// ---------------------------------------------------------------- gravity
template <int NV>
__global__ void __launch_bounds__(kThreads)
    k_gravity(const aph_block *__restrict__ blocks, int first, int ndim,
              double *__restrict__ cons, const double *__restrict__ prim, double g0, double g1,
              double g2, double dt) {
  const aph_block &bk = blocks[first + blockIdx.y];
  const int N0 = ext(bk, 0, ndim), N1 = ext(bk, 1, ndim), N2 = ext(bk, 2, ndim);
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int gx = gst(bk, 0, ndim), gy = gst(bk, 1, ndim), gz = gst(bk, 2, ndim);
  const int64_t nint = (int64_t)bk.nx[0] * bk.nx[1] * bk.nx[2];
  double *u = cons + (int64_t)NV * bk.cell_base;
  const double *w = prim + (int64_t)NV * bk.cell_base;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nint;
       t += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    decompose(t, bk.nx[0], bk.nx[1], i, j, k);
    const int64_t c = ((int64_t)(k + gz) * N1 + (j + gy)) * N0 + (i + gx);
    u[nc + c] += g0 * w[c] * dt;
    double de = g0 * w[nc + c];
    if (ndim >= 2) {
      u[2 * nc + c] += g1 * w[c] * dt;
      de += g1 * w[2 * nc + c];
    }
    if (ndim >= 3) {
      u[3 * nc + c] += g2 * w[c] * dt;
      de += g2 * w[3 * nc + c];
    }
    u[4 * nc + c] += de * w[c] * dt;
  }
}

// ---------------------------------------------------------------- ghost fill
template <int NV>
__global__ void __launch_bounds__(kThreads)
    k_exchange(const aph_block *__restrict__ blocks, int first, int ndim,
               double *__restrict__ cons) {
  const aph_block &bk = blocks[first + blockIdx.y];
  const int N[3] = {ext(bk, 0, ndim), ext(bk, 1, ndim), ext(bk, 2, ndim)};
  const int g[3] = {gst(bk, 0, ndim), gst(bk, 1, ndim), gst(bk, 2, ndim)};
  const int64_t nc = (int64_t)N[0] * N[1] * N[2];
  double *u = cons + (int64_t)NV * bk.cell_base;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
       c += (int64_t)gridDim.x * blockDim.x) {
    int idx[3];
    decompose(c, N[0], N[1], idx[0], idx[1], idx[2]);
    int o[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) o[d] = idx[d] < g[d] ? -1 : (idx[d] >= g[d] + bk.nx[d] ? 1 : 0);
    if (o[0] == 0 && o[1] == 0 && o[2] == 0) continue;
    const int nbi = bk.nbr[(o[2] + 1) * 9 + (o[1] + 1) * 3 + (o[0] + 1)];
    if (nbi < 0) continue;
    const aph_block &nk = blocks[nbi];
    const double *s = cons + (int64_t)NV * nk.cell_base;
    const int64_t src = ((int64_t)(idx[2] - o[2] * bk.nx[2]) * N[1] +
                         (idx[1] - o[1] * bk.nx[1])) * N[0] + (idx[0] - o[0] * bk.nx[0]);
#pragma unroll
    for (int v = 0; v < NV; ++v) u[v * nc + c] = s[v * nc + src];
  }
}

// Physical boundary on both faces of axis D (apophis_boundaries.cpp:77-90 and the
// upstream outflow rule), over the whole ghost slab.
template <int NV, int D>
__global__ void __launch_bounds__(kThreads)
    k_bc(const aph_block *__restrict__ blocks, int first, int ndim, double *__restrict__ cons) {
  const aph_block &bk = blocks[first + blockIdx.y];
  constexpr int sdn = (D == 0) ? 1 : (D == 1 ? 3 : 9);
  const bool inner = bk.nbr[13 - sdn] < 0, outer = bk.nbr[13 + sdn] < 0;
  if (!inner && !outer) return;
  const int N[3] = {ext(bk, 0, ndim), ext(bk, 1, ndim), ext(bk, 2, ndim)};
  const int g = bk.ng;
  const int64_t nc = (int64_t)N[0] * N[1] * N[2];
  double *u = cons + (int64_t)NV * bk.cell_base;
  // slab: 2*g layers along D (first g: inner, last g: outer) x the other two axes
  int S[3] = {N[0], N[1], N[2]};
  S[D] = 2 * g;
  const int64_t ns = (int64_t)S[0] * S[1] * S[2];
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ns;
       t += (int64_t)gridDim.x * blockDim.x) {
    int idx[3];
    decompose(t, S[0], S[1], idx[0], idx[1], idx[2]);
    const bool side = idx[D] >= g;  // false: inner, true: outer
    if (side ? !outer : !inner) continue;
    if (side) idx[D] = g + bk.nx[D] + (idx[D] - g);
    const int type = bk.bc[2 * D + (side ? 1 : 0)];
    if (type == APH_BC_AMR) continue;  // set from the other level (aph_amr_fill)
    const int ref = side ? g + bk.nx[D] - 1 : g;
    const int offset = 2 * ref + (side ? 1 : -1);
    int s[3] = {idx[0], idx[1], idx[2]};
    s[D] = (type == APH_BC_REFLECT) ? offset - idx[D] : ref;
    const int64_t dst = ((int64_t)idx[2] * N[1] + idx[1]) * N[0] + idx[0];
    const int64_t src = ((int64_t)s[2] * N[1] + s[1]) * N[0] + s[0];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      if (type == APH_BC_REFLECT)
        u[v * nc + dst] = ((v == 1 + D) ? -1.0 : 1.0) * u[v * nc + src];
      else
        u[v * nc + dst] = u[v * nc + src];
    }
  }
}

// ---------------------------------------------------------------- dt, sums
__global__ void k_set(double *p, double v) { *p = v; }

template <int NV>
__global__ void __launch_bounds__(kThreads)
    k_estimate_dt(const aph_block *__restrict__ blocks, int first, int ndim, aph_eos eos,
                  const double *__restrict__ prim, const double *__restrict__ cons,
                  double cfl, double *__restrict__ out) {
  const aph_block &bk = blocks[first + blockIdx.y];
  const int N0 = ext(bk, 0, ndim), N1 = ext(bk, 1, ndim), N2 = ext(bk, 2, ndim);
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int gx = gst(bk, 0, ndim), gy = gst(bk, 1, ndim), gz = gst(bk, 2, ndim);
  const int64_t nint = (int64_t)bk.nx[0] * bk.nx[1] * bk.nx[2];
  const double *w = prim + (int64_t)NV * bk.cell_base;
  const double *u = cons + (int64_t)NV * bk.cell_base;
  double m = 1.79769313486231570815e308;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nint;
       t += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    decompose(t, bk.nx[0], bk.nx[1], i, j, k);
    const int64_t c = ((int64_t)(k + gz) * N1 + (j + gy)) * N0 + (i + gx);
    double ww[5] = {w[c], w[nc + c], w[2 * nc + c], w[3 * nc + c], w[4 * nc + c]};
    m = fmin(m, dt_point(ww, u[4 * nc + c], eos, bk.dx, ndim));
  }
  workgroup_min_pos<kThreads>(out, cfl * m);
}

template <int NV>
__global__ void __launch_bounds__(kThreads)
    k_block_sum(const aph_block *__restrict__ blocks, int ndim, const double *__restrict__ cons,
                int var, double *__restrict__ partial) {
  const aph_block &bk = blocks[blockIdx.x];
  const int N0 = ext(bk, 0, ndim), N1 = ext(bk, 1, ndim), N2 = ext(bk, 2, ndim);
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int gx = gst(bk, 0, ndim), gy = gst(bk, 1, ndim), gz = gst(bk, 2, ndim);
  const int64_t nint = (int64_t)bk.nx[0] * bk.nx[1] * bk.nx[2];
  const double *u = cons + (int64_t)NV * bk.cell_base + (int64_t)var * nc;
  double s = 0.0;
  for (int64_t t = threadIdx.x; t < nint; t += blockDim.x) {
    int i, j, k;
    decompose(t, bk.nx[0], bk.nx[1], i, j, k);
    s += u[((int64_t)(k + gz) * N1 + (j + gy)) * N0 + (i + gx)] * bk.vol;
  }
  __shared__ double red[kThreads];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_sum_partials(const double *__restrict__ partial, int n, double *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < n; ++b) s += partial[b];
    *out = s;
  }
}

// The device address of this device's `scratch`, or null when its lookup failed.
double *noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

// ---------------------------------------------------------------- dispatch helpers
inline dim3 grid_for(int64_t work, int threads, int nblk, int64_t cap = 1024) {
  int64_t x = (work + threads - 1) / threads;
  x = std::max<int64_t>(1, std::min<int64_t>(x, cap));
  return dim3((unsigned)x, (unsigned)nblk, 1);
}

#define APH_NV_DISPATCH(nv, KCALL)                                                     \
  switch (nv) {                                                                        \
    case 5: { constexpr int NV = 5; KCALL; } break;                                    \
    case 6: { constexpr int NV = 6; KCALL; } break;                                    \
    case 12: { constexpr int NV = 12; KCALL; } break;                                  \
    case 14: { constexpr int NV = 14; KCALL; } break;                                  \
    default: return set_error(ctx, APH_E_INVALID, "unsupported nvar %d", nv);          \
  }

template <int RECON, int RS, int NV>
int launch_flux(aph_context *ctx, const aph_blockset *bs, int lo, int cnt,
                const double *prim, const double *gamma, double *f1, double *f2, double *f3) {
  const int nd = bs->ndim;
  auto g = [&](int d) {
    return dim3((unsigned)((bs->max_pencils[d] + 63) / 64), (unsigned)cnt, 1);
  };
  k_flux<RECON, RS, NV, 0><<<g(0), 64, 0, ctx->stream>>>(bs->d, lo, nd, prim, gamma, f1);
  if (nd >= 2)
    k_flux<RECON, RS, NV, 1><<<g(1), 64, 0, ctx->stream>>>(bs->d, lo, nd, prim, gamma, f2);
  if (nd >= 3)
    k_flux<RECON, RS, NV, 2><<<g(2), 64, 0, ctx->stream>>>(bs->d, lo, nd, prim, gamma, f3);
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

template <int RECON, int RS>
int launch_flux_nv(aph_context *ctx, const aph_blockset *bs, int lo, int cnt,
                   const double *prim, const double *gamma, double *f1, double *f2,
                   double *f3) {
  APH_NV_DISPATCH(bs->nvar,
                  return (launch_flux<RECON, RS, NV>(ctx, bs, lo, cnt, prim, gamma, f1, f2, f3)));
  return APH_OK;
}

}  // namespace
}  // namespace aph

using namespace aph;

extern "C" {

int aph_cons_to_prim(aph_context *ctx, const aph_blockset *bs, int32_t first_block,
                     int32_t nblocks, const aph_eos *eos, double *cons, double *prim,
                     double *gamma) {
  if (!ctx) return APH_E_INVALID;
  int lo, cnt;
  int rc = block_range(ctx, bs, first_block, nblocks, &lo, &cnt);
  if (rc) return rc;
  if (cnt == 0) return APH_OK;  // an empty range: nothing to read or write
  if (!eos || !cons || !prim || !gamma)
    return set_error(ctx, APH_E_INVALID, "aph_cons_to_prim: null pointer");
  dim3 gr = grid_for(bs->max_ncell, kThreads, cnt, 256);
  APH_NV_DISPATCH(bs->nvar, (k_cons_to_prim<NV><<<gr, kThreads, 0, ctx->stream>>>(
                                bs->d, lo, bs->ndim, *eos, cons, prim, gamma)));
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

int aph_calculate_fluxes(aph_context *ctx, const aph_blockset *bs, int32_t first_block,
                         int32_t nblocks, int32_t recon, int32_t rsolver,
                         const double *prim, const double *gamma, double *flux_x1,
                         double *flux_x2, double *flux_x3) {
  if (!ctx) return APH_E_INVALID;
  int lo, cnt;
  int rc = block_range(ctx, bs, first_block, nblocks, &lo, &cnt);
  if (rc) return rc;
  if (cnt == 0) return APH_OK;  // an empty range: nothing to read or write
  if (!prim || !gamma || !flux_x1 || (bs->ndim >= 2 && !flux_x2) ||
      (bs->ndim >= 3 && !flux_x3))
    return set_error(ctx, APH_E_INVALID, "aph_calculate_fluxes: null pointer");
  const int need = recon == APH_RECON_PPM ? 3 : (recon == APH_RECON_PLM ? 2 : 1);
  for (int b = lo; b < lo + cnt; ++b)
    if (bs->h[b].ng < need)  // hydro.cpp:72-75
      return set_error(ctx, APH_E_INVALID,
                       "Not enough ghost zones for reconstruction (block %d has %d, need %d)",
                       b, bs->h[b].ng, need);
#define APH_FLUX_CASE(R, S)                                                            \
  if (recon == R && rsolver == S)                                                      \
    return launch_flux_nv<R, S>(ctx, bs, lo, cnt, prim, gamma, flux_x1, flux_x2, flux_x3);
  APH_FLUX_CASE(RECON_PLM, RS_HLLE)
  APH_FLUX_CASE(RECON_PLM, RS_HLLC)
  APH_FLUX_CASE(RECON_PPM, RS_HLLE)
  APH_FLUX_CASE(RECON_PPM, RS_HLLC)
  APH_FLUX_CASE(RECON_DC, RS_HLLE)
  APH_FLUX_CASE(RECON_DC, RS_HLLC)
  APH_FLUX_CASE(RECON_PLM, RS_LHLLC)
  APH_FLUX_CASE(RECON_PPM, RS_LHLLC)
  APH_FLUX_CASE(RECON_DC, RS_LHLLC)
  APH_FLUX_CASE(RECON_DC, RS_NONE)  // the only "none" pairing registered (hydro.cpp:112)
#undef APH_FLUX_CASE
  return set_error(ctx, APH_E_INVALID, "unsupported (reconstruction=%d, riemann=%d)", recon,
                   rsolver);
}

int aph_update_with_flux_divergence(aph_context *ctx, const aph_blockset *bs,
                                    int32_t first_block, int32_t nblocks, double *u0,
                                    const double *u1, const double *flux_x1,
                                    const double *flux_x2, const double *flux_x3,
                                    double gam0, double gam1, double beta_dt) {
  if (!ctx) return APH_E_INVALID;
  int lo, cnt;
  int rc = block_range(ctx, bs, first_block, nblocks, &lo, &cnt);
  if (rc) return rc;
  if (cnt == 0) return APH_OK;  // an empty range: nothing to read or write
  if (!u0 || !u1 || !flux_x1 || (bs->ndim >= 2 && !flux_x2) || (bs->ndim >= 3 && !flux_x3))
    return set_error(ctx, APH_E_INVALID, "aph_update_with_flux_divergence: null pointer");
  dim3 gr = grid_for(bs->max_interior, kThreads, cnt, 256);
  APH_NV_DISPATCH(bs->nvar, (k_update<NV><<<gr, kThreads, 0, ctx->stream>>>(
                                bs->d, lo, bs->ndim, u0, u1, flux_x1, flux_x2, flux_x3, gam0,
                                gam1, beta_dt)));
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

int aph_apply_gravity(aph_context *ctx, const aph_blockset *bs, int32_t first_block,
                      int32_t nblocks, double *cons, const double *prim, const double g[3],
                      double dt) {
  if (!ctx) return APH_E_INVALID;
  int lo, cnt;
  int rc = block_range(ctx, bs, first_block, nblocks, &lo, &cnt);
  if (rc) return rc;
  if (cnt == 0) return APH_OK;  // an empty range: nothing to read or write
  if (!cons || !prim || !g) return set_error(ctx, APH_E_INVALID, "aph_apply_gravity: null");
  dim3 gr = grid_for(bs->max_interior, kThreads, cnt, 256);
  APH_NV_DISPATCH(bs->nvar, (k_gravity<NV><<<gr, kThreads, 0, ctx->stream>>>(
                                bs->d, lo, bs->ndim, cons, prim, g[0], g[1], g[2], dt)));
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

int aph_fill_ghosts(aph_context *ctx, const aph_blockset *bs, int32_t first_block,
                    int32_t nblocks, double *cons) {
  if (!ctx) return APH_E_INVALID;
  int lo, cnt;
  int rc = block_range(ctx, bs, first_block, nblocks, &lo, &cnt);
  if (rc) return rc;
  if (cnt == 0) return APH_OK;  // an empty range: nothing to read or write
  if (!cons) return set_error(ctx, APH_E_INVALID, "aph_fill_ghosts: null");
  dim3 gr = grid_for(bs->max_ncell, kThreads, cnt, 256);
  const int nd = bs->ndim;
  int64_t slab = 0;
  for (int d = 0; d < 3; ++d) slab = std::max<int64_t>(slab, (int64_t)bs->max_ext[0] *
                                                             bs->max_ext[1] * bs->max_ext[2] /
                                                             std::max(1, bs->max_ext[d]) * 8);
  dim3 gs = grid_for(slab, kThreads, cnt, 64);
  APH_NV_DISPATCH(bs->nvar, ({
    k_exchange<NV><<<gr, kThreads, 0, ctx->stream>>>(bs->d, lo, nd, cons);
    k_bc<NV, 0><<<gs, kThreads, 0, ctx->stream>>>(bs->d, lo, nd, cons);
    if (nd >= 2) k_bc<NV, 1><<<gs, kThreads, 0, ctx->stream>>>(bs->d, lo, nd, cons);
    if (nd >= 3) k_bc<NV, 2><<<gs, kThreads, 0, ctx->stream>>>(bs->d, lo, nd, cons);
  }));
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

}  // extern "C"

namespace aph {

// The boundary-condition passes of aph_fill_ghosts alone, over every block (a multilevel
// mesh runs them again after its cross-level fill).
int physical_bc(aph_context *ctx, const aph_blockset *bs, double *cons) {
  if (!bs->nblocks) return APH_OK;
  const int cnt = bs->nblocks, nd = bs->ndim;
  int64_t slab = 0;
  for (int d = 0; d < 3; ++d) slab = std::max<int64_t>(slab, (int64_t)bs->max_ext[0] *
                                                             bs->max_ext[1] * bs->max_ext[2] /
                                                             std::max(1, bs->max_ext[d]) * 8);
  dim3 gs = grid_for(slab, kThreads, cnt, 64);
  APH_NV_DISPATCH(bs->nvar, ({
    k_bc<NV, 0><<<gs, kThreads, 0, ctx->stream>>>(bs->d, 0, nd, cons);
    if (nd >= 2) k_bc<NV, 1><<<gs, kThreads, 0, ctx->stream>>>(bs->d, 0, nd, cons);
    if (nd >= 3) k_bc<NV, 2><<<gs, kThreads, 0, ctx->stream>>>(bs->d, 0, nd, cons);
  }));
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

}  // namespace aph

extern "C" {

int aph_estimate_timestep(aph_context *ctx, const aph_blockset *bs, int32_t first_block,
                          int32_t nblocks, const aph_eos *eos, const double *prim,
                          const double *cons, double cfl, double *dt_out) {
  if (!ctx) return APH_E_INVALID;
  int lo, cnt;
  int rc = block_range(ctx, bs, first_block, nblocks, &lo, &cnt);
  if (rc) return rc;
  if (!dt_out) return set_error(ctx, APH_E_INVALID, "aph_estimate_timestep: null pointer");
  if (cnt > 0 && (!eos || !prim || !cons))
    return set_error(ctx, APH_E_INVALID, "aph_estimate_timestep: null pointer");
  k_set<<<1, 1, 0, ctx->stream>>>(dt_out, 1.79769313486231570815e308);
  if (cnt > 0) {
    dim3 gr = grid_for(bs->max_interior, kThreads, cnt, 64);
    APH_NV_DISPATCH(bs->nvar, (k_estimate_dt<NV><<<gr, kThreads, 0, ctx->stream>>>(
                                  bs->d, lo, bs->ndim, *eos, prim, cons, cfl, dt_out)));
  }
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

int aph_global_sum(aph_context *ctx, const aph_blockset *bs, const double *cons,
                   int32_t var, double *out) {
  if (!ctx) return APH_E_INVALID;
  if (!bs || !out || var < 0 || var >= bs->nvar || (!cons && bs->nblocks > 0))
    return set_error(ctx, APH_E_INVALID, "aph_global_sum: bad arguments");
  if (bs->nblocks == 0) {
    k_set<<<1, 1, 0, ctx->stream>>>(out, 0.0);
    APH_LAUNCH_CHECK(ctx);
    return APH_OK;
  }
  if (ctx->d_scalar_n < bs->nblocks) {
    if (ctx->d_scalar) APH_HIP_CHECK(ctx, hipFree(ctx->d_scalar));
    ctx->d_scalar = nullptr;
    ctx->d_scalar_n = 0;
    if (hipMalloc(&ctx->d_scalar, sizeof(double) * bs->nblocks) != hipSuccess)
      return set_error(ctx, APH_E_NOMEM, "aph_global_sum: scratch allocation failed");
    ctx->d_scalar_n = bs->nblocks;
  }
  APH_NV_DISPATCH(bs->nvar, (k_block_sum<NV><<<bs->nblocks, kThreads, 0, ctx->stream>>>(
                                bs->d, bs->ndim, cons, var, ctx->d_scalar)));
  k_sum_partials<<<1, 1, 0, ctx->stream>>>(ctx->d_scalar, bs->nblocks, out);
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

}  // extern "C"

namespace aph {

// Set-up of aph_global_sum_noalloc, called by aph_context_create with the context's
// device current: looks up this device's address of `scratch` once.
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return APH_E_INVALID;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return APH_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    g_scratch_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
    return APH_E_HIP;
  }
  g_scratch_ptr[device] = static_cast<double *>(ptr);
  g_scratch_err[device] = hipSuccess;
  return APH_OK;
}

}  // namespace aph

extern "C" {

int aph_global_sum_noalloc(aph_context *ctx, const aph_blockset *bs, const double *cons,
                           int32_t var, double *out) {
  if (!ctx) return APH_E_INVALID;
  if (!bs || !out || var < 0 || var >= bs->nvar || (!cons && bs->nblocks > 0))
    return set_error(ctx, APH_E_INVALID, "aph_global_sum_noalloc: bad arguments");
  if (bs->nblocks > kNoAllocMaxBlocks)
    return set_error(ctx, APH_E_INVALID,
                     "aph_global_sum_noalloc: %d blocks exceed its fixed scratch of %d "
                     "partial sums (use aph_global_sum)", bs->nblocks, kNoAllocMaxBlocks);
  if (bs->nblocks == 0) {
    k_set<<<1, 1, 0, ctx->stream>>>(out, 0.0);
    APH_LAUNCH_CHECK(ctx);
    return APH_OK;
  }
  double *partial = noalloc_scratch(ctx->device);
  if (!partial)
    return set_error(ctx, APH_E_HIP,
                     "aph_global_sum_noalloc: no scratch address for device %d (%s)",
                     ctx->device,
                     ctx->device >= 0 && ctx->device < kMaxDevices
                         ? hipGetErrorString(g_scratch_err[ctx->device])
                         : "device index out of range");
  // every block writes its own partial before the final sum reads them, so nothing
  // left from an earlier call is read
  APH_NV_DISPATCH(bs->nvar, (k_block_sum<NV><<<bs->nblocks, kThreads, 0, ctx->stream>>>(
                                bs->d, bs->ndim, cons, var, partial)));
  k_sum_partials<<<1, 1, 0, ctx->stream>>>(partial, bs->nblocks, out);
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

}  // extern "C"
