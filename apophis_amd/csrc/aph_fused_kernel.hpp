// aph_fused_kernel.hpp -- the batched stage kernel k_fused and the initial time-step
// kernel, as templates over (reconstruction, Riemann solver, NV, ndim).  aph_fused.hip
// instantiates NV = 5 and 6 (hydro, hydro + one scalar) and holds the host side;
// aph_fused_wide.hip instantiates the composition counts (NV = 12: the ONe composition's
// 6 species + Ye; NV = 14: plus one level set, fuel and distance), so the two compile in
// parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "aph_device.hpp"
#include "aph_march.hpp"
#include "aph_runtime.hpp"

namespace aph {

struct TileCfg {
  int t[3];  // tile extent per axis (cells)
  int s[3];  // segment length per axis (cells one lane marches)
};

constexpr int kMaxClasses = 8;

struct FusedArgs {
  const aph_block *blocks;
  const int4 *items;  // {block | cls << 24, t0x, t0y, t0z}; block < 0: padding
  int nitems;
  int ndim;
  TileCfg cfg[kMaxClasses];
  aph_eos eos;
  double g[3];
  int grav;
  const double *gfield;  // per-cell acceleration (3 components, pack layout) or null: g
  const double *u_in;
  const double *u_n;
  double *u_out;
  const double *dt;
  double gam0, gam1, beta;
  double *dt_min;  // last stage: atomic min of cfl * dt candidate (reset before stage 1)
  double cfl;
  int has_un;      // u_n is a separate register (every stage but the first)
  int last;        // last stage of the step: estimate the next dt
  const aph_amr_face *amr_face;  // multilevel mesh: [block][6] interface roles, or null
  double *amr_buf;               // interface flux buffer (fine side writes, coarse reads)
  // one launch for every level: a block face's interface region (-1: none), per region
  // the count of fine tiles that finished writing it (never reset; compared against
  // epoch x expect), the fine tiles per region and launch, the launch epoch, and the flag
  // a wait that ran out sets
  const int *amr_region;
  unsigned long long *amr_count;
  const unsigned *amr_expect;
  const unsigned *amr_epoch;
  int *amr_err;
  // polls a multilevel wait makes before it gives up and sets amr_err; negative: every
  // wait gives up at once (forces the error path from the host, for its tests)
  int amr_wait_limit;
};

namespace {

// Bases (element offset of component 0) of the block and its two neighbours along
// axis D, or -1 for a physical boundary.
struct AxisNbrs {
  int64_t self, lo, hi;
  int bclo, bchi;
};

// Issues the loads of cell c of the pencil (neighbour block, reflected or clamped
// ghost) and returns whether its normal momentum must change sign (reflecting
// boundary).  The sign change is left to the caller: applied here it would make the
// loads complete before the function returns and defeat the march's prefetch.
template <int NV>
__device__ __forceinline__ bool load_cell(const FusedArgs &a, const aph_block &bk,
                                          const AxisNbrs &nb, int D, int n_d, int c,
                                          int64_t tbase, int64_t sd, int64_t nc, double *u) {
  int64_t base = nb.self;
  int idx = c;
  bool flip = false;
  if (c < 0) {
    if (nb.lo >= 0) {
      base = nb.lo;
      idx = c + n_d;
    } else if (nb.bclo == APH_BC_REFLECT) {
      idx = -1 - c;
      flip = true;
    } else if (nb.bclo == APH_BC_AMR) {
      idx = c;  // the block's own ghost cells, set from the other level before the stage
    } else {
      idx = 0;
    }
  } else if (c >= n_d) {
    if (nb.hi >= 0) {
      base = nb.hi;
      idx = c - n_d;
    } else if (nb.bchi == APH_BC_REFLECT) {
      idx = 2 * n_d - 1 - c;
      flip = true;
    } else if (nb.bchi == APH_BC_AMR) {
      idx = c;
    } else {
      idx = n_d - 1;
    }
  }
  const double *p = a.u_in + base + tbase + (int64_t)idx * sd;
#pragma unroll
  for (int v = 0; v < NV; ++v) u[v] = p[v * nc];
  return flip;
}

// Bounded wait of a coarse tile for the fine tiles that write interface region r in this
// launch (workgroup lane 0).  The items are ordered finest level first and every
// workgroup of the persistent grid is resident, so the writers are running or done; the
// bound only turns a broken plan into a flagged error instead of a hang.
__device__ __forceinline__ void amr_wait(const FusedArgs &a, int r) {
  const int limit = a.amr_wait_limit;
  if (limit < 0) {
    __hip_atomic_fetch_or(a.amr_err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  const unsigned long long target =
      (unsigned long long)(*a.amr_epoch) * (unsigned long long)a.amr_expect[r];
  int n = 0;
  while (__hip_atomic_load(a.amr_count + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
         target) {
    __builtin_amdgcn_s_sleep(4);
    ++n;
    // once any wait has run out the launch is lost: stop waiting at once
    if ((n & 255) == 0 &&
        __hip_atomic_load(a.amr_err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      break;
    if (n > limit) {
      __hip_atomic_fetch_or(a.amr_err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
  }
}

// Faces of the tile (origin t0) on the block's low / high boundary of axis D.
__device__ __forceinline__ bool tile_on_face(const TileCfg &cfg, const int *t0, int n_d,
                                             int D, int side) {
  return side ? (t0[D] + cfg.t[D] == n_d) : (t0[D] == 0);
}

template <int RECON, int RS, int NV, int NDIM, int D, bool AMR>
__device__ __forceinline__ void sweep(const FusedArgs &a, const aph_block &bk, int bidx,
                                      const TileCfg &cfg, const int *t0, double *du_s,
                                      double &dtmin) {
  constexpr bool LAST = (D == NDIM - 1);
  constexpr int A = (D == 0) ? 1 : 0;  // first transverse axis (fastest across lanes)
  constexpr int B = (D == 2) ? 1 : 2;  // second transverse axis
  const int ndim = NDIM;
  const int N0 = ext(bk, 0, ndim), N1 = ext(bk, 1, ndim), N2 = ext(bk, 2, ndim);
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int64_t st[3] = {1, N0, (int64_t)N0 * N1};
  const int g[3] = {gst(bk, 0, ndim), gst(bk, 1, ndim), gst(bk, 2, ndim)};
  constexpr int sdn = (D == 0) ? 1 : (D == 1 ? 3 : 9);
  AxisNbrs nb;
  nb.self = (int64_t)NV * bk.cell_base;
  {
    const int l = bk.nbr[13 - sdn], h = bk.nbr[13 + sdn];
    nb.lo = l >= 0 ? (int64_t)NV * a.blocks[l].cell_base : -1;
    nb.hi = h >= 0 ? (int64_t)NV * a.blocks[h].cell_base : -1;
    nb.bclo = bk.bc[2 * D];
    nb.bchi = bk.bc[2 * D + 1];
  }
  const int n_d = bk.nx[D];
  const int nseg = cfg.t[D] / cfg.s[D];
  const int units = cfg.t[A] * cfg.t[B] * nseg;
  const double area = bk.area[D];
  const double vol = bk.vol;
  const double beta_dt = a.beta * a.dt[0];
  if constexpr (AMR) {
    // a face to a finer region: wait until every fine tile has written its fluxes (the
    // fences write back / invalidate the XCD's L2, so only interface tiles take them)
    bool waits = false;
    for (int side = 0; side < 2; ++side) {
      const int fi = bidx * 6 + 2 * D + side;
      if (tile_on_face(cfg, t0, n_d, D, side) && a.amr_face[fi].role == 2) {
        waits = true;
        if (threadIdx.x == 0) amr_wait(a, a.amr_region[fi]);
      }
    }
    if (waits) {
      __syncthreads();
      __threadfence();  // acquire: the buffer is read after the counts were seen
    }
  }
  const int tx = cfg.t[0], ty = cfg.t[1];
  for (int unit = threadIdx.x; unit < units; unit += blockDim.x) {
    int seg, pa, pb;
    if (D == 0) {  // segments of one row on adjacent lanes
      seg = unit % nseg;
      const int p = unit / nseg;
      pa = p % cfg.t[A];
      pb = p / cfg.t[A];
    } else {  // adjacent lanes on adjacent x (coalesced)
      pa = unit % cfg.t[A];
      const int p = unit / cfg.t[A];
      seg = p % nseg;
      pb = p / nseg;
    }
    const int ca = t0[A] + pa, cb = t0[B] + pb;  // interior coordinates
    const int64_t tbase = (int64_t)(ca + g[A]) * st[A] + (int64_t)(cb + g[B]) * st[B] +
                          (int64_t)g[D] * st[D];
    const int f_lo = t0[D] + seg * cfg.s[D];
    const int f_hi = f_lo + cfg.s[D];
    double fprev[NV];
    bool flip = false;  // sign change of the fetched cell, applied by convert
    auto fetch = [&](int c, double *u) {
      flip = load_cell<NV>(a, bk, nb, D, n_d, c, tbase, st[D], nc, u);
    };
    auto convert = [&](const double *raw, double *dst) {
      double u[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) u[v] = raw[v];
      if (flip) u[1 + D] = -1.0 * u[1 + D];
      cons_to_prim_pt<NV>(u, a.eos, dst, dst + NV);
    };
    // last sweep of PLM / DC: the next cell's u_in, u_n and acceleration field, loaded one
    // step ahead (their latency overlaps the next face's reconstruction and Riemann
    // solve).  It takes 26 more registers: free for PLM, whose LDS use already caps it at
    // two waves per SIMD, but it would leave PPM (217 registers) one wave per SIMD.
    constexpr bool PREF = LAST && RECON != RECON_PPM;
    double pu[NV], pun[NV], pg[3];
    auto prefetch = [&](int m) {
      const int64_t cidx = tbase + (int64_t)m * st[D];
      const double *pin = a.u_in + nb.self + cidx;
#pragma unroll
      for (int v = 0; v < NV; ++v) pu[v] = pin[v * nc];
      if (a.has_un) {
        const double *pn = a.u_n + nb.self + cidx;
#pragma unroll
        for (int v = 0; v < NV; ++v) pun[v] = pn[v * nc];
      }
      if (a.gfield) {  // a per-cell field (monopole gravity)
        const double *gp = a.gfield + 3 * bk.cell_base + cidx;
        pg[0] = gp[0];
        pg[1] = gp[nc];
        pg[2] = gp[2 * nc];
      }
    };
    auto emit = [&](int f, const double *fo_in, const double *, const double *) {
      const double *fo = fo_in;
      double cu[NV], cun[NV], cg[3];
      if constexpr (PREF) {
        if (f > f_lo) {
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            cu[v] = pu[v];
            cun[v] = pun[v];
          }
          cg[0] = pg[0];
          cg[1] = pg[1];
          cg[2] = pg[2];
        }
        if (f < f_hi) prefetch(f);
      }
      double fc[NV];
      if (AMR && (f == 0 || f == n_d)) {
        // a block face to the other level: the fine side records its flux, the coarse
        // side takes the mean of the fine fluxes across (flux correction)
        const aph_amr_face &af = a.amr_face[bidx * 6 + 2 * D + (f == 0 ? 0 : 1)];
        constexpr int NCH = (NDIM == 3) ? 4 : 2;
        if (af.role == 1) {
          const int fa = af.qa + ca, fb = af.qb + cb;
          const int child = (NDIM == 3) ? (fa & 1) * 2 + (fb & 1) : (fa & 1);
          double *q = a.amr_buf + af.base +
                      ((int64_t)((fb >> 1) * af.na + (fa >> 1)) * NCH + child) * NV;
#pragma unroll
          for (int v = 0; v < NV; ++v) q[v] = fo_in[v];
        } else if (af.role == 2) {
          const double *q = a.amr_buf + af.base + ((int64_t)cb * af.na + ca) * NCH * NV;
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            double sum = q[v];
#pragma unroll
            for (int k = 1; k < NCH; ++k) sum += q[k * NV + v];
            fc[v] = sum * (1.0 / NCH);
          }
          fo = fc;
        }
      }
      if (f > f_lo) {
        const int m = f - 1;  // cell whose upper face is f
        int lc[3];
        lc[D] = m - t0[D];
        lc[A] = pa;
        lc[B] = pb;
        // the x position is permuted within its row by the row index (tx a power of two):
        // the x sweep's lanes write one cell of each of 64 rows, and rows 8 doubles apart
        // would otherwise fall on 4 of the 32 LDS banks
        const int row = lc[2] * ty + lc[1];
        const int l = row * tx + ((tx & (tx - 1)) ? lc[0] : (lc[0] ^ ((row >> 1) & (tx - 1))));
        const int tcells = tx * ty * cfg.t[2];
        if constexpr (!LAST) {
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            const double d = area * fo[v] - area * fprev[v];
            if (D == 0)
              du_s[v * tcells + l] = d;
            else
              du_s[v * tcells + l] = du_s[v * tcells + l] + d;
          }
        } else {
          const int64_t cidx = tbase + (int64_t)m * st[D];
          double u[NV], w[NV], gm[2];
          double g0 = a.g[0], g1 = a.g[1], g2 = a.g[2];
          if constexpr (PREF) {
#pragma unroll
            for (int v = 0; v < NV; ++v) u[v] = cu[v];
            if (a.gfield) {
              g0 = cg[0];
              g1 = cg[1];
              g2 = cg[2];
            }
          } else {  // PPM: each load where it is used (registers)
            const double *pin = a.u_in + nb.self + cidx;
#pragma unroll
            for (int v = 0; v < NV; ++v) u[v] = pin[v * nc];
          }
          if (a.grav)  // the gravity source needs the stage-start primitives
            cons_to_prim_pt<NV>(u, a.eos, w, gm);
          else
            floor_cons_pt<NV>(u, a.eos);
          double un[NV];
          if (a.has_un) {
            if constexpr (PREF) {
#pragma unroll
              for (int v = 0; v < NV; ++v) un[v] = cun[v];
            } else {
              const double *pn = a.u_n + nb.self + cidx;
#pragma unroll
              for (int v = 0; v < NV; ++v) un[v] = pn[v * nc];
            }
            floor_cons_pt<NV>(un, a.eos);
          } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) un[v] = u[v];
          }
          double out[NV];
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            double du;
            if (NDIM == 1)
              du = (area * fo[v] - area * fprev[v]);
            else
              du = du_s[v * tcells + l] + (area * fo[v] - area * fprev[v]);
            const double div = -du / vol;
            out[v] = a.gam0 * u[v] + a.gam1 * un[v] + beta_dt * div;
          }
          if (a.grav) {  // ApplyGravity with the stage-start primitives
            const double gdt = beta_dt;
            if (!PREF && a.gfield) {  // a per-cell field (monopole gravity)
              const double *gp = a.gfield + 3 * bk.cell_base + cidx;
              g0 = gp[0];
              g1 = gp[nc];
              g2 = gp[2 * nc];
            }
            out[1] += g0 * w[0] * gdt;
            double de = g0 * w[1];
            if (NDIM >= 2) {
              out[2] += g1 * w[0] * gdt;
              de += g1 * w[2];
            }
            if (NDIM >= 3) {
              out[3] += g2 * w[0] * gdt;
              de += g2 * w[3];
            }
            out[4] += de * w[0] * gdt;
          }
          double *po = a.u_out + nb.self + cidx;
#pragma unroll
          for (int v = 0; v < NV; ++v) po[v * nc] = out[v];
          if (a.last) {  // EstimateTimestep of the new state (after its floors)
            double wq[NV], gq[2];
            cons_to_prim_pt<NV>(out, a.eos, wq, gq);
            dtmin = fmin(dtmin, dt_point(wq, out[4], a.eos, bk.dx, NDIM));
          }
        }
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) fprev[v] = fo[v];
    };
    march<RECON, RS, NV, 1 + D, NV>(f_lo, f_hi, fetch, convert, emit);
  }
  if constexpr (AMR) {
    // a face to a coarser leaf: publish this tile's interface fluxes
    bool pub[2];
    for (int side = 0; side < 2; ++side)
      pub[side] = tile_on_face(cfg, t0, n_d, D, side) &&
                  a.amr_face[bidx * 6 + 2 * D + side].role == 1;
    if (pub[0] || pub[1]) {
      __threadfence();  // release: the buffer writes of every lane before the count
      __syncthreads();
      if (threadIdx.x == 0)
        for (int side = 0; side < 2; ++side)
          if (pub[side])
            __hip_atomic_fetch_add(a.amr_count + a.amr_region[bidx * 6 + 2 * D + side], 1ull,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// AMR: the multilevel-mesh variant (interface fluxes recorded / corrected at block faces
// to the other level); the single-level kernel carries none of that code.
// One kernel for workgroups of 64, 128 or 256 lanes (the autotuner's choice): the register
// budget does not depend on it (217 VGPRs for PPM/HLLC in 3D, two waves per SIMD).
template <int RECON, int RS, int NV, int NDIM, bool AMR = false>
__global__ void __launch_bounds__(256) k_fused(FusedArgs a) {
  extern __shared__ double du_s[];
  double dtmin = 1.79769313486231570815e308;
  for (int it = blockIdx.x; it < a.nitems; it += gridDim.x) {
    const int4 item = a.items[it];
    if (item.x < 0) continue;
    const int b = item.x & 0xFFFFFF;
    const int cls = item.x >> 24;
    const aph_block &bk = a.blocks[b];
    const TileCfg cfg = a.cfg[cls];
    const int t0[3] = {item.y, item.z, item.w};
    sweep<RECON, RS, NV, NDIM, 0, AMR>(a, bk, b, cfg, t0, du_s, dtmin);
    if constexpr (NDIM >= 2) {
      __syncthreads();
      sweep<RECON, RS, NV, NDIM, 1, AMR>(a, bk, b, cfg, t0, du_s, dtmin);
    }
    if constexpr (NDIM >= 3) {
      __syncthreads();
      sweep<RECON, RS, NV, NDIM, 2, AMR>(a, bk, b, cfg, t0, du_s, dtmin);
    }
    __syncthreads();
  }
  if (a.last && a.dt_min) {  // one fold per wave
    for (int o = 32; o > 0; o >>= 1) dtmin = fmin(dtmin, __shfl_xor(dtmin, o));
    if ((threadIdx.x & 63) == 0) fold_min_pos(a.dt_min, a.cfl * dtmin);
  }
}

// Initial time step of a raw conserved state (floors applied on the fly).
template <int NV, int NDIM>
__global__ void __launch_bounds__(256)
    k_plan_dt(const aph_block *__restrict__ blocks, aph_eos eos, const double *__restrict__ cons,
              double cfl, double *__restrict__ out) {
  const aph_block &bk = blocks[blockIdx.y];
  const int N0 = ext(bk, 0, NDIM), N1 = ext(bk, 1, NDIM), N2 = ext(bk, 2, NDIM);
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int gx = gst(bk, 0, NDIM), gy = gst(bk, 1, NDIM), gz = gst(bk, 2, NDIM);
  const int64_t nint = (int64_t)bk.nx[0] * bk.nx[1] * bk.nx[2];
  const double *u = cons + (int64_t)NV * bk.cell_base;
  double m = 1.79769313486231570815e308;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nint;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % bk.nx[0]);
    const int j = (int)((t / bk.nx[0]) % bk.nx[1]);
    const int k = (int)(t / ((int64_t)bk.nx[0] * bk.nx[1]));
    const int64_t c = ((int64_t)(k + gz) * N1 + (j + gy)) * N0 + (i + gx);
    double uu[NV], w[NV], gm[2];
#pragma unroll
    for (int v = 0; v < NV; ++v) uu[v] = u[v * nc + c];
    cons_to_prim_pt<NV>(uu, eos, w, gm);
    m = fmin(m, dt_point(w, uu[4], eos, bk.dx, NDIM));
  }
  workgroup_min_pos<256>(out, cfl * m);
}

}  // namespace

using KernelFn = void (*)(FusedArgs);
using PlanDtFn = void (*)(const aph_block *, aph_eos, const double *, double, double *);

// The stage kernel / initial time-step kernel for NV = 12 or 14 (aph_fused_wide.hip);
// null for an unsupported combination.
KernelFn pick_kernel_wide(int recon, int rs, int nv, int ndim);
// The multilevel-mesh stage kernel, NV = 5 or 6, ndim 2 or 3 (aph_fused_amr.hip).
KernelFn pick_kernel_amr(int recon, int rs, int nv, int ndim);
PlanDtFn pick_plan_dt_wide(int nv, int ndim);

}  // namespace aph
