// aph_adjoint.hip -- adjoint (reverse-mode) kernels of one fused RK2 stage, with frozen dt.
//
// Forward stage (aph_fused.hip):  U_out = gam0 U + gam1 U_n + beta dt (-(1/V) sum_d A_d
// [F(c+1/2) - F(c-1/2)]) + gravity(prim(U)), with F = Riemann(recon_L, recon_R) of prim(U)
// on the stencil (neighbour blocks / boundary rule beyond the block).  Given the cotangent
// L_out of U_out this computes
//   L_in = (gam0 [+ gam1 in stage 1]) L_out + (stencil part) [+ L_acc],
//   L_n  = gam1 L_out                                   (stage 2).
// Structure (deterministic, no atomics):
//   k_adj_axis<D>  one lane per (block, pencil, segment) per axis re-marches the forward
//                  (ConsToPrim, reconstruction, Riemann) over its segment, back-propagates
//                  the face cotangents mu_f = beta dt A/V (L_out(f) - L_out(f-1)) through
//                  the Riemann solver and the reconstruction of both adjacent cells, and
//                  writes the primitive-space cotangent of its own cells to Pbar (x writes,
//                  y and z add) and the cotangent of the <= 3 cells it reads beyond its
//                  segment on each side -- already mapped to conserved variables at that
//                  cell's (possibly reflected) state -- to a 6-slot spill record;
//   k_adj_gather   per interior cell: gam terms, gravity adjoint, J_prim^T (Pbar), plus the
//                  spills of the neighbouring segments / blocks / boundary images that read
//                  this cell, in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "aph_adjoint.hpp"
#include "aph_adjoint_host.hpp"
#include "aph_device.hpp"
#include "aph_runtime.hpp"

namespace aph {

// Segment length of the adjoint sweeps: the largest divisor of n not above cap; at least
// 3 (a spill reaches 3 cells), else the whole pencil.  cap is 8 unless the block set is
// too small to give every SIMD two waves with it (adj_prepare).
__host__ __device__ inline int adj_seg(int n, int cap) {
  for (int d = (n < cap ? n : cap); d >= 3; --d)
    if (n % d == 0) return d;
  return n;
}

struct AdjArgs {
  const aph_block *blocks;
  int nblocks;
  aph_eos eos;
  double g[3];
  int grav;
  const double *u_in;
  const double *lam_out;
  const double *lam_acc;
  double *lam_in;
  double *lam_n;
  const double *dt;
  double gam0, gam1, beta;
  int stage;
  double *pbar;                // NF x ncells, pack layout
  double *spill[3];            // per axis
  const int64_t *spill_base;   // [3][nblocks] element offsets
  int segcap[3];               // adj_seg cap per axis
};

namespace {

struct AxisMap {
  int64_t self, lo, hi;  // element base (component 0) of block / neighbours, -1 = none
  int bclo, bchi;
};

template <int NV>
__device__ __forceinline__ bool map_load(const AdjArgs &a, const AxisMap &m, int D, int n_d,
                                         int c, int64_t tbase, int64_t sd, int64_t nc,
                                         double *u) {
  int64_t base = m.self;
  int idx = c;
  bool flip = false;
  if (c < 0) {
    if (m.lo >= 0) {
      base = m.lo;
      idx = c + n_d;
    } else if (m.bclo == APH_BC_REFLECT) {
      idx = -1 - c;
      flip = true;
    } else {
      idx = 0;
    }
  } else if (c >= n_d) {
    if (m.hi >= 0) {
      base = m.hi;
      idx = c - n_d;
    } else if (m.bchi == APH_BC_REFLECT) {
      idx = 2 * n_d - 1 - c;
      flip = true;
    } else {
      idx = n_d - 1;
    }
  }
  const double *p = a.u_in + base + tbase + (int64_t)idx * sd;
#pragma unroll
  for (int v = 0; v < NV; ++v) u[v] = p[v * nc];
  if (flip) u[1 + D] = -1.0 * u[1 + D];
  return flip;
}

template <int RECON, int RS, int NV, int NDIM, int D>
__global__ void __launch_bounds__(64) k_adj_axis(AdjArgs a) {
  constexpr int GR = ReconTraits<RECON>::GR;
  constexpr int W = ReconTraits<RECON>::W;
  constexpr int NF = NV + 2;
  constexpr int IVX = 1 + D, IVY = 1 + IVX % 3, IVZ = 1 + (IVX + 1) % 3;
  constexpr int A = (D == 0) ? 1 : 0, B = (D == 2) ? 1 : 2;
  constexpr int sdn = (D == 0) ? 1 : (D == 1 ? 3 : 9);
  const int b = blockIdx.y;
  const aph_block &bk = a.blocks[b];
  const int n_d = bk.nx[D];
  const int S = adj_seg(n_d, a.segcap[D]);
  const int nseg = n_d / S;
  const int nA = bk.nx[A], nB = bk.nx[B];
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nA * nB * nseg) return;
  int seg, pa, pb;
  if (D == 0) {
    seg = u % nseg;
    const int p = u / nseg;
    pa = p % nA;
    pb = p / nA;
  } else {
    pa = u % nA;
    const int r = u / nA;
    seg = r % nseg;
    pb = r / nseg;
  }
  const int pencil = pa + pb * nA;
  const int N0 = ext(bk, 0, NDIM), N1 = ext(bk, 1, NDIM), N2 = ext(bk, 2, NDIM);
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int64_t st[3] = {1, N0, (int64_t)N0 * N1};
  const int g[3] = {gst(bk, 0, NDIM), gst(bk, 1, NDIM), gst(bk, 2, NDIM)};
  AxisMap m;
  m.self = (int64_t)NV * bk.cell_base;
  const int nl = bk.nbr[13 - sdn], nh = bk.nbr[13 + sdn];
  m.lo = nl >= 0 ? (int64_t)NV * a.blocks[nl].cell_base : -1;
  m.hi = nh >= 0 ? (int64_t)NV * a.blocks[nh].cell_base : -1;
  m.bclo = bk.bc[2 * D];
  m.bchi = bk.bc[2 * D + 1];
  const int64_t tbase = (int64_t)(pa + g[A]) * st[A] + (int64_t)(pb + g[B]) * st[B] +
                        (int64_t)g[D] * st[D];
  const int a0 = seg * S;
  const int fend = (seg == nseg - 1) ? n_d + 1 : a0 + S;  // faces [a0, fend)
  const double coef = a.beta * a.dt[0] * bk.area[D] / bk.vol;
  const double *lo_ptr = a.lam_out + m.self + tbase;
  double *pb_ptr = a.pbar + (int64_t)NF * bk.cell_base + tbase;
  double *sp = a.spill[D] + a.spill_base[D * a.nblocks + b] +
               ((int64_t)pencil * nseg + seg) * 6 * NV;
  // every halo cell is emitted exactly once: clear the record, then write each slot once
  for (int s = 0; s < 6 * NV; ++s) sp[s] = 0.0;

  auto load = [&](int c, double *dst) {
    double uu[NV];
    map_load<NV>(a, m, D, n_d, c, tbase, st[D], nc, uu);
    cons_to_prim_pt<NV>(uu, a.eos, dst, dst + NV);
  };
  // final primitive-space cotangent of cell c from this segment
  auto emit = [&](int c, const double *pbv) {
    if (c >= a0 && c < a0 + S) {
      double *q = pb_ptr + (int64_t)c * st[D];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        if (D == 0)
          q[f * nc] = pbv[f];
        else
          q[f * nc] += pbv[f];
      }
    } else {
      const int slot = (c < a0) ? c - (a0 - 3) : 3 + (c - (a0 + S));
      double uu[NV];
      const bool flip = map_load<NV>(a, m, D, n_d, c, tbase, st[D], nc, uu);
      double ub[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) ub[v] = 0.0;
      cons_to_prim_adj<NV>(uu, a.eos, pbv, pbv + NV, ub);
      if (flip) ub[1 + D] = -1.0 * ub[1 + D];
#pragma unroll
      for (int v = 0; v < NV; ++v) sp[slot * NV + v] = ub[v];
    }
  };

  // windows: P over cells [m-1-GR, m+GR] (W+1 cells), cotangents over [m-1-GR, m-1+GR]
  double win[W + 1][NF];
  double cw[W][NF];
#pragma unroll
  for (int s = 0; s < W; ++s)
#pragma unroll
    for (int f = 0; f < NF; ++f) cw[s][f] = 0.0;
  const int mstart = a0 - 1;
  // preload cells [mstart-GR, mstart+GR-1] into win[2..W]; win[1] (cell mstart-1-GR) is
  // never read, so the sweep reads exactly the forward's cells [a0-1-GR, fend-1+GR]
#pragma unroll
  for (int f = 0; f < NF; ++f) win[1][f] = 0.0;
#pragma unroll
  for (int s = 1; s < W; ++s) load(mstart - 1 - GR + s, win[s + 1]);
  double qlp[NF], qrbp[NF], lprev[NV];
#pragma unroll
  for (int f = 0; f < NF; ++f) qrbp[f] = 0.0;
#pragma unroll
  for (int v = 0; v < NV; ++v)
    lprev[v] = (a0 >= 1) ? lo_ptr[v * nc + (int64_t)(a0 - 1) * st[D]] : 0.0;
  int emitted = mstart - 1 - GR;  // next cell whose cotangent will be emitted
  for (int mm = mstart; mm <= fend - 1; ++mm) {
#pragma unroll
    for (int s = 0; s < W; ++s)
#pragma unroll
      for (int f = 0; f < NF; ++f) win[s][f] = win[s + 1][f];
    load(mm + GR, win[W]);  // win[k] = cell mm-1-GR+k
    // forward reconstruction of cell mm (stencil win[1..W])
    double ql[NF], qr[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      double col[W];
#pragma unroll
      for (int s = 0; s < W; ++s) col[s] = win[s + 1][f];
      recon1<RECON>(col, ql[f], qr[f]);
    }
    double qlb[NF], qrb[NF];  // cotangents of ql output of cell mm-1, qr output of cell mm
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      qlb[f] = 0.0;
      qrb[f] = 0.0;
    }
    if (mm >= a0) {  // face mm
      const int f = mm;
      double lcur[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v)
        lcur[v] = (f < n_d) ? lo_ptr[v * nc + (int64_t)f * st[D]] : 0.0;
      double mu[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) mu[v] = coef * (lcur[v] - lprev[v]);
      // forward Riemann flux (needed for the scalar upwind)
      const double wl[5] = {qlp[0], qlp[IVX], qlp[IVY], qlp[IVZ], qlp[4]};
      const double wr[5] = {qr[0], qr[IVX], qr[IVY], qr[IVZ], qr[4]};
      double flx[5];
      riemann<RS>(wl, wr, &qlp[NV], &qr[NV], flx);
      double fb[5] = {mu[0], mu[IVX], mu[IVY], mu[IVZ], mu[4]};
#pragma unroll
      for (int n = 5; n < NV; ++n) {  // F_s = F_0 * (F_0 >= 0 ? ql_s : qr_s)
        if (flx[0] >= 0.0) {
          fb[0] += mu[n] * qlp[n];
          qlb[n] += mu[n] * flx[0];
        } else {
          fb[0] += mu[n] * qr[n];
          qrb[n] += mu[n] * flx[0];
        }
      }
      double wlb[5] = {0, 0, 0, 0, 0}, wrb[5] = {0, 0, 0, 0, 0};
      double iflb[2] = {0, 0}, ifrb[2] = {0, 0};
      riemann_adj<RS>(wl, wr, &qlp[NV], &qr[NV], fb, wlb, wrb, iflb, ifrb);
      qlb[0] += wlb[0];
      qlb[IVX] += wlb[1];
      qlb[IVY] += wlb[2];
      qlb[IVZ] += wlb[3];
      qlb[4] += wlb[4];
      qlb[NV] += iflb[0];
      qlb[NV + 1] += iflb[1];
      qrb[0] += wrb[0];
      qrb[IVX] += wrb[1];
      qrb[IVY] += wrb[2];
      qrb[IVZ] += wrb[3];
      qrb[4] += wrb[4];
      qrb[NV] += ifrb[0];
      qrb[NV + 1] += ifrb[1];
#pragma unroll
      for (int v = 0; v < NV; ++v) lprev[v] = lcur[v];
      // reconstruction adjoint of cell mm-1: ql output from face mm, qr output from face mm-1
#pragma unroll
      for (int fld = 0; fld < NF; ++fld) {
        double col[W], colb[W];
#pragma unroll
        for (int s = 0; s < W; ++s) {
          col[s] = win[s][fld];
          colb[s] = 0.0;
        }
        recon1_adj<RECON>(col, qlb[fld], qrbp[fld], colb);
#pragma unroll
        for (int s = 0; s < W; ++s) cw[s][fld] += colb[s];
      }
      // cell mm-1-GR is complete: emit it and shift the cotangent window
      emit(emitted, cw[0]);
      ++emitted;
#pragma unroll
      for (int s = 0; s < W - 1; ++s)
#pragma unroll
        for (int fld = 0; fld < NF; ++fld) cw[s][fld] = cw[s + 1][fld];
#pragma unroll
      for (int fld = 0; fld < NF; ++fld) cw[W - 1][fld] = 0.0;
    } else {
      // mm == a0-1: no face yet; cells up to mm-1+GR stay pending.  The window of
      // cotangents starts at cell mm-GR from the next step on.
      emitted = mm - GR;
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      qlp[f] = ql[f];
      qrbp[f] = qrb[f];
    }
  }
  // last cell fend-1: its qr output feeds face fend-1 (owned), its ql output is not owned
  {
#pragma unroll
    for (int fld = 0; fld < NF; ++fld) {
      double col[W], colb[W];
#pragma unroll
      for (int s = 0; s < W; ++s) {
        col[s] = win[s + 1][fld];
        colb[s] = 0.0;
      }
      recon1_adj<RECON>(col, 0.0, qrbp[fld], colb);
      // win[s+1] = cell fend-1-GR+s; cotangent window cw[k] = cell `emitted`+k
#pragma unroll
      for (int s = 0; s < W; ++s) cw[s][fld] += colb[s];
    }
#pragma unroll
    for (int s = 0; s < W; ++s) emit(emitted + s, cw[s]);
  }
}

template <int NV>
__device__ __forceinline__ void add_slot(const AdjArgs &a, int d, int blk, int pencil,
                                         int seg, int nseg, int slot, double *acc) {
  const double *p = a.spill[d] + a.spill_base[d * a.nblocks + blk] +
                    (((int64_t)pencil * nseg + seg) * 6 + slot) * NV;
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] += p[v];
}

template <int NV, int NDIM>
__global__ void __launch_bounds__(256) k_adj_gather(AdjArgs a) {
  constexpr int NF = NV + 2;
  const int b = blockIdx.y;
  const aph_block &bk = a.blocks[b];
  const int N0 = ext(bk, 0, NDIM), N1 = ext(bk, 1, NDIM), N2 = ext(bk, 2, NDIM);
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int g[3] = {gst(bk, 0, NDIM), gst(bk, 1, NDIM), gst(bk, 2, NDIM)};
  const int64_t nint = (int64_t)bk.nx[0] * bk.nx[1] * bk.nx[2];
  const double beta_dt = a.beta * a.dt[0];
  const double gsum = (a.stage == 1) ? a.gam0 + a.gam1 : a.gam0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nint;
       t += (int64_t)gridDim.x * blockDim.x) {
    int cc[3];
    cc[0] = (int)(t % bk.nx[0]);
    cc[1] = (int)((t / bk.nx[0]) % bk.nx[1]);
    cc[2] = (int)(t / ((int64_t)bk.nx[0] * bk.nx[1]));
    const int64_t cell = ((int64_t)(cc[2] + g[2]) * N1 + (cc[1] + g[1])) * N0 + (cc[0] + g[0]);
    const int64_t o = (int64_t)NV * bk.cell_base + cell;
    double lo[NV], u[NV], acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      lo[v] = a.lam_out[o + v * nc];
      u[v] = a.u_in[o + v * nc];
      acc[v] = gsum * lo[v] + (a.lam_acc ? a.lam_acc[o + v * nc] : 0.0);
    }
    double pbv[NF];
    const double *pp = a.pbar + (int64_t)NF * bk.cell_base + cell;
#pragma unroll
    for (int f = 0; f < NF; ++f) pbv[f] = pp[f * nc];
    if (a.grav) {  // out[1+d] += g_d w0 bdt ; out[4] += (sum g_d w_{1+d}) w0 bdt
      double uu[NV], w[NV], gm[2];
#pragma unroll
      for (int v = 0; v < NV; ++v) uu[v] = u[v];
      cons_to_prim_pt<NV>(uu, a.eos, w, gm);
      double de = 0.0;
      for (int d = 0; d < NDIM; ++d) {
        de += a.g[d] * w[1 + d];
        pbv[0] += a.g[d] * beta_dt * lo[1 + d];
        pbv[1 + d] += a.g[d] * w[0] * beta_dt * lo[4];
      }
      pbv[0] += de * beta_dt * lo[4];
    }
    cons_to_prim_adj<NV>(u, a.eos, pbv, pbv + NV, acc);
    // spills of the segments / blocks / boundary images that read this cell
    for (int d = 0; d < NDIM; ++d) {
      const int Ad = (d == 0) ? 1 : 0, Bd = (d == 2) ? 1 : 2;
      const int n_d = bk.nx[d];
      const int S = adj_seg(n_d, a.segcap[d]), nseg = n_d / S;
      const int pencil = cc[Ad] + cc[Bd] * bk.nx[Ad];
      const int s = cc[d] / S, off = cc[d] % S;
      const int sdn = (d == 0) ? 1 : (d == 1 ? 3 : 9);
      if (off < 3) {  // from below
        if (s > 0) {
          add_slot<NV>(a, d, b, pencil, s - 1, nseg, 3 + off, acc);
        } else {
          const int nb = bk.nbr[13 - sdn];
          if (nb >= 0) {
            add_slot<NV>(a, d, nb, pencil, nseg - 1, nseg, 3 + off, acc);
          } else if (bk.bc[2 * d] == APH_BC_REFLECT) {
            add_slot<NV>(a, d, b, pencil, 0, nseg, 2 - off, acc);
          } else if (cc[d] == 0) {
            for (int sl = 0; sl < 3; ++sl) add_slot<NV>(a, d, b, pencil, 0, nseg, sl, acc);
          }
        }
      }
      if (off >= S - 3) {  // from above
        if (s < nseg - 1) {
          add_slot<NV>(a, d, b, pencil, s + 1, nseg, off - (S - 3), acc);
        } else {
          const int nb = bk.nbr[13 + sdn];
          if (nb >= 0) {
            add_slot<NV>(a, d, nb, pencil, 0, nseg, off - (S - 3), acc);
          } else if (bk.bc[2 * d + 1] == APH_BC_REFLECT) {
            add_slot<NV>(a, d, b, pencil, nseg - 1, nseg, 3 + (n_d - 1 - cc[d]), acc);
          } else if (cc[d] == n_d - 1) {
            for (int sl = 3; sl < 6; ++sl)
              add_slot<NV>(a, d, b, pencil, nseg - 1, nseg, sl, acc);
          }
        }
      }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) a.lam_in[o + v * nc] = acc[v];
    if (a.lam_n && a.stage == 2) {
#pragma unroll
      for (int v = 0; v < NV; ++v) a.lam_n[o + v * nc] = a.gam1 * lo[v];
    }
  }
}

}  // namespace

// host side ------------------------------------------------------------------------

using AxisFn = void (*)(AdjArgs);
using GatherFn = void (*)(AdjArgs);

template <int RECON, int RS, int NV, int ND>
void pick_axis(int d, AxisFn &f) {
  if (d == 0) f = k_adj_axis<RECON, RS, NV, ND, 0>;
  if constexpr (ND >= 2)
    if (d == 1) f = k_adj_axis<RECON, RS, NV, ND, 1>;
  if constexpr (ND >= 3)
    if (d == 2) f = k_adj_axis<RECON, RS, NV, ND, 2>;
}

template <int RECON, int RS, int NV>
AxisFn axis_nd(int nd, int d) {
  AxisFn f = nullptr;
  if (nd == 1) pick_axis<RECON, RS, NV, 1>(d, f);
  if (nd == 2) pick_axis<RECON, RS, NV, 2>(d, f);
  if (nd == 3) pick_axis<RECON, RS, NV, 3>(d, f);
  return f;
}

AxisFn pick_adj_axis(int recon, int rs, int nv, int nd, int d) {
#define APH_AX(R, S)                                                             \
  if (recon == R && rs == S) {                                                    \
    if (nv == 5) return axis_nd<R, S, 5>(nd, d);                                  \
    if (nv == 6) return axis_nd<R, S, 6>(nd, d);                                  \
  }
  APH_AX(RECON_PPM, RS_HLLC)
  APH_AX(RECON_PPM, RS_HLLE)
  APH_AX(RECON_PLM, RS_HLLC)
  APH_AX(RECON_PLM, RS_HLLE)
#undef APH_AX
  return nullptr;
}

bool adj_supported(int recon, int rs, int nv, int nd) {
  return pick_adj_axis(recon, rs, nv, nd, 0) != nullptr;
}

GatherFn pick_gather(int nv, int nd) {
  if (nv == 5) return nd == 1 ? k_adj_gather<5, 1> : (nd == 2 ? k_adj_gather<5, 2> : k_adj_gather<5, 3>);
  if (nv == 6) return nd == 1 ? k_adj_gather<6, 1> : (nd == 2 ? k_adj_gather<6, 2> : k_adj_gather<6, 3>);
  return nullptr;
}

int adj_prepare(aph_context *ctx, const aph_blockset *bs, AdjScratch &s) {
  if (s.pbar) return APH_OK;
  const int nv = bs->nvar, nf = nv + 2, nd = bs->ndim, nb = bs->nblocks;
  // segment cap per axis: the longest segments (fewest recomputed halo cells) that still
  // give >= 2 waves of 64 lanes per SIMD (MI355X: 256 CUs x 4 SIMDs)
  constexpr int64_t kTargetLanes = 2LL * 64 * 1024;
  for (int d = 0; d < nd; ++d) {
    const int A = (d == 0) ? 1 : 0, B = (d == 2) ? 1 : 2;
    int best = 8;
    int64_t best_lanes = -1;
    for (int cap = 8; cap >= 3; --cap) {
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
      tot[d] += (int64_t)k.nx[A] * k.nx[B] * nseg * 6 * nv;
    }
  if (hipMalloc(&s.pbar, sizeof(double) * nf * bs->ncells) != hipSuccess ||
      hipMalloc(&s.lam1, sizeof(double) * nv * bs->ncells) != hipSuccess ||
      hipMalloc(&s.lam0p, sizeof(double) * nv * bs->ncells) != hipSuccess ||
      hipMalloc(&s.spill_base, sizeof(int64_t) * base.size()) != hipSuccess)
    return set_error(ctx, APH_E_NOMEM, "adjoint scratch allocation failed");
  for (int d = 0; d < nd; ++d)
    if (hipMalloc(&s.spill[d], sizeof(double) * std::max<int64_t>(tot[d], 1)) != hipSuccess)
      return set_error(ctx, APH_E_NOMEM, "adjoint spill allocation failed");
  APH_HIP_CHECK(ctx, hipMemcpy(s.spill_base, base.data(), sizeof(int64_t) * base.size(),
                               hipMemcpyHostToDevice));
  APH_HIP_CHECK(ctx, hipMemset(s.pbar, 0, sizeof(double) * nf * bs->ncells));
  APH_HIP_CHECK(ctx, hipMemset(s.lam1, 0, sizeof(double) * nv * bs->ncells));
  APH_HIP_CHECK(ctx, hipMemset(s.lam0p, 0, sizeof(double) * nv * bs->ncells));
  s.ncells = bs->ncells;
  return APH_OK;
}

void adj_release(AdjScratch &s) {
  if (s.pbar) (void)hipFree(s.pbar);
  for (auto &p : s.spill)
    if (p) (void)hipFree(p);
  if (s.spill_base) (void)hipFree(s.spill_base);
  if (s.lam1) (void)hipFree(s.lam1);
  if (s.lam0p) (void)hipFree(s.lam0p);
  s = AdjScratch{};
}

int adj_stage(aph_context *ctx, const aph_blockset *bs, int recon, int rs, const aph_eos &eos,
              const double *g, int grav, AdjScratch &s, int stage, const double *u_in,
              const double *dt_dev, const double *lam_out, const double *lam_acc,
              double *lam_in, double *lam_n) {
  int rc = adj_prepare(ctx, bs, s);
  if (rc) return rc;
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
  a.gam0 = stage == 1 ? 0.0 : 0.5;
  a.gam1 = stage == 1 ? 1.0 : 0.5;
  a.beta = stage == 1 ? 1.0 : 0.5;
  a.stage = stage;
  a.pbar = s.pbar;
  for (int d = 0; d < 3; ++d) a.spill[d] = s.spill[d];
  a.spill_base = s.spill_base;
  for (int d = 0; d < 3; ++d) a.segcap[d] = s.segcap[d];
  const int nd = bs->ndim;
  for (int d = 0; d < nd; ++d) {
    AxisFn f = pick_adj_axis(recon, rs, bs->nvar, nd, d);
    if (!f) return set_error(ctx, APH_E_INVALID, "adjoint: unsupported combination");
    int64_t maxu = 0;
    for (const auto &k : bs->h) {
      const int A = (d == 0) ? 1 : 0, B = (d == 2) ? 1 : 2;
      maxu = std::max<int64_t>(maxu, (int64_t)k.nx[A] * k.nx[B] *
                                         (k.nx[d] / adj_seg(k.nx[d], s.segcap[d])));
    }
    dim3 gr((unsigned)((maxu + 63) / 64), (unsigned)bs->nblocks);
    hipLaunchKernelGGL(f, gr, dim3(64), 0, ctx->stream, a);
  }
  GatherFn gf = pick_gather(bs->nvar, nd);
  dim3 gg((unsigned)std::max<int64_t>(1, std::min<int64_t>(64, (bs->max_interior + 255) / 256)),
          (unsigned)bs->nblocks);
  hipLaunchKernelGGL(gf, gg, dim3(256), 0, ctx->stream, a);
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

}  // namespace aph
