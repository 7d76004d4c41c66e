// aph_adjoint_kernels.hpp -- the adjoint kernels of the fused stage as templates over
// (reconstruction, Riemann solver, NV, ndim), shared by aph_adjoint.hip (NV 5, 6) and
// aph_adjoint_wide.hip (the composition counts NV 12, 14), which instantiate them.
//
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
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

#include "aph_adjoint.hpp"
#include "aph_adjoint_host.hpp"
#include "aph_device.hpp"
#include "aph_runtime.hpp"

namespace aph {

// Segment length of the adjoint sweeps: the largest divisor of n not above cap; at least
// 3 (a spill reaches 3 cells), else the whole pencil.  cap is 8 unless the block set is
// too small to give every SIMD two waves with it (adj_prepare).
constexpr int kSplitSegMax = 16;  // longest segment of the split (PPM) form

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
  int stage;                   // 1: u_n is u_in (a step's first stage); > 1: lam_n written
  int lam_n_add;               // lam_n += gam1 L_out instead of lam_n = gam1 L_out
  double *pbar;                // per axis NF x ncells, pack layout (axis d at d * pbar_axis)
  int64_t pbar_axis;
  int64_t qb_axis;             // per axis offset of qlb / qrb (the axes run concurrently)
  double *prim, *prim_t;       // split form: NF x ncells, P of the interior cells in the
                               // pack / x-transposed layout
  double *lam_t;               // split form: NV x ncells, L_out x-transposed
  double *qlb, *qrb;           // split form: NF x ncells, cotangents of each cell's ql / qr
                               // output, in the layout of the axis being swept
  int pbar_xt;                 // the x axis' Pbar is in the x-transposed layout (split form)
  double *spill[3];            // per axis, [segment][slot][field][pencil] per block
  const int64_t *spill_base;   // [3][nblocks] element offsets
  int segcap[3];               // adj_seg cap per axis
  AdjAmr amr;                  // multilevel mesh (faces == nullptr: single level)
  double *contrib;             // multilevel: per ref of the ghost ops, NV cotangents
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
    } else if (m.bclo == APH_BC_AMR) {
      idx = c;  // the block's own ghost cells, set from the other level
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
    } else if (m.bchi == APH_BC_AMR) {
      idx = c;
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


// Split form ---------------------------------------------------------------------

// Where pencil coordinate c of the current block lives along axis D: the block itself,
// a neighbour, or a boundary image (reflect: mirrored, normal velocity negated; outflow:
// the edge cell).  Returns that block's cell_base and sets the index along D.
struct CellMap {
  int64_t self, lo, hi;  // block / neighbour cell_base, -1 = none
  int bclo, bchi;
};

__device__ __forceinline__ int64_t map_cell(const CellMap &m, int n_d, int c, int &idx,
                                            bool &flip) {
  int64_t base = m.self;
  idx = c;
  flip = false;
  if (c < 0) {
    if (m.lo >= 0) {
      base = m.lo;
      idx = c + n_d;
    } else if (m.bclo == APH_BC_REFLECT) {
      idx = -1 - c;
      flip = true;
    } else if (m.bclo == APH_BC_AMR) {
      idx = c;  // the block's own ghost cells (P from k_adj_amr_prim)
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
    } else if (m.bchi == APH_BC_AMR) {
      idx = c;
    } else {
      idx = n_d - 1;
    }
  }
  return base;
}

// Flux correction in reverse, at a block's boundary face f = 0 (side 0) or n (side 1) on a
// face to another level.  Forward: a face to a finer region (role 2) replaces its own flux
// by (sum of the 2^(ndim-1) fine fluxes across it) * 2^(1-ndim), so its own Riemann solve
// gets no cotangent; a face to a coarser leaf (role 1) is one of those fine fluxes, which
// the coarse cell beside it uses too: its cotangent gains 2^(1-ndim) times the coarse
// copy's, mu_c = beta dt A_c / V_c * (+-L_out of the coarse cell) (the coarse block's own
// copy of the face: the cell below gives -, the cell above +).  pa, pb: the pencil's
// transverse cells in this block.
template <int NV, int NDIM, int D>
__device__ __forceinline__ void amr_face_mu(const AdjArgs &a, int b, int side, int pa, int pb,
                                            double *mu) {
  constexpr int A = (D == 0) ? 1 : 0, B = (D == 2) ? 1 : 2;
  const aph_amr_face af = a.amr.faces[b * 6 + 2 * D + side];
  if (af.role == 2) {
#pragma unroll
    for (int v = 0; v < NV; ++v) mu[v] = 0.0;
    return;
  }
  if (af.role != 1) return;
  const aph_block &ck = a.blocks[a.amr.cblk[b * 6 + 2 * D + side]];
  const int N0 = ext(ck, 0, NDIM), N1 = ext(ck, 1, NDIM), N2 = ext(ck, 2, NDIM);
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int g[3] = {gst(ck, 0, NDIM), gst(ck, 1, NDIM), gst(ck, 2, NDIM)};
  int idx[3] = {0, 0, 0};
  idx[D] = side ? 0 : ck.nx[D] - 1;  // the coarse cell beside the face
  idx[A] = (af.qa + pa) >> 1;
  if (B < NDIM) idx[B] = (af.qb + pb) >> 1;
  const int64_t cell =
      ((int64_t)(idx[2] + g[2]) * N1 + (idx[1] + g[1])) * N0 + (idx[0] + g[0]);
  const double coef_c = a.beta * a.dt[0] * ck.area[D] / ck.vol;
  const double share = (NDIM == 3) ? 0.25 : 0.5;
  const double *lam = a.lam_out + (int64_t)NV * ck.cell_base + cell;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double muc = side ? coef_c * (lam[v * nc] - 0.0) : coef_c * (0.0 - lam[v * nc]);
    mu[v] = mu[v] + muc * share;
  }
}

template <int D>
__device__ __forceinline__ CellMap cell_map(const AdjArgs &a, const aph_block &bk) {
  constexpr int sdn = (D == 0) ? 1 : (D == 1 ? 3 : 9);
  CellMap m;
  m.self = bk.cell_base;
  const int nl = bk.nbr[13 - sdn], nh = bk.nbr[13 + sdn];
  m.lo = nl >= 0 ? a.blocks[nl].cell_base : -1;
  m.hi = nh >= 0 ? a.blocks[nh].cell_base : -1;
  m.bclo = bk.bc[2 * D];
  m.bchi = bk.bc[2 * D + 1];
  return m;
}

// Unit u of axis D -> (segment, pencil coordinates pa, pb), pa fastest: adjacent lanes
// take adjacent pencils.
__device__ __forceinline__ void split_geom(int u, int nA, int nseg, int &seg, int &pa,
                                           int &pb) {
  pa = u % nA;
  const int r = u / nA;
  seg = r % nseg;
  pb = r / nseg;
}

// Spill record of axis d, block blk: slot `slot` (0-2 below, 3-5 above the segment) of
// (pencil, segment), field 0; field f lies f * fstride further.  The fastest index is the
// one the writing lanes vary along, so a wave writes adjacent addresses: pencils for the
// split form and the one-march form's y and z axes ([segment][slot][field][pencil]),
// segments for the one-march form's x axis, whose adjacent lanes take the segments of one
// row ([slot][field][pencil][segment]).
template <int NF>
__device__ __forceinline__ double *spill_at(const AdjArgs &a, int d, int blk, int pencil,
                                            int seg, int slot, bool segfast,
                                            int64_t &fstride) {
  const aph_block &k = a.blocks[blk];
  const int A = (d == 0) ? 1 : 0, B = (d == 2) ? 1 : 2;
  const int64_t npen = (int64_t)k.nx[A] * k.nx[B];
  const int64_t base = a.spill_base[d * a.nblocks + blk];
  if (segfast) {
    const int64_t nseg = k.nx[d] / adj_seg(k.nx[d], a.segcap[d]);
    fstride = npen * nseg;
    return a.spill[d] + base + (int64_t)slot * NF * fstride + (int64_t)pencil * nseg + seg;
  }
  fstride = npen;
  return a.spill[d] + base + ((int64_t)(seg * 6 + slot) * NF) * npen + pencil;
}

// Offset of pencil (pa, pb) at axis coordinate 0 and the stride along axis D, in the pack
// layout (x fastest) or, XT, the x-transposed layout (y fastest, then z, then x) of the
// same ghost-inclusive N0 x N1 x N2 box.
struct Lay {
  int64_t base, sd;
};
template <int D, bool XT>
__device__ __forceinline__ Lay axis_lay(int pa, int pb, int N0, int N1, int N2,
                                        const int *g) {
  constexpr int A = (D == 0) ? 1 : 0, B = (D == 2) ? 1 : 2;
  int64_t st[3];
  if (XT) {
    st[0] = (int64_t)N1 * N2;
    st[1] = 1;
    st[2] = N1;
  } else {
    st[0] = 1;
    st[1] = N0;
    st[2] = (int64_t)N0 * N1;
  }
  Lay l;
  l.base = (int64_t)(pa + g[A]) * st[A] + (int64_t)(pb + g[B]) * st[B] + (int64_t)g[D] * st[D];
  l.sd = st[D];
  return l;
}

// P = (w[NV], gamma_c, gamma_e) = ConsToPrim(U) of every interior cell (with the floors, as
// the forward applies them when it reads a cell), in the pack and x-transposed layouts,
// and L_out in the x-transposed layout.  A workgroup takes 16 x 16 (x, y) tiles of one
// z-plane: it reads U and L_out and writes P along x, then writes the transposed copies
// along y through LDS, so every access runs over 16 consecutive doubles.
// prim / gather tiles: 32 cells along x by kPrimTileY / kGatherTileY along y (one row in
// 1D).  Measured on 64 RT instances: the transposing prim pass is fastest with whole
// 32 x 32 tiles (2.3 ms against 3.1 ms at 32 x 16 and 3.5 ms at 16 x 16), the gather with
// 32 x 16 (2.4 ms against 3.0 and 2.55): it holds 104 registers a lane, so 1024-lane
// workgroups leave one per CU with nothing to overlap its barrier.
// The wide composition counts (NV + 2 > 8 fields) take 8-row tiles, which keep the LDS tile
// of every field within 64 KB.
constexpr int prim_tile_y(int nf) { return nf > 8 ? 8 : 32; }
constexpr int gather_tile_y(int nf) { return nf > 8 ? 8 : 16; }
constexpr int kGatherTileYMarch = 8;  // the one-march form's gather (no LDS tile)

template <int NV, int NDIM>
__global__ void __launch_bounds__(1024) k_adj_prim(AdjArgs a) {
  constexpr int NF = NV + 2;
  constexpr int TX = 32, TY = (NDIM == 1) ? 1 : prim_tile_y(NF);
  // one 32 x 32 (x, y) tile per pass: rows of 32 doubles read and written along x, then
  // the transposed copies written along y
  __shared__ double tile[NF][TY][TX + 1];
  const int b = blockIdx.y;
  const aph_block &bk = a.blocks[b];
  const int N0 = ext(bk, 0, NDIM), N1 = ext(bk, 1, NDIM), N2 = ext(bk, 2, NDIM);
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int g[3] = {gst(bk, 0, NDIM), gst(bk, 1, NDIM), gst(bk, 2, NDIM)};
  const int nx0 = bk.nx[0], nx1 = bk.nx[1], nx2 = bk.nx[2];
  const int nti = (nx0 + TX - 1) / TX, ntj = (nx1 + TY - 1) / TY;
  const int ntiles = nti * ntj * nx2;
  const int64_t o = (int64_t)NV * bk.cell_base;
  double *pp = a.prim + (int64_t)NF * bk.cell_base;
  double *pt = a.prim_t + (int64_t)NF * bk.cell_base;
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int k = t / (nti * ntj), r = t % (nti * ntj);
    const int i0 = (r % nti) * TX, j0 = (r / nti) * TY;
    const int i = i0 + tx, j = j0 + ty;  // along x
    const bool in = i < nx0 && j < nx1;
    const int64_t cell = ((int64_t)(k + g[2]) * N1 + (j + g[1])) * N0 + (i + g[0]);
    // the transposed copies: thread q takes cell (i0 + q / TY, j0 + q % TY), adjacent y
    const int jl = threadIdx.x % TY, il = threadIdx.x / TY;
    const bool int_ = i0 + il < nx0 && j0 + jl < nx1;
    const int64_t cellt =
        ((int64_t)(i0 + il + g[0]) * N2 + (k + g[2])) * N1 + (j0 + jl + g[1]);
    if (in) {
      double u[NV], w[NV], gm[2];
#pragma unroll
      for (int v = 0; v < NV; ++v) u[v] = a.u_in[o + v * nc + cell];
      cons_to_prim_pt<NV>(u, a.eos, w, gm);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        pp[v * nc + cell] = w[v];
        tile[v][ty][tx] = w[v];
      }
      pp[NV * nc + cell] = gm[0];
      pp[(NV + 1) * nc + cell] = gm[1];
      tile[NV][ty][tx] = gm[0];
      tile[NV + 1][ty][tx] = gm[1];
    }
    __syncthreads();
    if (int_) {
#pragma unroll
      for (int f = 0; f < NF; ++f) pt[f * nc + cellt] = tile[f][jl][il];
    }
    __syncthreads();
    // L_out, the same way (the tile reused)
    if (in) {
#pragma unroll
      for (int v = 0; v < NV; ++v) tile[v][ty][tx] = a.lam_out[o + v * nc + cell];
    }
    __syncthreads();
    if (int_) {
#pragma unroll
      for (int v = 0; v < NV; ++v) a.lam_t[o + v * nc + cellt] = tile[v][jl][il];
    }
    __syncthreads();
  }
}

// Face adjoint of axis D: one lane per (block, pencil, segment) re-marches the forward
// reconstruction over its faces [a0, fend) from P and back-propagates the face cotangent
// mu_f = beta dt A/V (L_out(f) - L_out(f-1)) (only this block's cells use this block's
// copy of a face) through the Riemann solver and the passive-scalar upwind rule.  It
// writes QL_bar(f-1) (cotangent of the ql output of cell f-1) and QR_bar(f) (of the qr
// output of cell f), NF fields each; the block-local copies for the cells -1 and n go to
// the ghost slots.
template <int RECON, int RS, int NV, int NDIM, int D>
__device__ __forceinline__ void adj_face_body(const AdjArgs &a) {
  constexpr int GR = ReconTraits<RECON>::GR;
  constexpr int W = ReconTraits<RECON>::W;
  constexpr int NF = NV + 2;
  constexpr int IVX = 1 + D, IVY = 1 + IVX % 3, IVZ = 1 + (IVX + 1) % 3;
  constexpr int A = (D == 0) ? 1 : 0, B = (D == 2) ? 1 : 2;
  constexpr bool XT = (D == 0);
  const int b = blockIdx.y;
  const aph_block &bk = a.blocks[b];
  const int n_d = bk.nx[D];
  const int S = adj_seg(n_d, a.segcap[D]);
  const int nseg = n_d / S;
  const int nA = bk.nx[A], nB = bk.nx[B];
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nA * nB * nseg) return;
  int seg, pa, pb;
  split_geom(u, nA, nseg, seg, pa, pb);
  const int N0 = ext(bk, 0, NDIM), N1 = ext(bk, 1, NDIM), N2 = ext(bk, 2, NDIM);
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int g[3] = {gst(bk, 0, NDIM), gst(bk, 1, NDIM), gst(bk, 2, NDIM)};
  const CellMap m = cell_map<D>(a, bk);
  const Lay L = axis_lay<D, XT>(pa, pb, N0, N1, N2, g);
  const int a0 = seg * S;
  const int fend = (seg == nseg - 1) ? n_d + 1 : a0 + S;  // faces [a0, fend)
  const double coef = a.beta * a.dt[0] * bk.area[D] / bk.vol;
  const double *prim = XT ? a.prim_t : a.prim;
  const double *lo_ptr = (XT ? a.lam_t : a.lam_out) + (int64_t)NV * bk.cell_base + L.base;
  double *qlb_ptr = a.qlb + D * a.qb_axis + (int64_t)NF * bk.cell_base + L.base;
  double *qrb_ptr = a.qrb + D * a.qb_axis + (int64_t)NF * bk.cell_base + L.base;
  // fetch issues a cell's loads and returns whether it is a reflected image (whose normal
  // velocity changes sign); the sign is applied where the value is used, so the loads of
  // the next step stay in flight over this step's Riemann adjoint
  auto fetch = [&](int c, double *dst) -> bool {
    int idx;
    bool flip;
    const int64_t base = map_cell(m, n_d, c, idx, flip);
    const double *p = prim + (int64_t)NF * base + L.base + (int64_t)idx * L.sd;
#pragma unroll
    for (int f = 0; f < NF; ++f) dst[f] = p[f * nc];
    return flip;
  };
  auto load = [&](int c, double *dst) {
    if (fetch(c, dst)) dst[IVX] = -1.0 * dst[IVX];
  };
  double win[W][NF];  // win[k] = cell mm-GR+k
  // PPM as in the forward march: bit k of `plain` says every field of win[k] is +-0 or
  // 2^-900 <= |.| < 2^1000; faces over four such cells divide by 6 with div6_fast, and
  // each limited face value is formed once (fv: face mm-1/2)
  auto plain_cell = [](const double *c) -> unsigned {
    int lo = 0, hi = 0;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const int k = __builtin_amdgcn_frexp_exp(c[f]);
      lo = min(lo, k);
      hi = max(hi, k);
    }
    return (lo >= -899 && hi <= 1000) ? 1u : 0u;
  };
  auto faces = [&](unsigned pl, double *out) {  // face between win[2] and win[3]
    if (((pl >> 1) & 0xFu) == 0xFu) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
        out[f] = ppm_face<true>(win[1][f], win[2][f], win[3][f], win[4][f]);
    } else {
#pragma unroll
      for (int f = 0; f < NF; ++f)
        out[f] = ppm_face<false>(win[1][f], win[2][f], win[3][f], win[4][f]);
    }
  };
  unsigned plain = 0;
  double fv[NF];
#pragma unroll
  for (int s = 0; s < W - 1; ++s) {
    load(a0 - 1 - GR + s, win[s + 1]);
    if constexpr (RECON == RECON_PPM) plain |= plain_cell(win[s + 1]) << (s + 1);
  }
  if constexpr (RECON == RECON_PPM) faces(plain, fv);  // face a0-3/2
  double qlp[NF], lprev[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v)
    lprev[v] = (a0 >= 1) ? lo_ptr[v * nc + (int64_t)(a0 - 1) * L.sd] : 0.0;
  // the next step's cell and face cotangent, loaded one step ahead (the same addresses,
  // cells a0-1+GR .. fend-1+GR and faces a0 .. fend-1)
  double nxt[NF], lnxt[NV];
  bool nflip = fetch(a0 - 1 + GR, nxt);
#pragma unroll
  for (int v = 0; v < NV; ++v)
    lnxt[v] = (a0 < n_d) ? lo_ptr[v * nc + (int64_t)a0 * L.sd] : 0.0;
  for (int mm = a0 - 1; mm <= fend - 1; ++mm) {
#pragma unroll
    for (int s = 0; s < W - 1; ++s)
#pragma unroll
      for (int f = 0; f < NF; ++f) win[s][f] = win[s + 1][f];
#pragma unroll
    for (int f = 0; f < NF; ++f) win[W - 1][f] = nxt[f];
    if (nflip) win[W - 1][IVX] = -1.0 * win[W - 1][IVX];
    if (mm + 1 <= fend - 1) nflip = fetch(mm + 1 + GR, nxt);
    double ql[NF], qr[NF];
    if constexpr (RECON == RECON_PPM) {
      plain = (plain >> 1) | (plain_cell(win[W - 1]) << (W - 1));
      double up[NF];
      faces(plain, up);  // face mm+1/2
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        ppm_cell(win[0][f], win[1][f], win[2][f], win[3][f], win[4][f], fv[f], up[f], ql[f],
                 qr[f]);
        fv[f] = up[f];
      }
    } else {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        double col[W];
#pragma unroll
        for (int s = 0; s < W; ++s) col[s] = win[s][f];
        recon1<RECON>(col, ql[f], qr[f]);
      }
    }
    if (mm >= a0) {  // face mm between cells mm-1 (ql = qlp) and mm (qr)
      const int f = mm;
      double lcur[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) lcur[v] = lnxt[v];
      if (f + 1 <= fend - 1) {
#pragma unroll
        for (int v = 0; v < NV; ++v)
          lnxt[v] = (f + 1 < n_d) ? lo_ptr[v * nc + (int64_t)(f + 1) * L.sd] : 0.0;
      }
      double mu[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) mu[v] = coef * (lcur[v] - lprev[v]);
#pragma unroll
      for (int v = 0; v < NV; ++v) lprev[v] = lcur[v];
      if (a.amr.faces && ((f == 0 && m.bclo == APH_BC_AMR) || (f == n_d && m.bchi == APH_BC_AMR)))
        amr_face_mu<NV, NDIM, D>(a, b, f == 0 ? 0 : 1, pa, pb, mu);
      const double wl[5] = {qlp[0], qlp[IVX], qlp[IVY], qlp[IVZ], qlp[4]};
      const double wr[5] = {qr[0], qr[IVX], qr[IVY], qr[IVZ], qr[4]};
      double flx[5];
      riemann<RS>(wl, wr, &qlp[NV], &qr[NV], flx);  // forward flux for the scalar upwind
      double fb[5] = {mu[0], mu[IVX], mu[IVY], mu[IVZ], mu[4]};
      double qlb[NF], qrb[NF];
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        qlb[k] = 0.0;
        qrb[k] = 0.0;
      }
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
      double *pl = qlb_ptr + (int64_t)(f - 1) * L.sd;
      double *pr = qrb_ptr + (int64_t)f * L.sd;
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        pl[k * nc] = qlb[k];
        pr[k * nc] = qrb[k];
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) qlp[f] = ql[f];
  }
}

// Reconstruction (PPM) adjoint of axis D, one field per lane: lane (block, pencil,
// segment, field) takes the cells m of its segment (segment 0 also m = -1, the last also
// m = n) and applies the adjoint of field `fld`'s reconstruction of cell m to
// (QL_bar(m), QR_bar(m)), recomputing the forward from P.  Faces are shared as in the
// forward march: each step forms the limited value of face m+1/2 once (with its branch
// tape), runs the cell part's adjoint, and then propagates face m-1/2's cotangent --
// cell m-1's share plus cell m's -- once.  A face between two segments gets each
// segment's share from that segment's lane (the face adjoint is linear in its cotangent).
// The contributions to cells m-2 .. m+2 are summed in a window; the segment's own cells
// go to the axis' Pbar, the <= 3 cells on either side to the segment's
// 6-slot spill record.
template <int RECON, int NV, int NDIM, int D>
__device__ __forceinline__ void adj_recon_body(const AdjArgs &a) {
  static_assert(RECON == RECON_PPM, "the split form is the PPM adjoint");
  constexpr int GR = ReconTraits<RECON>::GR;
  constexpr int W = ReconTraits<RECON>::W;
  constexpr int NF = NV + 2;
  constexpr int IVX = 1 + D;
  constexpr int A = (D == 0) ? 1 : 0, B = (D == 2) ? 1 : 2;
  constexpr bool XT = (D == 0);
  const int b = blockIdx.y;
  const aph_block &bk = a.blocks[b];
  const int n_d = bk.nx[D];
  const int S = adj_seg(n_d, a.segcap[D]);
  const int nseg = n_d / S;
  const int nA = bk.nx[A], nB = bk.nx[B];
  const int per_field = nA * nB * nseg;
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= per_field * NF) return;
  const int fld = u / per_field;
  int seg, pa, pb;
  split_geom(u - fld * per_field, nA, nseg, seg, pa, pb);
  const int pencil = pa + pb * nA;
  const int N0 = ext(bk, 0, NDIM), N1 = ext(bk, 1, NDIM), N2 = ext(bk, 2, NDIM);
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int g[3] = {gst(bk, 0, NDIM), gst(bk, 1, NDIM), gst(bk, 2, NDIM)};
  const CellMap m = cell_map<D>(a, bk);
  const Lay L = axis_lay<D, XT>(pa, pb, N0, N1, N2, g);  // P, QL_bar, QR_bar and Pbar
  const Lay Lp = L;  // the x axis writes its Pbar x-transposed too: adjacent lanes, adjacent
                     // addresses (the gather transposes it back through LDS)
  const int a0 = seg * S;
  const int m_lo = (seg == 0) ? -1 : a0;
  const int m_hi = (seg == nseg - 1) ? n_d : a0 + S - 1;
  const double *prim = (XT ? a.prim_t : a.prim) + fld * nc + L.base;
  const double *qlb_ptr = a.qlb + D * a.qb_axis + (int64_t)NF * bk.cell_base + fld * nc + L.base;
  const double *qrb_ptr = a.qrb + D * a.qb_axis + (int64_t)NF * bk.cell_base + fld * nc + L.base;
  double *pb_ptr = a.pbar + D * a.pbar_axis + (int64_t)NF * bk.cell_base + fld * nc + Lp.base;
  int64_t fstr;
  double *sp = spill_at<NF>(a, D, b, pencil, seg, 0, false, fstr) + fld * fstr;
  const int64_t sstride = NF * fstr;  // one slot to the next
  // every halo cell is emitted exactly once: clear the record, then write each slot once
#pragma unroll
  for (int s = 0; s < 6; ++s) sp[s * sstride] = 0.0;
  const bool vel = (fld == IVX);
  auto load = [&](int c) -> double {
    int idx;
    bool flip;
    const int64_t base = map_cell(m, n_d, c, idx, flip);
    const double v = prim[(int64_t)NF * base + (int64_t)idx * L.sd];
    return (flip && vel) ? -1.0 * v : v;
  };
  // the next step's operands are loaded one step ahead; the sign change of a reflected
  // value is applied when it is used, so the loads stay in flight across a step
  auto fetch = [&](int c, double &v, bool &ng) {
    int idx;
    bool flip;
    const int64_t base = map_cell(m, n_d, c, idx, flip);
    v = prim[(int64_t)NF * base + (int64_t)idx * L.sd];
    ng = flip && vel;
  };
  // each axis writes its own Pbar (the gather sums the axes in order), one owner lane
  // per cell
  auto own = [&](int c) { return c >= a0 && c < a0 + S; };
  auto emit = [&](int c, double v) {
    if (own(c)) {
      pb_ptr[(int64_t)c * Lp.sd] = v;
    } else {
      const bool below = c < a0;
      const int slot = below ? c - (a0 - 3) : 3 + (c - (a0 + S));
      const bool image = below ? (c < 0 && m.lo < 0 && m.bclo == APH_BC_REFLECT)
                               : (c >= n_d && m.hi < 0 && m.bchi == APH_BC_REFLECT);
      sp[slot * sstride] = (image && vel) ? -1.0 * v : v;
    }
  };
  // bit k of `plain`: col[k] is +-0 or 2^-900 <= |.| < 2^1000, so a face over four such
  // cells may divide by 6 with div6_fast (exact there, as in the forward march)
  auto plain1 = [](double x) -> unsigned {
    const int k = __builtin_amdgcn_frexp_exp(x);
    return (k >= -899 && k <= 1000) ? 1u : 0u;
  };
  auto face = [](const double *c, unsigned pl) -> PpmFace {
    return ((pl >> 1) & 0xFu) == 0xFu ? ppm_face_taped<true>(c[1], c[2], c[3], c[4])
                                       : ppm_face_taped<false>(c[1], c[2], c[3], c[4]);
  };
  double col[W];  // col[k] = field value of cell mm-GR+k
  double cw[W];   // cw[k] = cotangent of cell mm-GR+k
  unsigned plain = 0;
#pragma unroll
  for (int s = 0; s < W - 1; ++s) {
    col[s + 1] = load(m_lo - GR + s);
    plain |= plain1(col[s + 1]) << (s + 1);
  }
#pragma unroll
  for (int s = 0; s < W; ++s) cw[s] = 0.0;
  PpmFace flo = face(col, plain);          // face mm-1/2 (cells mm-2 .. mm+1)
  double fb_lo = 0.0;                      // its cotangent from cell mm-1 (this lane's)
  double nv, nql, nqr;
  bool nng;
  fetch(m_lo + GR, nv, nng);
  nql = (m_lo <= n_d - 1) ? qlb_ptr[(int64_t)m_lo * L.sd] : 0.0;
  nqr = (m_lo >= 0) ? qrb_ptr[(int64_t)m_lo * L.sd] : 0.0;
  for (int mm = m_lo; mm <= m_hi; ++mm) {
    // (the first step's cell, m_lo - GR, is never the segment's own: no load needed)
#pragma unroll
    for (int s = 0; s < W - 1; ++s) col[s] = col[s + 1];
    col[W - 1] = nng ? -1.0 * nv : nv;
    const double ql_b = nql, qr_b = nqr;
    if (mm < m_hi) {
      fetch(mm + 1 + GR, nv, nng);
      nql = (mm + 1 <= n_d - 1) ? qlb_ptr[(int64_t)(mm + 1) * L.sd] : 0.0;
      nqr = qrb_ptr[(int64_t)(mm + 1) * L.sd];  // mm + 1 >= 0
    }
    plain = (plain >> 1) | (plain1(col[W - 1]) << (W - 1));
    const PpmFace fhi = face(col, plain);  // face mm+1/2 (cells mm-1 .. mm+2)
    double colb[W];
#pragma unroll
    for (int s = 0; s < W; ++s) colb[s] = 0.0;
    double flob, fhib;
    ppm_cell_adj(col, flo.dph, fhi.dph, ql_b, qr_b, colb, flob, fhib);
    ppm_face_adj(flo.t, fb_lo + flob, colb);  // face mm-1/2 -> cells mm-2 .. mm+1
    fb_lo = fhib;
    flo = fhi;
#pragma unroll
    for (int s = 0; s < W; ++s) cw[s] += colb[s];
    emit(mm - GR, cw[0]);  // no later cell or face reaches mm-GR
#pragma unroll
    for (int s = 0; s < W - 1; ++s) cw[s] = cw[s + 1];
    cw[W - 1] = 0.0;
  }
  ppm_face_adj(flo.t, fb_lo, cw);  // face m_hi+1/2 -> cells m_hi-1 .. m_hi+2 = cw[0..3]
#pragma unroll
  for (int s = 0; s < W - 1; ++s) emit(m_hi - GR + 1 + s, cw[s]);
}

// The split form's axes in one launch each (blockIdx.z = axis): they read the same
// primitives and write separate QL_bar / QR_bar, Pbar arrays and spill records.
template <int RECON, int RS, int NV, int NDIM>
__global__ void __launch_bounds__(64, 1) k_adj_face(AdjArgs a) {
  if (blockIdx.z == 0) {
    adj_face_body<RECON, RS, NV, NDIM, 0>(a);
  } else if constexpr (NDIM >= 2) {
    if (blockIdx.z == 1) {
      adj_face_body<RECON, RS, NV, NDIM, 1>(a);
    } else if constexpr (NDIM >= 3) {
      adj_face_body<RECON, RS, NV, NDIM, 2>(a);
    }
  }
}

template <int RECON, int NV, int NDIM>
__global__ void __launch_bounds__(64) k_adj_recon(AdjArgs a) {
  if (blockIdx.z == 0) {
    adj_recon_body<RECON, NV, NDIM, 0>(a);
  } else if constexpr (NDIM >= 2) {
    if (blockIdx.z == 1) {
      adj_recon_body<RECON, NV, NDIM, 1>(a);
    } else if constexpr (NDIM >= 3) {
      adj_recon_body<RECON, NV, NDIM, 2>(a);
    }
  }
}

// One-march form (PLM) -------------------------------------------------------------

template <int RECON, int RS, int NV, int NDIM, int D>
__device__ __forceinline__ void adj_axis_body(const AdjArgs &a) {
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
  double *pb_ptr = a.pbar + D * a.pbar_axis + (int64_t)NF * bk.cell_base + tbase;
  int64_t fstr;  // the x axis' lanes take the segments of one row: segment-fastest records
  double *sp = spill_at<NF>(a, D, b, pencil, seg, 0, D == 0, fstr);
  // every halo cell is emitted exactly once: clear the record, then write each slot once
  for (int s = 0; s < 6 * NF; ++s) sp[s * fstr] = 0.0;

  auto load = [&](int c, double *dst) {
    double uu[NV];
    map_load<NV>(a, m, D, n_d, c, tbase, st[D], nc, uu);
    cons_to_prim_pt<NV>(uu, a.eos, dst, dst + NV);
  };
  // final primitive-space cotangent of cell c from this segment: each axis writes its own
  // Pbar (the gather sums the axes in order), one owner lane per cell
  auto emit = [&](int c, const double *pbv) {
    if (c >= a0 && c < a0 + S) {
      double *q = pb_ptr + (int64_t)c * st[D];
#pragma unroll
      for (int f = 0; f < NF; ++f) q[f * nc] = pbv[f];
    } else {
      const bool below = c < a0;
      const int slot = below ? c - (a0 - 3) : 3 + (c - (a0 + S));
      const bool image = below ? (c < 0 && m.lo < 0 && m.bclo == APH_BC_REFLECT)
                               : (c >= n_d && m.hi < 0 && m.bchi == APH_BC_REFLECT);
#pragma unroll
      for (int f = 0; f < NF; ++f)
        sp[(slot * NF + f) * fstr] = (image && f == IVX) ? -1.0 * pbv[f] : pbv[f];
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
      if (a.amr.faces && ((f == 0 && m.bclo == APH_BC_AMR) || (f == n_d && m.bchi == APH_BC_AMR)))
        amr_face_mu<NV, NDIM, D>(a, b, f == 0 ? 0 : 1, pa, pb, mu);
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

// Every axis of the one-march form in one launch (blockIdx.z = axis): the axes write
// separate Pbar arrays and spill records, so they are independent, and a small block set
// fills the GPU twice (2D) or three times (3D) as deep as one axis alone.
template <int RECON, int RS, int NV, int NDIM>
__global__ void __launch_bounds__(64) k_adj_axes(AdjArgs a) {
  if (blockIdx.z == 0) {
    adj_axis_body<RECON, RS, NV, NDIM, 0>(a);
  } else if constexpr (NDIM >= 2) {
    if (blockIdx.z == 1) {
      adj_axis_body<RECON, RS, NV, NDIM, 1>(a);
    } else if constexpr (NDIM >= 3) {
      adj_axis_body<RECON, RS, NV, NDIM, 2>(a);
    }
  }
}

template <int NF>
__device__ __forceinline__ void add_slot(const AdjArgs &a, int d, int blk, int pencil,
                                         int seg, int nseg, int slot, double *acc) {
  (void)nseg;
  int64_t fstr;
  const double *p = spill_at<NF>(a, d, blk, pencil, seg, slot, !a.pbar_xt && d == 0, fstr);
#pragma unroll
  for (int v = 0; v < NF; ++v) acc[v] += p[v * fstr];
}

// A block's spill records along one axis, its layout resolved once (spill_at's arithmetic
// without the per-call block-table reads and divisions).
struct SpillView {
  const double *base;
  int64_t npen, nseg;
  bool segfast;
};

template <int NF>
__device__ __forceinline__ void add_slot_view(const SpillView &v, int pencil, int seg, int slot,
                                              double *acc) {
  const double *p;
  int64_t fstr;
  if (v.segfast) {
    fstr = v.npen * v.nseg;
    p = v.base + (int64_t)slot * NF * fstr + (int64_t)pencil * v.nseg + seg;
  } else {
    fstr = v.npen;
    p = v.base + ((int64_t)(seg * 6 + slot) * NF) * v.npen + pencil;
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) acc[f] += p[f * fstr];
}

// XT: the x axis' Pbar is x-transposed (the split form) and goes through an LDS tile; the
// one-march form's is in the pack layout and needs no LDS, so its workgroups are smaller
// (32 x 8) and hold no tile.
template <int NV, int NDIM, bool XT>
__global__ void __launch_bounds__(512) k_adj_gather(AdjArgs a) {
  constexpr int NF = NV + 2;
  constexpr int TX = 32, TY = (NDIM == 1) ? 1 : (XT ? gather_tile_y(NF) : kGatherTileYMarch);
  // the x axis' Pbar of one 32 x 16 (x, y) tile, read along y from the x-transposed
  // layout and used along x (split form); every other array is read along x
  using Tile = double[NF][TY][TX + 1];
  Tile *xtp = nullptr;
  if constexpr (XT) {
    __shared__ Tile xt;
    xtp = &xt;
  }
  const int b = blockIdx.y;
  const aph_block &bk = a.blocks[b];
  const int N0 = ext(bk, 0, NDIM), N1 = ext(bk, 1, NDIM), N2 = ext(bk, 2, NDIM);
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int g[3] = {gst(bk, 0, NDIM), gst(bk, 1, NDIM), gst(bk, 2, NDIM)};
  const int nx0 = bk.nx[0], nx1 = bk.nx[1], nx2 = bk.nx[2];
  const int nti = (nx0 + TX - 1) / TX, ntj = (nx1 + TY - 1) / TY;
  const int ntiles = nti * ntj * nx2;
  const double beta_dt = a.beta * a.dt[0];
  const double gsum = (a.stage == 1) ? a.gam0 + a.gam1 : a.gam0;
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const double *pbx = a.pbar + (int64_t)NF * bk.cell_base;
  // per axis, once: the segment length (and its shift when a power of two), the segment
  // count and this block's spill records
  int segS[NDIM], segShift[NDIM];
  SpillView own[NDIM];
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    const int A = (d == 0) ? 1 : 0, B = (d == 2) ? 1 : 2;
    segS[d] = adj_seg(bk.nx[d], a.segcap[d]);
    segShift[d] = (segS[d] & (segS[d] - 1)) == 0 ? __builtin_ctz(segS[d]) : -1;
    own[d].base = a.spill[d] + a.spill_base[d * a.nblocks + b];
    own[d].npen = (int64_t)bk.nx[A] * bk.nx[B];
    own[d].nseg = bk.nx[d] / segS[d];
    own[d].segfast = !XT && d == 0;
  }
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int kz = t / (nti * ntj), r = t % (nti * ntj);
    const int i0 = (r % nti) * TX, j0 = (r / nti) * TY;
    if constexpr (XT) {  // thread q reads cell (i0 + q / TY, j0 + q % TY): adjacent y
      const int jl = threadIdx.x % TY, il = threadIdx.x / TY;
      const int i = i0 + il, j = j0 + jl;
      if (i < nx0 && j < nx1) {
        const int64_t cellt = ((int64_t)(i + g[0]) * N2 + (kz + g[2])) * N1 + (j + g[1]);
#pragma unroll
        for (int f = 0; f < NF; ++f) (*xtp)[f][jl][il] = pbx[f * nc + cellt];
      }
      __syncthreads();
    }
    const int cc[3] = {i0 + tx, j0 + ty, kz};
    if (cc[0] < nx0 && cc[1] < nx1) {
      const int64_t cell =
          ((int64_t)(cc[2] + g[2]) * N1 + (cc[1] + g[1])) * N0 + (cc[0] + g[0]);
      const int64_t o = (int64_t)NV * bk.cell_base + cell;
      double lo[NV], u[NV], acc[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        lo[v] = a.lam_out[o + v * nc];
        u[v] = a.u_in[o + v * nc];
        acc[v] = gsum * lo[v] + (a.lam_acc ? a.lam_acc[o + v * nc] : 0.0);
      }
      // the axes' Pbar, summed x, then y, then z
      double pbv[NF];
      const double *pp = a.pbar + (int64_t)NF * bk.cell_base + cell;
      if constexpr (XT) {
#pragma unroll
        for (int f = 0; f < NF; ++f) pbv[f] = (*xtp)[f][ty][tx];
      } else {
#pragma unroll
        for (int f = 0; f < NF; ++f) pbv[f] = pp[f * nc];
      }
#pragma unroll
      for (int d = 1; d < NDIM; ++d)
#pragma unroll
        for (int f = 0; f < NF; ++f) pbv[f] = pbv[f] + pp[d * a.pbar_axis + f * nc];
      // spills (primitive space) of the segments / blocks / boundary images that read
      // this cell
#pragma unroll
      for (int d = 0; d < NDIM; ++d) {
        const int Ad = (d == 0) ? 1 : 0, Bd = (d == 2) ? 1 : 2;
        const int n_d = bk.nx[d];
        const int S = segS[d], nseg = (int)own[d].nseg;
        const int pencil = cc[Ad] + cc[Bd] * bk.nx[Ad];
        const int s = segShift[d] >= 0 ? cc[d] >> segShift[d] : cc[d] / S;
        const int off = cc[d] - s * S;
        const int sdn = (d == 0) ? 1 : (d == 1 ? 3 : 9);
        if (off < 3) {  // from below
          if (s > 0) {
            add_slot_view<NF>(own[d], pencil, s - 1, 3 + off, pbv);
          } else {
            const int nb = bk.nbr[13 - sdn];
            if (nb >= 0) {
              add_slot<NF>(a, d, nb, pencil, nseg - 1, nseg, 3 + off, pbv);
            } else if (bk.bc[2 * d] == APH_BC_REFLECT) {
              add_slot_view<NF>(own[d], pencil, 0, 2 - off, pbv);
            } else if (bk.bc[2 * d] != APH_BC_AMR && cc[d] == 0) {  // outflow images
              for (int sl = 0; sl < 3; ++sl) add_slot_view<NF>(own[d], pencil, 0, sl, pbv);
            }
          }
        }
        if (off >= S - 3) {  // from above
          if (s < nseg - 1) {
            add_slot_view<NF>(own[d], pencil, s + 1, off - (S - 3), pbv);
          } else {
            const int nb = bk.nbr[13 + sdn];
            if (nb >= 0) {
              add_slot<NF>(a, d, nb, pencil, 0, nseg, off - (S - 3), pbv);
            } else if (bk.bc[2 * d + 1] == APH_BC_REFLECT) {
              add_slot_view<NF>(own[d], pencil, nseg - 1, 3 + (n_d - 1 - cc[d]), pbv);
            } else if (bk.bc[2 * d + 1] != APH_BC_AMR && cc[d] == n_d - 1) {
              for (int sl = 3; sl < 6; ++sl) add_slot_view<NF>(own[d], pencil, nseg - 1, sl, pbv);
            }
          }
        }
      }
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
  #pragma unroll
      for (int v = 0; v < NV; ++v) a.lam_in[o + v * nc] = acc[v];
      if (a.lam_n && a.stage > 1) {
  #pragma unroll
        for (int v = 0; v < NV; ++v)
          a.lam_n[o + v * nc] = a.lam_n_add ? a.gam1 * lo[v] + a.lam_n[o + v * nc]
                                            : a.gam1 * lo[v];
      }
    }
    if constexpr (XT) __syncthreads();  // the tile is used before the next one is loaded
  }
}

// Multilevel meshes ------------------------------------------------------------------
// The ghost cells across a face to another level hold values the cross-level ops computed
// from interior cells (the caller ran aph_amr_fill on u_in).  The sweeps read them like
// any stencil cell and leave their primitive-space cotangents in the spill slots beyond
// the block; these kernels take them back to the interior cells the ops read.

// Block-local ghost-inclusive (i, j, k) of a component-0 offset of block b.
__device__ __forceinline__ void local_ijk(const aph_block &bk, int64_t off, int ndim,
                                          int &i, int &j, int &k) {
  const int N0 = ext(bk, 0, ndim), N1 = ext(bk, 1, ndim);
  i = (int)(off % N0);
  off /= N0;
  j = (int)(off % N1);
  k = (int)(off / N1);
}

// Split form: P = ConsToPrim(U) of every op's ghost cell, in the pack and x-transposed
// layouts (the sweeps read the face ghosts among them).
template <int NV, int NDIM>
__global__ void __launch_bounds__(256) k_adj_amr_prim(AdjArgs a) {
  constexpr int NF = NV + 2;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < a.amr.nops;
       t += (int64_t)gridDim.x * blockDim.x) {
    const aph_amr_op op = a.amr.ops[t];
    const aph_block &bk = a.blocks[a.amr.op_block[t]];
    const int N0 = ext(bk, 0, NDIM), N1 = ext(bk, 1, NDIM), N2 = ext(bk, 2, NDIM);
    const int64_t cell = op.dst - (int64_t)NV * bk.cell_base;
    int i, j, k;
    local_ijk(bk, cell, NDIM, i, j, k);
    double u[NV], w[NV], gm[2];
#pragma unroll
    for (int v = 0; v < NV; ++v) u[v] = a.u_in[op.dst + v * op.dst_stride];
    cons_to_prim_pt<NV>(u, a.eos, w, gm);
    double *pp = a.prim + (int64_t)NF * bk.cell_base + cell;
    double *pt = a.prim_t + (int64_t)NF * bk.cell_base + ((int64_t)i * N2 + k) * N1 + j;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const double x = f < NV ? w[f] : gm[f - NV];
      pp[f * op.dst_stride] = x;
      pt[f * op.dst_stride] = x;
    }
  }
}

__device__ __forceinline__ double sgn1(double x) { return x < 0.0 ? -1.0 : 1.0; }

// Per op: the ghost cell's primitive-space cotangent (a face ghost within three cells of
// the block: the spill slot of the one pencil that reads it; edge and corner ghosts are
// never read), taken to conserved space at the ghost state (J_prim^T), then back through
// the op: MEAN (sum * 1/count) gives each child 1/count of it; PROLONG c + sum_d
// (sign_d / 4) minmod(c - lo_d, hi_d - c) follows minmod's active branch (equal signs;
// the smaller magnitude, a on a tie; nothing when the signs differ) and the mirror bits.
// Every ref of every op gets its NV contributions written (zeros included).
template <int NV, int NDIM>
__global__ void __launch_bounds__(256) k_adj_amr_ghost(AdjArgs a) {
  constexpr int NF = NV + 2;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < a.amr.nops;
       t += (int64_t)gridDim.x * blockDim.x) {
    const aph_amr_op op = a.amr.ops[t];
    const int b = a.amr.op_block[t];
    const aph_block &bk = a.blocks[b];
    const int g[3] = {gst(bk, 0, NDIM), gst(bk, 1, NDIM), gst(bk, 2, NDIM)};
    int id[3];
    local_ijk(bk, op.dst - (int64_t)NV * bk.cell_base, NDIM, id[0], id[1], id[2]);
    int axis = -1, nout = 0, c = 0;
    for (int d = 0; d < NDIM; ++d) {
      const int q = id[d] - g[d];
      if (q < 0 || q >= bk.nx[d]) {
        ++nout;
        axis = d;
        c = q;
      }
    }
    double pbv[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) pbv[f] = 0.0;
    if (nout == 1 && c >= -3 && c < bk.nx[axis] + 3) {
      const int d = axis;
      const int Ad = (d == 0) ? 1 : 0, Bd = (d == 2) ? 1 : 2;
      const int n_d = bk.nx[d];
      const int nseg = n_d / adj_seg(n_d, a.segcap[d]);
      const int pencil = (id[Ad] - g[Ad]) + (Bd < NDIM ? (id[Bd] - g[Bd]) * bk.nx[Ad] : 0);
      if (c < 0)
        add_slot<NF>(a, d, b, pencil, 0, nseg, c + 3, pbv);
      else
        add_slot<NF>(a, d, b, pencil, nseg - 1, nseg, 3 + (c - n_d), pbv);
    }
    double u[NV], ub[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      u[v] = a.u_in[op.dst + v * op.dst_stride];
      ub[v] = 0.0;
    }
    cons_to_prim_adj<NV>(u, a.eos, pbv, pbv + NV, ub);
    const int kind = op.kind & 0xff, mirror = op.kind >> APH_AMR_MIRROR_BIT;
    auto value = [&](const aph_amr_val &vv, int comp) -> double {
      const aph_amr_ref *r = a.amr.refs + vv.first;
      double s = a.u_in[r[0].offset + comp * r[0].stride];
      for (int q = 1; q < vv.count; ++q) s += a.u_in[r[q].offset + comp * r[q].stride];
      return vv.count == 1 ? s : s * (1.0 / vv.count);
    };
    auto put = [&](const aph_amr_val &vv, int comp, double vb) {
      const double each = vv.count == 1 ? vb : vb * (1.0 / vv.count);
      for (int q = 0; q < vv.count; ++q) a.contrib[(int64_t)(vv.first + q) * NV + comp] = each;
    };
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      if (kind == APH_AMR_MEAN) {
        put(op.val[0], v, ub[v]);
        continue;
      }
      const double cv = value(op.val[0], v);
      double cb = ub[v];
      for (int d = 0; d < NDIM; ++d) {
        const bool mlo = (v == 1 + d) && (mirror & (1 << (2 * d)));
        const bool mhi = (v == 1 + d) && (mirror & (2 << (2 * d)));
        const double lo = mlo ? -value(op.val[1 + 2 * d], v) : value(op.val[1 + 2 * d], v);
        const double hi = mhi ? -value(op.val[2 + 2 * d], v) : value(op.val[2 + 2 * d], v);
        const double x = cv - lo, y = hi - cv;
        const double gb = ub[v] * (op.sign[d] * 0.25);
        double xb = 0.0, yb = 0.0;
        if (sgn1(x) == sgn1(y)) {  // minmod = sgn * min(|x|, |y|): the smaller one
          if (fabs(y) < fabs(x)) yb = gb;
          else xb = gb;
        }
        cb += xb - yb;
        put(op.val[1 + 2 * d], v, mlo ? xb : -xb);  // lo enters x with -1 (-lo mirrored)
        put(op.val[2 + 2 * d], v, mhi ? -yb : yb);
      }
      put(op.val[0], v, cb);
    }
  }
}

// lam_in += the contributions of the ghost ops, per interior cell they read, summed over
// its refs in ascending order (one lane per cell: no atomics, a fixed order).
template <int NV>
__global__ void __launch_bounds__(256) k_adj_amr_scatter(AdjArgs a) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < a.amr.ntgt;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t off = a.amr.tgt[t], st = a.amr.tgt_stride[t];
    const int64_t r0 = a.amr.row[t], r1 = a.amr.row[t + 1];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double s = a.contrib[(int64_t)a.amr.col[r0] * NV + v];
      for (int64_t r = r0 + 1; r < r1; ++r) s += a.contrib[(int64_t)a.amr.col[r] * NV + v];
      a.lam_in[off + v * st] += s;
    }
  }
}
}  // namespace

// host side: kernel pointers of one axis
using KFn = void (*)(AdjArgs);

// Kernels of one axis: PPM takes the split form (face, recon), PLM the one-march form.
struct AxisKernels {
  KFn face = nullptr, recon = nullptr, axis = nullptr;
};

template <int RECON, int RS, int NV, int ND, int D>
void set_axis(AxisKernels &k) {
  if constexpr (RECON == RECON_PPM) {
    k.face = k_adj_face<RECON, RS, NV, ND>;  // every axis in one launch
    k.recon = k_adj_recon<RECON, NV, ND>;
  } else {
    k.axis = k_adj_axes<RECON, RS, NV, ND>;  // every axis in one launch
  }
}

template <int RECON, int RS, int NV, int ND>
void pick_axis(int d, AxisKernels &k) {
  if (d == 0) set_axis<RECON, RS, NV, ND, 0>(k);
  if constexpr (ND >= 2)
    if (d == 1) set_axis<RECON, RS, NV, ND, 1>(k);
  if constexpr (ND >= 3)
    if (d == 2) set_axis<RECON, RS, NV, ND, 2>(k);
}

template <int RECON, int RS, int NV>
void axis_nd(int nd, int d, AxisKernels &k) {
  if (nd == 1) pick_axis<RECON, RS, NV, 1>(d, k);
  if (nd == 2) pick_axis<RECON, RS, NV, 2>(d, k);
  if (nd == 3) pick_axis<RECON, RS, NV, 3>(d, k);
}

}  // namespace aph
