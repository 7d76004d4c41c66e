// aph_march.hpp -- one lane sweeps one pencil: sliding-window reconstruction and
// Riemann solve, shared by the plain flux kernels and the fused stage kernel.
//
// The reference reconstructs cells [s-1, e+1] of a pencil into team scratch and then
// solves every face [s, e+1] (hydro.cpp:596-603, plm/ppm_simple.hpp).  Here one lane
// walks the pencil holding a window of W = 2*GR+1 cells of all fields in registers:
// each step loads one new cell, reconstructs the window's centre cell (ql at its upper
// face, qr at its lower face) and solves the face below it with the ql kept from the
// previous step.  Every cell is loaded once and reconstructed once per sweep, the
// face states never leave registers, and no team barrier is needed.
#pragma once

#include "aph_device.hpp"

namespace aph {

// Sweep faces [f_lo, f_hi] along one axis.  A cell (pencil coordinate c) enters the
// window in two parts: fetch(c, raw[NR]) issues its loads, convert(raw, dst[NF]) turns
// them into the NV primitive fields followed by (gamma_c, gamma_e).  The fetch of the
// next step's cell is issued before the current step's reconstruction and Riemann
// solve, so its memory latency is hidden behind them.
// emit(f, flux[NV], ql[NF], qr[NF]) receives the face flux in conserved-variable
// order (IDN, IM1, IM2, IM3, IEN, scalars...) for face f (between cells f-1 and f).
// IVX = 1 + sweep axis (the reference's ivx, hydro_hllc.hpp:157-158).
template <int RECON, int RS, int NV, int IVX, int NR, class Fetch, class Convert, class Emit>
__device__ __forceinline__ void march(int f_lo, int f_hi, Fetch &&fetch, Convert &&convert,
                                      Emit &&emit) {
  constexpr int GR = ReconTraits<RECON>::GR;
  constexpr int W = ReconTraits<RECON>::W;
  constexpr int NF = NV + 2;
  constexpr int IVY = 1 + IVX % 3;
  constexpr int IVZ = 1 + (IVX + 1) % 3;
  double win[W][NF];
  double qlp[NF];
  double fv[NF];  // PPM: limited value of the face below the cell being reconstructed
  double d2c[NF];  // PPM: second difference (ppm_d2) of the cell being reconstructed
  double raw[NR];
  // PPM: bit s is set when every field of window cell s is +-0 or 2^-900 <= |q| < 2^1000
  // (frexp exponent in [-899, 1000]); a face whose four cells all qualify may take
  // ppm_face<true>.  Non-finite values give exponent 0 and qualify: a state holding them
  // is already invalid.
  unsigned plain = 0;
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
#pragma unroll
  for (int s = 0; s < W - 1; ++s) {
    fetch(f_lo - 1 - GR + s, raw);
    convert(raw, win[s + 1]);
    if constexpr (RECON == RECON_PPM) plain |= plain_cell(win[s + 1]) << (s + 1);
  }
  fetch(f_lo - 1 + GR, raw);
  if constexpr (RECON == RECON_PPM) {
    // lower face of the first reconstructed cell (m = f_lo - 1): cells f_lo-3 .. f_lo,
    // with the second differences of cells f_lo-2 and f_lo-1 (the latter is carried)
    double d2lo[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      d2lo[f] = ppm_d2(win[1][f], win[2][f], win[3][f]);
      d2c[f] = ppm_d2(win[2][f], win[3][f], win[4][f]);
    }
    if (((plain >> 1) & 0xFu) == 0xFu) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
        fv[f] = ppm_face<true>(win[1][f], win[2][f], win[3][f], win[4][f], d2lo[f], d2c[f]);
    } else {
#pragma unroll
      for (int f = 0; f < NF; ++f)
        fv[f] = ppm_face<false>(win[1][f], win[2][f], win[3][f], win[4][f], d2lo[f], d2c[f]);
    }
  }
  for (int m = f_lo - 1; m <= f_hi; ++m) {
#pragma unroll
    for (int s = 0; s < W - 1; ++s)
#pragma unroll
      for (int f = 0; f < NF; ++f) win[s][f] = win[s + 1][f];
    convert(raw, win[W - 1]);
    if (m < f_hi) fetch(m + 1 + GR, raw);
    double ql[NF], qr[NF];
    if constexpr (RECON == RECON_PPM) {
      // face m+1/2 once (cells m-1 .. m+2 = win[1..4]); cell m's own face m-1/2 was the
      // previous step's upper face
      plain = (plain >> 1) | (plain_cell(win[W - 1]) << (W - 1));
      // second differences: cell m+1 (new), cell m (carried), cell m-1 (recomputed from
      // the window: three operations, cheaper in registers than carrying it)
      double d2n[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) d2n[f] = ppm_d2(win[2][f], win[3][f], win[4][f]);
      double up[NF];
      if (((plain >> 1) & 0xFu) == 0xFu) {
#pragma unroll
        for (int f = 0; f < NF; ++f)
          up[f] = ppm_face<true>(win[1][f], win[2][f], win[3][f], win[4][f], d2c[f], d2n[f]);
      } else {
#pragma unroll
        for (int f = 0; f < NF; ++f)
          up[f] = ppm_face<false>(win[1][f], win[2][f], win[3][f], win[4][f], d2c[f], d2n[f]);
      }
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        ppm_cell(win[0][f], win[1][f], win[2][f], win[3][f], win[4][f],
                 ppm_d2(win[0][f], win[1][f], win[2][f]), d2c[f], d2n[f], fv[f], up[f], ql[f],
                 qr[f]);
        fv[f] = up[f];
        d2c[f] = d2n[f];
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
    if (m >= f_lo) {
      const double wl[5] = {qlp[0], qlp[IVX], qlp[IVY], qlp[IVZ], qlp[4]};
      const double wr[5] = {qr[0], qr[IVX], qr[IVY], qr[IVZ], qr[4]};
      double flx[5];
      riemann<RS>(wl, wr, &qlp[NV], &qr[NV], flx);
      double fo[NV];
      fo[0] = flx[0];
      fo[IVX] = flx[1];
      fo[IVY] = flx[2];
      fo[IVZ] = flx[3];
      fo[4] = flx[4];
#pragma unroll
      for (int n = 5; n < NV; ++n)  // passive scalars: upwind by the mass flux (:606-615)
        fo[n] = (fo[0] >= 0.0) ? fo[0] * qlp[n] : fo[0] * qr[n];
      emit(m, fo, qlp, qr);
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) qlp[f] = ql[f];
  }
}

}  // namespace aph
