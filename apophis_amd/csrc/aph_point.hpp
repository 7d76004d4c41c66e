// aph_point.hpp -- point functions of the hydro flux path (FP64), callable from host
// and device code.  They restate the reference's arithmetic in its operand order; every
// caller compiles them with -ffp-contract=off so no multiply-add is fused.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define APH_HD __host__ __device__ __forceinline__
#else
#define APH_HD inline
#endif

#include "../../include/apophis_hydro.h"

namespace aph {
using std::fabs;
using std::fma;
using std::fmax;
using std::fmin;
using std::sqrt;

// Parthenon macros used by the reference (upstream): SIGN, SQR, TINY_NUMBER.
APH_HD double SIGN(double x) { return (x < 0.0) ? -1.0 : 1.0; }
APH_HD double SQR(double x) { return x * x; }
constexpr double TINY_NUMBER = 1.0e-20;
constexpr double C_LIGHT = 2.99792458e10;  // hydro_hlle.hpp:26, hydro_hllc.hpp:159
constexpr double C_LIGHT2 = C_LIGHT * C_LIGHT;
constexpr double EOS_SMALL = 10.0 * 2.2250738585072014e-308;  // singularity robust::SMALL()

// std::min / std::max semantics
APH_HD double mn(double a, double b) { return (b < a) ? b : a; }
APH_HD double mx(double a, double b) { return (a < b) ? b : a; }

constexpr int RECON_DC = APH_RECON_DC, RECON_PLM = APH_RECON_PLM, RECON_PPM = APH_RECON_PPM;
constexpr int RS_HLLE = APH_RIEMANN_HLLE, RS_HLLC = APH_RIEMANN_HLLC;
constexpr int RS_LHLLC = APH_RIEMANN_LHLLC, RS_NONE = APH_RIEMANN_NONE;

template <int RECON>
struct ReconTraits {
  static constexpr int GR = (RECON == RECON_PPM) ? 2 : ((RECON == RECON_PLM) ? 1 : 0);
  static constexpr int W = 2 * GR + 1;  // window of cells one reconstruction reads
};

// ------------------------------------------------------------ PLM
// src/recon/plm_simple.hpp:14-31
APH_HD void plm(double q_im1, double q_i, double q_ip1, double &ql_ip1,
                                    double &qr_i) {
  double dql = (q_i - q_im1);
  double dqr = (q_ip1 - q_i);
  double dq2 = dql * dqr;
  double dqm = 0.0;
  if (dq2 > 0.0) dqm = dq2 / (dql + dqr);
  ql_ip1 = q_i + dqm;
  qr_i = q_i - dqm;
}

// ------------------------------------------------------------ PPM
// src/recon/ppm_simple.hpp:31-161
APH_HD void ppm(double q_im2, double q_im1, double q_i, double q_ip1,
                                    double q_ip2, double &ql_ip1, double &qr_i) {
  const double C2 = 1.25;
  double qa = (q_i - q_im1);
  double qb = (q_ip1 - q_i);
  const double dd_im1 = 0.5 * qa + 0.5 * (q_im1 - q_im2);
  const double dd = 0.5 * qb + 0.5 * qa;
  const double dd_ip1 = 0.5 * (q_ip2 - q_ip1) + 0.5 * qb;
  double dph = 0.5 * (q_im1 + q_i) + (dd_im1 - dd) / 6.0;
  double dph_ip1 = 0.5 * (q_i + q_ip1) + (dd - dd_ip1) / 6.0;
  const double d2qc_im1 = q_im2 + q_i - 2.0 * q_im1;
  const double d2qc = q_im1 + q_ip1 - 2.0 * q_i;
  const double d2qc_ip1 = q_i + q_ip2 - 2.0 * q_ip1;
  double qa_tmp = dph - q_im1;
  double qb_tmp = q_i - dph;
  qa = 3.0 * (q_im1 + q_i - 2.0 * dph);
  qb = d2qc_im1;
  double qc = d2qc;
  double qd = 0.0;
  if (SIGN(qa) == SIGN(qb) && SIGN(qa) == SIGN(qc))
    qd = SIGN(qa) * mn(C2 * fabs(qb), mn(C2 * fabs(qc), fabs(qa)));
  double dph_tmp = 0.5 * (q_im1 + q_i) - qd / 6.0;
  if (qa_tmp * qb_tmp < 0.0) dph = dph_tmp;
  qa_tmp = dph_ip1 - q_i;
  qb_tmp = q_ip1 - dph_ip1;
  qa = 3.0 * (q_i + q_ip1 - 2.0 * dph_ip1);
  qb = d2qc;
  qc = d2qc_ip1;
  qd = 0.0;
  if (SIGN(qa) == SIGN(qb) && SIGN(qa) == SIGN(qc))
    qd = SIGN(qa) * mn(C2 * fabs(qb), mn(C2 * fabs(qc), fabs(qa)));
  double dphip1_tmp = 0.5 * (q_i + q_ip1) - qd / 6.0;
  if (qa_tmp * qb_tmp < 0.0) dph_ip1 = dphip1_tmp;
  const double d2qf = 6.0 * (dph + dph_ip1 - 2.0 * q_i);
  double qr = dph;
  double ql = dph_ip1;
  const double dqf_minus = q_i - qr;
  const double dqf_plus = ql - q_i;
  qa_tmp = dqf_minus * dqf_plus;
  qb_tmp = (q_ip1 - q_i) * (q_i - q_im1);
  qa = d2qc_im1;
  qb = d2qc;
  qc = d2qc_ip1;
  qd = d2qf;
  double qe = 0.0;
  if (SIGN(qa) == SIGN(qb) && SIGN(qa) == SIGN(qc) && SIGN(qa) == SIGN(qd))
    qe = SIGN(qd) * mn(mn(C2 * fabs(qa), C2 * fabs(qb)), mn(C2 * fabs(qc), fabs(qd)));
  qa = mx(fabs(q_im1), fabs(q_im2));
  qb = mx(mx(fabs(q_i), fabs(q_ip1)), fabs(q_ip2));
  double rho = 0.0;
  if (fabs(qd) > (1.0e-12) * mx(qa, qb)) rho = qe / qd;
  double tmp_m = q_i - rho * dqf_minus;
  double tmp_p = q_i + rho * dqf_plus;
  double tmp2_m = q_i - 2.0 * dqf_plus;
  double tmp2_p = q_i + 2.0 * dqf_minus;
  if ((qa_tmp <= 0.0 || qb_tmp <= 0.0)) {
    if (rho <= (1.0 - (1.0e-12))) {
      qr = tmp_m;
      ql = tmp_p;
    }
  } else {
    if (fabs(dqf_minus) >= 2.0 * fabs(dqf_plus)) qr = tmp2_m;
    if (fabs(dqf_plus) >= 2.0 * fabs(dqf_minus)) ql = tmp2_p;
  }
  qr_i = qr;
  ql_ip1 = ql;
}

// x / 6.0 correctly rounded, without a division instruction sequence, for x = +-0 or
// 2^-1019 <= |x| <= 2^1020 (other x: use the true division).  y = RN(1/6);
// q0 = RN(x*y) is within one ulp of x/6; r = 6*q0 - x is exact (one fma) and
// q0 - r*y = x/6 + e*(q0 - x/6) with |e| = |1 - 6y| < 2^-53, i.e. within 2^-52 ulp of x/6
// before its one rounding.  Since x/6 = (x/2)/3 and a normal quotient by 3 lies 0, 1/3 or
// 2/3 of an ulp from a representable number, it is never within 1/6 ulp of a rounding
// midpoint, so that rounding gives RN(x/6).  Signed zeros come out right.
APH_HD double div6_fast(double x) {
  constexpr double y = 1.0 / 6.0;
  const double q0 = x * y;
  const double r = fma(q0, 6.0, -x);
  return fma(-r, y, q0);
}

// Helpers that keep ppm_simple.hpp's results bitwise:  SIGN(a) == SIGN(b) is the same
// test as (a < 0) == (b < 0), SIGN(a) * m is (a < 0 ? -m : m) (the product by +-1 is
// exact), and for the non-negative operands of the limiter min/max(a, b) is fmin/fmax(a, b)
// (they differ only when an operand is NaN, so the results agree for finite data).
APH_HD bool neg(double x) { return x < 0.0; }
APH_HD double with_sign_of(double s, double m) { return neg(s) ? -m : m; }

// The second difference of cell i, q_{i-1} + q_{i+1} - 2 q_i, in the operand order of
// every d2qc* in ppm_simple.hpp:55-60 (so one value serves each face and cell that uses it).
APH_HD double ppm_d2(double q_im1, double q_i, double q_ip1) {
  return q_im1 + q_ip1 - 2.0 * q_i;
}

//
// FAST = true evaluates the two divisions by 6 with div6_fast.  It is exact when every
// q is +-0 or 2^-900 <= |q| < 2^1000 (a "plain" window): then every intermediate is a
// multiple of 2^-1008 below 2^1007, so both dividends are 0 or inside div6_fast's range.
// d2qc_im1, d2qc: ppm_d2 of cells i-1 and i.  The limited value (dph_tmp) is formed only
// where the limiter selects it; the reference forms it always and then selects, which
// gives the same bits.
template <bool FAST = false>
APH_HD double ppm_face(double q_im2, double q_im1, double q_i, double q_ip1, double d2qc_im1,
                       double d2qc) {
  const double C2 = 1.25;
  const double qa0 = (q_i - q_im1);
  const double qb0 = (q_ip1 - q_i);
  const double dd_im1 = 0.5 * qa0 + 0.5 * (q_im1 - q_im2);
  const double dd = 0.5 * qb0 + 0.5 * qa0;
  const double avg = 0.5 * (q_im1 + q_i);
  double dph = avg + (FAST ? div6_fast(dd_im1 - dd) : (dd_im1 - dd) / 6.0);
  const double qa_tmp = dph - q_im1;
  const double qb_tmp = q_i - dph;
  if (qa_tmp * qb_tmp < 0.0) {
    const double qa = 3.0 * (q_im1 + q_i - 2.0 * dph);
    double qd = 0.0;
    if (neg(qa) == neg(d2qc_im1) && neg(qa) == neg(d2qc))
      qd = with_sign_of(qa, fmin(C2 * fabs(d2qc_im1), fmin(C2 * fabs(d2qc), fabs(qa))));
    dph = avg - (FAST ? div6_fast(qd) : qd / 6.0);
  }
  return dph;
}

template <bool FAST = false>
APH_HD double ppm_face(double q_im2, double q_im1, double q_i, double q_ip1) {
  return ppm_face<FAST>(q_im2, q_im1, q_i, q_ip1, ppm_d2(q_im2, q_im1, q_i),
                        ppm_d2(q_im1, q_i, q_ip1));
}

// d2qc_im1, d2qc, d2qc_ip1: ppm_d2 of cells i-1, i, i+1.  The extremum branch's quantities
// (qe, the roundoff threshold and the division rho = qe/d2qf) are formed only where that
// branch is taken; elsewhere the reference's rho is computed but never used.
APH_HD void ppm_cell(double q_im2, double q_im1, double q_i, double q_ip1, double q_ip2,
                     double d2qc_im1, double d2qc, double d2qc_ip1, double dph,
                     double dph_ip1, double &ql_ip1, double &qr_i) {
  const double C2 = 1.25;
  const double d2qf = 6.0 * (dph + dph_ip1 - 2.0 * q_i);
  double qr = dph;
  double ql = dph_ip1;
  const double dqf_minus = q_i - qr;
  const double dqf_plus = ql - q_i;
  const double qa_tmp = dqf_minus * dqf_plus;
  const double qb_tmp = (q_ip1 - q_i) * (q_i - q_im1);
  if ((qa_tmp <= 0.0 || qb_tmp <= 0.0)) {
    const double qa = fmax(fabs(q_im1), fabs(q_im2));
    const double qb = fmax(fmax(fabs(q_i), fabs(q_ip1)), fabs(q_ip2));
    if (fabs(d2qf) > (1.0e-12) * fmax(qa, qb)) {
      double qe = 0.0;
      if (neg(d2qc_im1) == neg(d2qc) && neg(d2qc_im1) == neg(d2qc_ip1) &&
          neg(d2qc_im1) == neg(d2qf))
        qe = with_sign_of(d2qf, fmin(fmin(C2 * fabs(d2qc_im1), C2 * fabs(d2qc)),
                                     fmin(C2 * fabs(d2qc_ip1), fabs(d2qf))));
      const double rho = qe / d2qf;
      if (rho <= (1.0 - (1.0e-12))) {
        qr = q_i - rho * dqf_minus;
        ql = q_i + rho * dqf_plus;
      }
    } else {  // rho = 0: the limited values are q_i -/+ 0 * dqf
      qr = q_i - 0.0 * dqf_minus;
      ql = q_i + 0.0 * dqf_plus;
    }
  } else {
    if (fabs(dqf_minus) >= 2.0 * fabs(dqf_plus)) qr = q_i - 2.0 * dqf_plus;
    if (fabs(dqf_plus) >= 2.0 * fabs(dqf_minus)) ql = q_i + 2.0 * dqf_minus;
  }
  qr_i = qr;
  ql_ip1 = ql;
}

APH_HD void ppm_cell(double q_im2, double q_im1, double q_i, double q_ip1, double q_ip2,
                     double dph, double dph_ip1, double &ql_ip1, double &qr_i) {
  ppm_cell(q_im2, q_im1, q_i, q_ip1, q_ip2, ppm_d2(q_im2, q_im1, q_i), ppm_d2(q_im1, q_i, q_ip1),
           ppm_d2(q_i, q_ip1, q_ip2), dph, dph_ip1, ql_ip1, qr_i);
}

// One reconstruction over a window w[0..W-1] (cells m-GR .. m+GR) of field values.
template <int RECON>
APH_HD void recon1(const double *w, double &ql_ip1, double &qr_i) {
  if constexpr (RECON == RECON_PPM) {
    ppm(w[0], w[1], w[2], w[3], w[4], ql_ip1, qr_i);
  } else if constexpr (RECON == RECON_PLM) {
    plm(w[0], w[1], w[2], ql_ip1, qr_i);
  } else {
    ql_ip1 = w[0];
    qr_i = w[0];
  }
}

// ------------------------------------------------------------ sound speed
APH_HD double sound_speed(double gamma, double igm1, double rho,
                                              double p) {
  return C_LIGHT * sqrt(gamma / (1 + (igm1 + rho * C_LIGHT2 / p)));
}

// ------------------------------------------------------------ HLLE
// src/hydro/rsolvers/hydro_hlle.hpp:28-133; w = (rho, vn, vt1, vt2, p)
APH_HD void hlle(const double *wl, const double *wr, const double *ifl,
                                     const double *ifr, double *flx) {
  double gammal = ifl[0], gammar = ifr[0];
  double gm1l = ifl[1] - 1, gm1r = ifr[1] - 1;
  double igm1l = 1.0 / gm1l, igm1r = 1.0 / gm1r;
  double cl = sound_speed(gammal, igm1l, wl[0], wl[4]);
  double cr = sound_speed(gammar, igm1r, wr[0], wr[4]);
  double el = wl[4] * igm1l + 0.5 * wl[0] * (SQR(wl[1]) + SQR(wl[2]) + SQR(wl[3]));
  double er = wr[4] * igm1r + 0.5 * wr[0] * (SQR(wr[1]) + SQR(wr[2]) + SQR(wr[3]));
  double rhoa = .5 * (wl[0] + wr[0]);
  double ca = .5 * (cl + cr);
  double pmid = .5 * (wl[4] + wr[4] + (wl[1] - wr[1]) * rhoa * ca);
  double ql = (pmid <= wl[4])
                  ? 1.0
                  : (1.0 + (gammal + 1) / sqrt(2 * gammal) * (pmid / wl[4] - 1.0));
  double qr = (pmid <= wr[4])
                  ? 1.0
                  : (1.0 + (gammar + 1) / sqrt(2 * gammar) * (pmid / wr[4] - 1.0));
  double al = wl[1] - cl * ql;
  double ar = wr[1] + cr * qr;
  double bp = ar > 0.0 ? ar : 0.0;
  double bm = al < 0.0 ? al : 0.0;
  double vxl = wl[1] - bm;
  double vxr = wr[1] - bp;
  double fl[5], fr[5];
  fl[0] = wl[0] * vxl;
  fr[0] = wr[0] * vxr;
  fl[1] = wl[0] * wl[1] * vxl;
  fr[1] = wr[0] * wr[1] * vxr;
  fl[2] = wl[0] * wl[2] * vxl;
  fr[2] = wr[0] * wr[2] * vxr;
  fl[3] = wl[0] * wl[3] * vxl;
  fr[3] = wr[0] * wr[3] * vxr;
  fl[1] += wl[4];
  fr[1] += wr[4];
  fl[4] = el * vxl + wl[4] * wl[1];
  fr[4] = er * vxr + wr[4] * wr[1];
  double tmp = 0.0;
  if (bp != bm) tmp = 0.5 * (bp + bm) / (bp - bm);
#pragma unroll
  for (int n = 0; n < 5; ++n) flx[n] = 0.5 * (fl[n] + fr[n]) + (fl[n] - fr[n]) * tmp;
}

// ------------------------------------------------------------ HLLC / LHLLC
// src/hydro/rsolvers/hydro_hllc.hpp:161-282; with LOWMACH, hydro_lhllc.hpp:12-158, whose
// only change is the factor phi = chi (2 - chi), chi = max(|v_L|, |v_R|) (a velocity
// magnitude, not a Mach number, and not clipped), on the PVRS term of p*.
template <bool LOWMACH = false>
APH_HD void hllc(const double *wl, const double *wr, const double *ifl,
                                     const double *ifr, double *flx) {
  double gammal = ifl[0], gammar = ifr[0];
  double gm1l = ifl[1] - 1, gm1r = ifr[1] - 1;
  double igm1l = 1.0 / gm1l, igm1r = 1.0 / gm1r;
  double cl = sound_speed(gammal, igm1l, wl[0], wl[4]);
  double cr = sound_speed(gammar, igm1r, wr[0], wr[4]);
  double el = wl[4] * igm1l + 0.5 * wl[0] * (SQR(wl[1]) + SQR(wl[2]) + SQR(wl[3]));
  double er = wr[4] * igm1r + 0.5 * wr[0] * (SQR(wr[1]) + SQR(wr[2]) + SQR(wr[3]));
  double rhoa = .5 * (wl[0] + wr[0]);
  double ca = .5 * (cl + cr);
  double pvrs = .5 * (wr[1] - wl[1]) * rhoa * ca;
  if constexpr (LOWMACH) {
    const double machl = sqrt(wl[1] * wl[1] + wl[2] * wl[2] + wl[3] * wl[3]);
    const double machr = sqrt(wr[1] * wr[1] + wr[2] * wr[2] + wr[3] * wr[3]);
    const double chi = mx(machl, machr);
    pvrs = pvrs * (chi * (2.0 - chi));
  }
  double pstar = .5 * (wl[4] + wr[4]) - pvrs;
  // (gammal + 1) / (2 * gammar) is the same quotient on both sides: evaluated once
  double ql = 1.0, qr = 1.0;
  const bool shl = !(pstar <= wl[4]), shr = !(pstar <= wr[4]);
  if (shl || shr) {
    const double gfac = (gammal + 1) / (2 * gammar);
    if (shl) ql = sqrt(1.0 + gfac * (pstar / wl[4] - 1.0));
    if (shr) qr = sqrt(1.0 + gfac * (pstar / wr[4] - 1.0));
  }
  double sl = wl[1] - cl * ql;
  double sr = wr[1] + cr * qr;
  double bp = sr > 0.0 ? sr : (TINY_NUMBER);
  double bm = sl < 0.0 ? sl : -(TINY_NUMBER);
  double vxl = wl[1] - sl;
  double vxr = wr[1] - sr;
  double tl = wl[4] + vxl * wl[0] * wl[1];
  double tr = wr[4] + vxr * wr[0] * wr[1];
  double ml = wl[0] * vxl;
  double mr = -(wr[0] * vxr);
  double sstar = (tl - tr) / (ml + mr);
  double cp = (ml * tr + mr * tl) / (ml + mr);
  cp = cp > 0.0 ? cp : 0.0;
  vxl = wl[1] - bm;
  vxr = wr[1] - bp;
  double fl[5], fr[5];
  fl[0] = wl[0] * vxl;
  fr[0] = wr[0] * vxr;
  fl[1] = wl[0] * wl[1] * vxl + wl[4];
  fr[1] = wr[0] * wr[1] * vxr + wr[4];
  fl[2] = wl[0] * wl[2] * vxl;
  fr[2] = wr[0] * wr[2] * vxr;
  fl[3] = wl[0] * wl[3] * vxl;
  fr[3] = wr[0] * wr[3] * vxr;
  fl[4] = el * vxl + wl[4] * wl[1];
  fr[4] = er * vxr + wr[4] * wr[1];
  // the reference's two branches divide by one denominator each; selecting the operands
  // first gives every lane the same two quotients with two divisions instead of four
  const bool pos = (sstar >= 0.0);
  const double den = pos ? (sstar - bm) : (bp - sstar);
  const double uw = (pos ? sstar : -sstar) / den;
  const double ustar = (pos ? -bm : bp) / den;
  const double ul = pos ? uw : 0.0;
  const double ur = pos ? 0.0 : uw;
  flx[0] = ul * fl[0] + ur * fr[0];
  flx[1] = ul * fl[1] + ur * fr[1] + ustar * cp;
  flx[2] = ul * fl[2] + ur * fr[2];
  flx[3] = ul * fl[3] + ur * fr[3];
  flx[4] = ul * fl[4] + ur * fr[4] + ustar * cp * sstar;
}

template <int RS>
APH_HD void riemann(const double *wl, const double *wr, const double *ifl,
                                        const double *ifr, double *flx) {
  if constexpr (RS == RS_HLLC) {
    hllc<false>(wl, wr, ifl, ifr, flx);
  } else if constexpr (RS == RS_LHLLC) {
    hllc<true>(wl, wr, ifl, ifr, flx);
  } else if constexpr (RS == RS_HLLE) {
    hlle(wl, wr, ifl, ifr, flx);
  } else {  // none (hydro_none.hpp:17-35): the hydro fluxes are zero
#pragma unroll
    for (int n = 0; n < 5; ++n) flx[n] = 0.0;
  }
}

// ------------------------------------------------------------ ConsToPrim point
// src/hydro/hydro.cpp:474-531 with singularity IdealGas closed forms (upstream).
// u[NV] in (floored in place), w[NV] out, gm[2] out.
template <int NV>
APH_HD void cons_to_prim_pt(double *u, const aph_eos &eos, double *w,
                                                double *gm) {
  const double gm1e = eos.gm1;
  double u_d = u[0];
  u_d = (u_d > eos.dfloor) ? u_d : eos.dfloor;
  u[0] = u_d;
  w[0] = u_d;
  double di = 1.0 / u_d;
  w[1] = u[1] * di;
  w[2] = u[2] * di;
  w[3] = u[3] * di;
  double e_k = 0.5 * di * (SQR(u[1]) + SQR(u[2]) + SQR(u[3]));
  double u_e = u[4];
  double sie = (u_e - e_k) / u_d;
  double w_p = mx(EOS_SMALL, gm1e * u_d * sie);
  if (!(w_p > eos.pfloor)) {  // pressure floor, energy fix (:517-521)
    double gm1 = w_p / (u_d * e_k);
    u_e = ((eos.pfloor / gm1) + e_k);
    w_p = eos.pfloor;
    sie = (u_e - e_k) / u_d;
  }
  u[4] = u_e;
  w[4] = w_p;
  double bulk = mx(EOS_SMALL, gm1e * (gm1e + 1.0) * u_d * sie);
  gm[0] = bulk / w_p;
  gm[1] = w_p / (u_e - e_k) + 1;
#pragma unroll
  for (int n = 5; n < NV; ++n) w[n] = u[n] * di;
}

// The in-place part of cons_to_prim_pt only: the density floor and the pressure-floor
// energy fix, decided by the same arithmetic (so the same branch), without the
// primitives and adiabatic indices.
template <int NV>
APH_HD void floor_cons_pt(double *u, const aph_eos &eos) {
  const double gm1e = eos.gm1;
  double u_d = u[0];
  u_d = (u_d > eos.dfloor) ? u_d : eos.dfloor;
  u[0] = u_d;
  double di = 1.0 / u_d;
  double e_k = 0.5 * di * (SQR(u[1]) + SQR(u[2]) + SQR(u[3]));
  double sie = (u[4] - e_k) / u_d;
  double w_p = mx(EOS_SMALL, gm1e * u_d * sie);
  if (!(w_p > eos.pfloor)) {
    double gm1 = w_p / (u_d * e_k);
    u[4] = ((eos.pfloor / gm1) + e_k);
  }
}

// EstimateTimestep point (hydro.cpp:384-415): returns min over active axes of
// dx_d / (|v_d| + lambda).
APH_HD double dt_point(const double *w, double cons_en, const aph_eos &eos,
                                           const double *dx, int ndim) {
  const double gm1e = eos.gm1;
  double e_internal = (cons_en - 0.5 * w[0] * (SQR(w[1]) + SQR(w[2]) + SQR(w[3]))) / w[0];
  double bulkmod = mx(EOS_SMALL, gm1e * (gm1e + 1.0) * w[0] * e_internal);
  const double gamma_c = bulkmod / w[4];
  double lam = C_LIGHT * sqrt(gamma_c / (1 + (e_internal + C_LIGHT2) * w[0] / w[4]));
  double m = fmin(1.79769313486231570815e308, dx[0] / (fabs(w[1]) + lam));
  if (ndim > 1) m = fmin(m, dx[1] / (fabs(w[2]) + lam));
  if (ndim > 2) m = fmin(m, dx[2] / (fabs(w[3]) + lam));
  return m;
}

}  // namespace aph
