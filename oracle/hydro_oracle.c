/*
 * hydro_oracle.c -- CPU restatement (C99 + OpenMP) of the AlexHls/apophis hydro
 * flux path.  TEST INFRASTRUCTURE: this is the oracle the HIP library is checked
 * against, and the CPU baseline bench.py times.  Product code never links it.
 *
 * Build: gcc -O3 -ffp-contract=off -fopenmp -fPIC -shared (see oracle/Makefile).
 * Every expression keeps the reference's operand order so that, with contraction
 * off, results are bitwise reproducible against the GPU path.
 *
 * What pins each part is listed in tests/test_oracle_*.py; functions restated from
 * upstream Parthenon / singularity-eos (not vendored in the reference) are marked
 * "upstream".
 */
#include "hydro_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Parthenon macros used by the reference (upstream basic_types / defs). */
#define SIGN(x) (((x) < 0.0) ? -1.0 : 1.0)
#define SQR(x) ((x) * (x))
#define TINY_NUMBER 1.0e-20
/* hydro_hlle.hpp:26, hydro_hllc.hpp:159, hydro.cpp:353 */
#define C_LIGHT 2.99792458e10
/* singularity-eos robust::SMALL() = 10 * DBL_MIN (upstream) */
#define EOS_SMALL (10.0 * DBL_MIN)

/* std::min / std::max semantics ((b < a) ? b : a) and ((a < b) ? b : a). */
static inline double mn(double a, double b) { return (b < a) ? b : a; }
static inline double mx(double a, double b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------ layout */
static inline int ext(const aph_block *b, int d, int ndim) {
  return d < ndim ? b->nx[d] + 2 * b->ng : b->nx[d];
}
static inline int gst(const aph_block *b, int d, int ndim) { return d < ndim ? b->ng : 0; }
static inline long ncell(const aph_block *b, int ndim) {
  return (long)ext(b, 0, ndim) * ext(b, 1, ndim) * ext(b, 2, ndim);
}
/* pointer to component v of a field with nv components in block b */
static inline double *fld(double *f, const aph_block *b, int ndim, int nv, int v) {
  return f + (long)nv * b->cell_base + (long)v * ncell(b, ndim);
}
static inline const double *cfld(const double *f, const aph_block *b, int ndim, int nv,
                                 int v) {
  return f + (long)nv * b->cell_base + (long)v * ncell(b, ndim);
}

void orc_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* ------------------------------------------------------------ PLM (a1)
 * src/recon/plm_simple.hpp:14-31 -- van Leer harmonic-mean slope. */
void orc_plm(double q_im1, double q_i, double q_ip1, double *ql_ip1, double *qr_i) {
  double dql = (q_i - q_im1);
  double dqr = (q_ip1 - q_i);
  double dq2 = dql * dqr;
  double dqm = 0.0;
  if (dq2 > 0.0) {
    dqm = dq2 / (dql + dqr);
  }
  *ql_ip1 = q_i + dqm;
  *qr_i = q_i - dqm;
}

/* ------------------------------------------------------------ PPM (a3)
 * src/recon/ppm_simple.hpp:31-161 (CW84 interface values, CS08/MC11 limiter). */
void orc_ppm(double q_im2, double q_im1, double q_i, double q_ip1, double q_ip2,
             double *ql_ip1, double *qr_i) {
  const double C2 = 1.25; /* :36 */
  /* step 1 (:40-49) */
  double qa = (q_i - q_im1);
  double qb = (q_ip1 - q_i);
  const double dd_im1 = 0.5 * qa + 0.5 * (q_im1 - q_im2);
  const double dd = 0.5 * qb + 0.5 * qa;
  const double dd_ip1 = 0.5 * (q_ip2 - q_ip1) + 0.5 * qb;
  double dph = 0.5 * (q_im1 + q_i) + (dd_im1 - dd) / 6.0;
  double dph_ip1 = 0.5 * (q_i + q_ip1) + (dd - dd_ip1) / 6.0;
  /* step 2a (:56-93) */
  const double d2qc_im1 = q_im2 + q_i - 2.0 * q_im1;
  const double d2qc = q_im1 + q_ip1 - 2.0 * q_i;
  const double d2qc_ip1 = q_i + q_ip2 - 2.0 * q_ip1;
  /* i-1/2 */
  double qa_tmp = dph - q_im1;
  double qb_tmp = q_i - dph;
  qa = 3.0 * (q_im1 + q_i - 2.0 * dph);
  qb = d2qc_im1;
  double qc = d2qc;
  double qd = 0.0;
  if (SIGN(qa) == SIGN(qb) && SIGN(qa) == SIGN(qc)) {
    qd = SIGN(qa) * mn(C2 * fabs(qb), mn(C2 * fabs(qc), fabs(qa)));
  }
  double dph_tmp = 0.5 * (q_im1 + q_i) - qd / 6.0;
  if (qa_tmp * qb_tmp < 0.0) {
    dph = dph_tmp;
  }
  /* i+1/2 */
  qa_tmp = dph_ip1 - q_i;
  qb_tmp = q_ip1 - dph_ip1;
  qa = 3.0 * (q_i + q_ip1 - 2.0 * dph_ip1);
  qb = d2qc;
  qc = d2qc_ip1;
  qd = 0.0;
  if (SIGN(qa) == SIGN(qb) && SIGN(qa) == SIGN(qc)) {
    qd = SIGN(qa) * mn(C2 * fabs(qb), mn(C2 * fabs(qc), fabs(qa)));
  }
  double dphip1_tmp = 0.5 * (q_i + q_ip1) - qd / 6.0;
  if (qa_tmp * qb_tmp < 0.0) {
    dph_ip1 = dphip1_tmp;
  }
  const double d2qf = 6.0 * (dph + dph_ip1 - 2.0 * q_i);
  double qr = dph;
  double ql = dph_ip1;
  /* step 3 (:105-108) */
  const double dqf_minus = q_i - qr;
  const double dqf_plus = ql - q_i;
  /* step 4 (:113-160) */
  qa_tmp = dqf_minus * dqf_plus;
  qb_tmp = (q_ip1 - q_i) * (q_i - q_im1);
  qa = d2qc_im1;
  qb = d2qc;
  qc = d2qc_ip1;
  qd = d2qf;
  double qe = 0.0;
  if (SIGN(qa) == SIGN(qb) && SIGN(qa) == SIGN(qc) && SIGN(qa) == SIGN(qd)) {
    qe = SIGN(qd) * mn(mn(C2 * fabs(qa), C2 * fabs(qb)), mn(C2 * fabs(qc), fabs(qd)));
  }
  qa = mx(fabs(q_im1), fabs(q_im2));
  qb = mx(mx(fabs(q_i), fabs(q_ip1)), fabs(q_ip2));
  double rho = 0.0;
  if (fabs(qd) > (1.0e-12) * mx(qa, qb)) {
    rho = qe / qd;
  }
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
    if (fabs(dqf_minus) >= 2.0 * fabs(dqf_plus)) {
      qr = tmp2_m;
    }
    if (fabs(dqf_plus) >= 2.0 * fabs(dqf_minus)) {
      ql = tmp2_p;
    }
  }
  *qr_i = qr;
  *ql_ip1 = ql;
}

/* ------------------------------------------------------------ HLLE (a4)
 * src/hydro/rsolvers/hydro_hlle.hpp:28-133 (dead rhol/rhor/umid omitted). */
void orc_hlle(const double *wl, const double *wr, const double *ifl, const double *ifr,
              double *flx) {
  double gammal = ifl[0], gammar = ifr[0];
  double gm1l = ifl[1] - 1, gm1r = ifr[1] - 1;
  double igm1l = 1.0 / gm1l, igm1r = 1.0 / gm1r;
  double cl = C_LIGHT * sqrt(gammal / (1 + (igm1l + wl[0] * SQR(C_LIGHT) / wl[4])));
  double cr = C_LIGHT * sqrt(gammar / (1 + (igm1r + wr[0] * SQR(C_LIGHT) / wr[4])));
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
  for (int n = 0; n < 5; ++n) flx[n] = 0.5 * (fl[n] + fr[n]) + (fl[n] - fr[n]) * tmp;
}

/* ------------------------------------------------------------ HLLC (a5)
 * src/hydro/rsolvers/hydro_hllc.hpp:161-282. */
void orc_hllc(const double *wl, const double *wr, const double *ifl, const double *ifr,
              double *flx) {
  double gammal = ifl[0], gammar = ifr[0];
  double gm1l = ifl[1] - 1, gm1r = ifr[1] - 1;
  double igm1l = 1.0 / gm1l, igm1r = 1.0 / gm1r;
  double cl = C_LIGHT * sqrt(gammal / (1 + (igm1l + wl[0] * SQR(C_LIGHT) / wl[4])));
  double cr = C_LIGHT * sqrt(gammar / (1 + (igm1r + wr[0] * SQR(C_LIGHT) / wr[4])));
  double el = wl[4] * igm1l + 0.5 * wl[0] * (SQR(wl[1]) + SQR(wl[2]) + SQR(wl[3]));
  double er = wr[4] * igm1r + 0.5 * wr[0] * (SQR(wr[1]) + SQR(wr[2]) + SQR(wr[3]));
  double rhoa = .5 * (wl[0] + wr[0]);
  double ca = .5 * (cl + cr);
  double pstar = .5 * (wl[4] + wr[4]) - .5 * (wr[1] - wl[1]) * rhoa * ca;
  double ql = (pstar <= wl[4])
                  ? 1.0
                  : sqrt(1.0 + (gammal + 1) / (2 * gammar) * (pstar / wl[4] - 1.0));
  double qr = (pstar <= wr[4])
                  ? 1.0
                  : sqrt(1.0 + (gammal + 1) / (2 * gammar) * (pstar / wr[4] - 1.0));
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
  double ur, ul, ustar;
  if (sstar >= 0.0) {
    ul = sstar / (sstar - bm);
    ur = 0.0;
    ustar = -bm / (sstar - bm);
  } else {
    ul = 0.0;
    ur = -sstar / (bp - sstar);
    ustar = bp / (bp - sstar);
  }
  flx[0] = ul * fl[0] + ur * fr[0];
  flx[1] = ul * fl[1] + ur * fr[1] + ustar * cp;
  flx[2] = ul * fl[2] + ur * fr[2];
  flx[3] = ul * fl[3] + ur * fr[3];
  flx[4] = ul * fl[4] + ur * fr[4] + ustar * cp * sstar;
}

/* ----------------------------------------------------------- LHLLC (a5, f)
 * src/hydro/rsolvers/hydro_lhllc.hpp:12-158: HLLC with a low-Mach factor on the PVRS
 * middle pressure.  "mach" there is the velocity magnitude |v| (not |v|/c) and
 * phi = chi (2 - chi) is not clipped: both quirks are kept. */
void orc_lhllc(const double *wl, const double *wr, const double *ifl, const double *ifr,
               double *flx) {
  double gammal = ifl[0], gammar = ifr[0];
  double gm1l = ifl[1] - 1, gm1r = ifr[1] - 1;
  double igm1l = 1.0 / gm1l, igm1r = 1.0 / gm1r;
  double cl = C_LIGHT * sqrt(gammal / (1 + (igm1l + wl[0] * SQR(C_LIGHT) / wl[4])));
  double cr = C_LIGHT * sqrt(gammar / (1 + (igm1r + wr[0] * SQR(C_LIGHT) / wr[4])));
  double el = wl[4] * igm1l + 0.5 * wl[0] * (SQR(wl[1]) + SQR(wl[2]) + SQR(wl[3]));
  double er = wr[4] * igm1r + 0.5 * wr[0] * (SQR(wr[1]) + SQR(wr[2]) + SQR(wr[3]));
  double rhoa = .5 * (wl[0] + wr[0]);
  double ca = .5 * (cl + cr);
  double machl = sqrt(wl[1] * wl[1] + wl[2] * wl[2] + wl[3] * wl[3]);
  double machr = sqrt(wr[1] * wr[1] + wr[2] * wr[2] + wr[3] * wr[3]);
  double chi = mx(machl, machr);
  double phi = chi * (2.0 - chi);
  double pstar = .5 * (wl[4] + wr[4]) - .5 * (wr[1] - wl[1]) * rhoa * ca * phi;
  double ql = (pstar <= wl[4])
                  ? 1.0
                  : sqrt(1.0 + (gammal + 1) / (2 * gammar) * (pstar / wl[4] - 1.0));
  double qr = (pstar <= wr[4])
                  ? 1.0
                  : sqrt(1.0 + (gammal + 1) / (2 * gammar) * (pstar / wr[4] - 1.0));
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
  double ur, ul, ustar;
  if (sstar >= 0.0) {
    ul = sstar / (sstar - bm);
    ur = 0.0;
    ustar = -bm / (sstar - bm);
  } else {
    ul = 0.0;
    ur = -sstar / (bp - sstar);
    ustar = bp / (bp - sstar);
  }
  flx[0] = ul * fl[0] + ur * fr[0];
  flx[1] = ul * fl[1] + ur * fr[1] + ustar * cp;
  flx[2] = ul * fl[2] + ur * fr[2];
  flx[3] = ul * fl[3] + ur * fr[3];
  flx[4] = ul * fl[4] + ur * fr[4] + ustar * cp * sstar;
}

/* ------------------------------------------------------- ConsToPrim (a9)
 * src/hydro/hydro.cpp:450-532 with the singularity-eos IdealGas closed forms
 * (upstream): P = max(SMALL, gm1*rho*sie), K = max(SMALL, gm1*(gm1+1)*rho*sie). */
void orc_cons_to_prim(const aph_block *blocks, int nb, int ndim, int nvar,
                      const aph_eos *eos, double *cons, double *prim, double *gamma) {
  const double gm1e = eos->gm1, dfloor = eos->dfloor, pfloor = eos->pfloor;
#pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < nb; ++b) {
    const aph_block *bk = &blocks[b];
    const long nc = ncell(bk, ndim);
    double *u = fld(cons, bk, ndim, nvar, 0);
    double *w = fld(prim, bk, ndim, nvar, 0);
    double *gm = fld(gamma, bk, ndim, 2, 0);
    for (long c = 0; c < nc; ++c) {
      double u_d = u[c], u_m1 = u[nc + c], u_m2 = u[2 * nc + c], u_m3 = u[3 * nc + c];
      double u_e = u[4 * nc + c];
      u_d = (u_d > dfloor) ? u_d : dfloor;
      u[c] = u_d;
      w[c] = u_d;
      double di = 1.0 / u_d;
      w[nc + c] = u_m1 * di;
      w[2 * nc + c] = u_m2 * di;
      w[3 * nc + c] = u_m3 * di;
      double e_k = 0.5 * di * (SQR(u_m1) + SQR(u_m2) + SQR(u_m3));
      double w_p = mx(EOS_SMALL, gm1e * u_d * ((u_e - e_k) / u_d));
      double gm1 = w_p / (u_d * e_k);
      u_e = (w_p > pfloor) ? u_e : ((pfloor / gm1) + e_k);
      w_p = (w_p > pfloor) ? w_p : pfloor;
      u[4 * nc + c] = u_e;
      w[4 * nc + c] = w_p;
      double bulk = mx(EOS_SMALL, gm1e * (gm1e + 1.0) * u_d * ((u_e - e_k) / u_d));
      gm[c] = bulk / w_p;
      gm[nc + c] = w_p / (u_e - e_k) + 1;
      for (int n = 5; n < nvar; ++n) w[n * nc + c] = u[n * nc + c] * di;
    }
  }
}

/* ------------------------------------------------- CalculateFluxes (a6)
 * src/hydro/hydro.cpp:535-744.  For every pencil the reference visits, the cells
 * [s-1, e+1] along the sweep axis are reconstructed (ql at face c+1, qr at face c,
 * plm_simple.hpp:41-47, ppm_simple.hpp:223-234) and the Riemann problem is solved
 * at faces [s, e+1]; passive scalars are upwinded by the mass flux (:606-615). */
static void recon_line(int recon, const double *q, long st, double *ql_ip1,
                       double *qr_i) {
  if (recon == APH_RECON_PLM) {
    orc_plm(q[-st], q[0], q[st], ql_ip1, qr_i);
  } else if (recon == APH_RECON_PPM) {
    orc_ppm(q[-2 * st], q[-st], q[0], q[st], q[2 * st], ql_ip1, qr_i);
  } else { /* dc_simple.hpp:14-31 */
    *ql_ip1 = q[0];
    *qr_i = q[0];
  }
}

static void flux_block_dir(const aph_block *bk, int ndim, int nvar, int recon,
                           int rsolver, int dir, const double *prim, const double *gamma,
                           double *flux, double *scr) {
  const int Nx = ext(bk, 0, ndim), Ny = ext(bk, 1, ndim);
  const long nc = ncell(bk, ndim);
  const int g[3] = {gst(bk, 0, ndim), gst(bk, 1, ndim), gst(bk, 2, ndim)};
  const long st[3] = {1, Nx, (long)Nx * Ny};
  const double *w = cfld(prim, bk, ndim, nvar, 0);
  const double *gm = cfld(gamma, bk, ndim, 2, 0);
  double *f = fld(flux, bk, ndim, nvar, 0);
  /* loop limits, hydro.cpp:541-549, 623-627, 688 */
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = g[d];
    hi[d] = g[d] + bk->nx[d] - 1;
  }
  int tlo[3], thi[3];
  for (int d = 0; d < 3; ++d) {
    tlo[d] = lo[d];
    thi[d] = hi[d];
  }
  if (dir == 0) {
    if (ndim >= 2) { tlo[1] = lo[1] - 1; thi[1] = hi[1] + 1; }
    if (ndim >= 3) { tlo[2] = lo[2] - 1; thi[2] = hi[2] + 1; }
  } else if (dir == 1) {
    tlo[0] = lo[0] - 1; thi[0] = hi[0] + 1;
    if (ndim >= 3) { tlo[2] = lo[2] - 1; thi[2] = hi[2] + 1; }
  } else {
    tlo[0] = lo[0] - 1; thi[0] = hi[0] + 1;
    tlo[1] = lo[1] - 1; thi[1] = hi[1] + 1;
  }
  const int t1 = (dir == 0) ? 1 : 0; /* first transverse axis */
  const int t2 = (dir == 2) ? 1 : 2; /* second transverse axis */
  const int nf = nvar + 2;
  const int Nd = ext(bk, dir, ndim);
  double *ql = scr;                 /* [nf][Nd+1] */
  double *qr = scr + (long)nf * (Nd + 1);
  const int ivx = 1 + dir, ivy = 1 + (dir + 1) % 3, ivz = 1 + (dir + 2) % 3;
  for (int b2 = tlo[t2]; b2 <= thi[t2]; ++b2) {
    for (int b1 = tlo[t1]; b1 <= thi[t1]; ++b1) {
      long base = (long)b1 * st[t1] + (long)b2 * st[t2];
      for (int c = lo[dir] - 1; c <= hi[dir] + 1; ++c) {
        long off = base + (long)c * st[dir];
        for (int n = 0; n < nvar; ++n)
          recon_line(recon, w + (long)n * nc + off, st[dir], &ql[(long)n * (Nd + 1) + c + 1],
                     &qr[(long)n * (Nd + 1) + c]);
        for (int n = 0; n < 2; ++n)
          recon_line(recon, gm + (long)n * nc + off, st[dir],
                     &ql[(long)(nvar + n) * (Nd + 1) + c + 1],
                     &qr[(long)(nvar + n) * (Nd + 1) + c]);
      }
      for (int c = lo[dir]; c <= hi[dir] + 1; ++c) {
        double wl[5], wr[5], ifl[2], ifr[2], fx[5];
        const int comp[5] = {0, ivx, ivy, ivz, 4};
        for (int n = 0; n < 5; ++n) {
          wl[n] = ql[(long)comp[n] * (Nd + 1) + c];
          wr[n] = qr[(long)comp[n] * (Nd + 1) + c];
        }
        for (int n = 0; n < 2; ++n) {
          ifl[n] = ql[(long)(nvar + n) * (Nd + 1) + c];
          ifr[n] = qr[(long)(nvar + n) * (Nd + 1) + c];
        }
        if (rsolver == APH_RIEMANN_HLLC)
          orc_hllc(wl, wr, ifl, ifr, fx);
        else if (rsolver == APH_RIEMANN_LHLLC)
          orc_lhllc(wl, wr, ifl, ifr, fx);
        else if (rsolver == APH_RIEMANN_HLLE)
          orc_hlle(wl, wr, ifl, ifr, fx);
        else /* none: hydro_none.hpp:17-35 zeroes the hydro fluxes */
          for (int n = 0; n < 5; ++n) fx[n] = 0.0;
        long off = base + (long)c * st[dir];
        f[off] = fx[0];
        f[(long)ivx * nc + off] = fx[1];
        f[(long)ivy * nc + off] = fx[2];
        f[(long)ivz * nc + off] = fx[3];
        f[4 * nc + off] = fx[4];
        for (int n = 5; n < nvar; ++n) {
          if (f[off] >= 0.0)
            f[(long)n * nc + off] = f[off] * ql[(long)n * (Nd + 1) + c];
          else
            f[(long)n * nc + off] = f[off] * qr[(long)n * (Nd + 1) + c];
        }
      }
    }
  }
}

int orc_calculate_fluxes(const aph_block *blocks, int nb, int ndim, int nvar, int recon,
                         int rsolver, const double *prim, const double *gamma,
                         double *fx1, double *fx2, double *fx3) {
  if (recon != APH_RECON_DC && recon != APH_RECON_PLM && recon != APH_RECON_PPM) return -1;
  if (rsolver != APH_RIEMANN_HLLE && rsolver != APH_RIEMANN_HLLC &&
      rsolver != APH_RIEMANN_LHLLC && rsolver != APH_RIEMANN_NONE)
    return -1;
  const int need = recon == APH_RECON_PPM ? 3 : (recon == APH_RECON_PLM ? 2 : 1);
  for (int b = 0; b < nb; ++b)
    if (blocks[b].ng < need) return -2; /* hydro.cpp:72-75 */
  int nmax = 0;
  for (int b = 0; b < nb; ++b)
    for (int d = 0; d < 3; ++d) {
      int e = ext(&blocks[b], d, ndim);
      if (e > nmax) nmax = e;
    }
  double *fl[3] = {fx1, fx2, fx3};
#pragma omp parallel
  {
    double *scr = (double *)malloc(sizeof(double) * 2 * (size_t)(nvar + 2) * (nmax + 1));
    for (int dir = 0; dir < ndim; ++dir) {
#pragma omp for schedule(dynamic)
      for (int b = 0; b < nb; ++b)
        flux_block_dir(&blocks[b], ndim, nvar, recon, rsolver, dir, prim, gamma, fl[dir],
                       scr);
    }
    free(scr);
  }
  return 0;
}

/* --------------------------------------------- UpdateWithFluxDivergence (a7)
 * upstream Parthenon FluxDivHelper / UpdateWithFluxDivergence, uniform Cartesian:
 *   du  = A1*F(i+1) - A1*F(i);  du += A2*G(j+1) - A2*G(j);  du += A3*H(k+1) - A3*H(k)
 *   u0  = gam0*u0 + gam1*u1 + beta_dt * (-du / V)          (interior cells) */
void orc_update(const aph_block *blocks, int nb, int ndim, int nvar, double *u0,
                const double *u1, const double *fx1, const double *fx2, const double *fx3,
                double gam0, double gam1, double beta_dt) {
#pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < nb; ++b) {
    const aph_block *bk = &blocks[b];
    const int Nx = ext(bk, 0, ndim), Ny = ext(bk, 1, ndim);
    const long nc = ncell(bk, ndim);
    const long sy = Nx, sz = (long)Nx * Ny;
    const int gx = gst(bk, 0, ndim), gy = gst(bk, 1, ndim), gz = gst(bk, 2, ndim);
    for (int v = 0; v < nvar; ++v) {
      double *a = fld(u0, bk, ndim, nvar, v);
      const double *a1 = cfld(u1, bk, ndim, nvar, v);
      const double *f1 = cfld(fx1, bk, ndim, nvar, v);
      const double *f2 = ndim >= 2 ? cfld(fx2, bk, ndim, nvar, v) : NULL;
      const double *f3 = ndim >= 3 ? cfld(fx3, bk, ndim, nvar, v) : NULL;
      (void)nc;
      for (int k = gz; k < gz + bk->nx[2]; ++k)
        for (int j = gy; j < gy + bk->nx[1]; ++j)
          for (int i = gx; i < gx + bk->nx[0]; ++i) {
            long c = k * sz + j * sy + i;
            double du = (bk->area[0] * f1[c + 1] - bk->area[0] * f1[c]);
            if (ndim >= 2) du += bk->area[1] * f2[c + sy] - bk->area[1] * f2[c];
            if (ndim == 3) du += bk->area[2] * f3[c + sz] - bk->area[2] * f3[c];
            double div = -du / bk->vol;
            a[c] = gam0 * a[c] + gam1 * a1[c] + beta_dt * div;
          }
    }
  }
}

/* ------------------------------------------------------ ApplyGravity (a11)
 * src/gravity/gravity.cpp:386-407 with a uniform acceleration field g. */
void orc_apply_gravity(const aph_block *blocks, int nb, int ndim, int nvar, double *cons,
                       const double *prim, const double *g, double dt) {
#pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < nb; ++b) {
    const aph_block *bk = &blocks[b];
    const int Nx = ext(bk, 0, ndim), Ny = ext(bk, 1, ndim);
    const long nc = ncell(bk, ndim);
    const int gx = gst(bk, 0, ndim), gy = gst(bk, 1, ndim), gz = gst(bk, 2, ndim);
    double *u = fld(cons, bk, ndim, nvar, 0);
    const double *w = cfld(prim, bk, ndim, nvar, 0);
    for (int k = gz; k < gz + bk->nx[2]; ++k)
      for (int j = gy; j < gy + bk->nx[1]; ++j)
        for (int i = gx; i < gx + bk->nx[0]; ++i) {
          long c = ((long)k * Ny + j) * Nx + i;
          u[nc + c] += g[0] * w[c] * dt;
          double de = g[0] * w[nc + c];
          if (ndim >= 2) {
            u[2 * nc + c] += g[1] * w[c] * dt;
            de += g[1] * w[2 * nc + c];
          }
          if (ndim >= 3) {
            u[3 * nc + c] += g[2] * w[c] * dt;
            de += g[2] * w[3 * nc + c];
          }
          u[4 * nc + c] += de * w[c] * dt;
        }
  }
}

/* ----------------------------------------------------- ghost fill (a12)
 * Upstream boundary exchange (copy from the neighbour's interior for every ghost
 * cell whose neighbour exists, faces/edges/corners alike), then the physical
 * boundaries axis by axis in the order inner_x1, outer_x1, inner_x2, ... over the
 * whole ghost slab: outflow copies the last interior layer, reflect mirrors and
 * negates the normal momentum (apophis_boundaries.cpp:77-90). */
void orc_fill_ghosts(const aph_block *blocks, int nb, int ndim, int nvar, double *cons) {
#pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < nb; ++b) {
    const aph_block *bk = &blocks[b];
    const int N[3] = {ext(bk, 0, ndim), ext(bk, 1, ndim), ext(bk, 2, ndim)};
    const int g[3] = {gst(bk, 0, ndim), gst(bk, 1, ndim), gst(bk, 2, ndim)};
    const long nc = ncell(bk, ndim);
    double *u = fld(cons, bk, ndim, nvar, 0);
    /* exchange */
    for (int k = 0; k < N[2]; ++k)
      for (int j = 0; j < N[1]; ++j)
        for (int i = 0; i < N[0]; ++i) {
          const int idx[3] = {i, j, k};
          int o[3];
          for (int d = 0; d < 3; ++d)
            o[d] = idx[d] < g[d] ? -1 : (idx[d] >= g[d] + bk->nx[d] ? 1 : 0);
          if (o[0] == 0 && o[1] == 0 && o[2] == 0) continue;
          int nbi = bk->nbr[(o[2] + 1) * 9 + (o[1] + 1) * 3 + (o[0] + 1)];
          if (nbi < 0) continue;
          const aph_block *nk = &blocks[nbi];
          const long nnc = ncell(nk, ndim);
          const double *s = cfld(cons, nk, ndim, nvar, 0);
          long src = ((long)(k - o[2] * bk->nx[2]) * N[1] + (j - o[1] * bk->nx[1])) * N[0] +
                     (i - o[0] * bk->nx[0]);
          long dst = ((long)k * N[1] + j) * N[0] + i;
          for (int v = 0; v < nvar; ++v) u[v * nc + dst] = s[v * nnc + src];
        }
    /* physical boundaries */
    for (int d = 0; d < ndim; ++d)
      for (int side = 0; side < 2; ++side) {
        const int sd = (d == 0) ? 1 : (d == 1 ? 3 : 9);
        const int face = 13 + (side ? sd : -sd);
        if (bk->nbr[face] >= 0) continue;
        const int type = bk->bc[2 * d + side];
        if (type == APH_BC_AMR) continue; /* filled from the other level */
        const int ref = side ? g[d] + bk->nx[d] - 1 : g[d];
        const int offset = 2 * ref + (side ? 1 : -1);
        const int glo = side ? g[d] + bk->nx[d] : 0;
        const int ghi = side ? N[d] - 1 : g[d] - 1;
        for (int k = 0; k < N[2]; ++k)
          for (int j = 0; j < N[1]; ++j)
            for (int i = 0; i < N[0]; ++i) {
              int idx[3] = {i, j, k};
              if (idx[d] < glo || idx[d] > ghi) continue;
              int s[3] = {i, j, k};
              s[d] = (type == APH_BC_REFLECT) ? offset - idx[d] : ref;
              long dst = ((long)k * N[1] + j) * N[0] + i;
              long src = ((long)s[2] * N[1] + s[1]) * N[0] + s[0];
              for (int v = 0; v < nvar; ++v) {
                if (type == APH_BC_REFLECT)
                  u[v * nc + dst] = ((v == 1 + d) ? -1.0 : 1.0) * u[v * nc + src];
                else
                  u[v * nc + dst] = u[v * nc + src];
              }
            }
      }
  }
}

/* --------------------------------------------------- EstimateTimestep (a10)
 * src/hydro/hydro.cpp:369-419. */
double orc_estimate_timestep(const aph_block *blocks, int nb, int ndim, int nvar,
                             const aph_eos *eos, const double *prim, const double *cons,
                             double cfl) {
  const double gm1e = eos->gm1;
  double min_dt = DBL_MAX;
#pragma omp parallel for schedule(dynamic) reduction(min : min_dt)
  for (int b = 0; b < nb; ++b) {
    const aph_block *bk = &blocks[b];
    const int Nx = ext(bk, 0, ndim), Ny = ext(bk, 1, ndim);
    const long nc = ncell(bk, ndim);
    const int gx = gst(bk, 0, ndim), gy = gst(bk, 1, ndim), gz = gst(bk, 2, ndim);
    const double *w = cfld(prim, bk, ndim, nvar, 0);
    const double *u = cfld(cons, bk, ndim, nvar, 0);
    for (int k = gz; k < gz + bk->nx[2]; ++k)
      for (int j = gy; j < gy + bk->nx[1]; ++j)
        for (int i = gx; i < gx + bk->nx[0]; ++i) {
          long c = ((long)k * Ny + j) * Nx + i;
          double w0 = w[c], w1 = w[nc + c], w2 = w[2 * nc + c], w3 = w[3 * nc + c];
          double w4 = w[4 * nc + c];
          double e_internal = (u[4 * nc + c] - 0.5 * w0 * (SQR(w1) + SQR(w2) + SQR(w3))) / w0;
          double bulkmod = mx(EOS_SMALL, gm1e * (gm1e + 1.0) * w0 * e_internal);
          const double gamma_c = bulkmod / w4;
          double lam = C_LIGHT * sqrt(gamma_c / (1 + (e_internal + SQR(C_LIGHT)) * w0 / w4));
          min_dt = fmin(min_dt, bk->dx[0] / (fabs(w1) + lam));
          if (ndim > 1) min_dt = fmin(min_dt, bk->dx[1] / (fabs(w2) + lam));
          if (ndim > 2) min_dt = fmin(min_dt, bk->dx[2] / (fabs(w3) + lam));
        }
  }
  return cfl * min_dt;
}

/* ---------------------------------------------------- GlobalQuantHst
 * src/hydro/hydro.cpp:746-771 (serial sum, block-major order). */
double orc_global_sum(const aph_block *blocks, int nb, int ndim, int nvar,
                      const double *cons, int var) {
  double s = 0.0;
  for (int b = 0; b < nb; ++b) {
    const aph_block *bk = &blocks[b];
    const int Nx = ext(bk, 0, ndim), Ny = ext(bk, 1, ndim);
    const int gx = gst(bk, 0, ndim), gy = gst(bk, 1, ndim), gz = gst(bk, 2, ndim);
    const double *u = cfld(cons, bk, ndim, nvar, var);
    for (int k = gz; k < gz + bk->nx[2]; ++k)
      for (int j = gy; j < gy + bk->nx[1]; ++j)
        for (int i = gx; i < gx + bk->nx[0]; ++i)
          s += u[((long)k * Ny + j) * Nx + i] * bk->vol;
  }
  return s;
}

/* ------------------------------------------------------- RK step (a8, f)
 * ApophisDriver::MakeTaskCollection (apophis_driver.cpp:27-130) with the upstream
 * Parthenon low-storage coefficients of each integrator (integrator enum = main.hpp):
 *   rk1: gam0 = {0},          gam1 = {1},            beta = {1}
 *   rk2: gam0 = {0, 1/2},     gam1 = {1, 1/2},       beta = {1, 1/2}
 *   rk3: gam0 = {0, 1/4, 2/3}, gam1 = {1, 3/4, 1/3}, beta = {1, 1/4, 2/3}
 *   vl2: gam0 = {0, 0},       gam1 = {1, 1},         beta = {1/2, 1}
 * and the first stage's flux function: DC reconstruction for vl2 (hydro.cpp:114-139). */
int orc_integrator_stages(int integrator, double *gam0, double *gam1, double *beta) {
  switch (integrator) {
    case 1:
      gam0[0] = 0.0; gam1[0] = 1.0; beta[0] = 1.0;
      return 1;
    case 2:
      gam0[0] = 0.0; gam1[0] = 1.0; beta[0] = 1.0;
      gam0[1] = 0.5; gam1[1] = 0.5; beta[1] = 0.5;
      return 2;
    case 3:
      gam0[0] = 0.0; gam1[0] = 1.0; beta[0] = 1.0;
      gam0[1] = 0.25; gam1[1] = 0.75; beta[1] = 0.25;
      gam0[2] = 2.0 / 3.0; gam1[2] = 1.0 / 3.0; beta[2] = 2.0 / 3.0;
      return 3;
    case 4:
      gam0[0] = 0.0; gam1[0] = 1.0; beta[0] = 0.5;
      gam0[1] = 0.0; gam1[1] = 1.0; beta[1] = 1.0;
      return 2;
    default:
      return 0;
  }
}

double orc_step(const aph_block *blocks, int nb, int ndim, int nvar, int integrator,
                int recon, int rsolver, const aph_eos *eos, const double *g, double cfl,
                double dt, double *u0, double *u1, double *prim, double *gamma, double *fx1,
                double *fx2, double *fx3) {
  double gam0[3], gam1[3], beta[3];
  const int nstages = orc_integrator_stages(integrator, gam0, gam1, beta);
  long ntot = 0;
  for (int b = 0; b < nb; ++b) ntot += ncell(&blocks[b], ndim);
  memcpy(u1, u0, sizeof(double) * (size_t)nvar * ntot); /* :44-55 */
  const int grav = g && (g[0] != 0.0 || g[1] != 0.0 || g[2] != 0.0);
  for (int s = 0; s < nstages; ++s) {
    const int rec = (s == 0 && integrator == 4) ? APH_RECON_DC : recon;
    orc_calculate_fluxes(blocks, nb, ndim, nvar, rec, rsolver, prim, gamma, fx1, fx2,
                         fx3);                                           /* :70-72 */
    orc_update(blocks, nb, ndim, nvar, u0, u1, fx1, fx2, fx3, gam0[s], gam1[s],
               beta[s] * dt);                                            /* :79-82 */
    if (grav) orc_apply_gravity(blocks, nb, ndim, nvar, u0, prim, g, beta[s] * dt);
    orc_fill_ghosts(blocks, nb, ndim, nvar, u0);                         /* :94,103 */
    orc_cons_to_prim(blocks, nb, ndim, nvar, eos, u0, prim, gamma);      /* :110-111 */
  }
  return orc_estimate_timestep(blocks, nb, ndim, nvar, eos, prim, u0, cfl); /* :113-116 */
}

double orc_rk2_step(const aph_block *blocks, int nb, int ndim, int nvar, int recon,
                    int rsolver, const aph_eos *eos, const double *g, double cfl,
                    double dt, double *u0, double *u1, double *prim, double *gamma,
                    double *fx1, double *fx2, double *fx3) {
  return orc_step(blocks, nb, ndim, nvar, 2, recon, rsolver, eos, g, cfl, dt, u0, u1, prim,
                  gamma, fx1, fx2, fx3);
}

/* ------------------------------------------------ monopole gravity (f4)
 * src/gravity/gsolvers/monopole_gravity.hpp.  AccumulateMass (:114-153) bins the
 * interior cells' masses rho*V by the cell-centre radius into nrbin shells of width
 * rmax/nrbin; BuildRadialTables (:155-213) forms the enclosed mass m_enc and the radial
 * potential phi_rad; ComputeGravityVectorDirect (:215-302) gives each cell g and phi;
 * ApplyGravity (gravity.cpp:374-410) adds g*rho*dt to the momenta and (g.v)*rho*dt to
 * the energy.
 *
 * The reference sums a bin's masses with floating-point atomics, in no fixed order.
 * Here (and on the GPU) each mass is rounded to an integer multiple of 2^-s and summed
 * as a 128-bit integer, which any order adds up to the same value; s is chosen per
 * call from the largest mass so that no sum can overflow (2^125 bound), which leaves
 * about 100 bits below the largest mass.  A bin's mass is that sum rounded once to a
 * double: for masses that are multiples of 2^-s it is the correctly rounded exact sum
 * (math.fsum), and otherwise within (cells/2)*2^-s of it. */
typedef unsigned __int128 orc_u128;

/* rho*V*2^s rounded to nearest (ties to even) as a two's complement 128-bit integer. */
static orc_u128 mono_quantize(double dm, int s) {
  if (dm == 0.0) return 0;
  uint64_t bits;
  memcpy(&bits, &dm, sizeof bits);
  const int neg = (int)(bits >> 63);
  int ex = (int)((bits >> 52) & 0x7ff);
  uint64_t man = bits & ((1ull << 52) - 1);
  if (ex == 0)
    ex = 1; /* subnormal */
  else
    man |= 1ull << 52;
  const int k = ex - 1075 + s; /* |dm| * 2^s = man * 2^k */
  orc_u128 q;
  if (k >= 0) {
    q = (orc_u128)man << k;
  } else if (-k >= 64) {
    q = 0; /* man < 2^53 <= 2^(-k-1): below one half */
  } else {
    const int sh = -k;
    uint64_t m = man >> sh;
    const uint64_t rem = man & ((1ull << sh) - 1), half = 1ull << (sh - 1);
    if (rem > half || (rem == half && (m & 1))) ++m;
    q = m;
  }
  return neg ? (orc_u128)0 - q : q;
}

/* The 128-bit sum (two's complement) times 2^-s, rounded once to nearest even. */
static double mono_to_double(orc_u128 q, int s) {
  const int neg = (int)(q >> 127);
  if (neg) q = (orc_u128)0 - q;
  if (q == 0) return 0.0;
  int t = 127;
  while (!((q >> t) & 1)) --t;
  uint64_t m;
  int e = 0;
  if (t <= 52) {
    m = (uint64_t)q;
  } else {
    const int sh = t - 52;
    m = (uint64_t)(q >> sh);
    const orc_u128 rem = q & (((orc_u128)1 << sh) - 1), half = (orc_u128)1 << (sh - 1);
    if (rem > half || (rem == half && (m & 1))) ++m;
    if (m == (1ull << 53)) {
      m >>= 1;
      ++e;
    }
    e += sh;
  }
  const double r = ldexp((double)m, e - s);
  return neg ? -r : r;
}

/* Binary exponent bound of a mass: |dm| < 2^(e+1). */
static int mono_exponent(double dm) {
  uint64_t bits;
  memcpy(&bits, &dm, sizeof bits);
  const int ex = (int)((bits >> 52) & 0x7ff);
  return (ex == 0 ? 1 : ex) - 1023;
}

/* The scale: the sum of `ncells` masses below 2^(emax+1) stays below 2^125. */
int orc_mono_scale(int emax, long ncells) {
  if (emax == INT_MIN) return 0;
  int L = 0;
  while ((1L << L) < ncells) ++L;
  return 125 - L - (emax + 1);
}

/* Cell centre of interior cell (i, j, k) of a block, on every axis (the inactive axes'
 * single cell is centred on the domain). */
static void mono_centre(const aph_block *bk, int i, int j, int k, double *x, double *y,
                        double *z) {
  *x = bk->xmin[0] + (i + 0.5) * bk->dx[0];
  *y = bk->xmin[1] + (j + 0.5) * bk->dx[1];
  *z = bk->xmin[2] + (k + 0.5) * bk->dx[2];
}

/* AccumulateMass: bin masses of every interior cell with r < rmax (density from the
 * primitive field, i.e. after ConsToPrim's floor), summed as above.  Returns s. */
int orc_mono_bin_masses(const aph_block *blocks, int nb, int ndim, int nvar,
                        const double *prim, int nrbin, double rmax, double *mbin) {
  const double dr = rmax / nrbin;
  long ncells = 0;
  int emax = INT_MIN;
  for (int pass = 0; pass < 2; ++pass) {
    int s = 0;
    orc_u128 *acc = NULL;
    if (pass == 1) {
      s = orc_mono_scale(emax, ncells);
      acc = (orc_u128 *)calloc((size_t)nrbin, sizeof(orc_u128));
    }
    for (int b = 0; b < nb; ++b) {
      const aph_block *bk = &blocks[b];
      const int Nx = ext(bk, 0, ndim), Ny = ext(bk, 1, ndim);
      const int gx = gst(bk, 0, ndim), gy = gst(bk, 1, ndim), gz = gst(bk, 2, ndim);
      const double *w = cfld(prim, bk, ndim, nvar, 0);
      if (pass == 0) ncells += (long)bk->nx[0] * bk->nx[1] * bk->nx[2];
      for (int k = 0; k < bk->nx[2]; ++k)
        for (int j = 0; j < bk->nx[1]; ++j)
          for (int i = 0; i < bk->nx[0]; ++i) {
            double x, y, z;
            mono_centre(bk, i, j, k, &x, &y, &z);
            const double r = sqrt(x * x + y * y + z * z);
            if (r >= rmax) continue;
            const long c = ((long)(k + gz) * Ny + (j + gy)) * Nx + (i + gx);
            const double dm = w[c] * bk->vol;
            if (pass == 0) {
              if (dm != 0.0 && mono_exponent(dm) > emax) emax = mono_exponent(dm);
              continue;
            }
            int bin = (int)(r / dr);
            if (bin < 0) bin = 0;
            if (bin >= nrbin) bin = nrbin - 1;
            acc[bin] += mono_quantize(dm, s);
          }
    }
    if (pass == 1) {
      for (int i = 0; i < nrbin; ++i) mbin[i] = mono_to_double(acc[i], s);
      free(acc);
      return s;
    }
  }
  return 0;
}

/* BuildRadialTables (monopole_gravity.hpp:155-213): m_enc[0..nrbin], phi_rad[0..nrbin]
 * from the bin masses, both recurrences in the reference's order. */
void orc_mono_tables(int nrbin, int ndim, double rmax, double G, const double *mbin,
                     double *m_enc, double *phi_rad) {
  const double dr = rmax / nrbin;
  m_enc[0] = 0.0;
  for (int i = 0; i < nrbin; ++i) m_enc[i + 1] = m_enc[i] + mbin[i];
  const double mtot = m_enc[nrbin];
  phi_rad[nrbin] = (rmax > 0.0 && ndim == 3) ? -G * mtot / rmax : 0.0;
  for (int i = nrbin - 1; i >= 0; --i) {
    const double r_inner = i * dr, r_outer = (i + 1) * dr;
    double dphi = 0.0;
    if (ndim == 3) {
      const double m_inner = m_enc[i];
      const double shell_mass = m_enc[i + 1] - m_inner;
      if (r_inner == 0.0) {
        dphi = -G * shell_mass / (2.0 * r_outer);
      } else {
        const double denom = r_outer * r_outer * r_outer - r_inner * r_inner * r_inner;
        dphi = -G * (m_inner * (1.0 / r_inner - 1.0 / r_outer) +
                     shell_mass / denom *
                         (0.5 * (r_outer * r_outer - r_inner * r_inner) +
                          r_inner * r_inner * r_inner * (1.0 / r_outer - 1.0 / r_inner)));
      }
    } else {
      const double r_mid = 0.5 * (r_inner + r_outer);
      const double m_mid = 0.5 * (m_enc[i] + m_enc[i + 1]);
      double gr = 0.0;
      if (r_mid > 0.0) gr = (ndim == 1) ? -G * m_mid : -2.0 * G * m_mid / r_mid;
      dphi = gr * dr;
    }
    phi_rad[i] = phi_rad[i + 1] + dphi;
  }
}

/* ComputeGravityVectorDirect for one point (x, y, z: the cell centre, y and z zero on
 * inactive axes): g[3] and phi. */
void orc_mono_point(int nrbin, int ndim, double rmax, double G, const double *m_enc,
                    const double *phi_rad, double x, double y, double z, double *g,
                    double *phi) {
  const double dr = rmax / nrbin;
  const double r2 = x * x + y * y + z * z;
  const double r = sqrt(r2);
  if (r < 1e-10) {
    g[0] = g[1] = g[2] = 0.0;
    *phi = phi_rad[0];
    return;
  }
  double m_enclosed, phi_val;
  if (r >= rmax) {
    m_enclosed = m_enc[nrbin];
    phi_val = (ndim == 3) ? -G * m_enclosed / r : phi_rad[nrbin];
  } else {
    const double rf = r / dr;
    int bin = (int)rf;
    if (bin < 0) bin = 0;
    if (bin >= nrbin) bin = nrbin - 1;
    const double r_inner = bin * dr, r_outer = (bin + 1) * dr;
    const double shell_mass = m_enc[bin + 1] - m_enc[bin];
    double mass_frac;
    if (ndim == 1) {
      mass_frac = (r - r_inner) / dr;
    } else if (ndim == 2) {
      const double denom = r_outer * r_outer - r_inner * r_inner;
      mass_frac = (r * r - r_inner * r_inner) / denom;
    } else {
      const double r_inner3 = r_inner * r_inner * r_inner;
      const double r_outer3 = r_outer * r_outer * r_outer;
      mass_frac = (r * r * r - r_inner3) / (r_outer3 - r_inner3);
    }
    if (mass_frac < 0.0) mass_frac = 0.0;
    if (mass_frac > 1.0) mass_frac = 1.0;
    m_enclosed = m_enc[bin] + shell_mass * mass_frac;
    const double frac = rf - bin;
    phi_val = (1.0 - frac) * phi_rad[bin] + frac * phi_rad[bin + 1];
  }
  double g_r;
  if (ndim == 1)
    g_r = -G * m_enclosed;
  else if (ndim == 2)
    g_r = -2.0 * G * m_enclosed / r;
  else
    g_r = -G * m_enclosed / r2;
  const double rinv = 1.0 / r;
  g[0] = g_r * x * rinv;
  g[1] = (ndim >= 2) ? g_r * y * rinv : 0.0;
  g[2] = (ndim >= 3) ? g_r * z * rinv : 0.0;
  *phi = phi_val;
}

/* The gravity field (3 components) and phi (1 component, may be NULL) of every
 * interior cell, in the pack layout. */
void orc_mono_field(const aph_block *blocks, int nb, int ndim, int nrbin, double rmax,
                    double G, const double *m_enc, const double *phi_rad, double *grav,
                    double *phi) {
#pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < nb; ++b) {
    const aph_block *bk = &blocks[b];
    const int Nx = ext(bk, 0, ndim), Ny = ext(bk, 1, ndim);
    const long nc = ncell(bk, ndim);
    const int gx = gst(bk, 0, ndim), gy = gst(bk, 1, ndim), gz = gst(bk, 2, ndim);
    double *gf = fld(grav, bk, ndim, 3, 0);
    double *pf = phi ? fld(phi, bk, ndim, 1, 0) : NULL;
    for (int k = 0; k < bk->nx[2]; ++k)
      for (int j = 0; j < bk->nx[1]; ++j)
        for (int i = 0; i < bk->nx[0]; ++i) {
          double x, y, z, g[3], p;
          mono_centre(bk, i, j, k, &x, &y, &z);
          if (ndim < 2) y = 0.0;
          if (ndim < 3) z = 0.0;
          orc_mono_point(nrbin, ndim, rmax, G, m_enc, phi_rad, x, y, z, g, &p);
          const long c = ((long)(k + gz) * Ny + (j + gy)) * Nx + (i + gx);
          gf[c] = g[0];
          gf[nc + c] = g[1];
          gf[2 * nc + c] = g[2];
          if (pf) pf[c] = p;
        }
  }
}

/* ApplyGravity (gravity.cpp:374-410) with a per-cell field (3 components, pack layout). */
void orc_apply_gravity_field(const aph_block *blocks, int nb, int ndim, int nvar,
                             double *cons, const double *prim, const double *grav,
                             double dt) {
#pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < nb; ++b) {
    const aph_block *bk = &blocks[b];
    const int Nx = ext(bk, 0, ndim), Ny = ext(bk, 1, ndim);
    const long nc = ncell(bk, ndim);
    const int gx = gst(bk, 0, ndim), gy = gst(bk, 1, ndim), gz = gst(bk, 2, ndim);
    double *u = fld(cons, bk, ndim, nvar, 0);
    const double *w = cfld(prim, bk, ndim, nvar, 0);
    const double *gf = cfld(grav, bk, ndim, 3, 0);
    for (int k = gz; k < gz + bk->nx[2]; ++k)
      for (int j = gy; j < gy + bk->nx[1]; ++j)
        for (int i = gx; i < gx + bk->nx[0]; ++i) {
          const long c = ((long)k * Ny + j) * Nx + i;
          const double g0 = gf[c], g1 = gf[nc + c], g2 = gf[2 * nc + c];
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
}

/* One step of `integrator` with monopole gravity (apophis_driver.cpp:64-85: each stage
 * bins the stage-start primitive density, builds the tables and the field, then the
 * update and ApplyGravity with beta*dt).  grav: scratch of 3 components (pack layout);
 * m_enc / phi_rad (nrbin + 1 each) hold the last stage's tables on return. */
double orc_step_monopole(const aph_block *blocks, int nb, int ndim, int nvar, int integrator,
                         int recon, int rsolver, const aph_eos *eos, int nrbin, double rmax,
                         double G, double cfl, double dt, double *u0, double *u1,
                         double *prim, double *gamma, double *fx1, double *fx2, double *fx3,
                         double *grav, double *m_enc, double *phi_rad) {
  double gam0[3], gam1[3], beta[3];
  const int nstages = orc_integrator_stages(integrator, gam0, gam1, beta);
  long ntot = 0;
  for (int b = 0; b < nb; ++b) ntot += ncell(&blocks[b], ndim);
  memcpy(u1, u0, sizeof(double) * (size_t)nvar * ntot);
  double *mbin = (double *)malloc(sizeof(double) * (size_t)nrbin);
  for (int s = 0; s < nstages; ++s) {
    const int rec = (s == 0 && integrator == 4) ? APH_RECON_DC : recon;
    orc_mono_bin_masses(blocks, nb, ndim, nvar, prim, nrbin, rmax, mbin);
    orc_mono_tables(nrbin, ndim, rmax, G, mbin, m_enc, phi_rad);
    orc_mono_field(blocks, nb, ndim, nrbin, rmax, G, m_enc, phi_rad, grav, NULL);
    orc_calculate_fluxes(blocks, nb, ndim, nvar, rec, rsolver, prim, gamma, fx1, fx2, fx3);
    orc_update(blocks, nb, ndim, nvar, u0, u1, fx1, fx2, fx3, gam0[s], gam1[s],
               beta[s] * dt);
    orc_apply_gravity_field(blocks, nb, ndim, nvar, u0, prim, grav, beta[s] * dt);
    orc_fill_ghosts(blocks, nb, ndim, nvar, u0);
    orc_cons_to_prim(blocks, nb, ndim, nvar, eos, u0, prim, gamma);
  }
  free(mbin);
  return orc_estimate_timestep(blocks, nb, ndim, nvar, eos, prim, u0, cfl);
}
