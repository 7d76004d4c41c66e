// aph_adjoint.hpp -- reverse-mode (adjoint) point functions of the hydro flux path.
//
// Each function recomputes the forward point function of aph_point.hpp (same branches)
// and propagates output cotangents back to its inputs along the branch that was taken:
// the limiter's active branch in PLM/PPM (plm_simple.hpp:14-31, ppm_simple.hpp:31-161),
// the active wave-speed/contact branches of HLLE/HLLC (hydro_hlle.hpp:28-133,
// hydro_hllc.hpp:161-282) and ConsToPrim without active floors (hydro.cpp:474-531).
// min/max select their argument's derivative; |x| has derivative sign(x) and 0 at 0;
// thresholds and sign tests are piecewise constant.  Output cotangents are *added* into
// the input-cotangent arrays.
#pragma once

#include "aph_point.hpp"

namespace aph {

APH_HD double dabs(double x) { return (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : 0.0); }

// ------------------------------------------------------------------ PLM
APH_HD void plm_adj(const double *q, double qlb, double qrb, double *qb) {
  const double dql = q[1] - q[0], dqr = q[2] - q[1];
  const double dq2 = dql * dqr;
  qb[1] += qlb + qrb;
  const double dqmb = qlb - qrb;
  if (dq2 > 0.0) {  // dqm = dq2 / s: one reciprocal for its three quotients
    const double is = 1.0 / (dql + dqr);
    const double dq2b = dqmb * is;
    const double sb = -dq2b * (dq2 * is);
    const double dqlb = dq2b * dqr + sb;
    const double dqrb = dq2b * dql + sb;
    qb[1] += dqlb - dqrb;
    qb[0] -= dqlb;
    qb[2] += dqrb;
  }
}

// ------------------------------------------------------------------ PPM
// Cotangent of the limited extremum value S(a)*min(C2|b|, min(C2|c|, |a|)) (step 2a).
APH_HD void ppm_qd_adj(double a, double b, double c, double qdb, double &ab, double &bb,
                       double &cb) {
  const double C2 = 1.25;
  const double sa = SIGN(a);
  const double ea = fabs(a), eb = C2 * fabs(b), ec = C2 * fabs(c);
  // inner = mn(ec, ea): ea if ea < ec else ec ; outer = mn(eb, inner)
  const bool inner_is_a = ea < ec;
  const double inner = inner_is_a ? ea : ec;
  if (inner < eb) {
    if (inner_is_a)
      ab += qdb * sa * dabs(a);
    else
      cb += qdb * sa * C2 * dabs(c);
  } else {
    bb += qdb * sa * C2 * dabs(b);
  }
}

APH_HD void ppm_adj(const double *q, double qlb, double qrb, double *qb) {
  const double C2 = 1.25;
  const double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4];
  // ---- forward (same expressions as ppm())
  const double qa0 = (q2 - q1);
  const double qb0 = (q3 - q2);
  const double dd_im1 = 0.5 * qa0 + 0.5 * (q1 - q0);
  const double dd = 0.5 * qb0 + 0.5 * qa0;
  const double dd_ip1 = 0.5 * (q4 - q3) + 0.5 * qb0;
  const double dph0 = 0.5 * (q1 + q2) + (dd_im1 - dd) / 6.0;
  const double dphp0 = 0.5 * (q2 + q3) + (dd - dd_ip1) / 6.0;
  const double d2qc_im1 = q0 + q2 - 2.0 * q1;
  const double d2qc = q1 + q3 - 2.0 * q2;
  const double d2qc_ip1 = q2 + q4 - 2.0 * q3;
  // i-1/2
  const double A1 = 3.0 * (q1 + q2 - 2.0 * dph0);
  const bool s1 = (SIGN(A1) == SIGN(d2qc_im1) && SIGN(A1) == SIGN(d2qc));
  double qd1 = 0.0;
  if (s1) qd1 = SIGN(A1) * mn(C2 * fabs(d2qc_im1), mn(C2 * fabs(d2qc), fabs(A1)));
  const bool e1 = ((dph0 - q1) * (q2 - dph0) < 0.0);
  const double dph = e1 ? (0.5 * (q1 + q2) - qd1 / 6.0) : dph0;
  // i+1/2
  const double A2 = 3.0 * (q2 + q3 - 2.0 * dphp0);
  const bool s2 = (SIGN(A2) == SIGN(d2qc) && SIGN(A2) == SIGN(d2qc_ip1));
  double qd2 = 0.0;
  if (s2) qd2 = SIGN(A2) * mn(C2 * fabs(d2qc), mn(C2 * fabs(d2qc_ip1), fabs(A2)));
  const bool e2 = ((dphp0 - q2) * (q3 - dphp0) < 0.0);
  const double dphp = e2 ? (0.5 * (q2 + q3) - qd2 / 6.0) : dphp0;
  const double d2qf = 6.0 * (dph + dphp - 2.0 * q2);
  const double dqf_minus = q2 - dph;
  const double dqf_plus = dphp - q2;
  const double qa_tmp = dqf_minus * dqf_plus;
  const double qb_tmp = (q3 - q2) * (q2 - q1);
  const bool s4 = (SIGN(d2qc_im1) == SIGN(d2qc) && SIGN(d2qc_im1) == SIGN(d2qc_ip1) &&
                   SIGN(d2qc_im1) == SIGN(d2qf));
  double qe = 0.0;
  const double ea = C2 * fabs(d2qc_im1), eb = C2 * fabs(d2qc), ec = C2 * fabs(d2qc_ip1),
               ed = fabs(d2qf);
  const bool m1_is_b = eb < ea;
  const double m1 = m1_is_b ? eb : ea;
  const bool m2_is_d = ed < ec;
  const double m2 = m2_is_d ? ed : ec;
  const bool sel_m2 = m2 < m1;
  if (s4) qe = SIGN(d2qf) * (sel_m2 ? m2 : m1);
  const double amax = mx(fabs(q1), fabs(q0));
  const double bmax = mx(mx(fabs(q2), fabs(q3)), fabs(q4));
  const bool rho_on = fabs(d2qf) > (1.0e-12) * mx(amax, bmax);
  const double rho = rho_on ? qe / d2qf : 0.0;
  // ---- reverse
  double dphb = 0.0, dphpb = 0.0, dqf_minusb = 0.0, dqf_plusb = 0.0, rhob = 0.0;
  double q2b = 0.0;
  if ((qa_tmp <= 0.0 || qb_tmp <= 0.0)) {
    if (rho <= (1.0 - (1.0e-12))) {  // qr = q2 - rho*dqf_minus, ql = q2 + rho*dqf_plus
      q2b += qrb + qlb;
      rhob += -qrb * dqf_minus + qlb * dqf_plus;
      dqf_minusb += -qrb * rho;
      dqf_plusb += qlb * rho;
    } else {
      dphb += qrb;
      dphpb += qlb;
    }
  } else {
    if (fabs(dqf_minus) >= 2.0 * fabs(dqf_plus)) {  // qr = q2 - 2 dqf_plus
      q2b += qrb;
      dqf_plusb -= 2.0 * qrb;
    } else {
      dphb += qrb;
    }
    if (fabs(dqf_plus) >= 2.0 * fabs(dqf_minus)) {  // ql = q2 + 2 dqf_minus
      q2b += qlb;
      dqf_minusb += 2.0 * qlb;
    } else {
      dphpb += qlb;
    }
  }
  double d2qfb = 0.0, d2qc_im1b = 0.0, d2qcb = 0.0, d2qc_ip1b = 0.0;
  if (rho_on) {
    const double qeb = rhob / d2qf;
    d2qfb -= rhob * rho / d2qf;
    if (s4) {
      const double sg = SIGN(d2qf) * qeb;
      if (sel_m2) {
        if (m2_is_d)
          d2qfb += sg * dabs(d2qf);
        else
          d2qc_ip1b += sg * C2 * dabs(d2qc_ip1);
      } else {
        if (m1_is_b)
          d2qcb += sg * C2 * dabs(d2qc);
        else
          d2qc_im1b += sg * C2 * dabs(d2qc_im1);
      }
    }
  }
  // dqf_minus = q2 - dph ; dqf_plus = dphp - q2
  q2b += dqf_minusb - dqf_plusb;
  dphb -= dqf_minusb;
  dphpb += dqf_plusb;
  // d2qf = 6 (dph + dphp - 2 q2)
  dphb += 6.0 * d2qfb;
  dphpb += 6.0 * d2qfb;
  q2b -= 12.0 * d2qfb;
  double q0b = 0.0, q1b = 0.0, q3b = 0.0, q4b = 0.0;
  double dph0b = 0.0, dphp0b = 0.0;
  // i+1/2 face value
  if (e2) {  // dphp = 0.5 (q2 + q3) - qd2/6
    q2b += 0.5 * dphpb;
    q3b += 0.5 * dphpb;
    const double qd2b = -dphpb / 6.0;
    if (s2) {
      double A2b = 0.0;
      ppm_qd_adj(A2, d2qc, d2qc_ip1, qd2b, A2b, d2qcb, d2qc_ip1b);
      q2b += 3.0 * A2b;
      q3b += 3.0 * A2b;
      dphp0b -= 6.0 * A2b;
    }
  } else {
    dphp0b += dphpb;
  }
  // i-1/2 face value
  if (e1) {
    q1b += 0.5 * dphb;
    q2b += 0.5 * dphb;
    const double qd1b = -dphb / 6.0;
    if (s1) {
      double A1b = 0.0;
      ppm_qd_adj(A1, d2qc_im1, d2qc, qd1b, A1b, d2qc_im1b, d2qcb);
      q1b += 3.0 * A1b;
      q2b += 3.0 * A1b;
      dph0b -= 6.0 * A1b;
    }
  } else {
    dph0b += dphb;
  }
  // second differences
  q0b += d2qc_im1b;
  q2b += d2qc_im1b;
  q1b -= 2.0 * d2qc_im1b;
  q1b += d2qcb;
  q3b += d2qcb;
  q2b -= 2.0 * d2qcb;
  q2b += d2qc_ip1b;
  q4b += d2qc_ip1b;
  q3b -= 2.0 * d2qc_ip1b;
  // step 1
  double dd_im1b = dph0b / 6.0, ddb = -dph0b / 6.0 + dphp0b / 6.0, dd_ip1b = -dphp0b / 6.0;
  q1b += 0.5 * dph0b;
  q2b += 0.5 * dph0b + 0.5 * dphp0b;
  q3b += 0.5 * dphp0b;
  double qa0b = 0.5 * dd_im1b + 0.5 * ddb;
  double qb0b = 0.5 * ddb + 0.5 * dd_ip1b;
  q1b += 0.5 * dd_im1b;
  q0b -= 0.5 * dd_im1b;
  q4b += 0.5 * dd_ip1b;
  q3b -= 0.5 * dd_ip1b;
  q2b += qa0b - qb0b;
  q1b -= qa0b;
  q3b += qb0b;
  qb[0] += q0b;
  qb[1] += q1b;
  qb[2] += q2b;
  qb[3] += q3b;
  qb[4] += q4b;
}

// ---- face-shared PPM (ppm_face / ppm_cell of aph_point.hpp): the limited value of a
// face is formed once and its cotangent, summed over the two cells that use it, is
// propagated once.  The reverse sweep multiplies by 1/6 where the forward divides by 6
// (a cotangent needs no bitwise match; the forward values do, and are recomputed
// exactly).
constexpr double kSixth = 1.0 / 6.0;

// What the reverse of one face needs from its forward.
struct PpmFaceTape {
  double qa, d2qc_im1, d2qc;
  bool lim;  // the extremum-limited value dph_tmp was taken
  bool sgn;  // the three signs agreed (qd != 0)
};

// ppm_face of cells (q_im2 .. q_ip1), bitwise, recording its branches (returned with
// the value, so the tape stays in registers).
struct PpmFace {
  double dph;
  PpmFaceTape t;
};
template <bool FAST>
APH_HD PpmFace ppm_face_taped(double q_im2, double q_im1, double q_i, double q_ip1) {
  PpmFace r;
  PpmFaceTape &t = r.t;
  const double C2 = 1.25;
  const double qa0 = (q_i - q_im1);
  const double qb0 = (q_ip1 - q_i);
  const double dd_im1 = 0.5 * qa0 + 0.5 * (q_im1 - q_im2);
  const double dd = 0.5 * qb0 + 0.5 * qa0;
  const double avg = 0.5 * (q_im1 + q_i);
  double dph = avg + (FAST ? div6_fast(dd_im1 - dd) : (dd_im1 - dd) / 6.0);
  t.d2qc_im1 = q_im2 + q_i - 2.0 * q_im1;
  t.d2qc = q_im1 + q_ip1 - 2.0 * q_i;
  const double qa_tmp = dph - q_im1;
  const double qb_tmp = q_i - dph;
  t.qa = 3.0 * (q_im1 + q_i - 2.0 * dph);
  double qd = 0.0;
  t.sgn = (neg(t.qa) == neg(t.d2qc_im1) && neg(t.qa) == neg(t.d2qc));
  if (t.sgn)
    qd = with_sign_of(t.qa, fmin(C2 * fabs(t.d2qc_im1), fmin(C2 * fabs(t.d2qc), fabs(t.qa))));
  const double dph_tmp = avg - (FAST ? div6_fast(qd) : qd / 6.0);
  t.lim = (qa_tmp * qb_tmp < 0.0);
  if (t.lim) dph = dph_tmp;
  r.dph = dph;
  return r;
}

template <bool FAST>
APH_HD double ppm_face_tape(double q_im2, double q_im1, double q_i, double q_ip1,
                            PpmFaceTape &t) {
  const PpmFace r = ppm_face_taped<FAST>(q_im2, q_im1, q_i, q_ip1);
  t = r.t;
  return r.dph;
}

// Cotangent Fb of a face value -> cells (q_im2 .. q_ip1), added into qb[0..3].
APH_HD void ppm_face_adj(const PpmFaceTape &t, double Fb, double *qb) {
  double dph0b = 0.0, avgb = 0.0;
  double b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;  // q_im2, q_im1, q_i, q_ip1
  if (t.lim) {  // dph = avg - qd/6
    avgb += Fb;
    if (t.sgn) {
      // ppm_qd_adj(t.qa, t.d2qc_im1, t.d2qc, -Fb/6, ab, bb, cb) written with selects:
      // adding into one of three references by branch made the compiler index a stack
      // array on every face
      const double C2 = 1.25;
      const double qdb = -Fb * kSixth;
      const double sa = SIGN(t.qa);
      const double ea = fabs(t.qa), eb = C2 * fabs(t.d2qc_im1), ec = C2 * fabs(t.d2qc);
      const bool inner_is_a = ea < ec;
      const bool use_inner = (inner_is_a ? ea : ec) < eb;
      const double ab = (use_inner && inner_is_a) ? qdb * sa * dabs(t.qa) : 0.0;
      const double cb = (use_inner && !inner_is_a) ? qdb * sa * C2 * dabs(t.d2qc) : 0.0;
      const double bb = use_inner ? 0.0 : qdb * sa * C2 * dabs(t.d2qc_im1);
      // qa = 3 (q_im1 + q_i - 2 dph0)
      b1 += 3.0 * ab;
      b2 += 3.0 * ab;
      dph0b -= 6.0 * ab;
      // d2qc_im1 = q_im2 + q_i - 2 q_im1 ; d2qc = q_im1 + q_ip1 - 2 q_i
      b0 += bb;
      b2 += bb;
      b1 -= 2.0 * bb;
      b1 += cb;
      b3 += cb;
      b2 -= 2.0 * cb;
    }
  } else {
    dph0b += Fb;
  }
  // dph0 = avg + (dd_im1 - dd)/6 ; avg = (q_im1 + q_i)/2
  avgb += dph0b;
  const double dd_im1b = dph0b * kSixth, ddb = -dd_im1b;
  b1 += 0.5 * avgb;
  b2 += 0.5 * avgb;
  // dd_im1 = qa0/2 + (q_im1 - q_im2)/2 ; dd = qb0/2 + qa0/2 ; qa0 = q_i - q_im1 ;
  // qb0 = q_ip1 - q_i
  const double qa0b = 0.5 * dd_im1b + 0.5 * ddb;
  const double qb0b = 0.5 * ddb;
  b1 += 0.5 * dd_im1b;
  b0 -= 0.5 * dd_im1b;
  b2 += qa0b - qb0b;
  b1 -= qa0b;
  b3 += qb0b;
  qb[0] += b0;
  qb[1] += b1;
  qb[2] += b2;
  qb[3] += b3;
}

// Cotangents (ql_ip1b, qr_ib) of ppm_cell -> cells q[0..4] (added into qb) and the two
// face values (dphb, dph_ip1b, set).  The forward is recomputed bitwise from the given
// face values.
APH_HD void ppm_cell_adj(const double *q, double dph, double dph_ip1, double qlb, double qrb,
                         double *qb, double &dphb, double &dph_ip1b) {
  const double C2 = 1.25;
  const double q_im2 = q[0], q_im1 = q[1], q_i = q[2], q_ip1 = q[3], q_ip2 = q[4];
  const double d2qc_im1 = q_im2 + q_i - 2.0 * q_im1;
  const double d2qc = q_im1 + q_ip1 - 2.0 * q_i;
  const double d2qc_ip1 = q_i + q_ip2 - 2.0 * q_ip1;
  const double d2qf = 6.0 * (dph + dph_ip1 - 2.0 * q_i);
  const double dqf_minus = q_i - dph;
  const double dqf_plus = dph_ip1 - q_i;
  const double qa_tmp = dqf_minus * dqf_plus;
  const double qb_tmp = (q_ip1 - q_i) * (q_i - q_im1);
  const bool s4 = (neg(d2qc_im1) == neg(d2qc) && neg(d2qc_im1) == neg(d2qc_ip1) &&
                   neg(d2qc_im1) == neg(d2qf));
  const double ea = C2 * fabs(d2qc_im1), eb = C2 * fabs(d2qc), ec = C2 * fabs(d2qc_ip1),
               ed = fabs(d2qf);
  const bool m1_is_b = eb < ea;
  const double m1 = m1_is_b ? eb : ea;
  const bool m2_is_d = ed < ec;
  const double m2 = m2_is_d ? ed : ec;
  const bool sel_m2 = m2 < m1;
  double qe = 0.0;
  if (s4) qe = with_sign_of(d2qf, fmin(fmin(ea, eb), fmin(ec, ed)));
  const double qa = fmax(fabs(q_im1), fabs(q_im2));
  const double qbm = fmax(fmax(fabs(q_i), fabs(q_ip1)), fabs(q_ip2));
  const bool rho_on = fabs(d2qf) > (1.0e-12) * fmax(qa, qbm);
  const double rho = rho_on ? qe / d2qf : 0.0;
  // ---- reverse
  double dphb_ = 0.0, dphpb = 0.0, dqf_minusb = 0.0, dqf_plusb = 0.0, rhob = 0.0;
  double q2b = 0.0;
  if ((qa_tmp <= 0.0 || qb_tmp <= 0.0)) {
    if (rho <= (1.0 - (1.0e-12))) {  // qr = q_i - rho dqf_minus, ql = q_i + rho dqf_plus
      q2b += qrb + qlb;
      rhob += -qrb * dqf_minus + qlb * dqf_plus;
      dqf_minusb += -qrb * rho;
      dqf_plusb += qlb * rho;
    } else {
      dphb_ += qrb;
      dphpb += qlb;
    }
  } else {
    if (fabs(dqf_minus) >= 2.0 * fabs(dqf_plus)) {  // qr = q_i - 2 dqf_plus
      q2b += qrb;
      dqf_plusb -= 2.0 * qrb;
    } else {
      dphb_ += qrb;
    }
    if (fabs(dqf_plus) >= 2.0 * fabs(dqf_minus)) {  // ql = q_i + 2 dqf_minus
      q2b += qlb;
      dqf_minusb += 2.0 * qlb;
    } else {
      dphpb += qlb;
    }
  }
  double d2qfb = 0.0, d2qc_im1b = 0.0, d2qcb = 0.0, d2qc_ip1b = 0.0;
  if (rho_on) {  // rho = qe / d2qf
    const double qeb = rhob / d2qf;
    d2qfb -= qeb * rho;
    if (s4) {
      const double sg = with_sign_of(d2qf, qeb);
      if (sel_m2) {
        if (m2_is_d)
          d2qfb += sg * dabs(d2qf);
        else
          d2qc_ip1b += sg * C2 * dabs(d2qc_ip1);
      } else {
        if (m1_is_b)
          d2qcb += sg * C2 * dabs(d2qc);
        else
          d2qc_im1b += sg * C2 * dabs(d2qc_im1);
      }
    }
  }
  q2b += dqf_minusb - dqf_plusb;
  dphb_ -= dqf_minusb;
  dphpb += dqf_plusb;
  dphb_ += 6.0 * d2qfb;
  dphpb += 6.0 * d2qfb;
  q2b -= 12.0 * d2qfb;
  qb[0] += d2qc_im1b;
  qb[1] += d2qcb - 2.0 * d2qc_im1b;
  qb[2] += q2b + d2qc_im1b + d2qc_ip1b - 2.0 * d2qcb;
  qb[3] += d2qcb - 2.0 * d2qc_ip1b;
  qb[4] += d2qc_ip1b;
  dphb = dphb_;
  dph_ip1b = dphpb;
}

template <int RECON>
APH_HD void recon1_adj(const double *w, double qlb, double qrb, double *wb) {
  if constexpr (RECON == RECON_PPM) {
    ppm_adj(w, qlb, qrb, wb);
  } else if constexpr (RECON == RECON_PLM) {
    plm_adj(w, qlb, qrb, wb);
  } else {
    wb[0] += qlb + qrb;
  }
}

// ------------------------------------------------------------------ sound speed
// c = C*sqrt(g / (1 + (ig + rho*C^2/p)));  adds cotangents of (g, ig, rho, p)
// c = C sqrt(s), s = g / dl, dl = 1 + (ig + t), t = rho C^2 / p; c is the forward value, so
// ds = dc C^2 / (2 c) (sqrt(s) = c / C) needs no square root, and p and dl are inverted once
// each.
APH_HD void sound_speed_adj(double g, double ig, double rho, double p, double c, double cb,
                            double &gb, double &igb, double &rhob, double &pb) {
  const double ip = 1.0 / p;
  const double t = rho * C_LIGHT2 * ip;
  const double idl = 1.0 / (1 + (ig + t));
  const double s = g * idl;
  const double sb = cb * (0.5 * C_LIGHT2) / c;
  gb += sb * idl;
  const double dlb = -sb * s * idl;
  igb += dlb;
  rhob += dlb * C_LIGHT2 * ip;
  pb -= dlb * t * ip;
}

// ------------------------------------------------------------------ HLLC / LHLLC
// LOWMACH: the LHLLC solver (hydro_lhllc.hpp:70-79), whose PVRS term is scaled by
// phi = chi (2 - chi), chi = max(|v_L|, |v_R|); its derivative goes to the side that
// attains the max (the forward's std::max picks the left one on ties), and |v| = 0 has
// none.
template <bool LOWMACH = false>
APH_HD void hllc_adj(const double *wl, const double *wr, const double *ifl, const double *ifr,
                     const double *flxb, double *wlb, double *wrb, double *iflb,
                     double *ifrb) {
  // ---- forward
  const double gl = ifl[0], gr = ifr[0];
  const double gm1l = ifl[1] - 1, gm1r = ifr[1] - 1;
  const double igl = 1.0 / gm1l, igr = 1.0 / gm1r;
  const double cl = sound_speed(gl, igl, wl[0], wl[4]);
  const double cr = sound_speed(gr, igr, wr[0], wr[4]);
  const double ksql = (SQR(wl[1]) + SQR(wl[2]) + SQR(wl[3]));
  const double ksqr = (SQR(wr[1]) + SQR(wr[2]) + SQR(wr[3]));
  const double el = wl[4] * igl + 0.5 * wl[0] * ksql;
  const double er = wr[4] * igr + 0.5 * wr[0] * ksqr;
  const double rhoa = .5 * (wl[0] + wr[0]);
  const double ca = .5 * (cl + cr);
  const double h = .5 * (wr[1] - wl[1]);
  double pvrs = h * rhoa * ca;
  double machl = 0.0, machr = 0.0, chi = 0.0, phi = 1.0;
  if constexpr (LOWMACH) {
    machl = sqrt(wl[1] * wl[1] + wl[2] * wl[2] + wl[3] * wl[3]);
    machr = sqrt(wr[1] * wr[1] + wr[2] * wr[2] + wr[3] * wr[3]);
    chi = mx(machl, machr);
    phi = chi * (2.0 - chi);
    pvrs = pvrs * phi;
  }
  const double pstar = .5 * (wl[4] + wr[4]) - pvrs;
  const double K = (gl + 1) / (2 * gr);
  const bool shl = !(pstar <= wl[4]), shr = !(pstar <= wr[4]);
  const double ql = shl ? sqrt(1.0 + K * (pstar / wl[4] - 1.0)) : 1.0;
  const double qr = shr ? sqrt(1.0 + K * (pstar / wr[4] - 1.0)) : 1.0;
  const double sl = wl[1] - cl * ql;
  const double sr = wr[1] + cr * qr;
  const double bp = sr > 0.0 ? sr : (TINY_NUMBER);
  const double bm = sl < 0.0 ? sl : -(TINY_NUMBER);
  const double vxl = wl[1] - sl, vxr = wr[1] - sr;
  const double tl = wl[4] + vxl * wl[0] * wl[1];
  const double tr = wr[4] + vxr * wr[0] * wr[1];
  const double ml = wl[0] * vxl, mr = -(wr[0] * vxr);
  const double M = ml + mr;
  const double sstar = (tl - tr) / M;
  const double cp0 = (ml * tr + mr * tl) / M;
  const double cp = cp0 > 0.0 ? cp0 : 0.0;
  const double vxl2 = wl[1] - bm, vxr2 = wr[1] - bp;
  double fl[5], fr[5];
  fl[0] = wl[0] * vxl2;
  fr[0] = wr[0] * vxr2;
  fl[1] = wl[0] * wl[1] * vxl2 + wl[4];
  fr[1] = wr[0] * wr[1] * vxr2 + wr[4];
  fl[2] = wl[0] * wl[2] * vxl2;
  fr[2] = wr[0] * wr[2] * vxr2;
  fl[3] = wl[0] * wl[3] * vxl2;
  fr[3] = wr[0] * wr[3] * vxr2;
  fl[4] = el * vxl2 + wl[4] * wl[1];
  fr[4] = er * vxr2 + wr[4] * wr[1];
  const bool left = sstar >= 0.0;
  double ul, ur, ust, D;
  if (left) {
    D = sstar - bm;
    ul = sstar / D;
    ur = 0.0;
    ust = -bm / D;
  } else {
    D = bp - sstar;
    ul = 0.0;
    ur = -sstar / D;
    ust = bp / D;
  }
  // ---- reverse
  double ulb = 0, urb = 0, ustb = 0, cpb = 0, sstarb = 0, bmb = 0, bpb = 0;
  double flb[5], frb[5];
  for (int n = 0; n < 5; ++n) {
    ulb += flxb[n] * fl[n];
    urb += flxb[n] * fr[n];
    flb[n] = flxb[n] * ul;
    frb[n] = flxb[n] * ur;
  }
  ustb += flxb[1] * cp + flxb[4] * cp * sstar;
  cpb += flxb[1] * ust + flxb[4] * ust * sstar;
  sstarb += flxb[4] * ust * cp;
  // the reverse's quotients share one reciprocal per denominator (the cotangents are
  // checked to a tolerance; the forward values above, which pick the branches, are exact)
  const double iD2 = 1.0 / (D * D);
  if (left) {
    sstarb += (ustb - ulb) * bm * iD2;
    bmb += (ulb - ustb) * sstar * iD2;
  } else {
    sstarb += (ustb - urb) * bp * iD2;
    bpb += (urb - ustb) * sstar * iD2;
  }
  // fluxes along bm / bp
  double vxl2b = 0, vxr2b = 0, elb = 0, erb = 0;
  wlb[0] += flb[0] * vxl2 + flb[1] * wl[1] * vxl2 + flb[2] * wl[2] * vxl2 + flb[3] * wl[3] * vxl2;
  wrb[0] += frb[0] * vxr2 + frb[1] * wr[1] * vxr2 + frb[2] * wr[2] * vxr2 + frb[3] * wr[3] * vxr2;
  vxl2b += flb[0] * wl[0] + flb[1] * wl[0] * wl[1] + flb[2] * wl[0] * wl[2] +
           flb[3] * wl[0] * wl[3] + flb[4] * el;
  vxr2b += frb[0] * wr[0] + frb[1] * wr[0] * wr[1] + frb[2] * wr[0] * wr[2] +
           frb[3] * wr[0] * wr[3] + frb[4] * er;
  wlb[1] += flb[1] * wl[0] * vxl2 + flb[4] * wl[4];
  wrb[1] += frb[1] * wr[0] * vxr2 + frb[4] * wr[4];
  wlb[2] += flb[2] * wl[0] * vxl2;
  wrb[2] += frb[2] * wr[0] * vxr2;
  wlb[3] += flb[3] * wl[0] * vxl2;
  wrb[3] += frb[3] * wr[0] * vxr2;
  wlb[4] += flb[1] + flb[4] * wl[1];
  wrb[4] += frb[1] + frb[4] * wr[1];
  elb += flb[4] * vxl2;
  erb += frb[4] * vxr2;
  wlb[1] += vxl2b;
  bmb -= vxl2b;
  wrb[1] += vxr2b;
  bpb -= vxr2b;
  // contact pressure and speed
  double cp0b = (cp0 > 0.0) ? cpb : 0.0;
  const double iM = 1.0 / M;
  const double Nb = cp0b * iM;
  double Mb = -cp0b * cp0 * iM;
  double mlb = Nb * tr, mrb = Nb * tl, trb = Nb * ml, tlb = Nb * mr;
  tlb += sstarb * iM;
  trb -= sstarb * iM;
  Mb -= sstarb * sstar * iM;
  mlb += Mb;
  mrb += Mb;
  double vxlb = 0, vxrb = 0;
  wrb[0] -= mrb * vxr;
  vxrb -= mrb * wr[0];
  wlb[0] += mlb * vxl;
  vxlb += mlb * wl[0];
  wrb[4] += trb;
  vxrb += trb * wr[0] * wr[1];
  wrb[0] += trb * vxr * wr[1];
  wrb[1] += trb * vxr * wr[0];
  wlb[4] += tlb;
  vxlb += tlb * wl[0] * wl[1];
  wlb[0] += tlb * vxl * wl[1];
  wlb[1] += tlb * vxl * wl[0];
  double slb = 0, srb = 0;
  wlb[1] += vxlb;
  slb -= vxlb;
  wrb[1] += vxrb;
  srb -= vxrb;
  if (sl < 0.0) slb += bmb;
  if (sr > 0.0) srb += bpb;
  // wave speeds
  double clb = 0, crb = 0, qlb = 0, qrb = 0;
  wlb[1] += slb;
  clb -= slb * ql;
  qlb -= slb * cl;
  wrb[1] += srb;
  crb += srb * qr;
  qrb += srb * cr;
  double pstarb = 0, Kb = 0;
  if (shl) {
    const double Xb = qlb / (2.0 * ql), ip = 1.0 / wl[4];
    Kb += Xb * (pstar * ip - 1.0);
    pstarb += Xb * K * ip;
    wlb[4] -= Xb * K * pstar * ip * ip;
  }
  if (shr) {
    const double Xb = qrb / (2.0 * qr), ip = 1.0 / wr[4];
    Kb += Xb * (pstar * ip - 1.0);
    pstarb += Xb * K * ip;
    wrb[4] -= Xb * K * pstar * ip * ip;
  }
  double glb = 0.0, grb = 0.0;  // K = (gl + 1) / (2 gr) enters only through a shock
  if (shl || shr) {
    const double igr2 = 1.0 / (2 * gr);
    glb = Kb * igr2;
    grb = -glb * (gl + 1) * (2 * igr2);
  }
  // pstar = .5 (pl + pr) - h * rhoa * ca [* phi]
  wlb[4] += .5 * pstarb;
  wrb[4] += .5 * pstarb;
  double rawb = -pstarb;  // cotangent of h * rhoa * ca
  if constexpr (LOWMACH) {
    const double phib = rawb * (h * rhoa * ca);
    rawb = rawb * phi;
    const double chib = phib * (2.0 - 2.0 * chi);
    if (machl < machr) {
      if (machr > 0.0) {
        const double im = chib / machr;
        for (int k = 1; k <= 3; ++k) wrb[k] += im * wr[k];
      }
    } else if (machl > 0.0) {
      const double im = chib / machl;
      for (int k = 1; k <= 3; ++k) wlb[k] += im * wl[k];
    }
  }
  const double hb = rawb * rhoa * ca;
  const double rhoab = rawb * h * ca;
  const double cab = rawb * h * rhoa;
  wrb[1] += .5 * hb;
  wlb[1] -= .5 * hb;
  wlb[0] += .5 * rhoab;
  wrb[0] += .5 * rhoab;
  clb += .5 * cab;
  crb += .5 * cab;
  // energies
  double iglb = elb * wl[4], igrb = erb * wr[4];
  wlb[4] += elb * igl;
  wrb[4] += erb * igr;
  wlb[0] += elb * 0.5 * ksql;
  wrb[0] += erb * 0.5 * ksqr;
  for (int k = 1; k <= 3; ++k) {
    wlb[k] += elb * (0.5 * wl[0]) * 2.0 * wl[k];
    wrb[k] += erb * (0.5 * wr[0]) * 2.0 * wr[k];
  }
  // sound speeds
  sound_speed_adj(gl, igl, wl[0], wl[4], cl, clb, glb, iglb, wlb[0], wlb[4]);
  sound_speed_adj(gr, igr, wr[0], wr[4], cr, crb, grb, igrb, wrb[0], wrb[4]);
  iflb[1] += -iglb * igl * igl;
  ifrb[1] += -igrb * igr * igr;
  iflb[0] += glb;
  ifrb[0] += grb;
}

// ------------------------------------------------------------------ HLLE
APH_HD void hlle_adj(const double *wl, const double *wr, const double *ifl, const double *ifr,
                     const double *flxb, double *wlb, double *wrb, double *iflb,
                     double *ifrb) {
  // ---- forward
  const double gl = ifl[0], gr = ifr[0];
  const double gm1l = ifl[1] - 1, gm1r = ifr[1] - 1;
  const double igl = 1.0 / gm1l, igr = 1.0 / gm1r;
  const double cl = sound_speed(gl, igl, wl[0], wl[4]);
  const double cr = sound_speed(gr, igr, wr[0], wr[4]);
  const double ksql = (SQR(wl[1]) + SQR(wl[2]) + SQR(wl[3]));
  const double ksqr = (SQR(wr[1]) + SQR(wr[2]) + SQR(wr[3]));
  const double el = wl[4] * igl + 0.5 * wl[0] * ksql;
  const double er = wr[4] * igr + 0.5 * wr[0] * ksqr;
  const double rhoa = .5 * (wl[0] + wr[0]);
  const double ca = .5 * (cl + cr);
  const double pmid = .5 * (wl[4] + wr[4] + (wl[1] - wr[1]) * rhoa * ca);
  const double Fl = (gl + 1) / sqrt(2 * gl), Fr = (gr + 1) / sqrt(2 * gr);
  const bool shl = !(pmid <= wl[4]), shr = !(pmid <= wr[4]);
  const double ql = shl ? (1.0 + Fl * (pmid / wl[4] - 1.0)) : 1.0;
  const double qr = shr ? (1.0 + Fr * (pmid / wr[4] - 1.0)) : 1.0;
  const double al = wl[1] - cl * ql;
  const double ar = wr[1] + cr * qr;
  const double bp = ar > 0.0 ? ar : 0.0;
  const double bm = al < 0.0 ? al : 0.0;
  const double vxl = wl[1] - bm, vxr = wr[1] - bp;
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
  const bool split = (bp != bm);
  const double tmp = split ? 0.5 * (bp + bm) / (bp - bm) : 0.0;
  // ---- reverse: flx = 0.5 (fl + fr) + (fl - fr) tmp
  double tmpb = 0, flb[5], frb[5];
  for (int n = 0; n < 5; ++n) {
    flb[n] = flxb[n] * (0.5 + tmp);
    frb[n] = flxb[n] * (0.5 - tmp);
    tmpb += flxb[n] * (fl[n] - fr[n]);
  }
  double bpb = 0, bmb = 0;
  if (split) {
    const double iden = 1.0 / (bp - bm);
    // tmp = 0.5 (bp + bm) / den
    bpb += tmpb * ((0.5 - tmp) * iden);
    bmb += tmpb * ((0.5 + tmp) * iden);
  }
  double vxlb = 0, vxrb = 0, elb = 0, erb = 0;
  wlb[0] += flb[0] * vxl + flb[1] * wl[1] * vxl + flb[2] * wl[2] * vxl + flb[3] * wl[3] * vxl;
  wrb[0] += frb[0] * vxr + frb[1] * wr[1] * vxr + frb[2] * wr[2] * vxr + frb[3] * wr[3] * vxr;
  vxlb += flb[0] * wl[0] + flb[1] * wl[0] * wl[1] + flb[2] * wl[0] * wl[2] +
          flb[3] * wl[0] * wl[3] + flb[4] * el;
  vxrb += frb[0] * wr[0] + frb[1] * wr[0] * wr[1] + frb[2] * wr[0] * wr[2] +
          frb[3] * wr[0] * wr[3] + frb[4] * er;
  wlb[1] += flb[1] * wl[0] * vxl + flb[4] * wl[4];
  wrb[1] += frb[1] * wr[0] * vxr + frb[4] * wr[4];
  wlb[2] += flb[2] * wl[0] * vxl;
  wrb[2] += frb[2] * wr[0] * vxr;
  wlb[3] += flb[3] * wl[0] * vxl;
  wrb[3] += frb[3] * wr[0] * vxr;
  wlb[4] += flb[1] + flb[4] * wl[1];
  wrb[4] += frb[1] + frb[4] * wr[1];
  elb += flb[4] * vxl;
  erb += frb[4] * vxr;
  wlb[1] += vxlb;
  bmb -= vxlb;
  wrb[1] += vxrb;
  bpb -= vxrb;
  double alb = (al < 0.0) ? bmb : 0.0;
  double arb = (ar > 0.0) ? bpb : 0.0;
  double clb = 0, crb = 0, qlb = 0, qrb = 0;
  wlb[1] += alb;
  clb -= alb * ql;
  qlb -= alb * cl;
  wrb[1] += arb;
  crb += arb * qr;
  qrb += arb * cr;
  double pmidb = 0, Flb = 0, Frb = 0;
  if (shl) {
    const double ip = 1.0 / wl[4];
    Flb += qlb * (pmid * ip - 1.0);
    pmidb += qlb * Fl * ip;
    wlb[4] -= qlb * Fl * pmid * ip * ip;
  }
  if (shr) {
    const double ip = 1.0 / wr[4];
    Frb += qrb * (pmid * ip - 1.0);
    pmidb += qrb * Fr * ip;
    wrb[4] -= qrb * Fr * pmid * ip * ip;
  }
  // F = (g+1)/sqrt(2g): dF/dg = (1 - (g+1)/(2g)) / sqrt(2g), only through a shock
  double glb = 0.0, grb = 0.0;
  if (shl) {
    const double ir = 1.0 / sqrt(2 * gl);
    glb = Flb * ((1.0 - (gl + 1) * (0.5 / gl)) * ir);
  }
  if (shr) {
    const double ir = 1.0 / sqrt(2 * gr);
    grb = Frb * ((1.0 - (gr + 1) * (0.5 / gr)) * ir);
  }
  // pmid = .5 (pl + pr + (ul - ur) rhoa ca)
  wlb[4] += .5 * pmidb;
  wrb[4] += .5 * pmidb;
  const double du = wl[1] - wr[1];
  const double dub = .5 * pmidb * rhoa * ca;
  wlb[1] += dub;
  wrb[1] -= dub;
  const double rhoab = .5 * pmidb * du * ca;
  const double cab = .5 * pmidb * du * rhoa;
  wlb[0] += .5 * rhoab;
  wrb[0] += .5 * rhoab;
  clb += .5 * cab;
  crb += .5 * cab;
  double iglb = elb * wl[4], igrb = erb * wr[4];
  wlb[4] += elb * igl;
  wrb[4] += erb * igr;
  wlb[0] += elb * 0.5 * ksql;
  wrb[0] += erb * 0.5 * ksqr;
  for (int k = 1; k <= 3; ++k) {
    wlb[k] += elb * (0.5 * wl[0]) * 2.0 * wl[k];
    wrb[k] += erb * (0.5 * wr[0]) * 2.0 * wr[k];
  }
  sound_speed_adj(gl, igl, wl[0], wl[4], cl, clb, glb, iglb, wlb[0], wlb[4]);
  sound_speed_adj(gr, igr, wr[0], wr[4], cr, crb, grb, igrb, wrb[0], wrb[4]);
  iflb[1] += -iglb * igl * igl;
  ifrb[1] += -igrb * igr * igr;
  iflb[0] += glb;
  ifrb[0] += grb;
}

template <int RS>
APH_HD void riemann_adj(const double *wl, const double *wr, const double *ifl,
                        const double *ifr, const double *flxb, double *wlb, double *wrb,
                        double *iflb, double *ifrb) {
  if constexpr (RS == RS_HLLC)
    hllc_adj<false>(wl, wr, ifl, ifr, flxb, wlb, wrb, iflb, ifrb);
  else if constexpr (RS == RS_LHLLC)
    hllc_adj<true>(wl, wr, ifl, ifr, flxb, wlb, wrb, iflb, ifrb);
  else
    hlle_adj(wl, wr, ifl, ifr, flxb, wlb, wrb, iflb, ifrb);
}

// ------------------------------------------------------------------ ConsToPrim
// u: raw conserved state (before floors).  wb[NV], gmb[2]: cotangents of the primitive
// fields and of (gamma_c, gamma_e).  Adds the conserved-variable cotangent to ub.
// A floor that fires makes the floored quantity a constant (zero derivative).
template <int NV>
APH_HD void cons_to_prim_adj(const double *u, const aph_eos &eos, const double *wb,
                             const double *gmb, double *ub) {
  const double gm1e = eos.gm1;
  const bool dfl = !(u[0] > eos.dfloor);
  const double u_d = dfl ? eos.dfloor : u[0];
  const double di = 1.0 / u_d;
  const double S = (SQR(u[1]) + SQR(u[2]) + SQR(u[3]));
  const double e_k = 0.5 * di * S;
  const double u_e = u[4];
  const double sie = (u_e - e_k) / u_d;
  const double praw = gm1e * u_d * sie;
  const bool psmall = !(EOS_SMALL < praw);  // mx(SMALL, praw) picks SMALL
  const double w_p0 = psmall ? EOS_SMALL : praw;
  const bool pfl = !(w_p0 > eos.pfloor);
  double u_db = 0, u_eb = 0, e_kb = 0, dib = 0, sieb = 0;
  if (!pfl) {
    const double w_p = w_p0;
    const double g = gm1e * (gm1e + 1.0);
    const double braw = g * u_d * sie;
    const bool bsmall = !(EOS_SMALL < braw);
    const double bulk = bsmall ? EOS_SMALL : braw;
    const double ee = u_e - e_k;
    // gamma_e = w_p / ee + 1 (the reverse's quotients share one reciprocal per
    // denominator; the forward values above, which pick the floor branches, are exact)
    const double iee = 1.0 / ee, iwp = 1.0 / w_p;
    double w_pb = gmb[1] * iee;
    const double eeb = -w_pb * w_p * iee;
    u_eb += eeb;
    e_kb -= eeb;
    // gamma_c = bulk / w_p
    const double bulkb = gmb[0] * iwp;
    w_pb -= bulkb * bulk * iwp;
    if (!bsmall) {
      u_db += bulkb * g * sie;
      sieb += bulkb * g * u_d;
    }
    w_pb += wb[4];
    if (!psmall) {
      u_db += w_pb * gm1e * sie;
      sieb += w_pb * gm1e * u_d;
    }
    const double siebd = sieb * di;
    u_eb += siebd;
    e_kb -= siebd;
    u_db -= siebd * sie;
  }
  dib += e_kb * 0.5 * S;
  for (int k = 1; k <= 3; ++k) {
    ub[k] += e_kb * (0.5 * di) * 2.0 * u[k];
    ub[k] += wb[k] * di;
    dib += wb[k] * u[k];
  }
  for (int n = 5; n < NV; ++n) {
    ub[n] += wb[n] * di;
    dib += wb[n] * u[n];
  }
  u_db += wb[0];
  u_db -= dib * di * di;
  if (!dfl) ub[0] += u_db;
  ub[4] += u_eb;
}

}  // namespace aph
