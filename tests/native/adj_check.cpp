// Host build of the product's point and adjoint functions (apophis_amd/csrc/aph_point.hpp,
// aph_adjoint.hpp) with C entry points, so CPU tests can check every adjoint against
// central finite differences of its forward function.  Test infrastructure only.
#include "../../apophis_amd/csrc/aph_adjoint.hpp"

using namespace aph;

extern "C" {
void chk_ppm(const double *q, double *out) { ppm(q[0], q[1], q[2], q[3], q[4], out[0], out[1]); }
void chk_ppm_adj(const double *q, const double *ob, double *qb) { ppm_adj(q, ob[0], ob[1], qb); }
void chk_plm(const double *q, double *out) { plm(q[0], q[1], q[2], out[0], out[1]); }
void chk_plm_adj(const double *q, const double *ob, double *qb) { plm_adj(q, ob[0], ob[1], qb); }
// x = (wl[5], wr[5], ifl[2], ifr[2]) -> flux[5]
void chk_riemann(int kind, const double *x, double *f) {
  if (kind == RS_HLLC)
    riemann<RS_HLLC>(x, x + 5, x + 10, x + 12, f);
  else if (kind == RS_LHLLC)
    riemann<RS_LHLLC>(x, x + 5, x + 10, x + 12, f);
  else if (kind == RS_HLLE)
    riemann<RS_HLLE>(x, x + 5, x + 10, x + 12, f);
  else
    riemann<RS_NONE>(x, x + 5, x + 10, x + 12, f);
}
void chk_riemann_adj(int kind, const double *x, const double *fb, double *xb) {
  if (kind == RS_HLLC)
    hllc_adj<false>(x, x + 5, x + 10, x + 12, fb, xb, xb + 5, xb + 10, xb + 12);
  else if (kind == RS_LHLLC)
    hllc_adj<true>(x, x + 5, x + 10, x + 12, fb, xb, xb + 5, xb + 10, xb + 12);
  else
    hlle_adj(x, x + 5, x + 10, x + 12, fb, xb, xb + 5, xb + 10, xb + 12);
}
// The PPM march of the device kernels on one field along a pencil q[0..n): the limited
// value of each face formed once and shared by the two cells beside it, each cell's second
// difference formed once and carried (ppm_d2), div6_fast when fast != 0.  Writes
// (ql at face i+1, qr at face i) of cells i = 2 .. n-3 to out[2*(i-2) ..].
void chk_ppm_march(const double *q, long n, int fast, double *out) {
  if (n < 5) return;
  auto face = [&](long i, double dm, double d0) {  // face i-1/2: cells i-2 .. i+1
    return fast ? ppm_face<true>(q[i - 2], q[i - 1], q[i], q[i + 1], dm, d0)
                : ppm_face<false>(q[i - 2], q[i - 1], q[i], q[i + 1], dm, d0);
  };
  double d2c = ppm_d2(q[1], q[2], q[3]);                        // cell 2
  double fv = face(2, ppm_d2(q[0], q[1], q[2]), d2c);            // face 3/2
  for (long m = 2; m <= n - 3; ++m) {
    const double d2n = ppm_d2(q[m], q[m + 1], q[m + 2]);        // cell m+1
    const double up = face(m + 1, d2c, d2n);                     // face m+1/2
    ppm_cell(q[m - 2], q[m - 1], q[m], q[m + 1], q[m + 2], ppm_d2(q[m - 2], q[m - 1], q[m]),
             d2c, d2n, fv, up, out[2 * (m - 2)], out[2 * (m - 2) + 1]);
    fv = up;
    d2c = d2n;
  }
}
// u[6] -> (w[6], gamma[2]) ; eos = (gm1, cv, dfloor, pfloor)
void chk_c2p(const double *u, const double *e, double *out) {
  aph_eos eos{e[0], e[1], e[2], e[3]};
  double uu[6] = {u[0], u[1], u[2], u[3], u[4], u[5]};
  cons_to_prim_pt<6>(uu, eos, out, out + 6);
}
void chk_c2p_adj(const double *u, const double *e, const double *ob, double *ub) {
  aph_eos eos{e[0], e[1], e[2], e[3]};
  cons_to_prim_adj<6>(u, eos, ob, ob + 6, ub);
}
}
extern "C" void chk_ppm_shared(const double *q, double *out) {
  const double lo = ppm_face(q[0], q[1], q[2], q[3]);
  const double hi = ppm_face(q[1], q[2], q[3], q[4]);
  ppm_cell(q[0], q[1], q[2], q[3], q[4], lo, hi, out[0], out[1]);
}
extern "C" void chk_ppm_face(const double *q, long n, int fast, double *out) {
  for (long i = 0; i < n; ++i)
    out[i] = fast ? ppm_face<true>(q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3])
                  : ppm_face<false>(q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]);
}
extern "C" void chk_div6_fast(const double *x, long n, double *out) {
  for (long i = 0; i < n; ++i) out[i] = div6_fast(x[i]);
}
// u[5] floored in place by floor_cons_pt (out[0..4]) and by cons_to_prim_pt (out[5..9])
extern "C" void chk_floor(const double *u, const double *e, double *out) {
  aph_eos eos{e[0], e[1], e[2], e[3]};
  double a[5] = {u[0], u[1], u[2], u[3], u[4]}, b[5] = {u[0], u[1], u[2], u[3], u[4]};
  double w[5], g[2];
  floor_cons_pt<5>(a, eos);
  cons_to_prim_pt<5>(b, eos, w, g);
  for (int i = 0; i < 5; ++i) {
    out[i] = a[i];
    out[5 + i] = b[i];
  }
}
// Face-shared PPM adjoint of one cell: its two faces (taped), the cell part, then each
// face's cotangent through ppm_face_adj.  Also returns the taped face values (out[5..6]).
extern "C" void chk_ppm_split_adj(const double *q, const double *ob, double *qb) {
  PpmFaceTape tlo, thi;
  const double lo = ppm_face_tape<false>(q[0], q[1], q[2], q[3], tlo);
  const double hi = ppm_face_tape<false>(q[1], q[2], q[3], q[4], thi);
  double lob = 0.0, hib = 0.0;
  ppm_cell_adj(q, lo, hi, ob[0], ob[1], qb, lob, hib);
  ppm_face_adj(tlo, lob, qb);
  ppm_face_adj(thi, hib, qb + 1);
  qb[5] = lo;
  qb[6] = hi;
}
