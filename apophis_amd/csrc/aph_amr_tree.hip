// aph_amr_tree.hip -- the multilevel mesh tree on the host: from the leaf list of a block
// set (each block's level and logical position in a base lattice whose axes are periodic
// or end in physical faces) to the op lists that aph_amr_plan_create takes (cross-level
// ghost values and flux fixes).
//
// The reference gets these from Parthenon's mesh tree (neighbour finding across levels for
// the boundary exchange and the flux correction, apophis_driver.cpp:68,74-76,94).  Here the
// tree is a hash of leaves and of refined nodes (the ancestors of leaves); every ghost cell
// with no same-level neighbour in its direction becomes one op:
//   * over a refined node: MEAN of the 2^ndim children of the ghost cell (one level finer);
//   * over a coarser leaf: PROLONG from the coarse cell holding it, the slope stencil's
//     neighbours taken at the coarse level (a coarse cell, or the mean of its children);
// and every coarse face cell on a block face whose neighbour node is refined becomes one
// flux fix (the mean of the 2^(ndim-1) fine faces across it).  Ghost cells beyond a
// physical domain face get no op (the boundary conditions fill them); a slope neighbour
// beyond one is the coarse level's boundary image: the coarse cell itself (outflow) or
// its mirror image (reflect, flagged in the op's kind).  Lists are emitted in block
// order, ghost cells k, j, i ascending, children (cz, cy, cx) with x fastest, transverse
// axes ascending with the first slowest -- the order the Python statement of the same
// builder (apophis_amd/amr.py amr_plan) uses, so both give identical lists.
// Host code only; no device work.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/apophis_hydro.h"

namespace {

constexpr int kMaxLevel = 30;
constexpr int64_t kMaxCoord = (int64_t)1 << 19;  // nodes per axis the key can hold

struct BuildError {
  char msg[256];
};

[[noreturn]] void fail(const char *fmt, ...) {
  BuildError e;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(e.msg, sizeof(e.msg), fmt, ap);
  va_end(ap);
  throw e;
}

inline uint64_t key(int lv, int64_t x, int64_t y, int64_t z) {
  return ((uint64_t)lv << 58) | ((uint64_t)x << 38) | ((uint64_t)y << 19) | (uint64_t)z;
}

// Domain face codes (-x1, +x1, -x2, ...): NULL or APH_BC_PERIODIC on both faces of an
// axis make it periodic; otherwise each face is APH_BC_OUTFLOW or APH_BC_REFLECT.
void face_codes(const int32_t *bc, int nd, int per[3], int code[6]) {
  for (int d = 0; d < 3; ++d) {
    per[d] = 1;
    code[2 * d] = code[2 * d + 1] = APH_BC_PERIODIC;
    if (!bc || d >= nd) continue;
    const int lo = bc[2 * d], hi = bc[2 * d + 1];
    if (lo == APH_BC_PERIODIC && hi == APH_BC_PERIODIC) continue;
    if (lo == APH_BC_PERIODIC || hi == APH_BC_PERIODIC)
      fail("axis %d: a periodic face needs a periodic opposite face", d);
    for (int f : {lo, hi})
      if (f != APH_BC_OUTFLOW && f != APH_BC_REFLECT)
        fail("axis %d: face code %d is not periodic, outflow or reflect", d, f);
    per[d] = 0;
    code[2 * d] = lo;
    code[2 * d + 1] = hi;
  }
}

struct Tree {
  int nd = 0, ng = 0, nvar = 0, nb = 0;
  int nx[3] = {1, 1, 1}, nbase[3] = {1, 1, 1}, E[3] = {1, 1, 1};
  int per[3] = {1, 1, 1}, code[6] = {0, 0, 0, 0, 0, 0};
  const aph_block *blk = nullptr;
  const int32_t *level = nullptr;
  const int64_t *loc = nullptr;
  std::unordered_map<uint64_t, int> leaf;
  std::unordered_set<uint64_t> refined;
  std::vector<int64_t> base;  // nvar * cell_base per block
  int64_t ncell = 0;
  std::vector<aph_amr_ref> refs;

  int64_t nlat(int lv, int d) const { return d < nd ? (int64_t)nbase[d] << lv : 1; }
  // global cell index g of axis d at level lv, wrapped on a periodic axis (elsewhere it
  // stays as it is and may lie beyond the domain)
  int64_t wrap(int64_t g, int lv, int d) const {
    if (d >= nd) return 0;
    if (!per[d]) return g;
    const int64_t n = nlat(lv, d) * nx[d];
    return ((g % n) + n) % n;
  }
  bool inside(int lv, const int64_t G[3]) const {
    for (int d = 0; d < nd; ++d)
      if (!per[d] && (G[d] < 0 || G[d] >= nlat(lv, d) * nx[d])) return false;
    return true;
  }
  bool is_refined(int lv, const int64_t node[3]) const {
    return refined.count(key(lv, node[0], node[1], node[2])) != 0;
  }
  // component-0 offset of block b's cell at ghost-inclusive local index l
  int64_t cell(int b, const int64_t l[3]) const {
    return base[b] + (l[2] * E[1] + l[1]) * E[0] + l[0];
  }
  // the leaf of level lv holding global cell G (wrapped): its block and local index
  bool interior(int lv, const int64_t G[3], int &b, int64_t l[3]) const {
    int64_t W[3], node[3];
    for (int d = 0; d < 3; ++d) {
      W[d] = wrap(G[d], lv, d);
      node[d] = d < nd ? W[d] / nx[d] : 0;
    }
    if (!inside(lv, W)) return false;
    auto it = leaf.find(key(lv, node[0], node[1], node[2]));
    if (it == leaf.end()) return false;
    b = it->second;
    for (int d = 0; d < 3; ++d) l[d] = W[d] - node[d] * nx[d] + (d < nd ? ng : 0);
    return true;
  }
  aph_amr_val add(const int64_t *offs, int n) {
    aph_amr_val v;
    v.first = (int32_t)refs.size();
    v.count = n;
    for (int q = 0; q < n; ++q) refs.push_back(aph_amr_ref{offs[q], ncell});
    if (refs.size() > (size_t)INT32_MAX) fail("too many cell references");
    return v;
  }
  // level-lv value of global cell I: the leaf cell, or the mean of its children
  aph_amr_val value(int lv, const int64_t I[3]) {
    int b;
    int64_t l[3];
    if (interior(lv, I, b, l)) {
      const int64_t c = cell(b, l);
      return add(&c, 1);
    }
    int64_t W[3];
    for (int d = 0; d < 3; ++d) W[d] = wrap(I[d], lv, d);
    if (!inside(lv, W)) fail("cell of level %d beyond the domain", lv);
    int64_t kids[8];
    int nk = 0;
    for (int cz = 0; cz < (nd == 3 ? 2 : 1); ++cz)
      for (int cy = 0; cy < 2; ++cy)
        for (int cx = 0; cx < 2; ++cx) {
          const int64_t F[3] = {2 * W[0] + cx, 2 * W[1] + cy, nd == 3 ? 2 * W[2] + cz : 0};
          if (!interior(lv + 1, F, b, l))
            fail("cell (%lld, %lld, %lld) of level %d is neither a leaf cell nor the parent "
                 "of leaf cells (2:1 balance)", (long long)W[0], (long long)W[1],
                 (long long)W[2], lv);
          kids[nk++] = cell(b, l);
        }
    return add(kids, nk);
  }
};

void setup(Tree &t, const aph_block *blocks, int32_t nblocks, int32_t ndim, int32_t nvar,
           const int32_t nbase[3], const int32_t *level, const int64_t *loc,
           const int32_t *bc) {
  if (!blocks || nblocks < 1 || ndim < 2 || ndim > 3 || nvar < 1 || !nbase || !level || !loc)
    fail("bad arguments");
  face_codes(bc, ndim, t.per, t.code);
  t.nd = ndim;
  t.nvar = nvar;
  t.nb = nblocks;
  t.blk = blocks;
  t.level = level;
  t.loc = loc;
  t.ng = blocks[0].ng;
  for (int d = 0; d < 3; ++d) {
    t.nx[d] = d < ndim ? blocks[0].nx[d] : 1;
    t.nbase[d] = d < ndim ? nbase[d] : 1;
    if (t.nbase[d] < 1) fail("nbase[%d] = %d", d, nbase[d]);
    t.E[d] = t.nx[d] + (d < ndim ? 2 * t.ng : 0);
  }
  t.ncell = (int64_t)t.E[0] * t.E[1] * t.E[2];
  for (int b = 0; b < nblocks; ++b) {
    const aph_block &k = blocks[b];
    if (k.ng != t.ng) fail("block %d: ng %d differs from block 0", b, k.ng);
    for (int d = 0; d < ndim; ++d)
      if (k.nx[d] != t.nx[d] || k.nx[d] % 2 || k.nx[d] < 2 * k.ng)
        fail("block %d: nx must be even, at least 2*ng and the same for every block", b);
    if (level[b] < 0 || level[b] > kMaxLevel) fail("block %d: level %d", b, level[b]);
    int64_t c[3];
    for (int d = 0; d < 3; ++d) {
      c[d] = d < ndim ? loc[3 * b + d] : 0;
      if (c[d] < 0 || c[d] >= t.nlat(level[b], d) || c[d] >= kMaxCoord)
        fail("block %d: logical position outside level %d", b, level[b]);
    }
    if (!t.leaf.emplace(key(level[b], c[0], c[1], c[2]), b).second)
      fail("block %d: two blocks at the same node", b);
    t.base.push_back((int64_t)nvar * k.cell_base);
  }
  // refined nodes: the ancestors of the leaves
  for (int b = 0; b < nblocks; ++b) {
    int64_t c[3] = {t.nd > 0 ? loc[3 * b] : 0, t.nd > 1 ? loc[3 * b + 1] : 0,
                    t.nd > 2 ? loc[3 * b + 2] : 0};
    for (int lv = level[b] - 1; lv >= 0; --lv) {
      for (int d = 0; d < t.nd; ++d) c[d] >>= 1;
      const uint64_t k = key(lv, c[0], c[1], c[2]);
      if (t.leaf.count(k)) fail("block %d lies inside block %d", b, t.leaf[k]);
      if (!t.refined.insert(k).second) break;
    }
  }
}

void build(Tree &t, std::vector<aph_amr_op> &ops, std::vector<aph_flux_fix> &fixes) {
  const int nd = t.nd, g = t.ng;
  for (int b = 0; b < t.nb; ++b) {
    const int lv = t.level[b];
    const int64_t lc[3] = {t.loc[3 * b], nd > 1 ? t.loc[3 * b + 1] : 0,
                           nd > 2 ? t.loc[3 * b + 2] : 0};
    const int32_t *nbr = t.blk[b].nbr;
    for (int k = 0; k < t.E[2]; ++k)
      for (int j = 0; j < t.E[1]; ++j)
        for (int i = 0; i < t.E[0]; ++i) {
          const int64_t l[3] = {i, j, k};
          int o[3] = {0, 0, 0};
          for (int d = 0; d < nd; ++d) o[d] = l[d] < g ? -1 : (l[d] >= g + t.nx[d] ? 1 : 0);
          if (!o[0] && !o[1] && !o[2]) continue;
          if (nbr[(o[2] + 1) * 9 + (o[1] + 1) * 3 + (o[0] + 1)] >= 0) continue;
          int64_t Gw[3], node[3];
          for (int d = 0; d < 3; ++d) {
            Gw[d] = d < nd ? t.wrap(lc[d] * t.nx[d] + l[d] - g, lv, d) : 0;
            node[d] = d < nd ? Gw[d] / t.nx[d] : 0;
          }
          if (!t.inside(lv, Gw)) continue;  // beyond a physical face: boundary conditions
          aph_amr_op op;
          std::memset(&op, 0, sizeof(op));
          op.dst = t.cell(b, l);
          op.dst_stride = t.ncell;
          if (t.is_refined(lv, node)) {  // over a finer region: restriction
            op.kind = APH_AMR_MEAN;
            op.val[0] = t.value(lv, Gw);
          } else {  // over a coarser leaf: prolongation
            const int64_t I[3] = {Gw[0] / 2, nd > 1 ? Gw[1] / 2 : 0, nd > 2 ? Gw[2] / 2 : 0};
            int cb;
            int64_t cl[3];
            if (lv == 0 || !t.interior(lv - 1, I, cb, cl))
              fail("block %d: ghost cell (%d, %d, %d) is over neither a finer region nor a "
                   "coarser leaf (2:1 balance)", b, i, j, k);
            op.kind = APH_AMR_PROLONG;
            op.val[0] = t.value(lv - 1, I);
            for (int d = 0; d < nd; ++d) {
              op.sign[d] = Gw[d] % 2 == 0 ? -1 : 1;
              for (int side = 0; side < 2; ++side) {
                int64_t J[3] = {I[0], I[1], I[2]};
                J[d] += side ? 1 : -1;
                const int slot = 1 + 2 * d + side;
                if (t.inside(lv - 1, J)) {
                  op.val[slot] = t.value(lv - 1, J);
                } else {  // beyond a physical face: the coarse cell's boundary image
                  op.val[slot] = t.value(lv - 1, I);
                  if (t.code[2 * d + side] == APH_BC_REFLECT)
                    op.kind |= 1 << (APH_AMR_MIRROR_BIT + slot - 1);
                }
              }
            }
          }
          ops.push_back(op);
        }
  }
  for (int b = 0; b < t.nb; ++b) {
    const int lv = t.level[b];
    const int64_t lc[3] = {t.loc[3 * b], nd > 1 ? t.loc[3 * b + 1] : 0,
                           nd > 2 ? t.loc[3 * b + 2] : 0};
    for (int d = 0; d < nd; ++d)
      for (int side = 0; side < 2; ++side) {
        if (t.blk[b].bc[2 * d + side] != APH_BC_AMR) continue;
        int64_t across[3] = {lc[0], lc[1], lc[2]};
        across[d] = ((across[d] + (side ? 1 : -1)) % t.nlat(lv, d) + t.nlat(lv, d)) %
                    t.nlat(lv, d);
        if (!t.is_refined(lv, across)) continue;  // the other side is coarser
        int tr[2], ntr = 0;
        for (int q = 0; q < nd; ++q)
          if (q != d) tr[ntr++] = q;
        const int64_t Fd = 2 * (lc[d] * t.nx[d] + (side ? t.nx[d] : 0)) - (side ? 0 : 1);
        const int n0 = t.nx[tr[0]], n1 = ntr > 1 ? t.nx[tr[1]] : 1;
        for (int c0 = 0; c0 < n0; ++c0)
          for (int c1 = 0; c1 < n1; ++c1) {
            const int tc[2] = {c0, c1};
            int64_t l[3] = {0, 0, 0};
            l[d] = g + (side ? t.nx[d] : 0);
            for (int q = 0; q < ntr; ++q) l[tr[q]] = g + tc[q];
            int64_t cells[4];
            int nc = 0;
            for (int k0 = 0; k0 < 2; ++k0)
              for (int k1 = 0; k1 < (ntr > 1 ? 2 : 1); ++k1) {
                const int kid[2] = {k0, k1};
                int64_t F[3] = {0, 0, 0};
                F[d] = Fd;
                for (int q = 0; q < ntr; ++q)
                  F[tr[q]] = 2 * (lc[tr[q]] * t.nx[tr[q]] + tc[q]) + kid[q];
                int fb;
                int64_t fl[3];
                if (!t.interior(lv + 1, F, fb, fl))
                  fail("block %d face %d: fine face outside the finer level", b, 2 * d + side);
                if (!side) fl[d] += 1;  // upper face of the last fine cell below
                cells[nc++] = t.cell(fb, fl);
              }
            aph_flux_fix fx;
            std::memset(&fx, 0, sizeof(fx));
            fx.dst = t.cell(b, l);
            fx.dst_stride = t.ncell;
            fx.dir = d;
            fx.val = t.add(cells, nc);
            fixes.push_back(fx);
          }
      }
  }
}

}  // namespace

extern "C" int aph_amr_build(const aph_block *blocks, int32_t nblocks, int32_t ndim,
                             int32_t nvar, const int32_t nbase[3], const int32_t *level,
                             const int64_t *loc, const int32_t *bc, aph_amr_op *ops,
                             int64_t *nops, aph_flux_fix *fixes, int64_t *nfixes,
                             aph_amr_ref *refs, int64_t *nrefs, char *err, size_t err_len) {
  auto report = [&](int code, const char *msg) {
    if (err && err_len) snprintf(err, err_len, "aph_amr_build: %s", msg);
    return code;
  };
  if (!nops || !nfixes || !nrefs) return report(APH_E_INVALID, "null length");
  try {
    Tree t;
    setup(t, blocks, nblocks, ndim, nvar, nbase, level, loc, bc);
    std::vector<aph_amr_op> o;
    std::vector<aph_flux_fix> f;
    build(t, o, f);
    const int64_t cap[3] = {*nops, *nfixes, *nrefs};
    *nops = (int64_t)o.size();
    *nfixes = (int64_t)f.size();
    *nrefs = (int64_t)t.refs.size();
    if ((ops && cap[0] < *nops) || (fixes && cap[1] < *nfixes) || (refs && cap[2] < *nrefs))
      return report(APH_E_INVALID, "output arrays too short");
    if (ops && !o.empty()) std::memcpy(ops, o.data(), sizeof(aph_amr_op) * o.size());
    if (fixes && !f.empty()) std::memcpy(fixes, f.data(), sizeof(aph_flux_fix) * f.size());
    if (refs && !t.refs.empty())
      std::memcpy(refs, t.refs.data(), sizeof(aph_amr_ref) * t.refs.size());
  } catch (const BuildError &e) {
    return report(APH_E_INVALID, e.msg);
  } catch (const std::bad_alloc &) {
    return report(APH_E_NOMEM, "host allocation failed");
  }
  if (err && err_len) err[0] = 0;
  return APH_OK;
}

// ------------------------------------------------------------------ the tree itself
// Leaves are (level, logical position); the refined nodes are their ancestors.  An axis is
// periodic unless the block table is built with physical faces on it.  Children are
// ordered (cz, cy, cx) with x fastest, and the depth-first leaf order walks the base
// lattice z, y, x -- the order of apophis_amd/amr.py AmrMesh.
namespace {

struct Forest {
  int nd = 0;
  int64_t nbase[3] = {1, 1, 1};
  int per[3] = {1, 1, 1};
  std::unordered_set<uint64_t> refined;

  int64_t nlat(int lv, int d) const { return d < nd ? nbase[d] << lv : 1; }
  uint64_t k(int lv, const int64_t *c) const { return key(lv, c[0], c[1], c[2]); }
  bool is_ref(int lv, const int64_t *c) const { return refined.count(k(lv, c)) != 0; }
  // the node at offset o from c (wrapped on periodic axes); false beyond a physical face
  bool shift(int lv, const int64_t *c, const int *o, int64_t *out) const {
    bool in = true;
    for (int d = 0; d < 3; ++d) {
      const int64_t n = nlat(lv, d);
      out[d] = c[d] + o[d];
      if (per[d])
        out[d] = (out[d] % n + n) % n;
      else if (out[d] < 0 || out[d] >= n)
        in = false;
    }
    return in;
  }
  int nchildren() const { return nd == 3 ? 8 : 4; }
  void child(const int64_t *c, int i, int64_t *out) const {  // i in (cz, cy, cx) order
    const int cx = i & 1, cy = (i >> 1) & 1, cz = (i >> 2) & 1;
    out[0] = 2 * c[0] + cx;
    out[1] = 2 * c[1] + cy;
    out[2] = nd == 3 ? 2 * c[2] + cz : 0;
  }
  template <class F>
  void offsets(F f) const {
    for (int oz = (nd == 3 ? -1 : 0); oz <= (nd == 3 ? 1 : 0); ++oz)
      for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox) {
          if (!ox && !oy && !oz) continue;
          const int o[3] = {ox, oy, oz};
          f(o);
        }
  }
  // whether leaf (lv, c) touches a leaf at level lv + 2 or deeper
  bool touches_finer(int lv, const int64_t *c) const {
    bool hit = false;
    offsets([&](const int *o) {
      if (hit) return;
      int64_t p[3];
      if (!shift(lv, c, o, p) || !is_ref(lv, p)) return;
      for (int i = 0; i < nchildren() && !hit; ++i) {
        int64_t ch[3];
        child(p, i, ch);
        bool touches = true;  // the child of p on the side facing (lv, c)
        for (int d = 0; d < nd; ++d)
          if (o[d] != 0 && (ch[d] & 1) != (o[d] < 0 ? 1 : 0)) touches = false;
        if (touches && is_ref(lv + 1, ch)) hit = true;
      }
    });
    return hit;
  }
  template <class F>
  void leaves(F f) const {  // depth-first, base lattice z, y, x
    std::vector<std::pair<int, std::array<int64_t, 3>>> stack;
    for (int64_t z = nbase[2] - 1; z >= 0; --z)
      for (int64_t y = nbase[1] - 1; y >= 0; --y)
        for (int64_t x = nbase[0] - 1; x >= 0; --x) stack.push_back({0, {x, y, z}});
    while (!stack.empty()) {
      auto [lv, c] = stack.back();
      stack.pop_back();
      if (is_ref(lv, c.data())) {
        for (int i = nchildren() - 1; i >= 0; --i) {
          std::array<int64_t, 3> ch;
          child(c.data(), i, ch.data());
          stack.push_back({lv + 1, ch});
        }
      } else {
        f(lv, c.data());
      }
    }
  }
};

void forest_setup(Forest &t, int32_t ndim, const int32_t nbase[3], int32_t n, const int32_t *level,
                  const int64_t *loc) {
  if (ndim < 2 || ndim > 3 || !nbase || n < 1 || !level || !loc) fail("bad arguments");
  t.nd = ndim;
  for (int d = 0; d < 3; ++d) {
    t.nbase[d] = d < ndim ? nbase[d] : 1;
    if (t.nbase[d] < 1) fail("nbase[%d] = %d", d, nbase[d]);
  }
  std::unordered_set<uint64_t> leafset;
  for (int b = 0; b < n; ++b) {
    if (level[b] < 0 || level[b] > kMaxLevel) fail("leaf %d: level %d", b, level[b]);
    int64_t c[3];
    for (int d = 0; d < 3; ++d) {
      c[d] = d < ndim ? loc[3 * b + d] : 0;
      if (c[d] < 0 || c[d] >= t.nlat(level[b], d) || c[d] >= kMaxCoord)
        fail("leaf %d: logical position outside level %d", b, level[b]);
    }
    if (!leafset.insert(t.k(level[b], c)).second) fail("leaf %d: two leaves at one node", b);
  }
  for (int b = 0; b < n; ++b) {
    int64_t c[3] = {loc[3 * b], loc[3 * b + 1], ndim == 3 ? loc[3 * b + 2] : 0};
    for (int lv = level[b] - 1; lv >= 0; --lv) {
      for (int d = 0; d < ndim; ++d) c[d] >>= 1;
      const uint64_t kk = t.k(lv, c);
      if (leafset.count(kk)) fail("leaf %d lies inside another leaf", b);
      if (!t.refined.insert(kk).second) break;
    }
  }
  int64_t nroot = t.nbase[0] * t.nbase[1] * t.nbase[2];
  if ((int64_t)n != nroot + (int64_t)(t.nchildren() - 1) * (int64_t)t.refined.size())
    fail("the leaves do not cover the domain");
}

int emit_leaves(const Forest &t, int32_t *out_level, int64_t *out_loc, int32_t *out_n,
                int32_t cap) {
  int32_t m = 0;
  t.leaves([&](int lv, const int64_t *c) {
    if (out_level && m < cap) {
      out_level[m] = lv;
      for (int d = 0; d < 3; ++d) out_loc[3 * m + d] = c[d];
    }
    ++m;
  });
  *out_n = m;
  return (out_level && m > cap) ? -1 : 0;
}

}  // namespace

extern "C" int aph_amr_regrid_tree(int32_t ndim, const int32_t nbase[3], int32_t max_level,
                                   int32_t n, const int32_t *level, const int64_t *loc,
                                   const int32_t *tags, int32_t *out_level, int64_t *out_loc,
                                   int32_t *out_n, int32_t cap, char *err, size_t err_len) {
  auto report = [&](int code, const char *msg) {
    if (err && err_len) snprintf(err, err_len, "aph_amr_regrid_tree: %s", msg);
    return code;
  };
  if (!out_n || !tags || (out_level && !out_loc)) return report(APH_E_INVALID, "bad arguments");
  try {
    Forest t;
    forest_setup(t, ndim, nbase, n, level, loc);
    std::unordered_map<uint64_t, int> old_leaf;
    for (int b = 0; b < n; ++b) {
      const int64_t c[3] = {loc[3 * b], loc[3 * b + 1], ndim == 3 ? loc[3 * b + 2] : 0};
      old_leaf.emplace(t.k(level[b], c), b);
    }
    // refine the tagged leaves below max_level
    for (int b = 0; b < n; ++b)
      if (tags[b] == 1 && level[b] < max_level) {
        const int64_t c[3] = {loc[3 * b], loc[3 * b + 1], ndim == 3 ? loc[3 * b + 2] : 0};
        t.refined.insert(t.k(level[b], c));
      }
    // 2:1 balance: refine every leaf that touches a leaf two levels finer (a monotone
    // closure, so the order of the sweep does not change the result)
    for (bool changed = true; changed;) {
      changed = false;
      std::vector<std::pair<int, std::array<int64_t, 3>>> add;
      t.leaves([&](int lv, const int64_t *c) {
        if (t.touches_finer(lv, c)) add.push_back({lv, {c[0], c[1], c[2]}});
      });
      for (auto &a : add) changed |= t.refined.insert(t.k(a.first, a.second.data())).second;
    }
    // merges: parents of leaves tagged -1, finest level first, then by position (x, y, z)
    std::vector<std::pair<int, std::array<int64_t, 3>>> groups;
    {
      std::unordered_set<uint64_t> seen;
      for (int b = 0; b < n; ++b)
        if (level[b] > 0 && tags[b] == -1) {
          std::array<int64_t, 3> p = {loc[3 * b] >> 1, loc[3 * b + 1] >> 1,
                                      ndim == 3 ? loc[3 * b + 2] >> 1 : 0};
          if (seen.insert(t.k(level[b] - 1, p.data())).second) groups.push_back({level[b] - 1, p});
        }
    }
    std::sort(groups.begin(), groups.end(), [](const auto &a, const auto &b) {
      if (a.first != b.first) return a.first > b.first;
      return a.second < b.second;
    });
    for (const auto &P : groups) {
      const int lv = P.first + 1;
      bool ok = true;
      for (int i = 0; i < t.nchildren() && ok; ++i) {
        int64_t ch[3];
        t.child(P.second.data(), i, ch);
        auto it = old_leaf.find(t.k(lv, ch));
        if (it == old_leaf.end() || tags[it->second] != -1 || t.is_ref(lv, ch)) ok = false;
      }
      if (!ok) continue;
      const uint64_t pk = t.k(P.first, P.second.data());
      t.refined.erase(pk);
      if (t.touches_finer(P.first, P.second.data())) t.refined.insert(pk);
    }
    if (emit_leaves(t, out_level, out_loc, out_n, cap))
      return report(APH_E_INVALID, "output arrays too short");
  } catch (const BuildError &e) {
    return report(APH_E_INVALID, e.msg);
  } catch (const std::bad_alloc &) {
    return report(APH_E_NOMEM, "host allocation failed");
  }
  if (err && err_len) err[0] = 0;
  return APH_OK;
}

extern "C" int aph_amr_block_table(int32_t ndim, const int32_t nbase[3], const int32_t nx[3],
                                   int32_t ng, const double xmin[3], const double xmax[3],
                                   const int32_t *bc, int32_t mesh_id, int32_t n,
                                   const int32_t *level, const int64_t *loc, aph_block *out,
                                   char *err, size_t err_len) {
  auto report = [&](int code, const char *msg) {
    if (err && err_len) snprintf(err, err_len, "aph_amr_block_table: %s", msg);
    return code;
  };
  if (!nx || !xmin || !xmax || !out) return report(APH_E_INVALID, "bad arguments");
  try {
    Forest t;
    forest_setup(t, ndim, nbase, n, level, loc);
    int code[6];
    face_codes(bc, ndim, t.per, code);
    std::unordered_map<uint64_t, int> leaf;
    for (int b = 0; b < n; ++b) {
      const int64_t c[3] = {loc[3 * b], loc[3 * b + 1], ndim == 3 ? loc[3 * b + 2] : 0};
      leaf.emplace(t.k(level[b], c), b);
    }
    for (int d = 0; d < ndim; ++d)
      if (nx[d] < 2 * ng || nx[d] % 2) fail("nx[%d] must be even and at least 2*ng", d);
    for (int b = 0; b < n; ++b) {
      const int lv = level[b];
      const int64_t c[3] = {loc[3 * b], loc[3 * b + 1], ndim == 3 ? loc[3 * b + 2] : 0};
      // 2:1 balance: every neighbour node is a leaf, a refined node whose children facing
      // this leaf are leaves, or lies inside a leaf one level coarser
      t.offsets([&](const int *o) {
        int64_t p[3];
        if (!t.shift(lv, c, o, p)) return;  // beyond a physical face
        if (leaf.count(t.k(lv, p))) return;
        if (t.is_ref(lv, p)) {
          for (int i = 0; i < t.nchildren(); ++i) {
            int64_t ch[3];
            t.child(p, i, ch);
            bool touches = true;
            for (int d = 0; d < ndim; ++d)
              if (o[d] != 0 && (ch[d] & 1) != (o[d] < 0 ? 1 : 0)) touches = false;
            if (touches && t.is_ref(lv + 1, ch))
              fail("leaf %d touches level %d (2:1 balance)", b, lv + 2);
          }
          return;
        }
        int64_t q[3] = {p[0] >> 1, p[1] >> 1, ndim == 3 ? p[2] >> 1 : 0};
        if (lv == 0 || !leaf.count(t.k(lv - 1, q)))
          fail("leaf %d touches a level below %d (2:1 balance)", b, lv - 1);
      });
      aph_block &r = out[b];
      std::memset(&r, 0, sizeof(r));
      double dx[3];
      for (int d = 0; d < 3; ++d) {
        const int64_t nxd = d < ndim ? nx[d] : 1;
        const int64_t cells = t.nbase[d] * nxd * (d < ndim ? (int64_t)1 << lv : 1);
        dx[d] = (xmax[d] - xmin[d]) / (double)cells;
        r.nx[d] = (int32_t)nxd;
      }
      r.ng = ng;
      r.cell_base = 0;
      for (int d = 0; d < 3; ++d) r.dx[d] = dx[d];
      r.area[0] = dx[1] * dx[2];
      r.area[1] = dx[0] * dx[2];
      r.area[2] = dx[0] * dx[1];
      r.vol = dx[0] * dx[1] * dx[2];
      for (int d = 0; d < 3; ++d)
        r.xmin[d] = xmin[d] + (d < ndim ? (double)(c[d] * (int64_t)nx[d]) * dx[d] : 0.0);
      r.mesh_id = mesh_id;
      for (int q = 0; q < 27; ++q) r.nbr[q] = -1;
      t.offsets([&](const int *o) {
        int64_t p[3];
        if (!t.shift(lv, c, o, p)) return;  // beyond a physical face: nbr stays -1
        auto it = leaf.find(t.k(lv, p));
        r.nbr[(o[2] + 1) * 9 + (o[1] + 1) * 3 + (o[0] + 1)] = it == leaf.end() ? -1 : it->second;
      });
      for (int f = 0; f < 6; ++f) r.bc[f] = APH_BC_OUTFLOW;  // faces with a same-level neighbour
      for (int d = 0; d < ndim; ++d)
        for (int side = 0; side < 2; ++side) {
          int o[3] = {0, 0, 0};
          o[d] = side ? 1 : -1;
          int64_t p[3];
          if (!t.shift(lv, c, o, p))
            r.bc[2 * d + side] = code[2 * d + side];  // a physical domain face
          else if (r.nbr[(o[2] + 1) * 9 + (o[1] + 1) * 3 + (o[0] + 1)] < 0)
            r.bc[2 * d + side] = APH_BC_AMR;
        }
    }
  } catch (const BuildError &e) {
    return report(APH_E_INVALID, e.msg);
  } catch (const std::bad_alloc &) {
    return report(APH_E_NOMEM, "host allocation failed");
  }
  if (err && err_len) err[0] = 0;
  return APH_OK;
}

// ------------------------------------------------------------------ interior remap
// The interiors of a regridded mesh from the old one: a kept leaf copies its cells, any
// other new leaf takes its box of the old mesh's level values -- a leaf's cells, the mean
// of a refined node's children ((dz, dy, dx) order, scaled by 2^-d), or where a coarser
// leaf covers the box, c + sum_d (sign_d / 4) minmod(c - c_-d, c_+d - c) of the level above
// (the ghost exchange's rule), axes in order -- the same operations in the same order as
// apophis_amd/amr.py's sampler, so the bits agree.  Fields hold blocks of one size, in
// the pack layout.
namespace {

struct Remap {
  int nd = 0, nv = 0, ng = 0;
  int64_t nbase[3] = {1, 1, 1}, nx[3] = {1, 1, 1}, E[3] = {1, 1, 1}, ncell = 0;
  const double *u = nullptr;
  std::unordered_map<uint64_t, int> leaf;
  std::unordered_set<uint64_t> refined;

  int64_t nlat(int L, int d) const { return d < nd ? nbase[d] << L : 1; }
  static int64_t fdiv(int64_t a, int64_t b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
  double cell(int b, int v, int64_t i, int64_t j, int64_t k) const {
    const int64_t g0 = ng, g1 = nd > 1 ? ng : 0, g2 = nd > 2 ? ng : 0;
    return u[(int64_t)nv * ncell * b + v * ncell + ((k + g2) * E[1] + (j + g1)) * E[0] + (i + g0)];
  }
  struct Seg { int64_t node, first, end, off; };
  std::vector<Seg> segments(int L, int d, int64_t lo, int64_t n) const {
    std::vector<Seg> out;
    if (d >= nd) {
      out.push_back({0, 0, 1, 0});
      return out;
    }
    for (int64_t c = lo; c < lo + n;) {
      const int64_t node = fdiv(c, nx[d]);
      const int64_t end = std::min((node + 1) * nx[d], lo + n);
      const int64_t nl = nlat(L, d);
      out.push_back({((node % nl) + nl) % nl, c - node * nx[d], end - node * nx[d], c - lo});
      c = end;
    }
    return out;
  }
  // level-L values over cells lo .. lo + n (x, y, z) into out[v][z][y][x]
  void box(int L, const int64_t *lo, const int64_t *n, std::vector<double> &out) const {
    out.assign((size_t)nv * n[0] * n[1] * n[2], 0.0);
    const auto sx = segments(L, 0, lo[0], n[0]), sy = segments(L, 1, lo[1], n[1]),
               sz = segments(L, 2, lo[2], n[2]);
    auto at = [&](int v, int64_t z, int64_t y, int64_t x) -> double & {
      return out[(((size_t)v * n[2] + z) * n[1] + y) * n[0] + x];
    };
    for (const Seg &a2 : sz)
      for (const Seg &a1 : sy)
        for (const Seg &a0 : sx) {
          const Seg s[3] = {a0, a1, a2};
          const int64_t node[3] = {a0.node, a1.node, a2.node};
          int64_t cnt[3];
          for (int d = 0; d < 3; ++d) cnt[d] = s[d].end - s[d].first;
          auto it = leaf.find(key(L, node[0], node[1], node[2]));
          if (it != leaf.end()) {
            for (int v = 0; v < nv; ++v)
              for (int64_t z = 0; z < cnt[2]; ++z)
                for (int64_t y = 0; y < cnt[1]; ++y)
                  for (int64_t x = 0; x < cnt[0]; ++x)
                    at(v, s[2].off + z, s[1].off + y, s[0].off + x) =
                        cell(it->second, v, s[0].first + x, s[1].first + y, s[2].first + z);
            continue;
          }
          int64_t g0[3];
          for (int d = 0; d < 3; ++d) g0[d] = d < nd ? node[d] * nx[d] + s[d].first : 0;
          std::vector<double> sub;
          if (refined.count(key(L, node[0], node[1], node[2]))) {
            int64_t flo[3], fn[3];
            for (int d = 0; d < 3; ++d) {
              flo[d] = d < nd ? 2 * g0[d] : 0;
              fn[d] = d < nd ? 2 * cnt[d] : 1;
            }
            std::vector<double> fine;
            box(L + 1, flo, fn, fine);
            const int np = nd == 3 ? 8 : 4;
            auto f = [&](int v, int64_t z, int64_t y, int64_t x) {
              return fine[(((size_t)v * fn[2] + z) * fn[1] + y) * fn[0] + x];
            };
            for (int v = 0; v < nv; ++v)
              for (int64_t z = 0; z < cnt[2]; ++z)
                for (int64_t y = 0; y < cnt[1]; ++y)
                  for (int64_t x = 0; x < cnt[0]; ++x) {
                    double acc = 0.0;
                    int k = 0;
                    for (int dz = 0; dz < (nd == 3 ? 2 : 1); ++dz)
                      for (int dy = 0; dy < 2; ++dy)
                        for (int dx = 0; dx < 2; ++dx, ++k) {
                          const double q = f(v, nd == 3 ? 2 * z + dz : 0, 2 * y + dy, 2 * x + dx);
                          acc = k == 0 ? q : acc + q;
                        }
                    at(v, s[2].off + z, s[1].off + y, s[0].off + x) = acc * (1.0 / np);
                  }
          } else {
            prolong(L, g0, cnt, sub);
            for (int v = 0; v < nv; ++v)
              for (int64_t z = 0; z < cnt[2]; ++z)
                for (int64_t y = 0; y < cnt[1]; ++y)
                  for (int64_t x = 0; x < cnt[0]; ++x)
                    at(v, s[2].off + z, s[1].off + y, s[0].off + x) =
                        sub[(((size_t)v * cnt[2] + z) * cnt[1] + y) * cnt[0] + x];
          }
        }
  }
  // level-L cells g0 .. g0 + cnt under a coarser leaf, into out[v][z][y][x]
  void prolong(int L, const int64_t *g0, const int64_t *cnt, std::vector<double> &out) const {
    if (L == 0) fail("a level-0 region outside every level-0 node");
    int64_t clo[3], cn[3];
    for (int d = 0; d < 3; ++d) {
      clo[d] = d < nd ? fdiv(g0[d], 2) - 1 : 0;
      const int64_t chi = d < nd ? fdiv(g0[d] + cnt[d] - 1, 2) + 1 : 0;
      cn[d] = chi - clo[d] + 1;
    }
    std::vector<double> C;
    box(L - 1, clo, cn, C);
    auto c_at = [&](int v, const int64_t *I) {
      return C[(((size_t)v * cn[2] + I[2]) * cn[1] + I[1]) * cn[0] + I[0]];
    };
    auto sgn = [](double a) { return a < 0.0 ? -1.0 : 1.0; };
    out.assign((size_t)nv * cnt[0] * cnt[1] * cnt[2], 0.0);
    for (int v = 0; v < nv; ++v)
      for (int64_t z = 0; z < cnt[2]; ++z)
        for (int64_t y = 0; y < cnt[1]; ++y)
          for (int64_t x = 0; x < cnt[0]; ++x) {
            const int64_t G[3] = {g0[0] + x, g0[1] + y, g0[2] + z};
            int64_t I[3];
            for (int d = 0; d < 3; ++d) I[d] = d < nd ? fdiv(G[d], 2) - clo[d] : 0;
            const double c = c_at(v, I);
            double val = c;
            for (int d = 0; d < nd; ++d) {
              int64_t lo[3] = {I[0], I[1], I[2]}, hi[3] = {I[0], I[1], I[2]};
              lo[d] -= 1;
              hi[d] += 1;
              const double a = c - c_at(v, lo);
              const double b = c_at(v, hi) - c;
              const double fa = std::fabs(a), fb = std::fabs(b);
              const double g = (0.5 * (sgn(a) + sgn(b))) * (fb < fa ? fb : fa);
              const double sign = (((G[d] % 2) + 2) % 2 == 0) ? -1.0 : 1.0;
              val = val + (sign * 0.25) * g;
            }
            out[(((size_t)v * cnt[2] + z) * cnt[1] + y) * cnt[0] + x] = val;
          }
  }
};

}  // namespace

extern "C" int aph_amr_remap(int32_t ndim, const int32_t nbase[3], const int32_t nx[3],
                             int32_t ng, int32_t nvar, int32_t n_old, const int32_t *old_level,
                             const int64_t *old_loc, const double *old_u, int32_t n_new,
                             const int32_t *new_level, const int64_t *new_loc, double *new_u,
                             char *err, size_t err_len) {
  auto report = [&](int code, const char *msg) {
    if (err && err_len) snprintf(err, err_len, "aph_amr_remap: %s", msg);
    return code;
  };
  if (!nx || !old_u || !new_u || nvar < 1 || n_new < 1 || !new_level || !new_loc)
    return report(APH_E_INVALID, "bad arguments");
  try {
    Forest t;
    forest_setup(t, ndim, nbase, n_old, old_level, old_loc);
    Remap r;
    r.nd = ndim;
    r.nv = nvar;
    r.ng = ng;
    for (int d = 0; d < 3; ++d) {
      r.nbase[d] = t.nbase[d];
      r.nx[d] = d < ndim ? nx[d] : 1;
      r.E[d] = r.nx[d] + (d < ndim ? 2 * ng : 0);
    }
    r.ncell = r.E[0] * r.E[1] * r.E[2];
    r.u = old_u;
    r.refined = t.refined;
    for (int b = 0; b < n_old; ++b)
      r.leaf.emplace(key(old_level[b], old_loc[3 * b], old_loc[3 * b + 1],
                         ndim == 3 ? old_loc[3 * b + 2] : 0), b);
    std::vector<double> vals;
    for (int b = 0; b < n_new; ++b) {
      const int lv = new_level[b];
      const int64_t lc[3] = {new_loc[3 * b], new_loc[3 * b + 1], ndim == 3 ? new_loc[3 * b + 2] : 0};
      const int64_t lo[3] = {lc[0] * r.nx[0], lc[1] * r.nx[1], lc[2] * r.nx[2]};
      const int64_t n[3] = {r.nx[0], r.nx[1], r.nx[2]};
      r.box(lv, lo, n, vals);  // a kept leaf: its own cells
      const int64_t g1 = ndim > 1 ? ng : 0, g2 = ndim > 2 ? ng : 0;
      for (int v = 0; v < nvar; ++v)
        for (int64_t z = 0; z < n[2]; ++z)
          for (int64_t y = 0; y < n[1]; ++y)
            for (int64_t x = 0; x < n[0]; ++x)
              new_u[(int64_t)nvar * r.ncell * b + v * r.ncell +
                    ((z + g2) * r.E[1] + (y + g1)) * r.E[0] + (x + ng)] =
                  vals[(((size_t)v * n[2] + z) * n[1] + y) * n[0] + x];
    }
  } catch (const BuildError &e) {
    return report(APH_E_INVALID, e.msg);
  } catch (const std::bad_alloc &) {
    return report(APH_E_NOMEM, "host allocation failed");
  }
  if (err && err_len) err[0] = 0;
  return APH_OK;
}

// ------------------------------------------------------------------ Poisson operator
// PoissonEquation (src/gravity/gsolvers/poisson_gravity.hpp:31-388, D = 1) on the leaf
// cells of a 3D multilevel mesh whose every domain face is physical:
//   F = (phi_low - phi_high) / dx through each face of a block (CalculateFluxes,
//   :168-224), the value beyond a block face taken from the same level (the leaf cell, or
//   the mean of the 2^3 children of a refined cell: RestrictAverage) or prolongated from
//   the coarser leaf beyond (the phi field's refinement op, gravity.cpp:343-363: "Linear"
//   is ProlongateSharedLinear, c + sum_d (+-1/4) (c_+d - c_-d) / 2 on the uniform coarse
//   level; "Constant", c), and minus the cell inside beyond a domain face (the FixedFace
//   zero boundary, SetZeroDirichletBoundary, on the coarse level too);
//   with flux correction (:89-99) a coarse face over a finer region carries the mean of
//   its 2^2 fine faces;
//   (A phi)_c = -alpha phi_c + sum_d [F(c - 1/2) - F(c + 1/2)] A_d / V (FluxMultiplyMatrix,
//   :299-349), each block with its own dx, areas and volume.
// Rows are accumulated term by term in a fixed order (first appearance of a column keeps
// its place), so the coefficients do not depend on hashing.
#include "aph_poisson_tree.hpp"

namespace {

using Row = std::vector<std::pair<int64_t, double>>;

void acc(Row &r, int64_t c, double w) {
  for (auto &e : r)
    if (e.first == c) {
      e.second += w;
      return;
    }
  r.emplace_back(c, w);
}

struct PoissonTree {
  Tree t;
  int64_t ncb = 0;
  bool fc = false, constant = false;

  int64_t lin(int b, const int64_t l[3]) const {  // l: ghost-inclusive local index
    const int g = t.ng;
    return (int64_t)b * ncb + ((l[2] - g) * t.nx[1] + (l[1] - g)) * t.nx[0] + (l[0] - g);
  }
  int64_t lin_int(int b, int i, int j, int k) const {
    return (int64_t)b * ncb + ((int64_t)k * t.nx[1] + j) * t.nx[0] + i;
  }
  // level-lv value of global cell G (inside the domain): the leaf cell, or the mean of its
  // children over a refined cell; false under a coarser leaf
  bool value(int lv, const int64_t G[3], Row &out) const {
    out.clear();
    int b;
    int64_t l[3];
    if (t.interior(lv, G, b, l)) {
      out.emplace_back(lin(b, l), 1.0);
      return true;
    }
    const int64_t node[3] = {G[0] / t.nx[0], G[1] / t.nx[1], G[2] / t.nx[2]};
    if (!t.is_refined(lv, node)) return false;
    Row kid;
    for (int cz = 0; cz < 2; ++cz)
      for (int cy = 0; cy < 2; ++cy)
        for (int cx = 0; cx < 2; ++cx) {
          const int64_t F[3] = {2 * G[0] + cx, 2 * G[1] + cy, 2 * G[2] + cz};
          if (!value(lv + 1, F, kid))
            fail("cell (%lld, %lld, %lld) of level %d is neither a leaf cell nor the parent "
                 "of leaf cells (2:1 balance)", (long long)G[0], (long long)G[1],
                 (long long)G[2], lv);
          for (auto &e : kid) acc(out, e.first, 0.125 * e.second);
        }
    return true;
  }
  // the value at global cell G of level lv seen from a block of level lv (G beyond it);
  // inner: the cell inside next to G (for the domain-face image)
  void ghost(int lv, const int64_t G[3], int64_t inner, Row &out) const {
    out.clear();
    if (!t.inside(lv, G)) {
      out.emplace_back(inner, -1.0);
      return;
    }
    if (value(lv, G, out)) return;
    const int64_t I[3] = {G[0] / 2, G[1] / 2, G[2] / 2};
    Row c;
    if (lv == 0 || !value(lv - 1, I, c))
      fail("cell (%lld, %lld, %lld) of level %d is over neither a finer region nor a coarser "
           "leaf (2:1 balance)", (long long)G[0], (long long)G[1], (long long)G[2], lv);
    out = c;
    if (constant) return;
    Row nb[2];
    for (int d = 0; d < 3; ++d) {
      const double sign = G[d] % 2 == 0 ? -1.0 : 1.0;
      for (int s = 0; s < 2; ++s) {
        int64_t J[3] = {I[0], I[1], I[2]};
        J[d] += s ? 1 : -1;
        if (t.inside(lv - 1, J)) {
          value(lv - 1, J, nb[s]);
        } else {
          nb[s].clear();
          for (auto &e : c) nb[s].emplace_back(e.first, -e.second);
        }
      }
      // (hi - lo) / 2 * (+-1/4)
      for (auto &e : nb[1]) acc(out, e.first, sign * 0.25 * 0.5 * 1.0 * e.second);
      for (auto &e : nb[0]) acc(out, e.first, sign * 0.25 * 0.5 * -1.0 * e.second);
    }
  }
  // the value at interior-or-ghost index c along d (transverse interior a, b) of block b
  void cell(int b, int d, int64_t c, const int tt[2], Row &out) const {
    const int A = d == 0 ? 1 : 0, B = d == 2 ? 1 : 2;
    int64_t idx[3];
    idx[d] = c;
    idx[A] = tt[0];
    idx[B] = tt[1];
    if (c >= 0 && c < t.nx[d]) {
      out.clear();
      out.emplace_back(lin_int(b, (int)idx[0], (int)idx[1], (int)idx[2]), 1.0);
      return;
    }
    const int lv = t.level[b];
    int64_t G[3];
    for (int q = 0; q < 3; ++q) G[q] = t.loc[3 * b + q] * t.nx[q] + idx[q];
    int64_t in[3] = {idx[0], idx[1], idx[2]};
    in[d] = c < 0 ? 0 : t.nx[d] - 1;
    ghost(lv, G, lin_int(b, (int)in[0], (int)in[1], (int)in[2]), out);
  }
  // the plain flux through face f (0..nx) of block b along d at transverse (a, b)
  void face_flux(int b, int d, int f, const int tt[2], Row &out) const {
    Row lo, hi;
    cell(b, d, f - 1, tt, lo);
    cell(b, d, f, tt, hi);
    out = lo;
    for (auto &e : hi) acc(out, e.first, -e.second);
    const double inv = 1.0 / t.blk[b].dx[d];
    for (auto &e : out) e.second *= inv;
  }
  // the flux the operator uses: on a block face over a refined node with flux correction,
  // the mean of the fine faces across it
  void flux(int b, int d, int f, const int tt[2], Row &out) const {
    if (fc && (f == 0 || f == t.nx[d])) {
      const int side = f == t.nx[d];
      const int lv = t.level[b];
      int64_t p[3] = {t.loc[3 * b], t.loc[3 * b + 1], t.loc[3 * b + 2]};
      p[d] += side ? 1 : -1;
      if (p[d] >= 0 && p[d] < t.nlat(lv, d) && t.is_refined(lv, p)) {
        const int A = d == 0 ? 1 : 0, B = d == 2 ? 1 : 2;
        out.clear();
        Row fine;
        for (int ka = 0; ka < 2; ++ka)
          for (int kb = 0; kb < 2; ++kb) {
            int64_t G[3];
            G[A] = 2 * (t.loc[3 * b + A] * t.nx[A] + tt[0]) + ka;
            G[B] = 2 * (t.loc[3 * b + B] * t.nx[B] + tt[1]) + kb;
            G[d] = 2 * (t.loc[3 * b + d] * t.nx[d] + f) - (side ? 0 : 1);
            int fb;
            int64_t fl[3];
            if (!t.interior(lv + 1, G, fb, fl))
              fail("block %d face %d: fine face outside the finer level", b, 2 * d + side);
            const int ff = (int)(fl[d] - t.ng) + (side ? 0 : 1);
            const int ft[2] = {(int)(fl[A] - t.ng), (int)(fl[B] - t.ng)};
            face_flux(fb, d, ff, ft, fine);
            for (auto &e : fine) acc(out, e.first, 0.25 * e.second);
          }
        return;
      }
    }
    face_flux(b, d, f, tt, out);
  }
};

void poisson_assemble(PoissonTree &pt, double alpha, aph::PoissonTreeCsr &out) {
  const Tree &t = pt.t;
  const int64_t n = (int64_t)t.nb * pt.ncb;
  if (n > INT32_MAX) fail("too many cells for 32-bit column indices");
  out.n = n;
  out.rowptr.assign(1, 0);
  out.col.clear();
  out.val.clear();
  Row row, f;
  for (int b = 0; b < t.nb; ++b) {
    const aph_block &blk = t.blk[b];
    for (int k = 0; k < t.nx[2]; ++k)
      for (int j = 0; j < t.nx[1]; ++j)
        for (int i = 0; i < t.nx[0]; ++i) {
          const int idx[3] = {i, j, k};
          row.clear();
          row.emplace_back(pt.lin_int(b, i, j, k), -alpha);
          for (int d = 0; d < 3; ++d) {
            const int A = d == 0 ? 1 : 0, B = d == 2 ? 1 : 2;
            const double ar = blk.area[d] / blk.vol;
            const int tt[2] = {idx[A], idx[B]};
            pt.flux(b, d, idx[d], tt, f);
            for (auto &e : f) acc(row, e.first, ar * e.second);
            pt.flux(b, d, idx[d] + 1, tt, f);
            for (auto &e : f) acc(row, e.first, -ar * e.second);
          }
          for (auto &e : row) {
            out.col.push_back((int32_t)e.first);
            out.val.push_back(e.second);
          }
          out.rowptr.push_back((int64_t)out.col.size());
        }
  }
  // neighbour rows, (2 d + side) * n + c
  out.nb_rowptr.assign(1, 0);
  out.nb_col.clear();
  out.nb_val.clear();
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s)
      for (int b = 0; b < t.nb; ++b)
        for (int k = 0; k < t.nx[2]; ++k)
          for (int j = 0; j < t.nx[1]; ++j)
            for (int i = 0; i < t.nx[0]; ++i) {
              int q[3] = {i, j, k};
              const int64_t r = pt.lin_int(b, i, j, k);
              q[d] += s ? 1 : -1;
              if (q[d] >= 0 && q[d] < t.nx[d]) {
                out.nb_col.push_back((int32_t)pt.lin_int(b, q[0], q[1], q[2]));
                out.nb_val.push_back(1.0);
              } else {
                int64_t G[3];
                for (int a = 0; a < 3; ++a) G[a] = t.loc[3 * b + a] * t.nx[a] + q[a];
                if (t.inside(t.level[b], G)) {
                  pt.ghost(t.level[b], G, r, f);
                  for (auto &e : f) {
                    out.nb_col.push_back((int32_t)e.first);
                    out.nb_val.push_back(e.second);
                  }
                }
              }
              out.nb_rowptr.push_back((int64_t)out.nb_col.size());
            }
  out.wall.assign(t.nb, 0);
  for (int b = 0; b < t.nb; ++b)
    for (int d = 0; d < 3; ++d) {
      if (t.loc[3 * b + d] == 0) out.wall[b] |= 1 << (2 * d);
      if (t.loc[3 * b + d] == t.nlat(t.level[b], d) - 1) out.wall[b] |= 1 << (2 * d + 1);
    }
}

}  // namespace

int aph::poisson_tree_csr(const aph_block *blocks, int32_t nblocks, const int32_t nbase[3],
                          const int32_t *level, const int64_t *loc, int32_t flux_correct,
                          int32_t prolongation, double alpha, PoissonTreeCsr &out, char *err,
                          size_t err_len) {
  auto report = [&](int code, const char *msg) {
    if (err && err_len) snprintf(err, err_len, "%s", msg);
    return code;
  };
  if (prolongation != APH_POISSON_PROLONG_LINEAR && prolongation != APH_POISSON_PROLONG_CONSTANT)
    return report(APH_E_INVALID, "prolongation: APH_POISSON_PROLONG_LINEAR or _CONSTANT (the "
                                 "quadratic form is not built)");
  try {
    PoissonTree pt;
    const int32_t walls[6] = {APH_BC_OUTFLOW, APH_BC_OUTFLOW, APH_BC_OUTFLOW,
                              APH_BC_OUTFLOW, APH_BC_OUTFLOW, APH_BC_OUTFLOW};
    setup(pt.t, blocks, nblocks, 3, 1, nbase, level, loc, walls);
    pt.ncb = (int64_t)pt.t.nx[0] * pt.t.nx[1] * pt.t.nx[2];
    pt.fc = flux_correct != 0;
    pt.constant = prolongation == APH_POISSON_PROLONG_CONSTANT;
    // the block geometry must be the tree's: one base spacing, halved per level, and each
    // block's corner where its logical position puts it
    double lo[3], dx0[3];
    for (int d = 0; d < 3; ++d) {
      dx0[d] = blocks[0].dx[d] * (double)((int64_t)1 << level[0]);
      lo[d] = blocks[0].xmin[d] - (double)loc[d] * pt.t.nx[d] * blocks[0].dx[d];
    }
    for (int b = 0; b < nblocks; ++b)
      for (int d = 0; d < 3; ++d) {
        const double dx = dx0[d] / (double)((int64_t)1 << level[b]);
        const double x = lo[d] + (double)loc[3 * b + d] * pt.t.nx[d] * dx;
        if (std::fabs(blocks[b].dx[d] - dx) > 1e-12 * dx ||
            std::fabs(blocks[b].xmin[d] - x) > 1e-9 * dx)
          fail("block %d: spacing or corner does not match its level and logical position", b);
      }
    // a block on a domain face must have no neighbour across it (the periodic problem
    // needs a mean density subtracted; the tree itself has no wrap)
    for (int b = 0; b < nblocks; ++b)
      for (int d = 0; d < 3; ++d)
        for (int side = 0; side < 2; ++side) {
          const int64_t edge = side ? pt.t.nlat(level[b], d) - 1 : 0;
          int o[3] = {0, 0, 0};
          o[d] = side ? 1 : -1;
          const int slot = (o[2] + 1) * 9 + (o[1] + 1) * 3 + o[0] + 1;
          if (loc[3 * b + d] == edge && blocks[b].nbr[slot] >= 0)
            fail("block %d: every face of the domain must be a physical boundary", b);
        }
    poisson_assemble(pt, alpha, out);
  } catch (const BuildError &e) {
    return report(APH_E_INVALID, e.msg);
  } catch (const std::bad_alloc &) {
    return report(APH_E_NOMEM, "host allocation failed");
  }
  if (err && err_len) err[0] = 0;
  return APH_OK;
}

extern "C" int aph_poisson_tree_matrix(const aph_block *blocks, int32_t nblocks,
                                       const int32_t nbase[3], const int32_t *level,
                                       const int64_t *loc, int32_t flux_correct,
                                       int32_t prolongation, double diagonal_alpha,
                                       int64_t *rowptr, int32_t *col, double *val,
                                       int64_t *nnz, int64_t *nb_rowptr, int32_t *nb_col,
                                       double *nb_val, int64_t *nb_nnz, char *err,
                                       size_t err_len) {
  if (!nnz || !nb_nnz) {
    if (err && err_len) snprintf(err, err_len, "aph_poisson_tree_matrix: null length");
    return APH_E_INVALID;
  }
  aph::PoissonTreeCsr m;
  char msg[256];
  const int rc = aph::poisson_tree_csr(blocks, nblocks, nbase, level, loc, flux_correct,
                                       prolongation, diagonal_alpha, m, msg, sizeof msg);
  if (rc) {
    if (err && err_len) snprintf(err, err_len, "aph_poisson_tree_matrix: %s", msg);
    return rc;
  }
  const int64_t cap[2] = {*nnz, *nb_nnz};
  *nnz = (int64_t)m.col.size();
  *nb_nnz = (int64_t)m.nb_col.size();
  if ((col && cap[0] < *nnz) || (nb_col && cap[1] < *nb_nnz)) {
    if (err && err_len) snprintf(err, err_len, "aph_poisson_tree_matrix: arrays too short");
    return APH_E_INVALID;
  }
  if (rowptr) std::memcpy(rowptr, m.rowptr.data(), sizeof(int64_t) * m.rowptr.size());
  if (col) std::memcpy(col, m.col.data(), sizeof(int32_t) * m.col.size());
  if (val) std::memcpy(val, m.val.data(), sizeof(double) * m.val.size());
  if (nb_rowptr)
    std::memcpy(nb_rowptr, m.nb_rowptr.data(), sizeof(int64_t) * m.nb_rowptr.size());
  if (nb_col) std::memcpy(nb_col, m.nb_col.data(), sizeof(int32_t) * m.nb_col.size());
  if (nb_val) std::memcpy(nb_val, m.nb_val.data(), sizeof(double) * m.nb_val.size());
  if (err && err_len) err[0] = 0;
  return APH_OK;
}

// ------------------------------------------------------------ Poisson multigrid levels
// Coarser meshes for the multigrid preconditioner of the composite problem: the tree cut at
// levels Lmax-1, ..., 0 (every leaf deeper than L merged into its level-L ancestor), then
// the base blocks with nx/2, nx/4, ... cells per axis while the count stays even and at
// least 2.  Each is assembled by poisson_tree_csr from block tables made here (ghosts 1, no
// neighbours), so the operator is the same discretisation on the coarser cells.
namespace {

struct MeshDesc {
  int nx[3];
  std::vector<int32_t> level;
  std::vector<int64_t> loc;
  std::unordered_map<uint64_t, int> index;  // (level, node) -> block
};

std::vector<aph_block> make_blocks(const MeshDesc &m, const double lo[3], const double H0[3]) {
  std::vector<aph_block> out(m.level.size());
  int64_t base = 0;
  for (size_t b = 0; b < out.size(); ++b) {
    aph_block &k = out[b];
    std::memset(&k, 0, sizeof(k));
    k.ng = 1;
    double dx[3];
    for (int d = 0; d < 3; ++d) {
      k.nx[d] = m.nx[d];
      dx[d] = H0[d] / ((double)m.nx[d] * (double)((int64_t)1 << m.level[b]));
      k.dx[d] = dx[d];
      k.xmin[d] = lo[d] + (double)m.loc[3 * b + d] * m.nx[d] * dx[d];
    }
    k.area[0] = dx[1] * dx[2];
    k.area[1] = dx[0] * dx[2];
    k.area[2] = dx[0] * dx[1];
    k.vol = dx[0] * dx[1] * dx[2];
    for (int q = 0; q < 27; ++q) k.nbr[q] = -1;
    k.cell_base = base;
    base += (int64_t)(m.nx[0] + 2) * (m.nx[1] + 2) * (m.nx[2] + 2);
  }
  return out;
}

// the cell of mesh c holding each cell of mesh f (c is f cut one level lower or with half
// the cells per block axis)
std::vector<int32_t> cell_map(const MeshDesc &f, const MeshDesc &c) {
  const int64_t ncf = (int64_t)f.nx[0] * f.nx[1] * f.nx[2];
  const int64_t ncc = (int64_t)c.nx[0] * c.nx[1] * c.nx[2];
  std::vector<int32_t> out((size_t)(ncf * f.level.size()));
  for (size_t b = 0; b < f.level.size(); ++b) {
    const int lf = f.level[b];
    // the coarse block: the leaf of c at this level or an ancestor of it
    int lc = lf;
    int64_t node[3] = {f.loc[3 * b], f.loc[3 * b + 1], f.loc[3 * b + 2]};
    auto it = c.index.find(key(lc, node[0], node[1], node[2]));
    while (it == c.index.end()) {
      if (lc == 0) fail("multigrid level: a block has no coarse block");
      --lc;
      for (int d = 0; d < 3; ++d) node[d] >>= 1;
      it = c.index.find(key(lc, node[0], node[1], node[2]));
    }
    const int cb = it->second;
    for (int k = 0; k < f.nx[2]; ++k)
      for (int j = 0; j < f.nx[1]; ++j)
        for (int i = 0; i < f.nx[0]; ++i) {
          const int ijk[3] = {i, j, k};
          int64_t cl[3];
          for (int d = 0; d < 3; ++d) {
            // global index at level lf in f's cells, then at level lc in c's cells
            int64_t g = f.loc[3 * b + d] * f.nx[d] + ijk[d];
            g >>= (lf - lc);
            if (c.nx[d] != f.nx[d]) g >>= 1;
            cl[d] = g - c.loc[3 * cb + d] * c.nx[d];
            if (cl[d] < 0 || cl[d] >= c.nx[d]) fail("multigrid level: cell outside its block");
          }
          out[(size_t)(b * ncf + ((int64_t)k * f.nx[1] + j) * f.nx[0] + i)] =
              (int32_t)(cb * ncc + (cl[2] * c.nx[1] + cl[1]) * c.nx[0] + cl[0]);
        }
  }
  return out;
}

}  // namespace

int aph::poisson_tree_levels(const aph_block *blocks, int32_t nblocks, const int32_t nbase[3],
                             const int32_t *level, const int64_t *loc, int32_t flux_correct,
                             int32_t prolongation, double alpha, int64_t max_coarse,
                             std::vector<PoissonMgLevel> &out, char *err, size_t err_len) {
  auto report = [&](int code, const char *msg) {
    if (err && err_len) snprintf(err, err_len, "%s", msg);
    return code;
  };
  out.clear();
  try {
    if (nblocks < 1) fail("no blocks");
    MeshDesc fine;
    int lmax = 0;
    for (int d = 0; d < 3; ++d) fine.nx[d] = blocks[0].nx[d];
    fine.level.assign(level, level + nblocks);
    fine.loc.assign(loc, loc + 3 * (size_t)nblocks);
    for (int b = 0; b < nblocks; ++b) {
      lmax = std::max(lmax, (int)level[b]);
      fine.index[key(level[b], loc[3 * b], loc[3 * b + 1], loc[3 * b + 2])] = b;
    }
    double lo[3], H0[3];
    for (int d = 0; d < 3; ++d) {
      const double dx0 = blocks[0].dx[d] * (double)((int64_t)1 << level[0]);
      H0[d] = dx0 * fine.nx[d];
      lo[d] = blocks[0].xmin[d] - (double)loc[d] * fine.nx[d] * blocks[0].dx[d];
    }
    std::vector<MeshDesc> meshes;
    for (int L = lmax - 1; L >= 0; --L) {  // the tree cut at L
      MeshDesc m;
      for (int d = 0; d < 3; ++d) m.nx[d] = fine.nx[d];
      for (int b = 0; b < nblocks; ++b) {
        const int l = std::min((int)level[b], L);
        int64_t c[3];
        for (int d = 0; d < 3; ++d) c[d] = loc[3 * b + d] >> (level[b] - l);
        const uint64_t k = key(l, c[0], c[1], c[2]);
        if (m.index.count(k)) continue;
        m.index[k] = (int)m.level.size();
        m.level.push_back(l);
        for (int d = 0; d < 3; ++d) m.loc.push_back(c[d]);
      }
      meshes.push_back(std::move(m));
    }
    for (;;) {  // the base blocks with half the cells per axis
      const MeshDesc &last = meshes.empty() ? fine : meshes.back();
      const int64_t n = (int64_t)last.level.size() * last.nx[0] * last.nx[1] * last.nx[2];
      if (n <= max_coarse) break;
      bool ok = true;  // the halved count must stay even (the tree's prolongation)
      for (int d = 0; d < 3; ++d) ok = ok && last.nx[d] % 4 == 0;
      if (!ok) break;
      MeshDesc m = last;
      for (int d = 0; d < 3; ++d) m.nx[d] = last.nx[d] / 2;
      meshes.push_back(std::move(m));
    }
    const MeshDesc *prev = &fine;
    for (const MeshDesc &m : meshes) {
      PoissonMgLevel lv;
      std::vector<aph_block> bt = make_blocks(m, lo, H0);
      PoissonTreeCsr csr;
      char msg[256];
      const int rc = poisson_tree_csr(bt.data(), (int32_t)bt.size(), nbase, m.level.data(),
                                      m.loc.data(), flux_correct, prolongation, alpha, csr,
                                      msg, sizeof msg);
      if (rc) fail("multigrid level: %s", msg);
      lv.n = csr.n;
      lv.rowptr = std::move(csr.rowptr);
      lv.col = std::move(csr.col);
      lv.val = std::move(csr.val);
      lv.cmap = cell_map(*prev, m);
      std::vector<int64_t> cnt((size_t)lv.n + 1, 0);
      for (int32_t c : lv.cmap) ++cnt[(size_t)c + 1];
      lv.rptr.assign((size_t)lv.n + 1, 0);
      for (int64_t c = 0; c < lv.n; ++c) lv.rptr[c + 1] = lv.rptr[c] + cnt[c + 1];
      lv.rcol.assign(lv.cmap.size(), 0);
      lv.rval.assign(lv.cmap.size(), 0.0);
      std::vector<int64_t> at(lv.rptr.begin(), lv.rptr.end() - 1);
      for (size_t f = 0; f < lv.cmap.size(); ++f) {  // finer cells in ascending order
        const int32_t c = lv.cmap[f];
        lv.rcol[at[c]] = (int32_t)f;
        lv.rval[at[c]] = 1.0 / (double)cnt[(size_t)c + 1];
        ++at[c];
      }
      out.push_back(std::move(lv));
      prev = &m;
    }
  } catch (const BuildError &e) {
    return report(APH_E_INVALID, e.msg);
  } catch (const std::bad_alloc &) {
    return report(APH_E_NOMEM, "host allocation failed");
  }
  if (err && err_len) err[0] = 0;
  return APH_OK;
}
