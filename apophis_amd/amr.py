"""Two-level meshes for the plain form: a 2:1 refinement jump, as Parthenon's AMR produces.

The reference leaves multilevel meshes to Parthenon (apophis_driver.cpp:68,74-76,94):
* the boundary exchange of `cons` fills a coarse block's ghost cells over a refined region
  with the mean of the fine children (RestrictAverage) and a fine block's ghost cells over
  a coarse region by limited linear prolongation (ProlongateSharedMinMod);
* the flux correction replaces every coarse face flux on a coarse-fine face by the
  area-weighted mean of the fine fluxes across it, before the update, so the levels
  exchange exactly the same fluxes and the total is conserved.

Here the host resolves the mesh tree once into flat op lists (``amr_plan``), which the
library validates and runs as two gather kernels (``aph_amr_fill``,
``aph_amr_flux_correct``).  Same-level neighbours keep using the block table's ``nbr``
(``aph_fill_ghosts``); a face whose neighbour region is on the other level has nbr = -1 and
bc = APH_BC_AMR.  Meshes are periodic; each block keeps the same cells per block at every
level (spacing halves), so a coarse cell's children always lie in one fine block.

Prolongation of coarse cell c to the child on the lower / upper side (sign -1 / +1) of axis
d uses g_d = minmod(c - c_{-d}, c_{+d} - c) and gives c + sum_d sign_d g_d / 4, the
upstream form c + g'_d (x_child - x_c) with g'_d the minmod of the two one-sided slopes
and x_child - x_c = +-dx/4, written without the division by and product with dx.
"""
from __future__ import annotations

import ctypes
import itertools

import numpy as np

from . import _lib
from .mesh import BLOCK_DTYPE, BlockSet

BC_AMR = 3  # APH_BC_AMR
AMR_MEAN, AMR_PROLONG = 0, 1

AMR_REF_DTYPE = np.dtype([("offset", "<i8"), ("stride", "<i8")])
AMR_VAL_DTYPE = np.dtype([("first", "<i4"), ("count", "<i4")])
AMR_OP_DTYPE = np.dtype([("dst", "<i8"), ("dst_stride", "<i8"), ("kind", "<i4"),
                         ("sign", "<i4", (3,)), ("val", AMR_VAL_DTYPE, (7,))])
FLUX_FIX_DTYPE = np.dtype([("dst", "<i8"), ("dst_stride", "<i8"), ("dir", "<i4"),
                           ("reserved", "<i4"), ("val", AMR_VAL_DTYPE)])
assert AMR_REF_DTYPE.itemsize == 16 and AMR_OP_DTYPE.itemsize == 88
assert FLUX_FIX_DTYPE.itemsize == 32


class TwoLevelMesh:
    """Leaf blocks of a periodic mesh whose base lattice ``nbase`` has the blocks at the
    logical positions ``refined`` replaced by their 2^ndim children (level 1, the same
    ``block_nx`` cells at half the spacing).  ``blocks`` is a BLOCK_DTYPE table (nbr links
    same-level leaves, with periodic wrap; faces to the other level carry BC_AMR);
    ``level[b]`` and ``loc[b]`` give each leaf's level and logical position there."""

    def __init__(self, ndim, nbase, block_nx, ng, xmin, xmax, refined, mesh_id=0):
        if ndim not in (2, 3):
            raise ValueError("two-level meshes are 2D or 3D")
        self.ndim = ndim
        self.xmin = tuple(float(v) for v in xmin)
        self.xmax = tuple(float(v) for v in xmax)
        self.mesh_id = int(mesh_id)
        self.nbase = tuple(int(v) for v in nbase) + (1,) * (3 - len(nbase))
        self.nx = tuple(int(v) for v in block_nx) + (1,) * (3 - len(block_nx))
        self.ng = int(ng)
        if any(self.nx[d] < 2 * self.ng for d in range(ndim)):
            raise ValueError("blocks need at least 2*ng cells per axis")
        refined = {tuple(int(v) for v in r) + (0,) * (3 - len(r)) for r in refined}
        span = [float(xmax[d]) - float(xmin[d]) for d in range(3)]
        leaves = []  # (level, loc)
        for lz in range(self.nbase[2]):
            for ly in range(self.nbase[1]):
                for lx in range(self.nbase[0]):
                    if (lx, ly, lz) not in refined:
                        leaves.append((0, (lx, ly, lz)))
                        continue
                    for cz in (range(2) if ndim == 3 else (0,)):
                        for cy in range(2):
                            for cx in range(2):
                                leaves.append((1, (2 * lx + cx, 2 * ly + cy, 2 * lz + cz)))
        self.level = np.array([lv for lv, _ in leaves], dtype=np.int32)
        self.loc = np.array([lc for _, lc in leaves], dtype=np.int64)
        self.index = {(lv, lc): b for b, (lv, lc) in enumerate(leaves)}
        out = np.zeros(len(leaves), dtype=BLOCK_DTYPE)
        for b, (lv, lc) in enumerate(leaves):
            r = out[b]
            f = 2 ** lv
            dx = [span[d] / (self.nbase[d] * self.nx[d] * (f if d < ndim else 1))
                  for d in range(3)]
            r["nx"] = self.nx
            r["ng"] = self.ng
            r["dx"] = dx
            r["area"] = (dx[1] * dx[2], dx[0] * dx[2], dx[0] * dx[1])
            r["vol"] = dx[0] * dx[1] * dx[2]
            r["xmin"] = [float(xmin[d]) + (lc[d] * self.nx[d] * dx[d] if d < ndim else 0.0)
                         for d in range(3)]
            r["mesh_id"] = mesh_id
            nbr = np.full(27, -1, dtype=np.int32)
            for o in itertools.product((-1, 0, 1), repeat=3):
                if o == (0, 0, 0) or any(o[d] != 0 for d in range(ndim, 3)):
                    continue
                p = tuple((lc[d] + o[d]) % self.nlat(lv, d) for d in range(3))
                nbr[(o[2] + 1) * 9 + (o[1] + 1) * 3 + (o[0] + 1)] = self.index.get((lv, p), -1)
            r["nbr"] = nbr
            bc = [1] * 6  # outflow placeholder on faces with a same-level neighbour
            for d in range(ndim):
                for side in (0, 1):
                    o = [0, 0, 0]
                    o[d] = 1 if side else -1
                    if nbr[(o[2] + 1) * 9 + (o[1] + 1) * 3 + (o[0] + 1)] < 0:
                        bc[2 * d + side] = BC_AMR
            r["bc"] = bc
        self.blocks = out

    def nlat(self, level, d):
        """Blocks per axis d at a level."""
        return self.nbase[d] * (2 ** level if d < self.ndim else 1)

    def ncells_axis(self, level, d):
        return self.nlat(level, d) * self.nx[d]

    def block_set(self, nscalars=0):
        return BlockSet(self.ndim, [self.blocks], nscalars=nscalars)


class _Refs:
    def __init__(self):
        self.rows = []

    def add(self, cells):
        first = len(self.rows)
        self.rows.extend(cells)
        return (first, len(cells))


def _cell(bs, b, local, _cache={}):
    """(component-0 offset, stride) of block b's cell at ghost-inclusive local (i, j, k)."""
    key = id(bs)
    if key not in _cache or _cache[key][0] is not bs:
        _cache.clear()
        _cache[key] = (bs, bs.extents(), bs.blocks["cell_base"].astype(np.int64))
    _, ext, cb = _cache[key]
    N = ext[b]
    c = (int(local[2]) * int(N[1]) + int(local[1])) * int(N[0]) + int(local[0])
    return (bs.nvar * int(cb[b]) + c, int(bs.ncell[b]))


def amr_plan(mesh: TwoLevelMesh, bs: BlockSet):
    """(ops, fixes, refs) for ``bs`` (a block set of ``mesh``'s table)."""
    nd, g, nx = mesh.ndim, mesh.ng, mesh.nx
    refs = _Refs()
    act = range(nd)

    def wrap(level, G):
        return tuple(G[d] % mesh.ncells_axis(level, d) if d < nd else 0 for d in range(3))

    def interior(level, G):
        """(block, ghost-inclusive local index) of the level-`level` leaf holding global
        cell G, or None when that region is not a leaf at this level."""
        G = wrap(level, G)
        lc = tuple(G[d] // nx[d] if d < nd else 0 for d in range(3))
        b = mesh.index.get((level, lc))
        if b is None:
            return None
        return b, tuple(G[d] - lc[d] * nx[d] + (g if d < nd else 0) for d in range(3))

    def coarse_value(I):
        """Level-0 value of global coarse cell I: the cell, or the mean of its children."""
        hit = interior(0, I)
        if hit is not None:
            return refs.add([_cell(bs, *hit)])
        kids = []
        for dz in ((0, 1) if nd == 3 else (0,)):
            for dy in (0, 1):
                for dx_ in (0, 1):
                    F = (2 * I[0] + dx_, 2 * I[1] + dy, 2 * I[2] + dz if nd == 3 else 0)
                    hit = interior(1, F)
                    assert hit is not None, "two levels only"
                    kids.append(_cell(bs, *hit))
        return refs.add(kids)

    ops = []
    N = bs.extents()
    for b in range(bs.nblocks):
        lv = int(mesh.level[b])
        lc = mesh.loc[b]
        nbr = bs.blocks["nbr"][b]
        # ghost cells whose direction has no same-level neighbour (the rest is exchanged)
        E = [int(N[b][d]) for d in range(3)]
        kk, jj, ii = np.meshgrid(np.arange(E[2]), np.arange(E[1]), np.arange(E[0]),
                                 indexing="ij")
        o3 = [np.zeros_like(ii) for _ in range(3)]
        for d in act:
            c = (ii, jj, kk)[d]
            o3[d] = np.where(c < g, -1, np.where(c >= g + nx[d], 1, 0))
        ghost = (o3[0] != 0) | (o3[1] != 0) | (o3[2] != 0)
        need = ghost & (nbr[(o3[2] + 1) * 9 + (o3[1] + 1) * 3 + (o3[0] + 1)] < 0)
        for k, j, i in zip(kk[need].tolist(), jj[need].tolist(), ii[need].tolist()):
            loc = (i, j, k)
            G = tuple(int(lc[d]) * nx[d] + loc[d] - g if d < nd else 0
                      for d in range(3))
            dst, stride = _cell(bs, b, loc)
            rec = np.zeros(1, dtype=AMR_OP_DTYPE)[0]
            rec["dst"], rec["dst_stride"] = dst, stride
            if lv == 0:  # coarse ghost over a refined region: restriction
                rec["kind"] = AMR_MEAN
                rec["val"][0] = coarse_value(wrap(0, G))
            else:  # fine ghost over a coarse region: prolongation
                Gw = wrap(1, G)
                I = tuple(Gw[d] // 2 if d < nd else 0 for d in range(3))
                assert interior(0, I) is not None, "two levels only"
                rec["kind"] = AMR_PROLONG
                rec["val"][0] = coarse_value(I)
                for d in act:
                    e = [0, 0, 0]
                    e[d] = 1
                    rec["sign"][d] = -1 if Gw[d] % 2 == 0 else 1
                    rec["val"][1 + 2 * d] = coarse_value(
                        tuple(I[t] - e[t] for t in range(3)))
                    rec["val"][2 + 2 * d] = coarse_value(
                        tuple(I[t] + e[t] for t in range(3)))
            ops.append(rec)
    fixes = []
    for b in range(bs.nblocks):
        if mesh.level[b] != 0:
            continue
        lc = mesh.loc[b]
        bc = bs.blocks["bc"][b]
        for d in act:
            for side in (0, 1):
                if bc[2 * d + side] != BC_AMR:
                    continue
                tr = [t for t in act if t != d]
                for tcell in itertools.product(*[range(nx[t]) for t in tr]):
                    loc = [0, 0, 0]
                    loc[d] = g + (nx[d] if side else 0)  # the block's boundary face
                    for t, c in zip(tr, tcell):
                        loc[t] = g + c
                    dst, stride = _cell(bs, b, loc)
                    # fine cells beside the face on the fine side, and the face of each
                    Fd = 2 * (int(lc[d]) * nx[d] + (nx[d] if side else 0)) - (0 if side else 1)
                    cells = []
                    for kid in itertools.product((0, 1), repeat=len(tr)):
                        F = [0, 0, 0]
                        F[d] = Fd
                        for t, c, q in zip(tr, tcell, kid):
                            F[t] = 2 * (int(lc[t]) * nx[t] + c) + q
                        fb, floc = interior(1, tuple(F))
                        floc = list(floc)
                        if not side:
                            floc[d] += 1  # upper face of the last fine cell below
                        cells.append(_cell(bs, fb, floc))
                    rec = np.zeros(1, dtype=FLUX_FIX_DTYPE)[0]
                    rec["dst"], rec["dst_stride"], rec["dir"] = dst, stride, d
                    rec["val"] = refs.add(cells)
                    fixes.append(rec)
    return (np.array(ops, dtype=AMR_OP_DTYPE), np.array(fixes, dtype=FLUX_FIX_DTYPE),
            np.array(refs.rows, dtype=AMR_REF_DTYPE))


AMR_FACE_DTYPE = np.dtype([("base", "<i8"), ("role", "<i4"), ("qa", "<i4"), ("qb", "<i4"),
                           ("na", "<i4")])
assert AMR_FACE_DTYPE.itemsize == 24


def fused_faces(mesh: TwoLevelMesh, bs: BlockSet):
    """The fused form's interface table (aph_plan_set_amr): (faces [nblocks*6], level,
    buffer length).  Each coarse face to the fine level owns a region of the buffer with
    one slot per (coarse face cell, child) of nvar doubles, transverse axes ascending and
    the first slowest; each fine face there points into that region with its offset in
    fine cells."""
    nd, nx, nv = mesh.ndim, mesh.nx, bs.nvar
    nch = 2 ** (nd - 1)
    faces = np.zeros(bs.nblocks * 6, dtype=AMR_FACE_DTYPE)
    region = {}
    total = 0
    trans = lambda d: ((1 if d == 0 else 0), (1 if d == 2 else 2))
    for b in range(bs.nblocks):
        if mesh.level[b] != 0:
            continue
        for d in range(nd):
            A, B = trans(d)
            for side in (0, 1):
                if bs.blocks["bc"][b][2 * d + side] != BC_AMR:
                    continue
                region[(b, d, side)] = total
                f = faces[b * 6 + 2 * d + side]
                f["base"], f["role"], f["na"] = total, 2, nx[A]
                total += nx[A] * (nx[B] if B < nd else 1) * nch * nv
    for b in range(bs.nblocks):
        if mesh.level[b] != 1:
            continue
        lf = mesh.loc[b]
        for d in range(nd):
            A, B = trans(d)
            for side in (0, 1):
                if bs.blocks["bc"][b][2 * d + side] != BC_AMR:
                    continue
                # the coarse cell across the face, and the coarse block holding it
                G = [int(lf[t]) * nx[t] for t in range(3)]
                G[d] = int(lf[d]) * nx[d] + (nx[d] if side else -1)
                I = [(G[t] % mesh.ncells_axis(1, t)) // 2 if t < nd else 0 for t in range(3)]
                lc = tuple(I[t] // nx[t] if t < nd else 0 for t in range(3))
                cb = mesh.index[(0, lc)]
                f = faces[b * 6 + 2 * d + side]
                f["base"] = region[(cb, d, 1 - side)]
                f["role"] = 1
                f["na"] = nx[A]
                f["qa"] = int(lf[A]) * nx[A] - 2 * lc[A] * nx[A]
                f["qb"] = (int(lf[B]) * nx[B] - 2 * lc[B] * nx[B]) if B < nd else 0
    return faces, mesh.level.astype(np.int32), total


class AmrPlan:
    """The plan on the device (aph_amr_plan_create)."""

    def __init__(self, ctx, dbs, mesh: TwoLevelMesh, bs: BlockSet):
        self.ctx = ctx
        self.ops, self.fixes, self.refs = amr_plan(mesh, bs)
        self.h = ctypes.c_void_p()
        vp = ctypes.c_void_p
        ctx.check(_lib.lib().aph_amr_plan_create(
            ctx.h, dbs, self.ops.ctypes.data_as(vp), len(self.ops),
            self.fixes.ctypes.data_as(vp), len(self.fixes), self.refs.ctypes.data_as(vp),
            len(self.refs), ctypes.byref(self.h)))

    def fill(self, cons):
        self.ctx.check(_lib.lib().aph_amr_fill(self.ctx.h, self.h, ctypes.c_void_p(cons.data_ptr())))

    def flux_correct(self, flux):
        p = [ctypes.c_void_p(f.data_ptr()) if f is not None else None for f in flux]
        self.ctx.check(_lib.lib().aph_amr_flux_correct(self.ctx.h, self.h, *p))

    def __del__(self):
        try:
            if self.h:
                _lib.lib().aph_amr_plan_destroy(self.ctx.h, self.h)
        except Exception:
            pass


def regrid(mesh: TwoLevelMesh, bs: BlockSet, cons, tags):
    """One regrid of a two-level mesh from per-block tags (-1 / 0 / 1, as
    aph_refinement_tag returns them): a level-0 block tagged 1 is refined, a level-1
    sibling group whose 2^ndim blocks are all tagged -1 is merged back; nothing else
    changes (two levels: level-1 blocks do not refine further).  Returns
    (new_mesh, new_bs, new_cons) with the interiors remapped on the host:
    * kept blocks: copied;
    * merged groups: the mean of each coarse cell's children, (dz, dy, dx) order;
    * new children: c + sum_d (sign_d / 4) minmod(c - c_-d, c_+d - c) of their parent's
      cells, the neighbours taken at level 0 (a cell, or the mean of its children).
    Ghost cells are left to the next ghost fill.  Totals are conserved to round-off and
    refining then merging returns the coarse data."""
    nd, nx = mesh.ndim, mesh.nx
    nv = bs.nvar
    tags = np.asarray(tags)
    old_ref = {tuple(int(v) // 2 for v in mesh.loc[b])[:3] for b in range(bs.nblocks)
               if mesh.level[b] == 1}
    old_ref = {t if nd == 3 else (t[0], t[1], 0) for t in old_ref}
    groups = {}
    for b in range(bs.nblocks):
        if mesh.level[b] == 1:
            key = tuple(int(mesh.loc[b][d]) // 2 if d < nd else 0 for d in range(3))
            groups.setdefault(key, []).append(int(tags[b]))
    new_ref = set(old_ref)
    for key, ts in groups.items():
        if len(ts) == 2 ** nd and all(t == -1 for t in ts):
            new_ref.discard(key)
    for b in range(bs.nblocks):
        if mesh.level[b] == 0 and tags[b] == 1:
            new_ref.add(tuple(int(mesh.loc[b][d]) if d < nd else 0 for d in range(3)))
    refined = [t[:nd] for t in sorted(new_ref)]
    new = TwoLevelMesh(nd, mesh.nbase[:nd], nx[:nd], mesh.ng, mesh.xmin, mesh.xmax, refined,
                       mesh_id=mesh.mesh_id)
    nbs = new.block_set(nscalars=bs.nscalars)
    out = np.zeros(nbs.field_size(nv))
    # level-0 values over the whole domain (refined regions: the mean of their children)
    shape0 = tuple(mesh.ncells_axis(0, d) for d in (2, 1, 0))
    coarse = np.zeros((nv,) + shape0)
    fine = {}
    for b in range(bs.nblocks):
        lv, lc = int(mesh.level[b]), mesh.loc[b]
        v = bs.interior(cons, b, nv)
        if lv == 0:
            sl = tuple(slice(int(lc[d]) * nx[d], (int(lc[d]) + 1) * nx[d]) for d in (2, 1, 0))
            coarse[(slice(None),) + sl] = v
        else:
            fine[tuple(int(x) for x in lc)] = v
    for key in old_ref:
        kids = []
        for dz in ((0, 1) if nd == 3 else (0,)):
            for dy in (0, 1):
                for dxx in (0, 1):
                    k = (2 * key[0] + dxx, 2 * key[1] + dy, (2 * key[2] + dz) if nd == 3 else 0)
                    kids.append(fine[k])
        # the parent's cells: children of parent cell (i, j, k) are (2i+dx, 2j+dy, 2k+dz)
        # across the 2^nd child blocks; assemble the 2n fine cells first
        fz = 2 * nx[2] if nd == 3 else 1
        full = np.zeros((nv, fz, 2 * nx[1], 2 * nx[0]))
        idx = 0
        for dz in ((0, 1) if nd == 3 else (0,)):
            for dy in (0, 1):
                for dxx in (0, 1):
                    zs = slice(dz * nx[2], (dz + 1) * nx[2]) if nd == 3 else slice(0, 1)
                    full[:, zs, dy * nx[1]:(dy + 1) * nx[1], dxx * nx[0]:(dxx + 1) * nx[0]] = kids[idx]
                    idx += 1
        parts = []
        for dz in ((0, 1) if nd == 3 else (0,)):
            for dy in (0, 1):
                for dxx in (0, 1):
                    parts.append(full[:, dz::2, dy::2, dxx::2] if nd == 3 else full[:, :, dy::2, dxx::2])
        s = parts[0].copy()
        for q in parts[1:]:
            s = s + q
        s = s * (1.0 / len(parts))
        sl = tuple(slice(key[d] * nx[d], (key[d] + 1) * nx[d]) if d < nd else slice(0, 1)
                   for d in (2, 1, 0))
        coarse[(slice(None),) + sl] = s
    N0 = [mesh.ncells_axis(0, d) for d in range(3)]

    def sgn(x):
        return np.where(x < 0.0, -1.0, 1.0)

    for b in range(nbs.nblocks):
        lv, lc = int(new.level[b]), new.loc[b]
        dst = nbs.interior(out, b, nv)
        if lv == 0:
            sl = tuple(slice(int(lc[d]) * nx[d], (int(lc[d]) + 1) * nx[d]) for d in (2, 1, 0))
            dst[:] = coarse[(slice(None),) + sl]
            continue
        key = tuple(int(lc[d]) for d in range(3))
        if key in fine:  # a level-1 block of the old mesh, kept
            dst[:] = fine[key]
            continue
        # a new child: prolongate from level 0
        G = np.meshgrid(*[int(lc[d]) * nx[d] + np.arange(nx[d]) if d < nd else np.zeros(1, int)
                          for d in (2, 1, 0)], indexing="ij")
        G = [G[2], G[1], G[0]]  # x, y, z index arrays, shape (nz, ny, nx)
        I = [G[d] // 2 if d < nd else G[d] for d in range(3)]
        at = lambda J: coarse[(slice(None),) + tuple(J[d] % N0[d] for d in (2, 1, 0))]
        c = at(I)
        val = c.copy()
        for d in range(nd):
            lo = [I[t] - (1 if t == d else 0) for t in range(3)]
            hi = [I[t] + (1 if t == d else 0) for t in range(3)]
            a, bb = c - at(lo), at(hi) - c
            g = (0.5 * (sgn(a) + sgn(bb))) * np.minimum(np.abs(a), np.abs(bb))
            val = val + (np.where(G[d] % 2 == 0, -1.0, 1.0) * 0.25) * g
        dst[:] = val
    return new, nbs, out
