"""Multilevel meshes for the plain and fused forms: a block tree with 2:1 jumps, as
Parthenon's AMR and static refinement produce, whose axes are periodic or end in physical
(outflow / reflecting) faces.

The reference leaves multilevel meshes to Parthenon (apophis_driver.cpp:68,74-76,94):
* the boundary exchange of `cons` fills a block's ghost cells over a finer region with the
  mean of the fine children (RestrictAverage) and over a coarser region by limited linear
  prolongation (ProlongateSharedMinMod);
* the flux correction replaces every coarse face flux on a coarse-fine face by the
  area-weighted mean of the fine fluxes across it, before the update, so the levels
  exchange exactly the same fluxes and the total is conserved.

Here the host resolves the mesh tree once into flat op lists (``amr_plan``), which the
library validates and runs as two gather kernels (``aph_amr_fill``,
``aph_amr_flux_correct``).  Same-level neighbours keep using the block table's ``nbr``
(``aph_fill_ghosts``); a face whose neighbour region is on another level has nbr = -1 and
bc = APH_BC_AMR.  Leaves that touch differ by at most one level (2:1 balance across
faces, edges and corners), so every cross-level value is one level away.  Each block
keeps the same cells per block at every level (spacing halves), so a coarse cell's
children always lie in one fine block.

On a physical domain face the ghost cells beyond the domain are the boundary conditions'
(``aph_amr_fill`` runs them again after the cross-level values, so edge and corner ghosts
copy the cross-level ghosts beside them, Parthenon's order); a prolongation's slope
neighbour beyond the domain is the coarse level's boundary image, as Parthenon applies the
boundary conditions to the coarse buffer before it prolongates: the coarse cell itself
(outflow) or its mirror image with the normal momentum negated (reflect).  Adaptive
regrids (``regrid``) take periodic meshes; ``static_refinement`` builds the tree of a
deck's ``<parthenon/static_refinementN>`` blocks.

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
from .mesh import BC_OUTFLOW, BC_PERIODIC, BC_REFLECT, BLOCK_DTYPE, BlockSet, _bc_code

BC_AMR = 3  # APH_BC_AMR
AMR_MEAN, AMR_PROLONG = 0, 1
AMR_MIRROR_BIT = 8  # APH_AMR_MIRROR_BIT: value slot k mirrored at bit 8 + k - 1

AMR_REF_DTYPE = np.dtype([("offset", "<i8"), ("stride", "<i8")])
AMR_VAL_DTYPE = np.dtype([("first", "<i4"), ("count", "<i4")])
AMR_OP_DTYPE = np.dtype([("dst", "<i8"), ("dst_stride", "<i8"), ("kind", "<i4"),
                         ("sign", "<i4", (3,)), ("val", AMR_VAL_DTYPE, (7,))])
FLUX_FIX_DTYPE = np.dtype([("dst", "<i8"), ("dst_stride", "<i8"), ("dir", "<i4"),
                           ("reserved", "<i4"), ("val", AMR_VAL_DTYPE)])
assert AMR_REF_DTYPE.itemsize == 16 and AMR_OP_DTYPE.itemsize == 88
assert FLUX_FIX_DTYPE.itemsize == 32


class AmrMesh:
    """Leaf blocks of a periodic block tree, as Parthenon's AMR builds it: a base lattice
    ``nbase`` of level-0 blocks, where a refined node at level L is replaced by its 2^ndim
    children at level L+1 (the same ``block_nx`` cells at half the spacing).  ``refined``
    lists the refined nodes as (level, loc); it holds every ancestor of its members, and
    the leaves are 2:1 balanced: leaves that touch (across a face, an edge or a corner,
    with periodic wrap) differ by at most one level.  ``max_level`` is the deepest level a
    regrid may create (numlevel - 1).

    ``blocks`` is a BLOCK_DTYPE table, leaves in depth-first order (base lattice z, y, x;
    children z, y, x): nbr links same-level leaves, faces with no same-level neighbour
    carry BC_AMR.  ``level[b]``, ``loc[b]`` give each leaf's level and logical position;
    ``index[(level, loc)]`` is the leaf's block number.

    ``bc``: six domain face codes or names (-x1, +x1, -x2, +x2, -x3, +x3), "periodic" on
    both faces of a periodic axis, else "outflow" / "reflect"; None: every axis periodic.
    Across a physical face there is no neighbour (and no wrap)."""

    def __init__(self, ndim, nbase, block_nx, ng, xmin, xmax, refined, mesh_id=0,
                 max_level=None, bc=None):
        if ndim not in (2, 3):
            raise ValueError("multilevel meshes are 2D or 3D")
        self.ndim = ndim
        self.bc = _face_codes(bc, ndim)
        self.periodic = tuple(self.bc[2 * d] == BC_PERIODIC for d in range(3))
        self.xmin = tuple(float(v) for v in xmin)
        self.xmax = tuple(float(v) for v in xmax)
        self.mesh_id = int(mesh_id)
        self.nbase = tuple(int(v) for v in nbase) + (1,) * (3 - len(nbase))
        self.nx = tuple(int(v) for v in block_nx) + (1,) * (3 - len(block_nx))
        self.ng = int(ng)
        if any(self.nx[d] < 2 * self.ng for d in range(ndim)):
            raise ValueError("blocks need at least 2*ng cells per axis")
        if any(self.nx[d] % 2 for d in range(ndim)):
            raise ValueError("multilevel meshes need an even number of cells per block axis")
        self.refined = set()
        for lv, lc in refined:
            lc = tuple(int(v) for v in lc) + (0,) * (3 - len(lc))
            self.refined.add((int(lv), lc))
        for lv, lc in self.refined:
            if lv < 0 or any(not 0 <= lc[d] < self.nlat(lv, d) for d in range(3)):
                raise ValueError(f"refined node {(lv, lc)} is outside the mesh")
            if lv > 0 and (lv - 1, self.parent(lc)) not in self.refined:
                raise ValueError(f"refined node {(lv, lc)} has an unrefined parent")
        deepest = 1 + max((lv for lv, _ in self.refined), default=-1)
        self.max_level = deepest if max_level is None else int(max_level)
        if self.max_level < deepest:
            raise ValueError(f"the tree is {deepest} levels deep, max_level {self.max_level}")
        leaves = []  # (level, loc), depth first

        def visit(lv, lc):
            if (lv, lc) not in self.refined:
                leaves.append((lv, lc))
                return
            for c in self.children(lc):
                visit(lv + 1, c)

        for lz in range(self.nbase[2]):
            for ly in range(self.nbase[1]):
                for lx in range(self.nbase[0]):
                    visit(0, (lx, ly, lz))
        self.level = np.array([lv for lv, _ in leaves], dtype=np.int32)
        self.loc = np.array([lc for _, lc in leaves], dtype=np.int64).reshape(-1, 3)
        self.index = {(lv, lc): b for b, (lv, lc) in enumerate(leaves)}
        self.blocks = self._native_blocks()

    # -- tree geometry
    def nlat(self, level, d):
        """Nodes per axis d at a level."""
        return self.nbase[d] * (2 ** level if d < self.ndim else 1)

    def _lattice(self, level):
        cache = self.__dict__.setdefault("_lat", {})
        n = cache.get(level)
        if n is None:
            n = cache[level] = tuple(self.nlat(level, d) for d in range(3))
        return n

    def ncells_axis(self, level, d):
        return self.nlat(level, d) * self.nx[d]

    @property
    def nlevels(self):
        return int(self.level.max()) + 1 if len(self.level) else 1

    def offsets(self):
        """The 3^ndim - 1 neighbour offsets."""
        if getattr(self, "_offsets", None) is None:
            self._offsets = [o for o in itertools.product((-1, 0, 1), repeat=3)
                             if o != (0, 0, 0) and all(o[d] == 0 for d in range(self.ndim, 3))]
        return self._offsets

    def shift(self, level, lc, o):
        """The node at offset o from lc (wrapped on periodic axes), or None beyond a
        physical domain face."""
        n = self._lattice(level)
        out = []
        for d in range(3):
            c = lc[d] + o[d]
            if self.periodic[d]:
                c %= n[d]
            elif not 0 <= c < n[d]:
                return None
            out.append(c)
        return tuple(out)

    def parent(self, lc):
        return tuple(lc[d] // 2 if d < self.ndim else 0 for d in range(3))

    def children(self, lc):
        """The 2^ndim child positions, z, y, x order (x fastest)."""
        nd = self.ndim
        return [(2 * lc[0] + cx, 2 * lc[1] + cy, 2 * lc[2] + cz if nd == 3 else 0)
                for cz in ((0, 1) if nd == 3 else (0,)) for cy in (0, 1) for cx in (0, 1)]

    def is_node(self, level, lc):
        """Whether (level, lc) is a leaf or a refined node of the tree."""
        return level == 0 or (level - 1, self.parent(lc)) in self.refined

    def touching_children(self, lc, o):
        """Children of node lc that touch the node at offset -o from it (the neighbour
        that sees lc at offset o)."""
        out = []
        for c in self.children(lc):
            if all(o[d] == 0 or (c[d] % 2) == (1 if o[d] < 0 else 0)
                   for d in range(self.ndim)):
                out.append(c)
        return out

    def _native_blocks(self):
        """The block table from the library's tree (aph_amr_block_table), which checks 2:1
        balance; byte for byte the table of ``block_table_py``."""
        vp = ctypes.c_void_p
        n = len(self.level)
        out = np.zeros(n, dtype=BLOCK_DTYPE)
        nbase = np.array(self.nbase, dtype=np.int32)
        nx = np.array(self.nx, dtype=np.int32)
        xmin = np.array(self.xmin, dtype=np.float64)
        xmax = np.array(self.xmax, dtype=np.float64)
        level = np.ascontiguousarray(self.level, dtype=np.int32)
        loc = np.ascontiguousarray(self.loc, dtype=np.int64)
        bc = np.array(self.bc, dtype=np.int32)
        err = ctypes.create_string_buffer(512)
        rc = _lib.lib().aph_amr_block_table(
            self.ndim, nbase.ctypes.data_as(vp), nx.ctypes.data_as(vp), self.ng,
            xmin.ctypes.data_as(vp), xmax.ctypes.data_as(vp), bc.ctypes.data_as(vp),
            self.mesh_id, n, level.ctypes.data_as(vp), loc.ctypes.data_as(vp),
            out.ctypes.data_as(vp), err, len(err))
        if rc != _lib.APH_OK:
            raise ValueError(err.value.decode())
        return out

    def block_set(self, nscalars=0):
        return BlockSet(self.ndim, [self.blocks], nscalars=nscalars)


class TwoLevelMesh(AmrMesh):
    """An AmrMesh of two levels: the base blocks at the logical positions ``refined`` are
    replaced by their 2^ndim children (level 1), and regrids stay within the two levels."""

    def __init__(self, ndim, nbase, block_nx, ng, xmin, xmax, refined, mesh_id=0, bc=None):
        super().__init__(ndim, nbase, block_nx, ng, xmin, xmax,
                         [(0, tuple(r)) for r in refined], mesh_id=mesh_id, max_level=1,
                         bc=bc)


def _face_codes(bc, ndim):
    """Six domain face codes (0 periodic, 1 outflow, 2 reflect) from codes or names."""
    if bc is None:
        return (BC_PERIODIC,) * 6
    if len(bc) != 6:
        raise ValueError("bc needs six face codes (-x1, +x1, -x2, +x2, -x3, +x3)")
    codes = [_bc_code(b) for b in bc]
    for d in range(3):
        if d >= ndim:
            codes[2 * d] = codes[2 * d + 1] = BC_PERIODIC
            continue
        lo, hi = codes[2 * d], codes[2 * d + 1]
        if (lo == BC_PERIODIC) != (hi == BC_PERIODIC):
            raise ValueError(f"axis {d}: a periodic face needs a periodic opposite face")
        if any(c not in (BC_PERIODIC, BC_OUTFLOW, BC_REFLECT) for c in (lo, hi)):
            raise ValueError(f"axis {d}: faces must be periodic, outflow or reflect")
    return tuple(codes)


def static_refinement(ndim, nbase, block_nx, xmin, xmax, regions, bc=None):
    """The refined nodes of a mesh with static refinement regions, as Parthenon builds its
    tree from ``<parthenon/static_refinementN>`` blocks (Mesh constructor; the rule
    Athena++ introduced): for a region (rxmin, rxmax, level) every node of that level whose
    right edge lies beyond rxmin, up to the first whose right edge reaches rxmax, on each
    active axis, widened to whole sibling pairs (an odd first index steps down, an even
    last one up), is added with its ancestors; then every leaf that touches a leaf two
    levels finer is refined until none does (2:1 balance across faces, edges and corners;
    no wrap across physical faces).  Node edges are xmin + (l + 1) (xmax - xmin) /
    (nbase 2^level)."""
    nd = int(ndim)
    nb = tuple(int(v) for v in nbase) + (1,) * (3 - len(nbase))
    codes = _face_codes(bc, nd)
    per = [codes[2 * d] == BC_PERIODIC for d in range(3)]
    ref = set()

    def parent(lc):
        return tuple(lc[d] // 2 if d < nd else 0 for d in range(3))

    def add_node(lv, lc):  # the node and its ancestors become refined above it
        for L in range(lv - 1, -1, -1):
            lc = parent(lc)
            ref.add((L, lc))

    for rmin, rmax, level in regions:
        level = int(level)
        if level < 1:
            continue
        rng = []
        for d in range(3):
            if d >= nd:
                rng.append((0, 0))
                continue
            n = nb[d] << level
            w = (float(xmax[d]) - float(xmin[d])) / n
            right = lambda l: float(xmin[d]) + (l + 1) * w
            lo = 0
            while lo < n and not right(lo) > float(rmin[d]):
                lo += 1
            hi = lo
            while hi < n and not right(hi) >= float(rmax[d]):
                hi += 1
            if lo % 2 == 1:
                lo -= 1
            if hi % 2 == 0:
                hi += 1
            rng.append((lo, min(hi, n - 1)))
        for z in range(rng[2][0], rng[2][1] + 1):
            for y in range(rng[1][0], rng[1][1] + 1):
                for x in range(rng[0][0], rng[0][1] + 1):
                    add_node(level, (x, y, z))

    offsets = [o for o in itertools.product((-1, 0, 1), repeat=3)
               if o != (0, 0, 0) and all(o[d] == 0 for d in range(nd, 3))]

    def shift(lv, lc, o):
        out = []
        for d in range(3):
            n = nb[d] << lv if d < nd else 1
            c = lc[d] + o[d]
            if per[d]:
                c %= n
            elif not 0 <= c < n:
                return None
            out.append(c)
        return tuple(out)

    def children(lc):
        return [(2 * lc[0] + cx, 2 * lc[1] + cy, 2 * lc[2] + cz if nd == 3 else 0)
                for cz in ((0, 1) if nd == 3 else (0,)) for cy in (0, 1) for cx in (0, 1)]

    def leaves():
        stack = [(0, (x, y, z)) for z in range(nb[2]) for y in range(nb[1])
                 for x in range(nb[0])]
        while stack:
            lv, lc = stack.pop()
            if (lv, lc) in ref:
                stack.extend((lv + 1, c) for c in children(lc))
            else:
                yield lv, lc

    def touches_finer(lv, lc):
        for o in offsets:
            p = shift(lv, lc, o)
            if p is None or (lv, p) not in ref:
                continue
            for c in children(p):
                if all(o[d] == 0 or (c[d] % 2) == (1 if o[d] < 0 else 0) for d in range(nd)) \
                        and (lv + 1, c) in ref:
                    return True
        return False

    changed = True
    while changed:
        changed = False
        for lv, lc in list(leaves()):
            if touches_finer(lv, lc):
                ref.add((lv, lc))
                changed = True
    return ref


def mesh_from_pin(pin, ng):
    """The mesh of an input deck's ``<parthenon/mesh>``, ``<parthenon/meshblock>`` and
    ``<parthenon/static_refinementN>`` blocks (pin: nested dict {block: {key: value}} with
    the deck's names and values): nx1..3, x1min..x3max, ix1_bc..ox3_bc (periodic, outflow,
    reflecting) and numlevel of the mesh; the block size; each static refinement block's
    x1min..x3max and level.  Returns an AmrMesh (with no static refinement block, its single
    level).  ng: the ghost layers (the reconstruction's, hydro.cpp:56-67)."""
    m = pin["parthenon/mesh"]
    mb = pin.get("parthenon/meshblock", {})
    nx = [int(m.get(f"nx{d + 1}", 1)) for d in range(3)]
    nd = 3 if nx[2] > 1 else (2 if nx[1] > 1 else 1)
    bnx = [int(mb.get(f"nx{d + 1}", nx[d])) for d in range(3)]
    if any(nx[d] % bnx[d] for d in range(nd)):
        raise ValueError("the mesh size must be a multiple of the block size")
    nbase = [nx[d] // bnx[d] for d in range(nd)]
    xmin = [float(m.get(f"x{d + 1}min", 0.0 if d < nd else -0.5)) for d in range(3)]
    xmax = [float(m.get(f"x{d + 1}max", 1.0 if d < nd else 0.5)) for d in range(3)]
    names = {"reflecting": "reflect"}
    bc = []
    for d in range(3):
        for side in ("i", "o"):
            b = str(m.get(f"{side}x{d + 1}_bc", "periodic"))
            bc.append(names.get(b, b))
    regions = []
    for block, keys in pin.items():
        if not block.startswith("parthenon/static_refinement"):
            continue
        lo = [float(keys.get(f"x{d + 1}min", xmin[d])) for d in range(3)]
        hi = [float(keys.get(f"x{d + 1}max", xmax[d])) for d in range(3)]
        regions.append((lo, hi, int(keys["level"])))
    numlevel = int(m.get("numlevel", 1))
    if any(r[2] > numlevel - 1 for r in regions):
        raise ValueError("a static refinement level beyond numlevel - 1")
    ref = static_refinement(nd, nbase, bnx[:nd], xmin, xmax, regions, bc)
    return AmrMesh(nd, nbase, bnx[:nd], ng, xmin, xmax, ref, max_level=max(numlevel - 1, 0),
                   bc=bc)


def block_table_py(mesh: AmrMesh):
    """The block table of a mesh's leaves, stated in Python (the library builds the same
    bytes, aph_amr_block_table): 2:1 balance checked; dx, areas, volume and xmin per
    leaf; nbr to same-level leaves (wrapping on periodic axes); the domain's code on faces
    on a physical domain face, BC_AMR on other faces without a same-level neighbour."""
    ndim = mesh.ndim
    for lv, lc in mesh.index:
        for o in mesh.offsets():
            p = mesh.shift(lv, lc, o)
            if p is None or (lv, p) in mesh.index:
                continue
            if (lv, p) in mesh.refined:
                for c in mesh.touching_children(p, o):
                    if (lv + 1, c) in mesh.refined:
                        raise ValueError(f"leaf {(lv, lc)} touches level {lv + 2} "
                                         "(2:1 balance)")
                continue
            if (lv - 1, mesh.parent(p)) not in mesh.index:
                raise ValueError(f"leaf {(lv, lc)} touches a level below {lv - 1} "
                                 "(2:1 balance)")
    span = [float(mesh.xmax[d]) - float(mesh.xmin[d]) for d in range(3)]
    out = np.zeros(len(mesh.level), dtype=BLOCK_DTYPE)
    for b in range(len(mesh.level)):
        lv, lc = int(mesh.level[b]), tuple(int(v) for v in mesh.loc[b])
        r = out[b]
        f = 2 ** lv
        dx = [span[d] / (mesh.nbase[d] * mesh.nx[d] * (f if d < ndim else 1))
              for d in range(3)]
        r["nx"] = mesh.nx
        r["ng"] = mesh.ng
        r["dx"] = dx
        r["area"] = (dx[1] * dx[2], dx[0] * dx[2], dx[0] * dx[1])
        r["vol"] = dx[0] * dx[1] * dx[2]
        r["xmin"] = [float(mesh.xmin[d]) + (lc[d] * mesh.nx[d] * dx[d] if d < ndim else 0.0)
                     for d in range(3)]
        r["mesh_id"] = mesh.mesh_id
        nbr = np.full(27, -1, dtype=np.int32)
        for o in mesh.offsets():
            p = mesh.shift(lv, lc, o)
            if p is not None:
                nbr[(o[2] + 1) * 9 + (o[1] + 1) * 3 + (o[0] + 1)] = mesh.index.get((lv, p), -1)
        r["nbr"] = nbr
        bc = [1] * 6  # outflow placeholder on faces with a same-level neighbour
        for d in range(ndim):
            for side in (0, 1):
                o = [0, 0, 0]
                o[d] = 1 if side else -1
                if mesh.shift(lv, lc, o) is None:
                    bc[2 * d + side] = mesh.bc[2 * d + side]  # a physical domain face
                elif nbr[(o[2] + 1) * 9 + (o[1] + 1) * 3 + (o[0] + 1)] < 0:
                    bc[2 * d + side] = BC_AMR
        r["bc"] = bc
    return out


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


def amr_plan(mesh: AmrMesh, bs: BlockSet):
    """(ops, fixes, refs) for ``bs`` (a block set of ``mesh``'s table).

    Every ghost cell inside the domain whose direction has no same-level neighbour gets
    one op:
    * over a refined region (the leaves there are one level finer): MEAN of the 2^ndim
      children of the ghost cell;
    * over a coarser leaf: PROLONG from the coarse cell holding it, the slope stencil's
      neighbours taken at the coarse level (a coarse cell, or the mean of its children);
      a neighbour beyond a physical domain face is the coarse cell itself, mirrored
      (kind bit AMR_MIRROR_BIT + slot - 1) on a reflecting face.
    Every coarse face cell on a face to a refined region gets one flux fix (the mean of
    the 2^(ndim-1) fine faces across it).  2:1 balance keeps every value one level away."""
    nd, g, nx = mesh.ndim, mesh.ng, mesh.nx
    refs = _Refs()
    act = range(nd)

    def wrap(level, G):
        return tuple((G[d] % mesh.ncells_axis(level, d) if mesh.periodic[d] else G[d])
                     if d < nd else 0 for d in range(3))

    def inside(level, G):
        return all(mesh.periodic[d] or 0 <= G[d] < mesh.ncells_axis(level, d)
                   for d in range(nd))

    def node(G):
        return tuple(G[d] // nx[d] if d < nd else 0 for d in range(3))

    def interior(level, G):
        """(block, ghost-inclusive local index) of the level-`level` leaf holding global
        cell G, or None when that region is not a leaf at this level."""
        G = wrap(level, G)
        if not inside(level, G):
            return None
        lc = node(G)
        b = mesh.index.get((level, lc))
        if b is None:
            return None
        return b, tuple(G[d] - lc[d] * nx[d] + (g if d < nd else 0) for d in range(3))

    def value(level, I):
        """Level-`level` value of global cell I: the leaf cell, or the mean of its 2^ndim
        children, which are leaf cells one level finer."""
        hit = interior(level, I)
        if hit is not None:
            return refs.add([_cell(bs, *hit)])
        I = wrap(level, I)
        kids = []
        for F in mesh.children(I):
            hit = interior(level + 1, F)
            if hit is None:
                raise ValueError(f"cell {I} of level {level} is neither a leaf cell nor "
                                 "the parent of leaf cells (2:1 balance)")
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
            Gw = wrap(lv, tuple(int(lc[d]) * nx[d] + loc[d] - g if d < nd else 0
                                for d in range(3)))
            if not inside(lv, Gw):
                continue  # beyond a physical face: the boundary conditions fill it
            dst, stride = _cell(bs, b, loc)
            rec = np.zeros(1, dtype=AMR_OP_DTYPE)[0]
            rec["dst"], rec["dst_stride"] = dst, stride
            if (lv, node(Gw)) in mesh.refined:  # over a finer region: restriction
                rec["kind"] = AMR_MEAN
                rec["val"][0] = value(lv, Gw)
            else:  # over a coarser leaf: prolongation
                I = tuple(Gw[d] // 2 if d < nd else 0 for d in range(3))
                if lv == 0 or interior(lv - 1, I) is None:
                    raise ValueError(f"block {b}: ghost cell {loc} is over neither a "
                                     "finer region nor a coarser leaf (2:1 balance)")
                kind = AMR_PROLONG
                rec["val"][0] = value(lv - 1, I)
                for d in act:
                    rec["sign"][d] = -1 if Gw[d] % 2 == 0 else 1
                    for side in (0, 1):
                        J = tuple(I[t] + ((1 if side else -1) if t == d else 0)
                                  for t in range(3))
                        slot = 1 + 2 * d + side
                        if inside(lv - 1, J):
                            rec["val"][slot] = value(lv - 1, J)
                        else:  # beyond a physical face: the coarse cell's boundary image
                            rec["val"][slot] = value(lv - 1, I)
                            if mesh.bc[2 * d + side] == BC_REFLECT:
                                kind |= 1 << (AMR_MIRROR_BIT + slot - 1)
                rec["kind"] = kind
            ops.append(rec)
    fixes = []
    for b in range(bs.nblocks):
        lv = int(mesh.level[b])
        lc = mesh.loc[b]
        bc = bs.blocks["bc"][b]
        for d in act:
            for side in (0, 1):
                if bc[2 * d + side] != BC_AMR:
                    continue
                o = [0, 0, 0]
                o[d] = 1 if side else -1
                if (lv, mesh.shift(lv, lc, o)) not in mesh.refined:
                    continue  # the other side is coarser: this block is the fine side
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
                        fb, floc = interior(lv + 1, tuple(F))
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


def fused_faces(mesh: AmrMesh, bs: BlockSet):
    """The fused form's interface table (aph_plan_set_amr): (faces [nblocks*6], level,
    buffer length).  Each face to a finer region (role 2) owns a region of the buffer
    with one slot per (coarse face cell, child) of nvar doubles, transverse axes
    ascending and the first slowest; each face to a coarser leaf (role 1) points into the
    region of the coarse leaf's face across it, with its offset in fine cells.  A block
    can hold faces of both roles (a level between two others)."""
    nd, nx, nv = mesh.ndim, mesh.nx, bs.nvar
    nch = 2 ** (nd - 1)
    faces = np.zeros(bs.nblocks * 6, dtype=AMR_FACE_DTYPE)
    region = {}
    total = 0
    trans = lambda d: ((1 if d == 0 else 0), (1 if d == 2 else 2))

    def across(b, d, side):
        b, d, side = int(b), int(d), int(side)
        o = [0, 0, 0]
        o[d] = 1 if side else -1
        return mesh.shift(int(mesh.level[b]), tuple(int(v) for v in mesh.loc[b]), o)

    bc = bs.blocks["bc"]
    # (block, axis, side, the other side is refined) of every face on a level change
    amr_faces = [(b, d, side, (int(mesh.level[b]), across(b, d, side)) in mesh.refined)
                 for b, d, side in zip(*np.nonzero(bc[:, :2 * nd].reshape(-1, nd, 2) == BC_AMR))]
    for b, d, side, finer in amr_faces:
        if not finer:
            continue
        A, B = trans(d)
        region[(b, d, side)] = total
        f = faces[b * 6 + 2 * d + side]
        f["base"], f["role"], f["na"] = total, 2, nx[A]
        total += nx[A] * (nx[B] if B < nd else 1) * nch * nv
    for b, d, side, finer in amr_faces:
        if finer:
            continue
        lv = int(mesh.level[b])
        A, B = trans(d)
        lf = mesh.loc[b]
        # the coarse cell across the face, and the coarse leaf holding it
        G = [int(lf[t]) * nx[t] for t in range(3)]
        G[d] = int(lf[d]) * nx[d] + (nx[d] if side else -1)
        I = [(G[t] % mesh.ncells_axis(lv, t)) // 2 if t < nd else 0 for t in range(3)]
        lc = tuple(I[t] // nx[t] if t < nd else 0 for t in range(3))
        cb = mesh.index[(lv - 1, lc)]
        f = faces[b * 6 + 2 * d + side]
        f["base"] = region[(cb, d, 1 - side)]
        f["role"] = 1
        f["na"] = nx[A]
        f["qa"] = int(lf[A]) * nx[A] - 2 * lc[A] * nx[A]
        f["qb"] = (int(lf[B]) * nx[B] - 2 * lc[B] * nx[B]) if B < nd else 0
    return faces, mesh.level.astype(np.int32), total


def native_plan(mesh: AmrMesh, bs: BlockSet):
    """(ops, fixes, refs) of ``amr_plan``, built by the library's tree (aph_amr_build):
    the same lists, in milliseconds rather than seconds on large meshes."""
    L = _lib.lib()
    vp = ctypes.c_void_p
    nb = bs.nblocks
    blocks = np.ascontiguousarray(bs.blocks)
    nbase = np.array(mesh.nbase, dtype=np.int32)
    level = np.ascontiguousarray(mesh.level, dtype=np.int32)
    loc = np.ascontiguousarray(mesh.loc, dtype=np.int64)
    bc = np.array(mesh.bc, dtype=np.int32)
    err = ctypes.create_string_buffer(512)
    n = [ctypes.c_int64(0) for _ in range(3)]
    args = lambda o, f, r: (blocks.ctypes.data_as(vp), nb, mesh.ndim, bs.nvar,
                            nbase.ctypes.data_as(vp), level.ctypes.data_as(vp),
                            loc.ctypes.data_as(vp), bc.ctypes.data_as(vp), o,
                            ctypes.byref(n[0]), f, ctypes.byref(n[1]), r,
                            ctypes.byref(n[2]), err, len(err))
    # one build when the arrays are long enough: they are sized from the last build of a
    # mesh of this shape (a regrid changes the lists a little); a build into arrays too
    # short reports the exact lengths, and the second build uses them
    shape = (mesh.ndim, bs.nvar, tuple(int(v) for v in mesh.nx), int(bs.blocks["ng"][0]),
             mesh.bc)
    hint = _PLAN_HINT.get(shape)
    cap = None if hint is None else [int(1.25 * h) + 64 for h in hint]
    for _ in range(3):
        c = cap or [0, 0, 0]
        ops = np.zeros(c[0], dtype=AMR_OP_DTYPE)
        fixes = np.zeros(c[1], dtype=FLUX_FIX_DTYPE)
        refs = np.zeros(c[2], dtype=AMR_REF_DTYPE)
        for q in range(3):
            n[q].value = c[q]
        ptr = [None] * 3 if cap is None else [a.ctypes.data_as(vp) for a in (ops, fixes, refs)]
        rc = L.aph_amr_build(*args(*ptr))
        got = [n[q].value for q in range(3)]
        short = any(g > k for g, k in zip(got, c))
        if rc == _lib.APH_OK and cap is not None:
            break
        if rc != _lib.APH_OK and not (cap is not None and short):
            raise _lib.HydroError(rc, err.value.decode())
        cap = got
    _PLAN_HINT[shape] = got
    return ops[:got[0]], fixes[:got[1]], refs[:got[2]]


_PLAN_HINT = {}  # (ndim, nvar, nx, ng) -> list lengths of the last native build


def poisson_matrix(mesh: AmrMesh, bs: BlockSet, flux_correct=False, prolongation=0,
                   alpha=0.0):
    """The composite Poisson operator the device solver of a multilevel mesh uses, from the
    library's host assembly (aph_poisson_tree_matrix): (rowptr, col, val) over the leaf
    cells and (nb_rowptr, nb_col, nb_val), the six neighbour rows per cell."""
    L = _lib.lib()
    vp = ctypes.c_void_p
    blocks = np.ascontiguousarray(bs.blocks)
    nbase = np.array(mesh.nbase, dtype=np.int32)
    level = np.ascontiguousarray(mesh.level, dtype=np.int32)
    loc = np.ascontiguousarray(mesh.loc, dtype=np.int64)
    err = ctypes.create_string_buffer(512)
    nnz, nbn = ctypes.c_int64(0), ctypes.c_int64(0)
    n = bs.nblocks * int(np.prod(mesh.nx))
    head = (blocks.ctypes.data_as(vp), bs.nblocks, nbase.ctypes.data_as(vp),
            level.ctypes.data_as(vp), loc.ctypes.data_as(vp), int(bool(flux_correct)),
            int(prolongation), float(alpha))
    rc = L.aph_poisson_tree_matrix(*head, None, None, None, ctypes.byref(nnz), None, None,
                                   None, ctypes.byref(nbn), err, len(err))
    if rc != _lib.APH_OK:
        raise _lib.HydroError(rc, err.value.decode())
    rowptr = np.zeros(n + 1, dtype=np.int64)
    col = np.zeros(nnz.value, dtype=np.int32)
    val = np.zeros(nnz.value)
    nbptr = np.zeros(6 * n + 1, dtype=np.int64)
    nbcol = np.zeros(nbn.value, dtype=np.int32)
    nbval = np.zeros(nbn.value)
    rc = L.aph_poisson_tree_matrix(*head, *(a.ctypes.data_as(vp) for a in (rowptr, col, val)),
                                   ctypes.byref(nnz),
                                   *(a.ctypes.data_as(vp) for a in (nbptr, nbcol, nbval)),
                                   ctypes.byref(nbn), err, len(err))
    if rc != _lib.APH_OK:
        raise _lib.HydroError(rc, err.value.decode())
    return rowptr, col, val, nbptr, nbcol, nbval


class AmrPlan:
    """The plan on the device (aph_amr_plan_create), its lists built by the library's
    tree (aph_amr_build)."""

    def __init__(self, ctx, dbs, mesh: AmrMesh, bs: BlockSet):
        self.ctx = ctx
        self.mesh = mesh
        self.ops, self.fixes, self.refs = native_plan(mesh, bs)
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


def _restrict(a, nd):
    """Mean of the 2^nd children of every cell of a level array (nv, Z, Y, X), summed in
    (dz, dy, dx) order, x fastest, and scaled by 2^-nd."""
    parts = [a[:, dz::2, dy::2, dx::2] if nd == 3 else a[:, :, dy::2, dx::2]
             for dz in ((0, 1) if nd == 3 else (0,)) for dy in (0, 1) for dx in (0, 1)]
    s = parts[0].copy()
    for q in parts[1:]:
        s = s + q
    return s * (1.0 / len(parts))


def _segments(lo, n, nx, nlat):
    """The node-aligned pieces of the cell range [lo, lo + n) of one axis (periodic):
    (node, first, end, offset) -- cells [first, end) of node `node`, at `offset` in the
    range."""
    out = []
    c = lo
    while c < lo + n:
        node = c // nx
        end = min((node + 1) * nx, lo + n)
        out.append((node % nlat, c - node * nx, end - node * nx, c - lo))
        c = end
    return out


class _Sampler:
    """Values of a mesh's state at any level over any box of cells: a leaf's cells where
    a leaf of that level holds them, the mean of their children where the node is refined,
    and the minmod prolongation of the level above where a coarser leaf covers them (the
    ghost exchange's rule) -- each box built from the leaves around it only."""

    def __init__(self, mesh: AmrMesh, bs: BlockSet, cons):
        self.m, self.nd, self.nv = mesh, mesh.ndim, bs.nvar
        self.leaf = {(int(mesh.level[b]), tuple(int(v) for v in mesh.loc[b])):
                     bs.interior(cons, b, bs.nvar) for b in range(bs.nblocks)}

    def box(self, L, lo, n):
        """(nv, n_z, n_y, n_x) values of level L over cells lo .. lo + n (x, y, z)."""
        m, nd, nx = self.m, self.nd, self.m.nx
        out = np.empty((self.nv, n[2], n[1], n[0]))
        segs = [_segments(lo[d], n[d], nx[d], m.nlat(L, d)) if d < nd else [(0, 0, 1, 0)]
                for d in range(3)]
        for sz in segs[2]:
            for sy in segs[1]:
                for sx in segs[0]:
                    sg = (sx, sy, sz)
                    node = tuple(q[0] for q in sg)
                    cnt = [q[2] - q[1] for q in sg]
                    dst = (slice(None),) + tuple(slice(sg[d][3], sg[d][3] + cnt[d])
                                                 for d in (2, 1, 0))
                    v = self.leaf.get((L, node))
                    if v is not None:
                        out[dst] = v[(slice(None),) + tuple(slice(sg[d][1], sg[d][2])
                                                            for d in (2, 1, 0))]
                        continue
                    g0 = [node[d] * nx[d] + sg[d][1] if d < nd else 0 for d in range(3)]
                    if (L, node) in m.refined:
                        fine = self.box(L + 1, [2 * g0[d] for d in range(3)],
                                        [2 * cnt[d] if d < nd else 1 for d in range(3)])
                        out[dst] = _restrict(fine, nd)
                    else:
                        out[dst] = self._prolong(L, g0, cnt)
        return out

    def _prolong(self, L, g0, cnt):
        """Level-L cells g0 .. g0 + cnt under a coarser leaf: c + sum_d (sign_d / 4)
        minmod(c - c_-d, c_+d - c) of level L - 1, axes in order."""
        nd = self.nd
        if L == 0:
            raise ValueError("a level-0 region outside every level-0 node")
        clo = [g0[d] // 2 - 1 if d < nd else 0 for d in range(3)]
        chi = [(g0[d] + cnt[d] - 1) // 2 + 1 if d < nd else 0 for d in range(3)]
        C = self.box(L - 1, clo, [chi[d] - clo[d] + 1 for d in range(3)])
        G = [np.arange(g0[d], g0[d] + cnt[d]) if d < nd else np.zeros(1, int) for d in range(3)]
        I = [G[d] // 2 - clo[d] for d in range(3)]
        ix = lambda J: (slice(None), J[2][:, None, None], J[1][None, :, None], J[0][None, None, :])
        c = C[ix(I)]
        val = c
        for d in range(nd):
            lo = list(I)
            hi = list(I)
            lo[d], hi[d] = I[d] - 1, I[d] + 1
            a = c - C[ix(lo)]
            b = C[ix(hi)] - c
            g = (0.5 * (np.where(a < 0.0, -1.0, 1.0) + np.where(b < 0.0, -1.0, 1.0))) * \
                np.minimum(np.abs(a), np.abs(b))
            shape = [1, 1, 1, 1]
            shape[3 - d] = len(G[d])
            sign = np.where(G[d] % 2 == 0, -1.0, 1.0).reshape(shape)
            val = val + (sign * 0.25) * g
        return val


def regrid_tree_py(mesh: AmrMesh, tags):
    """The refined nodes after one regrid (the rules of ``regrid``), stated in Python; the
    library's aph_amr_regrid_tree gives the same tree."""
    _need_periodic(mesh)
    tags = np.asarray(tags)
    ref = set(mesh.refined)
    leaves = [(int(mesh.level[b]), tuple(int(v) for v in mesh.loc[b]))
              for b in range(len(mesh.level))]
    for b, (lv, lc) in enumerate(leaves):
        if tags[b] == 1 and lv < mesh.max_level:
            ref.add((lv, lc))

    def tree_leaves(ref):
        out = []
        stack = [(0, (x, y, z)) for z in range(mesh.nbase[2]) for y in range(mesh.nbase[1])
                 for x in range(mesh.nbase[0])]
        while stack:
            lv, lc = stack.pop()
            if (lv, lc) in ref:
                stack.extend((lv + 1, c) for c in mesh.children(lc))
            else:
                out.append((lv, lc))
        return out

    def touches_finer(lv, lc, ref):
        """Whether leaf (lv, lc) of tree `ref` touches a leaf at level lv + 2 or deeper."""
        for o in mesh.offsets():
            p = mesh.shift(lv, lc, o)
            if (lv, p) in ref and any((lv + 1, c) in ref
                                      for c in mesh.touching_children(p, o)):
                return True
        return False

    changed = True
    while changed:
        changed = False
        for lv, lc in tree_leaves(ref):
            if touches_finer(lv, lc, ref):
                ref.add((lv, lc))
                changed = True
    old_index = mesh.index
    groups = sorted({(lv - 1, mesh.parent(lc)) for b, (lv, lc) in enumerate(leaves)
                     if lv > 0 and tags[b] == -1}, key=lambda P: (-P[0], P[1]))
    for P in groups:
        lv = P[0] + 1
        kids = mesh.children(P[1])
        if any((lv, c) not in old_index or tags[old_index[(lv, c)]] != -1 or (lv, c) in ref
               for c in kids):
            continue
        ref.discard(P)
        if touches_finer(P[0], P[1], ref):
            ref.add(P)
    return ref


def _ancestors(levels, locs, nd):
    ref = set()
    for lv, lc in zip(levels, locs):
        c = tuple(int(v) for v in lc)
        for L in range(int(lv) - 1, -1, -1):
            c = tuple(c[d] // 2 if d < nd else 0 for d in range(3))
            if (L, c) in ref:
                break
            ref.add((L, c))
    return ref


def _need_periodic(mesh: AmrMesh):
    if not all(mesh.periodic):
        raise ValueError("adaptive regrids take periodic meshes; this mesh has physical "
                         "faces (static refinement builds its tree once)")


def regrid_tree(mesh: AmrMesh, tags):
    """The refined nodes after one regrid, from the library's tree (aph_amr_regrid_tree)."""
    level, loc = regrid_leaves(mesh, tags)
    return _ancestors(level, loc, mesh.ndim)


def regrid_leaves(mesh: AmrMesh, tags):
    """The leaves (level [n], loc [n, 3]) after one regrid, in the mesh's leaf order, from
    the library's tree (aph_amr_regrid_tree); equal to (mesh.level, mesh.loc) when the
    tags change nothing."""
    _need_periodic(mesh)
    vp = ctypes.c_void_p
    L = _lib.lib()
    n = len(mesh.level)
    nbase = np.array(mesh.nbase, dtype=np.int32)
    level = np.ascontiguousarray(mesh.level, dtype=np.int32)
    loc = np.ascontiguousarray(mesh.loc, dtype=np.int64)
    t = np.ascontiguousarray(np.asarray(tags), dtype=np.int32)
    err = ctypes.create_string_buffer(512)
    m = ctypes.c_int32(0)
    args = lambda ol, oc, cap: (mesh.ndim, nbase.ctypes.data_as(vp), mesh.max_level, n,
                                level.ctypes.data_as(vp), loc.ctypes.data_as(vp),
                                t.ctypes.data_as(vp), ol, oc, ctypes.byref(m), cap, err,
                                len(err))
    rc = L.aph_amr_regrid_tree(*args(None, None, 0))
    if rc != _lib.APH_OK:
        raise ValueError(err.value.decode())
    out_level = np.zeros(m.value, dtype=np.int32)
    out_loc = np.zeros((m.value, 3), dtype=np.int64)
    rc = L.aph_amr_regrid_tree(*args(out_level.ctypes.data_as(vp), out_loc.ctypes.data_as(vp),
                                     m.value))
    if rc != _lib.APH_OK:
        raise ValueError(err.value.decode())
    return out_level, out_loc


def regrid(mesh: AmrMesh, bs: BlockSet, cons, tags, leaves=None):
    """One regrid from per-block tags (-1 / 0 / 1, as aph_refinement_tag returns them),
    Parthenon's rules on a periodic tree:
    * a leaf tagged 1 below ``mesh.max_level`` is refined;
    * then every leaf that touches a leaf two levels finer is refined (2:1 balance);
    * a sibling group whose 2^ndim leaves were all tagged -1 (and none was refined) is
      merged into its parent, unless the parent would touch a leaf two levels finer;
      groups are taken finest level first, then by position, so a merge on a finer level
      can clear the way for one on a coarser level in the same regrid (a block loses at
      most one level per regrid).
    Returns (new_mesh, new_bs, new_cons): a kept leaf keeps its data, a merged parent
    takes the mean of its children (dz, dy, dx order), a new child the minmod prolongation
    of its parent's level (the neighbours a cell, a mean of children or a prolongation
    from the level above, as the ghost exchange fills them).  Ghost cells are left to the next ghost fill.  Totals are conserved to
    round-off and refining then merging returns the data."""
    nd = mesh.ndim
    _need_periodic(mesh)
    if leaves is None:
        leaves = regrid_leaves(mesh, tags)
    ref = _ancestors(leaves[0], leaves[1], nd)
    new = AmrMesh(nd, mesh.nbase[:nd], mesh.nx[:nd], mesh.ng, mesh.xmin, mesh.xmax, ref,
                  mesh_id=mesh.mesh_id, max_level=mesh.max_level)
    nbs = new.block_set(nscalars=bs.nscalars)
    return new, nbs, remap(mesh, bs, cons, new, nbs)


def remap(mesh: AmrMesh, bs: BlockSet, cons, new: AmrMesh, nbs: BlockSet):
    """The interiors of ``new`` (block set ``nbs``) from the state ``cons`` of ``mesh``,
    by the library (aph_amr_remap): the bits of ``remap_py``."""
    vp = ctypes.c_void_p
    # the library takes every old block as nx^ndim interior cells with ng ghosts and reads
    # nvar * ncell values per block: check that the arrays are laid out that way
    for s_, m_ in ((bs, mesh), (nbs, new)):
        if s_.nblocks != len(m_.level):
            raise ValueError("remap: block set and mesh differ in block count")
        if s_.nblocks and ((s_.blocks["ng"] != mesh.ng).any() or
                           (s_.blocks["nx"] != np.asarray(mesh.nx)[None, :]).any()):
            raise ValueError("remap: every block must have the mesh's nx and ng")
    if np.size(cons) != bs.field_size(bs.nvar):
        raise ValueError(f"remap: cons has {np.size(cons)} values, the block set needs "
                         f"{bs.field_size(bs.nvar)}")
    out = np.zeros(nbs.field_size(bs.nvar))
    u = np.ascontiguousarray(cons, dtype=np.float64)
    args = []
    for m_ in (mesh, new):
        args.append((np.ascontiguousarray(m_.level, dtype=np.int32),
                     np.ascontiguousarray(m_.loc, dtype=np.int64)))
    nbase = np.array(mesh.nbase, dtype=np.int32)
    nx = np.array(mesh.nx, dtype=np.int32)
    err = ctypes.create_string_buffer(512)
    rc = _lib.lib().aph_amr_remap(
        mesh.ndim, nbase.ctypes.data_as(vp), nx.ctypes.data_as(vp), mesh.ng, bs.nvar,
        len(mesh.level), args[0][0].ctypes.data_as(vp), args[0][1].ctypes.data_as(vp),
        u.ctypes.data_as(vp), len(new.level), args[1][0].ctypes.data_as(vp),
        args[1][1].ctypes.data_as(vp), out.ctypes.data_as(vp), err, len(err))
    if rc != _lib.APH_OK:
        raise ValueError(err.value.decode())
    return out


def remap_py(mesh: AmrMesh, bs: BlockSet, cons, new: AmrMesh, nbs: BlockSet):
    """``remap`` stated in Python with the level sampler."""
    nd = mesh.ndim
    nv, nx = bs.nvar, mesh.nx
    old = _Sampler(mesh, bs, cons)
    out = np.zeros(nbs.field_size(nv))
    groups = {}  # new children of an old leaf: prolongated together
    for b in range(nbs.nblocks):
        lv, lc = int(new.level[b]), tuple(int(v) for v in new.loc[b])
        dst = nbs.interior(out, b, nv)
        src = old.leaf.get((lv, lc))
        if src is not None:
            dst[:] = src
        elif lv > 0 and (lv - 1, new.parent(lc)) in old.leaf:
            groups.setdefault((lv, new.parent(lc)), []).append((b, lc))
        else:
            dst[:] = old.box(lv, [lc[d] * nx[d] for d in range(3)], list(nx))
    for (lv, pc), kids in groups.items():
        g0 = [2 * pc[d] * nx[d] if d < nd else 0 for d in range(3)]
        v = old._prolong(lv, g0, [2 * nx[d] if d < nd else 1 for d in range(3)])
        for b, lc in kids:
            sl = tuple(slice(lc[d] * nx[d] - g0[d], (lc[d] + 1) * nx[d] - g0[d])
                       for d in (2, 1, 0))
            nbs.interior(out, b, nv)[:] = v[(slice(None),) + sl]
    return out
