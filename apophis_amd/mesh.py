"""Block-set construction: the host side of the meshblock layout.

The reference gets its blocks from Parthenon's Mesh (uniform block size per pack,
``src/hydro/hydro.cpp:537-540``).  Here a block set is a flat table of
``aph_block`` records (``include/apophis_hydro.h``) built from one or more
*lattice meshes*: a mesh is a Bx x By x Bz arrangement of equal blocks with a
boundary condition per domain face; several independent meshes (uneven AMR-like
block collections, batched instances) concatenate into one block set.

Field storage is the pack layout ``[block][var][k][j][i]`` with ghosts; the helpers
here convert between that flat layout and per-block numpy views.
"""
from __future__ import annotations

import numpy as np

BC_PERIODIC = 0
BC_OUTFLOW = 1  # APH_BC_OUTFLOW
BC_REFLECT = 2  # APH_BC_REFLECT
_BC_NAMES = {"periodic": BC_PERIODIC, "outflow": BC_OUTFLOW, "reflect": BC_REFLECT,
             "reflecting": BC_REFLECT}

# numpy mirror of struct aph_block (240 bytes, no padding)
BLOCK_DTYPE = np.dtype([
    ("nx", "<i4", (3,)), ("ng", "<i4"), ("cell_base", "<i8"),
    ("dx", "<f8", (3,)), ("area", "<f8", (3,)), ("vol", "<f8"), ("xmin", "<f8", (3,)),
    ("nbr", "<i4", (27,)), ("bc", "<i4", (6,)), ("mesh_id", "<i4"),
])
assert BLOCK_DTYPE.itemsize == 240

EOS_DTYPE = np.dtype([("gm1", "<f8"), ("cv", "<f8"), ("dfloor", "<f8"), ("pfloor", "<f8")])

# hydro.cpp:295-296: sqrt(1024 * FLT_MIN)
DEFAULT_FLOOR = float(np.sqrt(1024.0 * np.finfo(np.float32).tiny))


def _bc_code(name):
    if isinstance(name, (int, np.integer)):
        return int(name)
    try:
        return _BC_NAMES[name]
    except KeyError:
        raise ValueError(f"unknown boundary condition {name!r}") from None


def lattice_mesh(ndim, nblocks, block_nx, ng, xmin, xmax, bc, mesh_id=0):
    """Blocks of one uniform lattice mesh.

    ndim: 1..3; nblocks: blocks per axis (3-tuple, 1 on inactive axes);
    block_nx: interior cells per block per axis (1 on inactive axes);
    xmin/xmax: domain extent per axis (inactive axes set the coordinate widths, as
    Parthenon's UniformCartesian does); bc: 6 names for faces -x1,+x1,-x2,+x2,-x3,+x3.
    Returns a BLOCK_DTYPE array with block-local neighbour indices (0-based within
    this mesh) and cell_base unset.
    """
    nb = tuple(int(v) for v in nblocks)
    bnx = tuple(int(v) for v in block_nx)
    for d in range(ndim, 3):
        if nb[d] != 1 or bnx[d] != 1:
            raise ValueError("inactive axes must have one block of one cell")
    codes = [_bc_code(b) for b in bc]
    total = nb[0] * nb[1] * nb[2]
    out = np.zeros(total, dtype=BLOCK_DTYPE)
    dx = [(float(xmax[d]) - float(xmin[d])) / (nb[d] * bnx[d]) for d in range(3)]
    area = (dx[1] * dx[2], dx[0] * dx[2], dx[0] * dx[1])
    vol = dx[0] * dx[1] * dx[2]

    def bid(lx, ly, lz):
        return (lz * nb[1] + ly) * nb[0] + lx

    for lz in range(nb[2]):
        for ly in range(nb[1]):
            for lx in range(nb[0]):
                r = out[bid(lx, ly, lz)]
                r["nx"] = bnx
                r["ng"] = ng
                r["dx"] = dx
                r["area"] = area
                r["vol"] = vol
                r["xmin"] = [float(xmin[0]) + lx * bnx[0] * dx[0],
                             float(xmin[1]) + ly * bnx[1] * dx[1],
                             float(xmin[2]) + lz * bnx[2] * dx[2]]
                r["mesh_id"] = mesh_id
                r["bc"] = [c if c != BC_PERIODIC else BC_OUTFLOW for c in codes]
                nbr = np.full(27, -1, dtype=np.int32)
                loc = (lx, ly, lz)
                for oz in (-1, 0, 1):
                    for oy in (-1, 0, 1):
                        for ox in (-1, 0, 1):
                            o = (ox, oy, oz)
                            p = list(loc)
                            ok = True
                            for d in range(3):
                                if o[d] == 0:
                                    continue
                                if d >= ndim:
                                    ok = False
                                    break
                                p[d] += o[d]
                                if 0 <= p[d] < nb[d]:
                                    continue
                                face = 2 * d + (1 if o[d] > 0 else 0)
                                if codes[face] == BC_PERIODIC:
                                    p[d] %= nb[d]
                                else:
                                    ok = False
                                    break
                            if ok:
                                nbr[(oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)] = bid(*p)
                r["nbr"] = nbr
    return out


class BlockSet:
    """A flat table of blocks (one or more meshes) plus layout helpers."""

    def __init__(self, ndim, meshes, nscalars=0):
        self.ndim = int(ndim)
        self.nscalars = int(nscalars)
        self.nvar = 5 + self.nscalars
        parts = []
        base = 0
        for m in meshes:
            m = m.copy()
            nbr = m["nbr"]
            nbr[nbr >= 0] += base
            m["nbr"] = nbr
            parts.append(m)
            base += len(m)
        self.blocks = np.concatenate(parts) if parts else np.zeros(0, dtype=BLOCK_DTYPE)
        self.nblocks = len(self.blocks)
        ext = self.extents()
        self._ext = ext  # fixed with ncell and cell_base below
        nc = ext.prod(axis=1)
        self.ncell = nc
        cb = np.zeros(self.nblocks, dtype=np.int64)
        if self.nblocks:
            cb[1:] = np.cumsum(nc)[:-1]
        self.blocks["cell_base"] = cb
        self.ncells_total = int(nc.sum())

    def extents(self):
        """Entire (ghosted) extents per block as an (nblocks, 3) array [Nx, Ny, Nz]."""
        nx = self.blocks["nx"].astype(np.int64)
        ng = self.blocks["ng"].astype(np.int64)
        ext = nx.copy()
        for d in range(self.ndim):
            ext[:, d] += 2 * ng
        return ext

    def interior_cells(self):
        return int(self.blocks["nx"].astype(np.int64).prod(axis=1).sum())

    def field_size(self, nv):
        return nv * self.ncells_total

    def block_view(self, field, b, nv):
        """View of block b of a flat field with nv components as (nv, Nz, Ny, Nx)."""
        Nx, Ny, Nz = (int(v) for v in self._ext[b])
        base = nv * int(self.blocks["cell_base"][b])
        n = int(self.ncell[b])
        return field[base: base + nv * n].reshape(nv, Nz, Ny, Nx)

    def interior_slices(self, b):
        g = int(self.blocks["ng"][b])
        nx = self.blocks["nx"][b]
        sl = []
        for d in (2, 1, 0):
            if d < self.ndim:
                sl.append(slice(g, g + int(nx[d])))
            else:
                sl.append(slice(0, 1))
        return tuple(sl)

    def interior(self, field, b, nv):
        """Interior of block b as (nv, nz, ny, nx) (a view)."""
        return self.block_view(field, b, nv)[(slice(None),) + self.interior_slices(b)]

    def cell_centers(self, b):
        """Cell-centre coordinates (x, y, z) of block b's interior, each (nz, ny, nx)."""
        r = self.blocks[b]
        c = []
        for d in range(3):
            n = int(r["nx"][d])
            c.append(r["xmin"][d] + (np.arange(n) + 0.5) * r["dx"][d])
        z, y, x = np.meshgrid(c[2], c[1], c[0], indexing="ij")
        return x, y, z

    def _uniform_shape(self):
        nx, ng = self.blocks["nx"], self.blocks["ng"]
        if self.nblocks == 0 or not ((nx == nx[0]).all() and (ng == ng[0]).all()):
            raise ValueError("blocks of one shape (and ghost width) required")
        return [int(v) for v in nx[0]], int(ng[0])

    def interior_index(self, nv=None):
        """Flat pack indices of every interior cell as (nblocks, nv, nz*ny*nx), for block
        sets whose blocks share one shape."""
        nv = self.nvar if nv is None else int(nv)
        n, ng = self._uniform_shape()
        Nx, Ny, Nz = (int(v) for v in self.extents()[0])
        g = [ng if d < self.ndim else 0 for d in range(3)]
        k, j, i = np.meshgrid(np.arange(n[2]) + g[2], np.arange(n[1]) + g[1],
                              np.arange(n[0]) + g[0], indexing="ij")
        local = ((k * Ny + j) * Nx + i).reshape(-1).astype(np.int64)
        base = nv * self.blocks["cell_base"].astype(np.int64)
        comp = np.arange(nv, dtype=np.int64) * (Nx * Ny * Nz)
        return base[:, None, None] + comp[None, :, None] + local[None, None, :]

    def cell_centers_all(self):
        """cell_centers of every block at once, each (nblocks, nz, ny, nx), for block sets
        whose blocks share one shape (same arithmetic, so the same values)."""
        n, _ = self._uniform_shape()
        c = []
        for d in range(3):
            c.append(self.blocks["xmin"][:, d][:, None] +
                     (np.arange(n[d]) + 0.5)[None, :] * self.blocks["dx"][:, d][:, None])
        x = np.broadcast_to(c[0][:, None, None, :], (self.nblocks, n[2], n[1], n[0]))
        y = np.broadcast_to(c[1][:, None, :, None], (self.nblocks, n[2], n[1], n[0]))
        z = np.broadcast_to(c[2][:, :, None, None], (self.nblocks, n[2], n[1], n[0]))
        return x, y, z

    def interior_mask(self):
        """Boolean mask over one component (ncells_total) selecting interior cells."""
        m = np.zeros(self.ncells_total, dtype=bool)
        for b in range(self.nblocks):
            v = self.block_view(m, b, 1)
            v[(0,) + self.interior_slices(b)] = True
        return m

    def size_classes(self):
        """Distinct interior edge lengths (cubic/square blocks) present in the set."""
        return sorted(set(int(v) for v in self.blocks["nx"][:, 0]))


def make_eos(gm1, cv=1.5, dfloor=DEFAULT_FLOOR, pfloor=DEFAULT_FLOOR):
    e = np.zeros(1, dtype=EOS_DTYPE)
    e["gm1"], e["cv"], e["dfloor"], e["pfloor"] = gm1, cv, dfloor, pfloor
    return e
