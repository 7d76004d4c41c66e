"""CPU restatement of the reference's Poisson self-gravity (src/gravity/gsolvers/
poisson_gravity.hpp) on a multilevel 3D mesh with physical boundaries on every face (the
static-refinement decks inputs/evrards_collapse.in and inputs/grav_test.in).

TEST INFRASTRUCTURE: imported only by tests/ as the checker.

The discrete problem is the one PoissonEquation defines on the mesh's leaf cells (D = 1):
* flux through a face between cells a (low side) and b (high side) of one block,
  F = (phi_a - phi_b) / dx (CalculateFluxes, :168-224), the value beyond a block face taken
  from the same level (the neighbouring leaf, or the mean of the 2^3 children of a refined
  cell), or prolongated from the coarser leaf beyond (the phi field's registered
  refinement op, gravity.cpp:343-363: "Linear", Parthenon's ProlongateSharedLinear,
  c + sum_d (+-1/4) (c_+d - c_-d) / 2 on a uniform coarse level; "Constant", c), with the
  solver's FixedFace zero boundary on every domain face (SetZeroDirichletBoundary: the
  ghost holds minus the cell inside, on the coarse level too);
* with flux_correct (:89-99), a coarse face over a finer region takes the mean of its
  2^2 fine faces;
* (A phi)_c = -alpha phi_c + sum_d [F(c - 1/2) A - F(c + 1/2) A] / V (FluxMultiplyMatrix,
  :299-349), with each block's own dx, areas and volume;
* right-hand side (ComputeRhs, :502-589): 4 pi G rho, minus 2 phi_b A_f / (dx V) on every
  face on the domain boundary, phi_b the multipole potential of the nine scaled moments
  (ComputeMPCoeff / ScaleMPCoeff, :430-500, summed over every leaf cell with its volume) at
  the face centre.
The operator is assembled as a sparse matrix and solved by Jacobi-preconditioned GMRES to
a relative residual of 1e-14, far inside the tolerance the reference's iterative solvers
(Parthenon's MG / BiCGSTAB) stop at.  The field
(ComputeGravityVector, :591-625) is g = -grad phi by centred differences over the same
ghost values, with 2 phi_b - phi (MultipoleGravityBC) at the domain faces.

Pins (tests/test_oracle_poisson_amr.py): on a single-level mesh the assembled problem and
its solution equal the DST solve of oracle/poisson_oracle.py; the operator annihilates
linear potentials away from the domain faces on every multilevel mesh (prolongation and
flux correction are exact for them; a sign or index slip at a level jump is not); with
flux correction the fluxes through the interfaces balance (the column sums of V A vanish
except at the domain faces); the reference's gravity accuracy criteria for its SMR case.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import poisson_oracle as P


class Problem:
    """The composite Poisson operator of a multilevel mesh (amr.AmrMesh, 3D, physical faces
    on every side) and its block set bs (blocks in the mesh's leaf order)."""

    def __init__(self, mesh, bs, alpha=0.0, flux_correct=False, prolongation="Linear"):
        if mesh.ndim != 3 or any(mesh.periodic):
            raise ValueError("3D meshes with physical faces only")
        if prolongation not in ("Linear", "Constant"):
            raise ValueError("prolongation Linear or Constant")
        self.mesh, self.bs, self.alpha = mesh, bs, float(alpha)
        self.fc, self.prol = bool(flux_correct), prolongation
        self.nx = [int(v) for v in mesh.nx]
        self.ncb = int(np.prod(self.nx))
        self.n = bs.nblocks * self.ncb
        self.leaf = {(int(mesh.level[b]), tuple(int(v) for v in mesh.loc[b])): b
                     for b in range(bs.nblocks)}
        self._value = lru_cache(maxsize=None)(self._value_uncached)
        self.A = self._assemble()

    # -- leaf values at any level
    def lin(self, b, i, j, k):
        nx = self.nx
        return b * self.ncb + (k * nx[1] + j) * nx[0] + i

    def inside(self, L, G):
        return all(0 <= G[d] < self.mesh.ncells_axis(L, d) for d in range(3))

    def _value_uncached(self, L, G):
        """{leaf index: weight} of level L's value at global cell G: the leaf cell, the mean
        of its children over a refined node, None under a coarser leaf."""
        nx = self.nx
        node = tuple(G[d] // nx[d] for d in range(3))
        b = self.leaf.get((L, node))
        if b is not None:
            return {self.lin(b, *(G[d] - node[d] * nx[d] for d in range(3))): 1.0}
        if (L, node) not in self.mesh.refined:
            return None
        out = {}
        for cz in (0, 1):
            for cy in (0, 1):
                for cx in (0, 1):
                    kid = self._value(L + 1, (2 * G[0] + cx, 2 * G[1] + cy, 2 * G[2] + cz))
                    for c, w in kid.items():
                        out[c] = out.get(c, 0.0) + 0.125 * w
        return out

    def ghost(self, L, G, inner_cell):
        """{leaf: weight} of the value at global cell G of level L seen from a block of
        level L (G beyond the block)."""
        if not self.inside(L, G):  # FixedFace zero: minus the cell inside
            return {inner_cell: -1.0}
        v = self._value(L, G)
        if v is not None:
            return v
        I = tuple(g // 2 for g in G)  # the coarser leaf's cell
        c = self._value(L - 1, I)
        if self.prol == "Constant":
            return dict(c)
        out = dict(c)
        for d in range(3):
            sign = -1.0 if G[d] % 2 == 0 else 1.0
            nb = []
            for s in (-1, 1):
                J = list(I)
                J[d] += s
                J = tuple(J)
                nb.append(self._value(L - 1, J) if self.inside(L - 1, J)
                          else {k: -w for k, w in c.items()})
            for part, f in ((nb[1], 1.0), (nb[0], -1.0)):  # (hi - lo) / 2 * (+-1/4)
                for k, w in part.items():
                    out[k] = out.get(k, 0.0) + sign * 0.25 * 0.5 * f * w
        return out

    # -- the operator
    def _face_flux(self, b, d, f, t):
        """{leaf: weight} of the flux F = (phi_low - phi_high) / dx of block b's face f
        (0..n) along axis d at transverse interior cells t = (ta, tb), and whether it lies
        on the domain boundary."""
        mesh, nx = self.mesh, self.nx
        L = int(mesh.level[b])
        lc = [int(v) for v in mesh.loc[b]]
        dx = float(self.bs.blocks["dx"][b][d])
        A, B = [q for q in range(3) if q != d]

        def cell(c):
            idx = [0, 0, 0]
            idx[d], idx[A], idx[B] = c, t[0], t[1]
            if 0 <= c < nx[d]:
                return {self.lin(b, *idx): 1.0}
            G = tuple(lc[q] * nx[q] + idx[q] for q in range(3))
            inner = list(idx)
            inner[d] = 0 if c < 0 else nx[d] - 1
            return self.ghost(L, G, self.lin(b, *inner))

        lo, hi = cell(f - 1), cell(f)
        out = dict(lo)
        for k, w in hi.items():
            out[k] = out.get(k, 0.0) - w
        inv = 1.0 / dx
        return {k: w * inv for k, w in out.items()}

    def _assemble(self):
        mesh, bs, nx = self.mesh, self.bs, self.nx
        rows, cols, vals = [], [], []
        flux = {}  # (b, d, f, t) -> row dict
        for b in range(bs.nblocks):
            for d in range(3):
                A, B = [q for q in range(3) if q != d]
                for ta in range(nx[A]):
                    for tb in range(nx[B]):
                        for f in range(nx[d] + 1):
                            flux[(b, d, f, (ta, tb))] = self._face_flux(b, d, f, (ta, tb))
        if self.fc:  # coarse faces over a finer region: the mean of the fine faces
            for b in range(bs.nblocks):
                L = int(mesh.level[b])
                lc = [int(v) for v in mesh.loc[b]]
                for d in range(3):
                    A, B = [q for q in range(3) if q != d]
                    for side in (0, 1):
                        o = [0, 0, 0]
                        o[d] = 1 if side else -1
                        p = mesh.shift(L, tuple(lc), o)
                        if p is None or (L, p) not in mesh.refined:
                            continue
                        f = nx[d] if side else 0
                        for ta in range(nx[A]):
                            for tb in range(nx[B]):
                                acc = {}
                                for ka in (0, 1):
                                    for kb in (0, 1):
                                        G = [0, 0, 0]
                                        G[A] = 2 * (lc[A] * nx[A] + ta) + ka
                                        G[B] = 2 * (lc[B] * nx[B] + tb) + kb
                                        G[d] = 2 * (lc[d] * nx[d] + f) - (0 if side else 1)
                                        node = tuple(G[q] // nx[q] for q in range(3))
                                        fb = self.leaf[(L + 1, node)]
                                        loc = [G[q] - node[q] * nx[q] for q in range(3)]
                                        ff = loc[d] if side else loc[d] + 1
                                        for k, w in flux[(fb, d, ff, (loc[A], loc[B]))].items():
                                            acc[k] = acc.get(k, 0.0) + 0.25 * w
                                flux[(b, d, f, (ta, tb))] = acc
        for b in range(bs.nblocks):
            blk = bs.blocks[b]
            vol = float(blk["vol"])
            for k in range(nx[2]):
                for j in range(nx[1]):
                    for i in range(nx[0]):
                        r = self.lin(b, i, j, k)
                        row = {r: -self.alpha}
                        idx = (i, j, k)
                        for d in range(3):
                            A, B = [q for q in range(3) if q != d]
                            ar = float(blk["area"][d]) / vol
                            t = (idx[A], idx[B])
                            for f, s in ((idx[d], ar), (idx[d] + 1, -ar)):
                                for c, w in flux[(b, d, f, t)].items():
                                    row[c] = row.get(c, 0.0) + s * w
                        for c, w in row.items():
                            rows.append(r)
                            cols.append(c)
                            vals.append(w)
        return sp.csr_matrix((vals, (rows, cols)), shape=(self.n, self.n))

    # -- right-hand side, solve, field
    def centres(self):
        x, y, z = [], [], []
        for b in range(self.bs.nblocks):
            cx, cy, cz = self.bs.cell_centers(b)
            x.append(cx.ravel())
            y.append(cy.ravel())
            z.append(cz.ravel())
        return np.concatenate(x), np.concatenate(y), np.concatenate(z)

    def leaf_values(self, field, nv, comp):
        return np.concatenate([self.bs.interior(field, b, nv)[comp].ravel()
                               for b in range(self.bs.nblocks)])

    def mp_coeffs(self, rho, G):
        x, y, z = self.centres()
        vol = np.repeat(self.bs.blocks["vol"].astype(float), self.ncb)
        s = rho * vol
        r2 = x * x + y * y + z * z
        mom = [s, s * y, s * z, s * x, s * x * y, s * y * z, s * (3.0 * z * z - r2), s * z * x,
               s * 0.5 * (x * x - y * y)]
        return np.array([math.fsum(m) * P.SCALE[n] * G for n, m in enumerate(mom)])

    def _walls(self):
        """Per leaf cell and domain face: (cell, face centre (x, y, z), A_f / (dx V))."""
        mesh, bs, nx = self.mesh, self.bs, self.nx
        out = []
        for b in range(bs.nblocks):
            blk = bs.blocks[b]
            L = int(mesh.level[b])
            lc = [int(v) for v in mesh.loc[b]]
            x, y, z = bs.cell_centers(b)
            cen = (x, y, z)
            for d in range(3):
                for side in (0, 1):
                    if not (lc[d] == (mesh.nlat(L, d) - 1 if side else 0)):
                        continue
                    c = nx[d] - 1 if side else 0
                    face = blk["xmin"][d] + (nx[d] if side else 0) * blk["dx"][d]
                    w = float(blk["area"][d]) / (float(blk["dx"][d]) * float(blk["vol"]))
                    sl = [slice(None)] * 3
                    sl[2 - d] = slice(c, c + 1)
                    pos = [cen[q][tuple(sl)].ravel() for q in range(3)]
                    pos[d] = np.full_like(pos[d], face)
                    kk, jj, ii = np.meshgrid(np.arange(nx[2]), np.arange(nx[1]),
                                             np.arange(nx[0]), indexing="ij")
                    ijk = (ii, jj, kk)
                    cells = self.lin(b, *(ijk[q][tuple(sl)].ravel() for q in range(3)))
                    out.append((cells, pos, w, d, side))
        return out

    def rhs(self, rho, G, coef):
        b = 4.0 * np.pi * G * rho
        for cells, pos, w, _, _ in self._walls():
            b[cells] -= 2.0 * P.multipole_potential(coef, *pos) * w
        return b

    def solve(self, rhs, rtol=1e-14):
        """A phi = rhs by GMRES with the Jacobi preconditioner, to rtol of |rhs| (a direct
        factorisation of the 3D operator fills in too much to be quick)."""
        dinv = 1.0 / self.A.diagonal()
        M = spla.LinearOperator(self.A.shape, matvec=lambda v: dinv * v)
        x, info = spla.gmres(self.A, rhs, M=M, rtol=rtol, atol=0.0, restart=200,
                             maxiter=200)
        if info != 0:
            raise RuntimeError(f"GMRES did not converge ({info})")
        return x

    def _grad_rows(self):
        """Sparse (lo, hi) per axis: the values at a cell's two neighbours along d as linear
        maps of the leaf values; a neighbour beyond the domain is left out (its value is
        2 phi_b - phi, added by field)."""
        if getattr(self, "_grad", None) is not None:
            return self._grad
        mesh, bs, nx = self.mesh, self.bs, self.nx
        mats = []
        for d in range(3):
            parts = []
            for s in (-1, 1):
                rows, cols, vals = [], [], []
                for b in range(bs.nblocks):
                    L = int(mesh.level[b])
                    lc = [int(v) for v in mesh.loc[b]]
                    for k in range(nx[2]):
                        for j in range(nx[1]):
                            for i in range(nx[0]):
                                q = [i, j, k]
                                r = self.lin(b, i, j, k)
                                q[d] += s
                                if 0 <= q[d] < nx[d]:
                                    rows.append(r)
                                    cols.append(self.lin(b, *q))
                                    vals.append(1.0)
                                    continue
                                Gc = tuple(lc[t] * nx[t] + q[t] for t in range(3))
                                if not self.inside(L, Gc):
                                    continue
                                for c, w in self.ghost(L, Gc, r).items():
                                    rows.append(r)
                                    cols.append(c)
                                    vals.append(w)
                parts.append(sp.csr_matrix((vals, (rows, cols)), shape=(self.n, self.n)))
            mats.append(parts)
        self._grad = mats
        return mats

    def field(self, phi, coef):
        """g = -grad phi (3, n) by centred differences: across block faces the same values
        as the operator (same level, refined mean, prolongation), 2 phi_b - phi at the
        domain faces (MultipoleGravityBC)."""
        grad = self._grad_rows()
        lo = [grad[d][0] @ phi for d in range(3)]
        hi = [grad[d][1] @ phi for d in range(3)]
        for cells, pos, w, d, side in self._walls():
            v = 2.0 * P.multipole_potential(coef, *pos) - phi[cells]
            (hi if side else lo)[d][cells] = v
        dxs = np.repeat(self.bs.blocks["dx"].astype(float), self.ncb, axis=0)
        return np.stack([-(hi[d] - lo[d]) / (2.0 * dxs[:, d]) for d in range(3)])

    def gravity(self, prim, G):
        rho = self.leaf_values(prim, self.bs.nvar, 0)
        coef = self.mp_coeffs(rho, G)
        phi = self.solve(self.rhs(rho, G, coef))
        return coef, phi, self.field(phi, coef)
