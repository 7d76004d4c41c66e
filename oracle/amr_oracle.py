"""Multilevel meshes in the oracle: cross-level ghost values and flux correction.

TEST INFRASTRUCTURE: imported only by tests/ as the checker.

The reference takes these from Parthenon (apophis_driver.cpp:68,74-76,94), which is not in
the reference tree; what is restated here is its published behaviour for a Cartesian
mesh with 2:1 jumps (RestrictAverage, ProlongateSharedMinMod, flux correction by the
mean of the fine fluxes), with the operation order the build defines
(apophis_amd/amr.py, include/apophis_hydro.h): restriction sums the 2^d children in
(dz, dy, dx) order, x fastest, and scales by 2^-d;
prolongation is c + sum_d (sign_d / 4) minmod(c - c_-d, c_+d - c) accumulated in axis
order; a corrected coarse face is the mean of its fine faces, the transverse child index
of the lower axis varying slowest.  Parity with the reference is unpinned (no Parthenon
source or output is available); tests/test_amr.py pins these functions by exactness on
linear fields, conservation, uniform states and an op-by-op cross-check of the host plan.

This module works from whole-level arrays (every cell of the domain at each level, the
refined ones holding the restriction of their children), not from the plan's op lists.

Physical domain faces (outflow, reflect): ghost cells beyond the domain are left to the
boundary conditions, which run again after the cross-level values (Parthenon fills the
neighbour regions first and the physical boundaries last, so an edge or corner ghost beyond
the domain copies the cross-level ghost beside it); a prolongation's slope neighbour beyond
the domain is the coarse level's boundary image (Parthenon applies the boundary conditions
to the coarse buffer before it prolongates): the coarse cell itself for outflow, its mirror
image with the normal momentum negated for reflect.  tests/test_amr_physical.py pins these
by uniform and mirror-symmetric states and by linear fields away from the domain faces.

refinement_tag restates the reference's own refinement criteria
(src/refinement/gradient.cpp, other.cpp) with their index ranges; tests/test_refinement.py
pins it by closed forms (linear pressure and velocity fields, uniform states) and the
ranges' edge cells.
"""
from __future__ import annotations

import numpy as np

from . import oracle as orc

BC_REFLECT, BC_AMR = 2, 3


def _sgn(x):
    return np.where(x < 0.0, -1.0, 1.0)


def _minmod(a, b):
    return (0.5 * (_sgn(a) + _sgn(b))) * np.minimum(np.abs(a), np.abs(b))


def _nzyx(mesh, level):
    return tuple(mesh.ncells_axis(level, d) for d in (2, 1, 0))


def _kids(a, nd):
    """The 2^nd strided child views of a level array, (dz, dy, dx) order."""
    zs = (0, 1) if nd == 3 else (0,)
    return [a[:, dz::2, dy::2, dx::2] if nd == 3 else a[:, :, dy::2, dx::2]
            for dz in zs for dy in (0, 1) for dx in (0, 1)]


def level_arrays(mesh, bs, u, nv):
    """[level] -> (nv, Nz, Ny, Nx) array of that level's cells over the whole domain: leaf
    cells hold their values, cells of a refined region the mean of their children (the
    deepest levels first), cells under a coarser leaf NaN."""
    nl = int(mesh.level.max()) + 1
    arrs = [np.full((nv,) + _nzyx(mesh, L), np.nan) for L in range(nl)]
    nx = mesh.nx
    for b in range(bs.nblocks):
        lv = int(mesh.level[b])
        lc = mesh.loc[b]
        sl = tuple(slice(int(lc[d]) * nx[d], (int(lc[d]) + 1) * nx[d]) for d in (2, 1, 0))
        arrs[lv][(slice(None),) + sl] = bs.interior(u, b, nv)
    nd = mesh.ndim
    for L in range(nl - 2, -1, -1):
        kids = _kids(arrs[L + 1], nd)
        s = kids[0].copy()
        for k in kids[1:]:
            s = s + k
        s = s * (1.0 / len(kids))
        refined = ~np.isnan(s)
        arrs[L][refined] = s[refined]
    return arrs


def amr_fill(mesh, bs, u):
    """Ghost cells of u on faces / edges / corners whose neighbour region is on another
    level (nbr = -1), after the same-level exchange: over a finer region the level's
    restriction of it, over a coarser leaf the prolongation of the level above."""
    nv, nd, g, nx = bs.nvar, mesh.ndim, mesh.ng, mesh.nx
    arrs = level_arrays(mesh, bs, u, nv)
    ext = bs.extents()
    for b in range(bs.nblocks):
        lv = int(mesh.level[b])
        lc = mesh.loc[b]
        nbr = bs.blocks["nbr"][b]
        E = [int(ext[b][d]) for d in range(3)]
        kk, jj, ii = np.meshgrid(np.arange(E[2]), np.arange(E[1]), np.arange(E[0]),
                                 indexing="ij")
        loc = [ii, jj, kk]
        o = [np.zeros_like(ii) for _ in range(3)]
        for d in range(nd):
            o[d] = np.where(loc[d] < g, -1, np.where(loc[d] >= g + nx[d], 1, 0))
        ghost = (o[0] != 0) | (o[1] != 0) | (o[2] != 0)
        code = (o[2] + 1) * 9 + (o[1] + 1) * 3 + (o[0] + 1)
        NL = [mesh.ncells_axis(lv, d) for d in range(3)]
        sel = ghost & (nbr[code] < 0)
        for d in range(nd):  # beyond a physical face: the boundary conditions fill it
            if not mesh.periodic[d]:
                G = int(lc[d]) * nx[d] + loc[d] - g
                sel &= (G >= 0) & (G < NL[d])
        if not sel.any():
            continue
        Gw = [(int(lc[d]) * nx[d] + loc[d][sel] - g) % NL[d] if d < nd
              else np.zeros(sel.sum(), int) for d in range(3)]
        view = bs.block_view(u, b, nv)
        own = arrs[lv][(slice(None),) + (Gw[2], Gw[1], Gw[0])]
        finer = ~np.isnan(own[0])  # a refined region: the level holds its restriction
        val = own.copy()
        coarser = ~finer
        if coarser.any():
            assert lv > 0, "a level-0 ghost over no level-0 region"
            C = arrs[lv - 1]
            NC = [mesh.ncells_axis(lv - 1, d) for d in range(3)]
            Gc = [Gw[d][coarser] for d in range(3)]
            I = [Gc[d] // 2 if d < nd else np.zeros_like(Gc[d]) for d in range(3)]
            at = lambda J: C[(slice(None),) + tuple(J[d] % NC[d] for d in (2, 1, 0))]
            c = at(I)
            pv = c.copy()
            def nb_val(d, side):
                """The slope neighbour of I along d; beyond a physical face, the coarse
                cell's boundary image."""
                J = [I[t] + ((1 if side else -1) if t == d else 0) for t in range(3)]
                if mesh.periodic[d]:
                    return at(J)
                out = (J[d] < 0) | (J[d] >= NC[d])
                J[d] = np.where(out, I[d], J[d])
                v = at(J)
                if mesh.bc[2 * d + side] == BC_REFLECT:
                    v[1 + d] = np.where(out, -v[1 + d], v[1 + d])
                return v

            for d in range(nd):
                gd = _minmod(c - nb_val(d, 0), nb_val(d, 1) - c)
                sign = np.where(Gc[d] % 2 == 0, -1.0, 1.0)
                pv = pv + (sign * 0.25) * gd
            val[:, coarser] = pv
        assert not np.isnan(val).any(), "a cross-level value two levels away"
        view[(slice(None),) + (kk[sel], jj[sel], ii[sel])] = val


def flux_correct(mesh, bs, fx):
    """Coarse fluxes on faces shared with a finer level <- the mean of the fine fluxes
    across each (fx: the three flux arrays)."""
    nv, nd, g, nx = bs.nvar, mesh.ndim, mesh.ng, mesh.nx
    nl = int(mesh.level.max()) + 1
    for L in range(nl - 1):
        N1 = [mesh.ncells_axis(L + 1, d) for d in range(3)]
        for d in range(nd):
            # level-(L+1) faces along d over the whole domain: face G between cells G-1, G
            shape = [N1[t] if t != d else N1[t] + 1 for t in range(3)]
            ff = np.full((nv, shape[2], shape[1], shape[0]), np.nan)
            for b in range(bs.nblocks):
                if mesh.level[b] != L + 1:
                    continue
                lc = mesh.loc[b]
                view = bs.block_view(fx[d], b, nv)
                sl_src, sl_dst = [], []
                for t in (2, 1, 0):
                    if t >= nd:
                        sl_src.append(slice(0, 1))
                        sl_dst.append(slice(0, 1))
                        continue
                    lo = int(lc[t]) * nx[t]
                    if t == d:
                        sl_src.append(slice(g, g + nx[t] + 1))
                        sl_dst.append(slice(lo, lo + nx[t] + 1))
                    else:
                        sl_src.append(slice(g, g + nx[t]))
                        sl_dst.append(slice(lo, lo + nx[t]))
                ff[(slice(None),) + tuple(sl_dst)] = view[(slice(None),) + tuple(sl_src)]
            # periodic: face N1[d] is face 0
            edge = [slice(None)] * 4
            edge[3 - d] = slice(0, 1)
            last = [slice(None)] * 4
            last[3 - d] = slice(shape[d] - 1, shape[d])
            f0, fN = ff[tuple(edge)], ff[tuple(last)]
            ff[tuple(edge)] = np.where(np.isnan(f0), fN, f0)
            tr = [t for t in range(nd) if t != d]
            for b in range(bs.nblocks):
                if mesh.level[b] != L:
                    continue
                lc = mesh.loc[b]
                bc = bs.blocks["bc"][b]
                view = bs.block_view(fx[d], b, nv)
                for side in (0, 1):
                    if bc[2 * d + side] != BC_AMR:
                        continue
                    Fd = 2 * (int(lc[d]) * nx[d] + (nx[d] if side else 0)) % N1[d]
                    # coarse face cells (transverse interior) and their fine children
                    tc = np.meshgrid(*[np.arange(nx[t]) for t in tr], indexing="ij")
                    acc = None
                    for kid in np.ndindex(*([2] * len(tr))):
                        F = [None, None, None]
                        F[d] = np.full(tc[0].shape, Fd)
                        for t, c, q in zip(tr, tc, kid):
                            F[t] = 2 * (int(lc[t]) * nx[t] + c) + q
                        for t in range(nd, 3):
                            F[t] = np.zeros(tc[0].shape, int)
                        v = ff[(slice(None),) + (F[2], F[1], F[0])]
                        acc = v.copy() if acc is None else acc + v
                    if np.isnan(acc).all():
                        continue  # the other side is coarser: this block is the fine side
                    assert not np.isnan(acc).any(), "a face half on the finer level"
                    acc = acc * (1.0 / 2 ** len(tr))
                    dst = [None, None, None]
                    dst[d] = np.full(tc[0].shape, g + (nx[d] if side else 0))
                    for t, c in zip(tr, tc):
                        dst[t] = g + c
                    for t in range(nd, 3):
                        dst[t] = np.zeros(tc[0].shape, int)
                    view[(slice(None),) + (dst[2], dst[1], dst[0])] = acc


class AmrState:
    """The oracle State of a problem on a multilevel mesh: the stage sequence of the plain
    driver (monopole or Poisson gravity when the problem has it, fluxes, flux correction,
    update, gravity source, ghost exchange + cross-level fill + boundary conditions,
    ConsToPrim) composed from the oracle's task functions, in the order orc_step_monopole
    runs them."""

    def __init__(self, prob, mesh, flux_correction=True):
        self.prob, self.mesh, self.bs = prob, mesh, prob.bs
        self.correct = flux_correction
        self.u0 = prob.cons.copy()
        self.u1 = np.zeros_like(self.u0)
        self.prim = np.zeros_like(self.u0)
        self.gamma = np.zeros(self.bs.field_size(2))
        self.fx = [np.zeros_like(self.u0) for _ in range(3)]
        self.mono = getattr(prob, "monopole", None)  # (nrbin, rmax, G) or None
        # (G, alpha, tolerance, max_it[, flux_correct, prolongation]) or None: the composite
        # Poisson problem of the mesh (poisson_amr_oracle), solved to 1e-14
        self.poisson = getattr(prob, "poisson", None)
        self._pp = None
        if self.mono is not None or self.poisson is not None:
            self.grav = np.zeros(self.bs.field_size(3))
            self.phi = np.zeros(self.bs.field_size(1))
        self.g = tuple(float(v) for v in getattr(prob, "g", (0.0, 0.0, 0.0)))
        self._ghosts()

    def _ghosts(self):
        orc.fill_ghosts(self.bs, self.u0)
        amr_fill(self.mesh, self.bs, self.u0)
        if not all(self.mesh.periodic):  # corners beyond the domain: after the fill
            orc.fill_ghosts(self.bs, self.u0)
        orc.cons_to_prim(self.bs, self.prob.eos, self.u0, self.prim, self.gamma)

    def gravity_field(self):
        """The gravity field of the current primitives: monopole (bins, tables, per-cell
        field) or Poisson (coefficients, right-hand side, solve, field)."""
        if self.poisson is not None:
            from . import poisson_amr_oracle as PA

            G, alpha = self.poisson[0], self.poisson[1]
            fc = self.poisson[4] if len(self.poisson) > 4 else False
            prol = self.poisson[5] if len(self.poisson) > 5 else "Linear"
            if self._pp is None:
                self._pp = PA.Problem(self.mesh, self.bs, alpha, fc, prol)
            pr = self._pp
            _, phi, g = pr.gravity(self.prim, G)
            nx = pr.nx
            for b in range(self.bs.nblocks):
                sl = slice(b * pr.ncb, (b + 1) * pr.ncb)
                gv = self.bs.interior(self.grav, b, 3)
                for d in range(3):
                    gv[d] = g[d, sl].reshape(nx[2], nx[1], nx[0])
                self.bs.interior(self.phi, b, 1)[0] = phi[sl].reshape(nx[2], nx[1], nx[0])
            return self.grav
        nrbin, rmax, G = self.mono
        mbin, _ = orc.mono_bin_masses(self.bs, self.prim, nrbin, rmax)
        m_enc, phi_rad = orc.mono_tables(nrbin, self.bs.ndim, rmax, G, mbin)
        self.grav, self.phi = orc.mono_field(self.bs, nrbin, rmax, G, m_enc, phi_rad)
        return self.grav

    def estimate_dt(self):
        return orc.estimate_timestep(self.bs, self.prob.eos, self.prim, self.u0,
                                     self.prob.cfl)

    def step(self, dt):
        p, bs = self.prob, self.bs
        g0, g1, be = orc.integrator_stages(p.integrator)
        self.u1[:] = self.u0
        for s in range(len(g0)):
            recon = "dc" if (p.integrator == "vl2" and s == 0) else p.recon
            field = self.mono is not None or self.poisson is not None
            if field:
                self.gravity_field()
            orc.calculate_fluxes(bs, recon, p.riemann, self.prim, self.gamma, *self.fx)
            if self.correct:
                flux_correct(self.mesh, bs, self.fx)
            orc.update(bs, self.u0, self.u1, *self.fx, g0[s], g1[s], be[s] * dt)
            if field:
                orc.apply_gravity_field(bs, self.u0, self.prim, self.grav, be[s] * dt)
            elif any(v != 0.0 for v in self.g):
                orc.apply_gravity(bs, self.u0, self.prim, np.array(self.g), be[s] * dt)
            self._ghosts()
        return self.estimate_dt()


def refinement_tag(bs, prim, crit, p0, p1):
    """Per block (tag, value) of the reference's refinement criteria
    (src/refinement/gradient.cpp:11-84, other.cpp:11-34) on prim with its ghost cells:
    crit 1 pressure gradient, 2 xy-velocity gradient, 3 max density (see
    include/apophis_hydro.h for the ranges; the density's i range runs to ib.e + 1)."""
    nd = bs.ndim
    tags = np.zeros(bs.nblocks, dtype=np.int32)
    vals = np.zeros(bs.nblocks)
    for b in range(bs.nblocks):
        w = bs.block_view(prim, b, bs.nvar)  # (nv, Nz, Ny, Nx)
        g = int(bs.blocks["ng"][b])
        nx = [int(v) for v in bs.blocks["nx"][b]]
        gg = [g if d < nd else 0 for d in range(3)]
        lo = [gg[d] for d in range(3)]
        hi = [gg[d] + nx[d] - 1 for d in range(3)]
        if crit == 1:
            for d in range(nd):
                lo[d] -= 1
                hi[d] += 1
        elif crit == 2:
            for d in (0, 1):
                lo[d] -= 1
                hi[d] += 1
        else:
            hi[0] += 1
        K = np.arange(lo[2], hi[2] + 1)[:, None, None]
        J = np.arange(lo[1], hi[1] + 1)[None, :, None]
        I = np.arange(lo[0], hi[0] + 1)[None, None, :]
        if crit == 1:
            if nd == 1:
                continue
            p = w[4]
            a = 0.5 * (p[K, J, I + 1] - p[K, J, I - 1])
            c = 0.5 * (p[K, J + 1, I] - p[K, J - 1, I])
            s2 = a * a + c * c
            if nd == 3:
                e = 0.5 * (p[K + 1, J, I] - p[K - 1, J, I])
                s2 = s2 + e * e
            q = np.sqrt(s2) / p[K, J, I]
        elif crit == 2:
            vgy = np.abs(w[2][K, J, I + 1] - w[2][K, J, I - 1]) * 0.5
            vgx = np.abs(w[1][K, J + 1, I] - w[1][K, J - 1, I]) * 0.5
            q = np.sqrt(vgx * vgx + vgy * vgy)
        else:
            q = w[0][K, J, I] + 0.0 * I  # broadcast to the range
        m = 0.0
        for v in q.ravel():  # std::max, in any order
            m = v if m < v else m
        vals[b] = m
        if crit == 1:
            tags[b] = 1 if m > p0 else (-1 if m < 0.25 * p0 else 0)
        elif crit == 2:
            tags[b] = 1 if m > p0 else (-1 if m < 0.5 * p0 else 0)
        else:
            tags[b] = 1 if m > p1 else (-1 if m < p0 else 0)
    return tags, vals
