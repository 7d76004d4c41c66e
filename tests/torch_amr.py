"""FP64 PyTorch restatement of a step on a multilevel mesh, for reference gradients of the
multilevel adjoint through torch.autograd.

TEST INFRASTRUCTURE: used only by tests/ as the checker.

A stage, as the plain multilevel driver runs it (apophis_driver.cpp:68-111 with
Parthenon's multilevel exchange and flux correction): the conserved state of every block
with its ghost cells filled -- the same-level exchange, the boundary conditions, the
cross-level ops of the host plan (apophis_amd/amr.py amr_plan; MEAN and PROLONG with the
minmod slope and the mirror bits), the boundary conditions again -- then per axis
ConsToPrim, the reconstruction and the Riemann solver of tests/torch_ref.py on every face
of every block's interior pencils, the flux correction of the plan's fixes (a coarse face
takes (sum of its fine faces) * 2^(1-ndim)), and the update with each block's areas and
volume.  Every block has the same shape (a multilevel mesh keeps its cells per block), so
the blocks are one batch axis.  minmod is written as sgn * min(|a|, |b|) with the smaller
magnitude chosen by a comparison (|b| < |a| takes b, so a tie takes a) and |x| = x < 0 ?
-x : x, the branches the adjoint kernels follow.  Floors are assumed inactive and gravity
absent.  tests/test_torch_amr.py pins a step of this restatement to the oracle's
(oracle/amr_oracle.py AmrState) bitwise.
"""
from __future__ import annotations

import numpy as np
import torch

from apophis_amd import amr as A
from tests import torch_ref as R


class Multilevel:
    def __init__(self, mesh, bs, recon, riemann, eos, integrator="rk2"):
        self.mesh, self.bs = mesh, bs
        self.recon, self.riemann, self.eos, self.integrator = recon, riemann, eos, integrator
        nd, nv, nb = bs.ndim, bs.nvar, bs.nblocks
        self.nd, self.nv, self.nb = nd, nv, nb
        ext = bs.extents()
        self.E = [int(v) for v in ext[0]]
        self.nc = int(np.prod(self.E))
        self.g = int(bs.blocks["ng"][0])
        self.n = [int(v) for v in bs.blocks["nx"][0]]
        g, n, E, nc = self.g, self.n, self.E, self.nc
        G = [g if d < nd else 0 for d in range(3)]
        self.G = G
        blocks = bs.blocks

        def off(b, i, j, k):  # component-0 offset of a ghost-inclusive cell of block b
            return nv * b * nc + (k * E[1] + j) * E[0] + i

        # interior cells, in (b, k, j, i) order
        kk, jj, ii = np.meshgrid(np.arange(n[2]) + G[2], np.arange(n[1]) + G[1],
                                 np.arange(n[0]) + G[0], indexing="ij")
        self.int0 = np.array([off(b, ii, jj, kk).ravel() for b in range(nb)]).ravel()
        # same-level exchange and the per-axis physical boundary rules
        exch_d, exch_s = [], []
        bcs = {d: ([], [], []) for d in range(nd)}
        for b in range(nb):
            K, J, I = np.meshgrid(np.arange(E[2]), np.arange(E[1]), np.arange(E[0]),
                                  indexing="ij")
            loc = [I.ravel(), J.ravel(), K.ravel()]
            o = [np.zeros_like(loc[0]) for _ in range(3)]
            for d in range(nd):
                o[d] = np.where(loc[d] < g, -1, np.where(loc[d] >= g + n[d], 1, 0))
            ghost = (o[0] != 0) | (o[1] != 0) | (o[2] != 0)
            code = (o[2] + 1) * 9 + (o[1] + 1) * 3 + (o[0] + 1)
            nbr = blocks["nbr"][b][code]
            sel = ghost & (nbr >= 0)
            src = off(nbr[sel], loc[0][sel] - o[0][sel] * n[0], loc[1][sel] - o[1][sel] * n[1],
                      loc[2][sel] - o[2][sel] * n[2])
            exch_d.append(off(b, loc[0][sel], loc[1][sel], loc[2][sel]))
            exch_s.append(src)
            for d in range(nd):  # the slab of axis d on each physical face, all other cells
                sdn = (1, 3, 9)[d]
                for side in (0, 1):
                    if blocks["nbr"][b][13 + (sdn if side else -sdn)] >= 0:
                        continue
                    typ = int(blocks["bc"][b][2 * d + side])
                    if typ == A.BC_AMR:
                        continue
                    m = (loc[d] >= g + n[d]) if side else (loc[d] < g)
                    ref = g + n[d] - 1 if side else g
                    s = [x[m].copy() for x in loc]
                    if typ == 2:  # reflect: mirror image
                        s[d] = 2 * ref + (1 if side else -1) - loc[d][m]
                    else:
                        s[d] = np.full_like(s[d], ref)
                    bcs[d][0].append(off(b, loc[0][m], loc[1][m], loc[2][m]))
                    bcs[d][1].append(off(b, *s))
                    bcs[d][2].append(np.full(m.sum(), typ == 2))
        cat = lambda x: np.concatenate(x) if x else np.zeros(0, np.int64)
        self.exch = (cat(exch_d), cat(exch_s))
        self.bcs = [(cat(bcs[d][0]), cat(bcs[d][1]), cat(bcs[d][2]).astype(bool))
                    for d in range(nd)]
        # the cross-level ops and flux fixes of the plan
        ops, fixes, refs = A.amr_plan(mesh, bs)
        self.ops, self.fixes, self.refs = ops, fixes, refs
        # per op and value slot: up to 8 ref offsets (padded) and the count
        nsl = 1 + 2 * nd
        self.op_dst = ops["dst"].astype(np.int64)
        cnt = ops["val"]["count"][:, :nsl].astype(np.int64)
        first = ops["val"]["first"][:, :nsl].astype(np.int64)
        idx = np.zeros((len(ops), nsl, 8), np.int64)
        for q in range(8):
            r = np.where(q < cnt, first + q, first)
            idx[:, :, q] = refs["offset"][np.clip(r, 0, max(len(refs) - 1, 0))] if len(refs) else 0
        self.op_idx, self.op_cnt = idx, cnt
        self.op_kind = (ops["kind"] & 0xFF).astype(np.int64)
        self.op_mirror = (ops["kind"] >> A.AMR_MIRROR_BIT).astype(np.int64)
        self.op_sign = ops["sign"][:, :nd].astype(np.float64)
        # flux fixes as positions in the per-axis face tensors (b, k, j, i) of faces
        self.fix = []
        for d in range(nd):
            fx = fixes[fixes["dir"] == d]
            dpos = np.array([self._face_pos(int(x), d) for x in fx["dst"]]).reshape(-1, 4)
            srcs = []
            for f in fx:
                rr = refs[f["val"]["first"]: f["val"]["first"] + f["val"]["count"]]
                srcs.append([self._face_pos(int(x), d) for x in rr["offset"]])
            self.fix.append((dpos, np.array(srcs).reshape(len(fx), -1, 4)))
        self.area = torch.from_numpy(blocks["area"].astype(np.float64))
        self.vol = torch.from_numpy(blocks["vol"].astype(np.float64))

    def _face_pos(self, off0, d):
        """(b, k, j, i) of a flux-array comp-0 offset in the face tensor of axis d (face index
        along d from 0, interior transverse cells from 0)."""
        nv, nc, E, G = self.nv, self.nc, self.E, self.G
        b, c = divmod(off0, nv * nc)
        i, r = c % E[0], c // E[0]
        j, k = r % E[1], r // E[1]
        return (b, k - G[2], j - G[1], i - G[0])

    # ---------------------------------------------------------------- ghost fill
    def _bc(self, u, d):
        dst, src, refl = self.bcs[d]
        if not len(dst):
            return u
        nv, nc = self.nv, self.nc
        out = u.clone()
        for v in range(nv):
            val = u[torch.from_numpy(src + v * nc)]
            if v == 1 + d and refl.any():
                val = torch.where(torch.from_numpy(refl), -1.0 * val, val)
            out[torch.from_numpy(dst + v * nc)] = val
        return out

    def _amr(self, u):
        nv, nc, nd = self.nv, self.nc, self.nd
        if not len(self.op_dst):
            return u
        out = u.clone()
        cnt = torch.from_numpy(np.maximum(self.op_cnt, 1))  # MEAN ops leave slots 1.. empty
        for v in range(nv):
            def value(k):
                ix = torch.from_numpy(self.op_idx[:, k, 0] + v * nc)
                s = u[ix]
                for q in range(1, 8):
                    nxt = s + u[torch.from_numpy(self.op_idx[:, k, q] + v * nc)]
                    s = torch.where(q < cnt[:, k], nxt, s)
                return torch.where(cnt[:, k] == 1, s, s * (1.0 / cnt[:, k].to(s.dtype)))

            c = value(0)
            o = c
            for d in range(nd):
                lo, hi = value(1 + 2 * d), value(2 + 2 * d)
                if v == 1 + d:
                    mlo = torch.from_numpy((self.op_mirror & (1 << (2 * d))) != 0)
                    mhi = torch.from_numpy((self.op_mirror & (2 << (2 * d))) != 0)
                    lo = torch.where(mlo, -lo, lo)
                    hi = torch.where(mhi, -hi, hi)
                x, y = c - lo, hi - c
                sx = torch.where(x < 0.0, -1.0, 1.0).to(x.dtype)
                sy = torch.where(y < 0.0, -1.0, 1.0).to(x.dtype)
                ax = torch.where(x < 0.0, -x, x)
                ay = torch.where(y < 0.0, -y, y)
                mm = torch.where(ay < ax, ay, ax)
                gd = (0.5 * (sx + sy)) * mm
                o = o + (torch.from_numpy(self.op_sign[:, d]) * 0.25) * gd
            o = torch.where(torch.from_numpy(self.op_kind == A.AMR_PROLONG), o, c)
            out[torch.from_numpy(self.op_dst + v * nc)] = o
        return out

    def fill(self, interior):
        """Flat ghost-filled state from the interior (NV, nb, nz, ny, nx)."""
        nv, nc, nb = self.nv, self.nc, self.nb
        u = torch.zeros(nv * nc * nb, dtype=torch.float64)
        vals = interior.permute(1, 0, 2, 3, 4).reshape(nb, nv, -1)  # (b, v, cells)
        idx = torch.from_numpy(self.int0.reshape(nb, 1, -1) +
                               (np.arange(nv) * nc).reshape(1, nv, 1))
        u = u.index_put((idx.reshape(-1),), vals.reshape(-1))
        d_, s_ = self.exch
        if len(d_):
            u2 = u.clone()
            for v in range(nv):
                u2[torch.from_numpy(d_ + v * nc)] = u[torch.from_numpy(s_ + v * nc)]
            u = u2
        for d in range(self.nd):
            u = self._bc(u, d)
        u = self._amr(u)
        for d in range(self.nd):
            u = self._bc(u, d)
        return u

    # ---------------------------------------------------------------- stage
    def stage(self, x, un, gam0, gam1, beta_dt, recon):
        nv, nb, nd, g = self.nv, self.nb, self.nd, self.g
        E, n, G = self.E, self.n, self.G
        gm1e, dfloor, pfloor = self.eos
        u = self.fill(x).reshape(nb, nv, E[2], E[1], E[0]).permute(1, 0, 2, 3, 4)
        w, gm = R.cons_to_prim(u, gm1e, dfloor, pfloor)
        q = torch.cat([w, gm])  # (NF, nb, Nz, Ny, Nx)
        du = None
        for d in range(nd):
            ax = 4 - d
            qd = q
            for t in range(nd):
                if t != d:
                    qd = qd.narrow(4 - t, G[t], n[t])
            nn = n[d]
            tk = lambda k: qd.narrow(ax, g - 1 + k, nn + 2)  # offset k of cells -1 .. n
            ql, qr = R.recon_fn(recon, tk)
            WL = ql.narrow(ax, 0, nn + 1)
            WR = qr.narrow(ax, 1, nn + 1)
            ivx, ivy, ivz = 1 + d, 1 + (1 + d) % 3, 1 + (2 + d) % 3
            wl = [WL[0], WL[ivx], WL[ivy], WL[ivz], WL[4]]
            wr = [WR[0], WR[ivx], WR[ivy], WR[ivz], WR[4]]
            f = R.riemann_fn(self.riemann)(wl, wr, WL[nv:], WR[nv:])
            F = [None] * nv
            F[0], F[ivx], F[ivy], F[ivz], F[4] = f
            for m in range(5, nv):
                F[m] = torch.where(F[0] >= 0.0, F[0] * WL[m], F[0] * WR[m])
            F = torch.stack(F)  # (NV, nb, faces along d / interior elsewhere)
            dpos, spos = self.fix[d]
            if len(dpos):
                F2 = F.clone()
                ix = lambda P: (slice(None), torch.from_numpy(P[:, 0]),
                                torch.from_numpy(P[:, 1]), torch.from_numpy(P[:, 2]),
                                torch.from_numpy(P[:, 3]))
                s = F[ix(spos[:, 0])]
                for c in range(1, spos.shape[1]):
                    s = s + F[ix(spos[:, c])]
                F2[ix(dpos)] = s * (1.0 / spos.shape[1])
                F = F2
            ar = self.area[:, d].reshape(1, nb, 1, 1, 1)
            dd = ar * F.narrow(ax, 1, nn) - ar * F.narrow(ax, 0, nn)
            du = dd if du is None else du + dd
        div = -du / self.vol.reshape(1, nb, 1, 1, 1)
        return gam0 * x + gam1 * un + beta_dt * div

    def step(self, u, dt):
        x = u
        for s, (g0, g1, b) in enumerate(R.STAGES[self.integrator]):
            r = "dc" if (self.integrator == "vl2" and s == 0) else self.recon
            x = self.stage(x, u, g0, g1, b * dt, r)
        return x

    def interior(self, flat):
        """(NV, nb, nz, ny, nx) interior of a pack-layout field (numpy)."""
        nv, nb, n = self.nv, self.nb, self.n
        v = flat[self.int0.reshape(nb, 1, -1) + (np.arange(nv) * self.nc).reshape(1, nv, 1)]
        return torch.from_numpy(np.ascontiguousarray(
            v.reshape(nb, nv, n[2], n[1], n[0]).transpose(1, 0, 2, 3, 4)))
