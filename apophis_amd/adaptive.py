"""Adaptive evolution on a multilevel mesh (refinement = adaptive).

The reference tags blocks with its refinement criterion once per cycle
(apophis_driver.cpp:120-127, hydro.cpp:303-333) and leaves the rest of the cycle to
Parthenon's EvolutionDriver:
* at start-up (Mesh::Initialize) the problem is generated, tagged, the tagged blocks are
  refined and the problem is generated again on the new mesh, until no block refines;
* after every step each block's tag is turned into a refinement flag (a derefine tag
  only counts once the block has been tagged so `derefine_count` times in a row), the
  tree is regridded with 2:1 balance, and dt is re-estimated on the new mesh
  (dt <- min(2 dt, estimate)).

Here the criterion runs on the device (aph_refinement_tag), the tree and its plans are
rebuilt on the host (apophis_amd/amr.py), and the steps run in the fused or plain form.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import amr, hydro


def device_tags(ctx, pkg, mesh, bs, cons, md=None, plan=None):
    """Refinement tags of a state (host array or device tensor of cons): ghost exchange,
    cross-level fill, ConsToPrim and the package's criterion, on the device.  md / plan
    (a MeshData and amr.AmrPlan of this mesh) are reused when given."""
    if md is None:
        md = hydro.MeshData(ctx, bs)
    if plan is None:
        plan = amr.AmrPlan(ctx, md.dbs, mesh, bs)
    md.cons.copy_(torch.from_numpy(cons) if isinstance(cons, np.ndarray) else cons)
    hydro.fill_ghosts(md)
    plan.fill(md.cons)
    hydro.cons_to_prim(md, pkg.param("eos"))
    tags, _ = hydro.refinement_tag(md, pkg)
    return tags.cpu().numpy()


def initial_mesh(mesh, pgen, tagger, max_rounds=32):
    """Parthenon's adaptive start: generate the problem (pgen(mesh, bs) -> cons), tag it
    (tagger(mesh, bs, cons) -> tags), refine the blocks tagged 1 (no derefinement, 2:1
    balance kept), and generate again on the new mesh, until no block refines.  Returns
    (mesh, bs, cons)."""
    for _ in range(max_rounds):
        bs = mesh.block_set()
        cons = pgen(mesh, bs)
        tags = np.maximum(np.asarray(tagger(mesh, bs, cons)), 0)
        new, _, _ = amr.regrid(mesh, bs, cons, tags)
        if new.refined == mesh.refined:
            return mesh, bs, cons
        mesh = new
    raise RuntimeError("initial refinement did not settle")


class DerefineCounter:
    """Parthenon's derefinement delay (MeshRefinement, parthenon/mesh/derefine_count,
    default 10): a leaf tagged -1 counts up and becomes a derefine flag on the `count`-th
    -1 tag in a row (deref_count_ is incremented, then compared with the threshold); any
    other tag resets it.  Leaves are keyed by
    (level, loc), so counters survive regrids of other blocks; new leaves start at 0."""

    def __init__(self, count=10):
        self.count = int(count)
        self._mesh = None
        self._keys = []    # (level, loc) of the last mesh's leaves, in block order
        self._c = np.zeros(0, dtype=np.int64)   # their counts

    @property
    def seen(self):
        """(level, loc) -> count, for the leaves with a nonzero count."""
        return {k: int(v) for k, v in zip(self._keys, self._c) if v}

    def flags(self, mesh, tags):
        tags = np.asarray(tags)
        if mesh is not self._mesh:  # a new block order: carry the counts over by key
            seen = self.seen
            keys = [(int(mesh.level[b]), tuple(int(v) for v in mesh.loc[b]))
                    for b in range(len(mesh.level))]
            self._c = np.array([seen.get(k, 0) for k in keys], dtype=np.int64)
            self._keys, self._mesh = keys, mesh
        c = self._c
        neg = tags < 0
        out = np.where(tags > 0, 1, 0).astype(np.int32)
        # MeshRefinement's rule: the count goes up first, and the leaf is flagged once it
        # reaches `count` (the count-th -1 tag in a row)
        c_new = np.where(neg, c + 1, 0)
        out[neg & (c_new >= self.count)] = -1
        c[:] = np.minimum(c_new, self.count)
        return out


def _may_change(mesh, flags):
    """False when no regrid rule can apply: no leaf is flagged to refine and no sibling
    group has all of its 2^ndim leaves flagged to derefine (merges need the whole group;
    balance only acts on refinements and merges)."""
    if (flags > 0).any():
        return True
    neg = flags < 0
    if not neg.any():
        return False
    par = np.concatenate([mesh.level[neg, None] - 1, mesh.loc[neg] >> 1], axis=1)
    _, cnt = np.unique(par, axis=0, return_counts=True)
    return bool((cnt == 2 ** mesh.ndim).any())


class AdaptiveDriver:
    """Cycles of one step + tag + regrid on an adaptive multilevel mesh.

    form "fused": the batched plan (one launch per level and stage); "plain": the task
    sequence of ApophisDriver.  Each regrid builds new plans for the new block set.
    ``stats`` accumulates cycles, zone-cycles, regrids and the host seconds spent in
    regrids (tree, remap and plan rebuilds)."""

    def __init__(self, ctx, pkg, mesh, bs, cons, form="fused", derefine_count=10,
                 tlim=np.inf):
        if form not in ("fused", "plain"):
            raise ValueError("form is 'fused' or 'plain'")
        self.ctx, self.pkg, self.form = ctx, pkg, form
        self.mesh, self.bs = mesh, bs
        self.counter = DerefineCounter(derefine_count)
        self.tlim = float(tlim)
        self.time = 0.0
        self.stats = dict(cycles=0, zone_cycles=0, regrids=0, regrid_s=0.0)
        self._build(cons)
        self.dt = min(self._estimate(), self.tlim)

    # -- plans of the current mesh
    def _build(self, cons):
        p = self.pkg
        if self.form == "fused":
            self.fd = hydro.FusedDriver(self.ctx, self.bs, cons, int(p.param("reconstruction")),
                                        int(p.param("riemann_solver")), p.param("eos"),
                                        p.param("cfl"), integrator=p.param("integrator").name,
                                        amr_mesh=self.mesh)
            # the tagging state shares the plan's block set and cross-level plan
            self.tag_md = hydro.MeshData(self.ctx, self.bs, dbs=self.fd.device_bs)
        else:
            self.md = hydro.MeshData(self.ctx, self.bs, cons)
            self.plan = amr.AmrPlan(self.ctx, self.md.dbs, self.mesh, self.bs)
            self.drv = hydro.ApophisDriver(p, self.md, amr=self.plan)
            self._est = self.drv.initialize()

    def _estimate(self):
        if self.form == "fused":
            return float(self.fd.initial_dt().item())
        return self._est

    def state(self):
        """The current conserved state (device tensor in pack layout)."""
        return self.fd.u0 if self.form == "fused" else self.md.cons

    def _tags(self):
        if self.form == "fused":
            return device_tags(self.ctx, self.pkg, self.mesh, self.bs, self.fd.u0,
                               md=self.tag_md, plan=self.fd.amr)
        tags, _ = hydro.refinement_tag(self.md, self.pkg)  # prim is current after a step
        return tags.cpu().numpy()

    def cycle(self, tlim=None):
        """One step at the current dt (clipped to end at tlim, default the driver's), then
        the refinement check; returns the dt used."""
        dt = self.dt
        end = self.tlim if tlim is None else float(tlim)
        if end - self.time < dt:
            dt = end - self.time
        if self.form == "fused":
            self.fd.set_dt(dt)
            self.fd.step(fixed_dt=True)
            est = float(self.fd.initial_dt().item())
            self.fd.status()  # synchronised already: raise if a multilevel wait ran out
        else:
            self._est = self.drv.step(dt)
            est = self._est
        self.time += dt
        self.stats["cycles"] += 1
        self.stats["zone_cycles"] += int(self.bs.blocks["nx"].prod(axis=1).sum())
        flags = self.counter.flags(self.mesh, self._tags())
        flags[(flags > 0) & (self.mesh.level >= self.mesh.max_level)] = 0
        flags[(flags < 0) & (self.mesh.level == 0)] = 0
        if _may_change(self.mesh, flags):
            t0 = time.perf_counter()
            leaves = amr.regrid_leaves(self.mesh, flags)  # the tree alone first
            if not (np.array_equal(leaves[0], self.mesh.level) and
                    np.array_equal(leaves[1], self.mesh.loc)):
                u = self.state().cpu().numpy()
                new, nbs, ncons = amr.regrid(self.mesh, self.bs, u, flags, leaves=leaves)
                self.mesh, self.bs = new, nbs
                self._drop()
                self._build(ncons)
                est = self._estimate()
                self.stats["regrids"] += 1
            self.stats["regrid_s"] += time.perf_counter() - t0
        self.dt = min(2.0 * dt, est)
        return dt

    def _drop(self):
        if self.form == "fused":
            self.fd.close()
            del self.fd, self.tag_md
        else:
            del self.drv, self.plan, self.md

    def evolve(self, ncycles=None, tlim=None):
        tl = self.tlim if tlim is None else float(tlim)
        n = 0
        while (ncycles is None or n < ncycles) and self.time < tl * (1 - 1e-14):
            self.cycle(tl)
            n += 1
        torch.cuda.synchronize()
        return n
