"""Differentiable rollouts of the batched fused form (torch.autograd over the C ABI).

``rollout(u0, stepper, nsteps, checkpoint_every)`` advances the conserved state nsteps
RK2 steps with a frozen dt through ``aph_plan_step`` and returns the final state; its
backward pass runs the hand-written adjoint kernels (``aph_adjoint_step``) with step
checkpointing: the forward keeps every ``checkpoint_every``-th state, the backward
recomputes the states of one checkpoint segment at a time and walks them backwards.

Also here: the Kelvin-Helmholtz and batched Rayleigh-Taylor initial states as torch
functions of their perturbation amplitudes (so gradients reach them through autograd),
and the losses of those workloads (scalar variance, mixing-layer width).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .hydro import FusedDriver, _ptr


class Stepper:
    """A FusedDriver used as a frozen-dt step function on caller-provided states."""

    def __init__(self, driver: FusedDriver, dt: float):
        self.drv = driver
        self.dt = float(dt)
        self.drv.set_dt(self.dt)
        self.scratch = torch.empty_like(driver.u0)

    def step_(self, u):
        """Advance u (device tensor, pack layout) one RK2 step in place."""
        d = self.drv
        d.ctx.check(_lib.lib().aph_plan_step(d.ctx.h, d.plan, _ptr(u), _ptr(self.scratch),
                                             _ptr(d.dt), _ptr(d.time), d.cfl, d.tlim, 1))

    def stage1(self):
        """The stage-1 state of the last step_ / stage1_ (valid until the next call)."""
        return self.scratch

    def stage1_(self, u):
        """Only the first RK stage of a step from u (u is not modified); see stage1()."""
        d = self.drv
        d.ctx.check(_lib.lib().aph_fused_stage(d.ctx.h, d.plan, 1, _ptr(u), None,
                                               _ptr(self.scratch), _ptr(d.dt)))

    def adjoint_step_(self, u_start, lam, stage1=None):
        """lam <- (d step(u_start) / d u_start)^T lam, in place.  With stage1 (the step's
        stage-1 state, e.g. a copy of stage1() after step_(u_start)) no stage is
        recomputed."""
        d = self.drv
        if stage1 is None:
            d.ctx.check(_lib.lib().aph_adjoint_step(d.ctx.h, d.plan, _ptr(u_start),
                                                    _ptr(self.scratch), _ptr(d.dt), _ptr(lam)))
        else:
            d.ctx.check(_lib.lib().aph_adjoint_step_stored(d.ctx.h, d.plan, _ptr(u_start),
                                                           _ptr(stage1), _ptr(d.dt),
                                                           _ptr(lam)))


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, stepper, nsteps, every, keep_stage1):
        u = u0.detach().clone().contiguous()
        checkpoints = []
        for n in range(nsteps):
            if n % every == 0:
                checkpoints.append(u.clone())
            stepper.step_(u)
        ctx.stepper, ctx.nsteps, ctx.every = stepper, nsteps, every
        ctx.keep_stage1 = keep_stage1
        ctx.checkpoints = checkpoints
        return u

    @staticmethod
    def backward(ctx, grad):
        st, nsteps, every = ctx.stepper, ctx.nsteps, ctx.every
        lam = grad.detach().clone().contiguous()
        mask = interior_mask_tensor(st.drv.bs, lam.device)
        lam.mul_(mask)  # ghost entries carry no cotangent
        for k in range(len(ctx.checkpoints) - 1, -1, -1):
            n0 = k * every
            n1 = min(nsteps, n0 + every)
            # recompute the segment's step inputs (and, if kept, their stage-1 states)
            states, stage1 = [], []
            u = ctx.checkpoints[k].clone()
            for n in range(n0, n1):
                states.append(u.clone())
                if n < n1 - 1:
                    st.step_(u)
                elif ctx.keep_stage1:
                    st.stage1_(u)  # the last input needs only its stage 1
                if ctx.keep_stage1:
                    stage1.append(st.stage1().clone())
            for n in range(n1 - 1, n0 - 1, -1):
                s1 = stage1[n - n0] if ctx.keep_stage1 else None
                st.adjoint_step_(states[n - n0], lam, stage1=s1)
        return lam * mask, None, None, None, None


def rollout(u0, stepper, nsteps, checkpoint_every=20, keep_stage1=True):
    """Differentiable nsteps-step rollout.  The backward pass keeps every checkpoint_every
    -th state; with keep_stage1 it also keeps the stage-1 states of the segment being
    reversed (twice the segment's memory) so that no stage is computed twice."""
    return _Rollout.apply(u0, stepper, int(nsteps), int(checkpoint_every), bool(keep_stage1))


def interior_mask_tensor(bs, device):
    """1.0 on interior cells, 0.0 on ghost storage, over all nvar components (cached on
    the block set object)."""
    cache = bs.__dict__.setdefault("_mask_tensors", {})
    key = str(device)
    if key not in cache:
        m = np.tile(bs.interior_mask(), bs.nvar).astype(np.float64)
        cache[key] = torch.from_numpy(m).to(device)
    return cache[key]


# ------------------------------------------------------------------ KH as f(a)
def kh_state(bs, amps, a=0.01, sigma2=0.01, drat=2.0, vflow=0.5, p0=2.5, gm1=2.0 / 3.0):
    """Kelvin-Helmholtz conserved state (problems.kh2d) as a differentiable torch
    function of the amplitudes (a 1-D float64 tensor on the target device)."""
    dev = amps.device
    parts = []
    cols = []
    for b in range(bs.nblocks):
        x, y, _ = bs.cell_centers(b)
        parts.append((x[0], y[0]))
    x = torch.from_numpy(np.stack([p[0] for p in parts])).to(dev)
    y = torch.from_numpy(np.stack([p[1] for p in parts])).to(dev)
    w = (torch.tanh((torch.abs(y) - 0.25) / a) + 1.0) * 0.5
    rho = w + (1.0 - w) * drat
    vx = vflow * w - (1.0 - w) * vflow
    env = torch.exp(-((torch.abs(y) - 0.25) ** 2) / sigma2)
    vy = torch.zeros_like(x)
    for k in range(amps.shape[0]):
        vy = vy + amps[k] * torch.sin(2.0 * np.pi * (k + 1) * x) * env
    en = p0 / gm1 + 0.5 * rho * (vx * vx + vy * vy)
    fields = [rho, rho * vx, rho * vy, torch.zeros_like(x), en, rho * (1.0 - w)]
    # scatter into the pack layout (interior cells)
    idx = torch.from_numpy(interior_index(bs)).to(dev)
    out = torch.zeros(bs.field_size(bs.nvar), dtype=torch.float64, device=dev)
    stacked = torch.stack([f.reshape(bs.nblocks, -1) for f in fields], dim=1)  # (nb, nv, cells)
    return out.index_put((idx.reshape(-1),), stacked.reshape(-1))


def interior_index(bs):
    """Flat pack indices of the interior cells, shaped (nblocks, nvar, cells_per_block)
    for blocks of equal shape."""
    return bs.interior_index()


def scalar_variance(u, bs):
    """Domain-integrated variance of the passive-scalar concentration s = (rho s)/rho:
    sum V (s - <s>)^2 with <s> the volume average, over interior cells."""
    idx = torch.from_numpy(interior_index(bs)).to(u.device)
    v = u[idx]  # (nb, nv, cells)
    s = v[:, 5] / v[:, 0]
    vol = torch.from_numpy(bs.blocks["vol"].astype(np.float64)).to(u.device)[:, None]
    mean = (vol * s).sum() / (vol.expand_as(s)).sum()
    return (vol * (s - mean) ** 2).sum()


# ------------------------------------------------------------------ RT batch as f(a)
def rt_state(bs, amps, g=-0.1, gm1=0.4, width=0.0):
    """Batched Rayleigh-Taylor conserved state (problems.rt_batch) as a differentiable
    torch function of the amplitudes, a (ninst, 6) float64 tensor on the target device;
    block b takes the amplitudes of its instance mesh_id[b]."""
    dev = amps.device
    xa, ya, _ = bs.cell_centers_all()
    x = torch.from_numpy(np.ascontiguousarray(xa)).to(dev)
    y = torch.from_numpy(np.ascontiguousarray(ya)).to(dev)
    mid = torch.from_numpy(bs.blocks["mesh_id"].astype(np.int64)).to(dev)
    a = amps[mid]  # (nblocks, 6)
    if width > 0.0:
        rho = 1.5 + 0.5 * torch.tanh(y / width)
        p = 2.5 + g * (1.5 * y + 0.5 * width * torch.log(torch.cosh(y / width)))
    else:
        rho = torch.where(y < 0.0, torch.ones_like(y), 2.0 * torch.ones_like(y))
        p = 2.5 + g * rho * y
    vy = torch.zeros_like(x)
    for k in range(amps.shape[1]):
        vy = vy + a[:, k, None, None, None] * torch.cos(2.0 * np.pi * (k + 1) * x / 0.5) \
            * torch.exp(-((y / 0.05) ** 2))
    zero = torch.zeros_like(x)
    fields = [rho, zero, rho * vy, zero, p / gm1 + 0.5 * rho * vy * vy]
    idx = torch.from_numpy(bs.interior_index()).to(dev)
    out = torch.zeros(bs.field_size(bs.nvar), dtype=torch.float64, device=dev)
    stacked = torch.stack([f.reshape(bs.nblocks, -1) for f in fields], dim=1)
    return out.index_put((idx.reshape(-1),), stacked.reshape(-1))


def mixing_width(u, bs, rho_light=1.0, rho_heavy=2.0, lx=0.5, per_instance=False):
    """Mixing-layer width of each instance, W = (1/Lx) * sum_cells 4 f (1 - f) dx dy
    with the heavy-fluid fraction f = (rho - rho_light) / (rho_heavy - rho_light); the
    sum over instances unless per_instance."""
    dev = u.device
    idx = torch.from_numpy(bs.interior_index()[:, 0, :]).to(dev)
    f = (u[idx] - rho_light) / (rho_heavy - rho_light)  # (nblocks, cells)
    area = torch.from_numpy((bs.blocks["dx"][:, 0] * bs.blocks["dx"][:, 1]).astype(np.float64))
    wb = (4.0 * f * (1.0 - f)).sum(dim=1) * area.to(dev) / lx  # per block
    if not per_instance:
        return wb.sum()
    mid = torch.from_numpy(bs.blocks["mesh_id"].astype(np.int64)).to(dev)
    n = int(bs.blocks["mesh_id"].max()) + 1
    return torch.zeros(n, dtype=torch.float64, device=dev).index_add(0, mid, wb)
