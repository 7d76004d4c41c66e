"""GPU checks of the batched Rayleigh-Taylor workload (instances = independent meshes in
one block set, one fused launch per stage, gravity + reflecting walls):

* batching is exact: after N frozen-dt steps every instance of the batch equals the same
  instance run alone, bitwise;
* the gradient of the summed mixing-layer width with respect to the (ninst, 6) mode
  amplitudes, through rt_state -> rollout (adjoint kernels, checkpointing) -> loss,
  equals the per-instance gradients computed alone (1e-13 relative: only the order of
  the final torch reductions differs);
* dW/du0 against torch autograd through the FP64 restatement (tests/torch_ref.py) of
  the same mesh over 40 PPM+HLLC steps: relative 1e-10;
* d(W)/d(a) against central finite differences with PLM, whose limiter is continuous
  (piecewise smooth): FD approaches the adjoint as h shrinks (O(h) kink crossings), to
  1e-4 at h = 1e-7.  The reference PPM limiter is discontinuous (tests/test_point_exact
  .py), so FD of a PPM rollout does not converge when a jump lies within +-h; PPM
  gradients are checked against autograd.
"""
import numpy as np
import pytest
import torch

from apophis_amd import problems

pytestmark = pytest.mark.gpu

GRIDS = ((32, 96), (64, 192))


def _stepper(ctx, p, dt):
    from apophis_amd import adjoint, hydro

    fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl, g=p.g)
    return adjoint.Stepper(fd, dt)


def _single(p, i, block=16):
    nx, ny = p.meta["grids"][i]
    return problems.rt2d(nx=nx, ny=ny, nblk=(nx // block, ny // block),
                         amps=p.meta["amps"][i], mesh_id=0)


def _dt(ctx, p):
    from apophis_amd import hydro

    fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl, g=p.g)
    return float(fd.initial_dt().item())


def test_batched_forward_equals_instances_alone(gpu_ctx):
    print("[case] RT batch forward vs instances alone", flush=True)
    p = problems.rt_batch(ninst=3, grids=GRIDS, block=16, seed=21)
    dt = _dt(gpu_ctx, p)
    st = _stepper(gpu_ctx, p, dt)
    u = torch.from_numpy(p.cons.copy()).cuda()
    for _ in range(20):
        st.step_(u)
    got = u.cpu().numpy()
    off = 0
    for i in range(3):
        q = _single(p, i)
        sq = _stepper(gpu_ctx, q, dt)
        v = torch.from_numpy(q.cons.copy()).cuda()
        for _ in range(20):
            sq.step_(v)
        ref = v.cpu().numpy()
        for b in range(q.bs.nblocks):
            a = p.bs.interior(got, off + b, 5)
            c = q.bs.interior(ref, b, 5)
            assert np.array_equal(a.view(np.uint64), c.view(np.uint64)), (i, b)
        off += q.bs.nblocks


def _width_grad(ctx, p, amps, dt, nsteps, every, width=0.0):
    from apophis_amd import adjoint

    st = _stepper(ctx, p, dt)
    a = torch.as_tensor(amps, dtype=torch.float64).cuda().requires_grad_(True)
    u0 = adjoint.rt_state(p.bs, a, width=width)
    loss = adjoint.mixing_width(adjoint.rollout(u0, st, nsteps, every), p.bs)
    loss.backward()
    return float(loss.item()), a.grad.detach().cpu().numpy()


def test_batched_gradient_equals_instances_alone(gpu_ctx):
    print("[case] RT batch gradient vs instances alone", flush=True)
    p = problems.rt_batch(ninst=3, grids=GRIDS, block=16, seed=22)
    dt = _dt(gpu_ctx, p)
    loss, grad = _width_grad(gpu_ctx, p, p.meta["amps"], dt, 30, 10)
    assert np.isfinite(grad).all() and np.abs(grad).max() > 0
    tot = 0.0
    for i in range(3):
        q = _single(p, i)
        li, gi = _width_grad(gpu_ctx, q, p.meta["amps"][i:i + 1], dt, 30, 10)
        tot += li
        np.testing.assert_allclose(grad[i], gi[0], rtol=1e-13, atol=1e-13 * np.abs(gi).max())
    assert abs(loss - tot) <= 1e-13 * abs(tot)


def test_width_gradient_matches_torch_autograd(gpu_ctx):
    """dW/du0 of one RT instance (sharp interface, 2x6 blocks, 40 steps, checkpoint every
    10) from the adjoint kernels against torch autograd through the FP64 restatement
    (tests/torch_ref.py) of the same mesh: relative 1e-10."""
    from apophis_amd import adjoint
    from tests import torch_ref

    print("[case] RT d(width)/d(u0) vs torch autograd", flush=True)
    p = problems.rt2d(nx=32, ny=96, nblk=(2, 6), seed=24)
    dt = _dt(gpu_ctx, p)
    nsteps = 40
    eos = (float(p.eos["gm1"][0]), float(p.eos["dfloor"][0]), float(p.eos["pfloor"][0]))
    dx = p.bs.blocks["dx"][0]
    u0 = torch.from_numpy(torch_ref.lattice_to_global(p.bs, p.cons)).requires_grad_(True)
    u = u0
    for _ in range(nsteps):
        u = torch_ref.rk2_step_bc(u, dt, "ppm", "hllc", eos, dx, 2, ("periodic", "reflect"),
                                  g=p.g)
    f = u[0] - 1.0
    (4.0 * f * (1.0 - f)).sum().mul(dx[0] * dx[1] / 0.5).backward()
    want = u0.grad.numpy()
    st = _stepper(gpu_ctx, p, dt)
    x = torch.from_numpy(p.cons.copy()).cuda().requires_grad_(True)
    adjoint.mixing_width(adjoint.rollout(x, st, nsteps, 10), p.bs).backward()
    got = torch_ref.lattice_to_global(p.bs, x.grad.cpu().numpy())
    scale = np.abs(want).max(axis=(1, 2, 3), keepdims=True)
    err = np.max(np.abs(got - want) / np.maximum(scale, 1e-300))
    print(f"  max rel err {err:.3e}", flush=True)
    assert err <= 1e-10


def test_c5_instance_width_gradient_matches_fp64_autograd(gpu_ctx):
    """One instance of the C5 set at its coarsest grid (128x384 in blocks of 32^2,
    PPM+HLLC, g = -0.1, reflecting y walls) over the C5 rollout (300 steps at the frozen
    CFL dt, checkpoints every 20): the gradient of the mixing width with respect to its six
    mode amplitudes through the adjoint kernels against FP64 torch autograd through
    tests/torch_ref.py on the device (torch.utils.checkpoint every 20 steps).  The
    restatement equals the oracle bitwise on the CPU, but torch's device arithmetic for it is
    not bitwise the kernels': the 300-step forwards agree to 1.4e-9 (asserted 1e-7), the
    gradients to 1e-12 of the largest component (measured 2.4e-14).  PPM's limiter is
    discontinuous, so central FD does not converge here; this is the check."""
    from torch.utils.checkpoint import checkpoint

    from apophis_amd import adjoint
    from tests import torch_ref

    print("[case] RT C5 instance 128x384, 300 steps: d(width)/d(a) vs FP64 autograd", flush=True)
    p = problems.rt_batch(ninst=1, grids=((128, 384),), block=32, seed=11)
    dt = _dt(gpu_ctx, p)
    st = _stepper(gpu_ctx, p, dt)
    amps = torch.from_numpy(p.meta["amps"]).cuda()
    a = amps.clone().requires_grad_(True)
    u_n = adjoint.rollout(adjoint.rt_state(p.bs, a), st, 300, 20)
    adjoint.mixing_width(u_n, p.bs).backward()
    grad = a.grad.detach().cpu().numpy()
    gidx = torch.from_numpy(torch_ref.lattice_to_global(
        p.bs, np.arange(p.bs.field_size(p.bs.nvar), dtype=np.float64)).astype(np.int64)).cuda()
    eos = (float(p.eos["gm1"][0]), float(p.eos["dfloor"][0]), float(p.eos["pfloor"][0]))
    dx = p.bs.blocks["dx"][0]

    def segment(u):
        for _ in range(20):
            u = torch_ref.rk2_step_bc(u, dt, "ppm", "hllc", eos, dx, 2,
                                      ("periodic", "reflect"), g=p.g)
        return u

    b = amps.clone().requires_grad_(True)
    u = adjoint.rt_state(p.bs, b)[gidx]
    for _ in range(15):
        u = checkpoint(segment, u, use_reentrant=False)
    fwd = float((u.detach() - u_n.detach()[gidx]).abs().max())
    print(f"  forward max |diff| {fwd:.2e}", flush=True)
    assert fwd <= 1e-7
    f = u[0] - 1.0
    (4.0 * f * (1.0 - f)).sum().mul(dx[0] * dx[1] / 0.5).backward()
    want = b.grad.detach().cpu().numpy()
    err = float(np.abs(grad - want).max() / np.abs(want).max())
    print(f"  grad {grad}\n  autograd {want}\n  max rel {err:.2e}", flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("width", [0.0, 0.05])
def test_width_gradient_vs_central_fd(gpu_ctx, width):
    print(f"[case] RT batch d(width)/d(a) vs FD, interface width {width}", flush=True)
    p = problems.rt_batch(ninst=2, grids=((32, 96),), block=16, seed=23, width=width,
                          recon="plm")
    dt = _dt(gpu_ctx, p)
    nsteps, every = 60, 20
    amps = p.meta["amps"]
    loss, grad = _width_grad(gpu_ctx, p, amps, dt, nsteps, every, width)
    errs = {}
    for h in (1e-4, 1e-5, 1e-6, 1e-7):
        fd = np.zeros_like(amps)
        for i in range(amps.shape[0]):
            for k in range(amps.shape[1]):
                e = np.zeros_like(amps)
                e[i, k] = h
                lp, _ = _width_grad(gpu_ctx, p, amps + e, dt, nsteps, every, width)
                lm, _ = _width_grad(gpu_ctx, p, amps - e, dt, nsteps, every, width)
                fd[i, k] = (lp - lm) / (2 * h)
        errs[h] = np.abs(grad - fd).max() / np.abs(fd).max()
        print(f"  h={h:g}: loss {loss:.10e} max rel {errs[h]:.2e}", flush=True)
    # sharp interface: 7e-7 at h = 1e-7; the smooth one (the limiter acts across its
    # whole width): 2.7e-5 at h = 1e-7; both fall as h does
    assert errs[1e-7] <= (1e-6 if width == 0.0 else 1e-4)
    assert errs[1e-7] < errs[1e-5] < errs[1e-4]
