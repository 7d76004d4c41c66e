"""The adjoint of multilevel meshes on the GPU (aph_adjoint_step on a plan with
aph_plan_set_amr): two frozen-dt steps of the batched form reversed with the hand-written
kernels -- the sweeps reading cross-level ghost cells, the flux correction in reverse (a
coarse face's own Riemann solve gets no cotangent, each fine face across it a 2^(1-ndim)
share of the coarse copy's), the ghost cotangents taken back through the cross-level ops
(minmod's active branch, the mirror bits) -- against torch autograd through
tests/torch_amr.py (itself equal to the oracle's multilevel step bitwise), 1e-10 relative;
the stored and recomputed stage-1 forms give the same bits."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle import amr_oracle as ao
from tests import test_amr, test_amr_physical, torch_amr

pytestmark = pytest.mark.gpu

CASES = [("2d_L", "plm", "hllc", "rk2", test_amr), ("2d_ppm", "ppm", "hllc", "rk2", test_amr),
         ("3d", "plm", "hlle", "rk2", test_amr), ("2d_3lev", "plm", "hllc", "rk2", test_amr),
         ("2d_3lev_ppm", "ppm", "hllc", "vl2", test_amr),
         ("2d_out", "plm", "hllc", "rk3", test_amr_physical),
         ("2d_ref", "ppm", "lhllc", "rk2", test_amr_physical),
         ("3d_ref", "ppm", "hllc", "rk2", test_amr_physical)]


@pytest.mark.parametrize("name,recon,riemann,integ,mod", CASES,
                         ids=[f"{c[0]}-{c[1]}-{c[2]}-{c[3]}" for c in CASES])
def test_multilevel_adjoint_matches_autograd(gpu_ctx, name, recon, riemann, integ, mod):
    from apophis_amd import adjoint, hydro

    print(f"[case] multilevel adjoint {name} {recon} {riemann} {integ}", flush=True)
    m = mod._mesh(name)
    p = mod._problem(m, recon=recon)
    p = dataclasses.replace(p, riemann=riemann, integrator=integ)
    dt = ao.AmrState(p, m).estimate_dt()
    eos = (float(p.eos["gm1"][0]), float(p.eos["dfloor"][0]), float(p.eos["pfloor"][0]))
    tm = torch_amr.Multilevel(m, p.bs, recon, riemann, eos, integrator=integ)
    nsteps = 2
    rng = np.random.default_rng(11)
    lam_flat = np.zeros_like(p.cons)
    mask = np.tile(p.bs.interior_mask(), p.bs.nvar)
    lam_flat[mask] = rng.normal(size=int(mask.sum()))
    # reference: d(sum lam * U_N) / dU_0 through the restatement
    u0 = tm.interior(p.cons).clone().requires_grad_(True)
    u = u0
    for _ in range(nsteps):
        u = tm.step(u, dt)
    (u * tm.interior(lam_flat)).sum().backward()
    want = u0.grad.numpy()
    # GPU: forward steps and adjoint steps of the multilevel plan
    fd = hydro.FusedDriver(gpu_ctx, p.bs, p.cons, recon, riemann, p.eos, p.cfl,
                           integrator=integ, amr_mesh=m)
    st = adjoint.Stepper(fd, dt)
    x = torch.from_numpy(p.cons.copy()).cuda()
    states, stage1 = [], []
    for _ in range(nsteps):
        states.append(x.clone())
        st.step_(x)
    torch.cuda.synchronize()
    assert torch.equal(tm.interior(x.cpu().numpy()), u.detach())  # the forward, bitwise
    for n in range(nsteps):
        st.stage1_(states[n])
        stage1.append(st.stage1().clone())
    lam = torch.from_numpy(lam_flat).cuda()
    lam2 = lam.clone()
    for n in range(nsteps - 1, -1, -1):
        st.adjoint_step_(states[n], lam)
        st.adjoint_step_(states[n], lam2, stage1=stage1[n])
    torch.cuda.synchronize()
    assert torch.equal(lam, lam2)
    got = tm.interior(lam.cpu().numpy()).numpy()
    scale = np.abs(want).max(axis=(1, 2, 3, 4), keepdims=True)
    err = np.max(np.abs(got - want) / np.maximum(scale, 1e-300))
    print(f"  max rel err {err:.3e}", flush=True)
    assert err <= 1e-10
