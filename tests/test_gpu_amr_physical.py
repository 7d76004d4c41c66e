"""Multilevel meshes with physical domain faces on the GPU through the C ABI: the
cross-level fill followed by the boundary conditions (edges and corners beyond the domain
included) against the oracle bitwise; plain and fused steps on outflow, reflecting and
mixed meshes (2D, 3D, two and three levels, PLM and PPM) against the oracle's steps
bitwise; the reference's static-refinement gravity cases with the monopole solver:
Evrard's collapse on the deck's three-level mesh (steps bitwise against the oracle, the
fused and plain forms to t = 0.8 equal, the reference's profile checks) and the gravity
accuracy test's SMR limits."""
import dataclasses

import numpy as np
import pytest
import torch

from apophis_amd import amr, hydro, problems
from oracle import amr_oracle as ao
from oracle import oracle as orc
from tests.monopole_checks import (evrard_exact, evrard_profiles, evrard_verdict,
                                   sphere_errors_max)
from tests.test_amr_physical import MESHES, _mesh, _problem, _totals

pytestmark = pytest.mark.gpu

CASES = [("2d_out", "plm", "hllc"), ("2d_ref", "ppm", "hllc"), ("2d_ppm_mixed", "ppm", "hlle"),
         ("2d_3lev_ref", "plm", "hllc"), ("3d_out", "plm", "hlle"), ("3d_ref", "ppm", "hllc")]


def _pin(p, ng):
    return {"parthenon/time": {"cfl": p.cfl, "integrator": p.integrator},
            "parthenon/mesh": {"nghost": ng},
            "hydro": {"reconstruction": p.recon, "riemann": p.riemann},
            "eos": {"gm1": float(p.eos["gm1"][0])}}


@pytest.mark.parametrize("name", list(MESHES))
def test_fill_with_walls_matches_oracle(gpu_ctx, name):
    print(f"[case] amr fill with walls {name}", flush=True)
    m = _mesh(name)
    bs = m.block_set()
    u = np.random.default_rng(5).normal(size=bs.field_size(5))
    md = hydro.MeshData(gpu_ctx, bs, u)
    plan = amr.AmrPlan(gpu_ctx, md.dbs, m, bs)
    hydro.fill_ghosts(md)
    plan.fill(md.cons)
    want = u.copy()
    orc.fill_ghosts(bs, want)
    ao.amr_fill(m, bs, want)
    orc.fill_ghosts(bs, want)  # the boundary conditions again: edges and corners
    torch.cuda.synchronize()
    assert np.array_equal(md.cons.cpu().numpy(), want)


@pytest.mark.parametrize("name,recon,riemann", CASES)
def test_plain_steps_with_walls_match_oracle(gpu_ctx, name, recon, riemann):
    print(f"[case] amr walls plain steps {name} {recon} {riemann}", flush=True)
    m = _mesh(name)
    p = _problem(m, recon=recon)
    p.riemann = riemann
    ref = ao.AmrState(p, m)
    md = hydro.MeshData(gpu_ctx, p.bs, p.cons)
    plan = amr.AmrPlan(gpu_ctx, md.dbs, m, p.bs)
    drv = hydro.ApophisDriver(hydro.initialize_hydro(_pin(p, m.ng)), md, amr=plan)
    dt = drv.initialize()
    assert dt == ref.estimate_dt()
    for _ in range(2):
        assert drv.step(dt) == ref.step(dt)
    torch.cuda.synchronize()
    assert np.array_equal(md.cons.cpu().numpy(), ref.u0)


@pytest.mark.parametrize("name,recon,riemann", CASES)
def test_fused_steps_with_walls_match_oracle(gpu_ctx, name, recon, riemann):
    print(f"[case] amr walls fused steps {name} {recon} {riemann}", flush=True)
    m = _mesh(name)
    p = _problem(m, recon=recon)
    p.riemann = riemann
    ref = ao.AmrState(p, m)
    dt = ref.estimate_dt()
    fd = hydro.FusedDriver(gpu_ctx, p.bs, p.cons, recon, riemann, p.eos, p.cfl, amr_mesh=m)
    fd.initial_dt()
    torch.cuda.synchronize()
    assert float(fd.dt.item()) == dt
    fd.set_dt(dt)
    for _ in range(2):
        fd.step(fixed_dt=True)
        ref.step(dt)
    torch.cuda.synchronize()
    mask = np.tile(p.bs.interior_mask(), 5)
    assert np.array_equal(fd.u0.cpu().numpy()[mask], ref.u0[mask])


def test_closed_reflecting_box_conserves_on_gpu(gpu_ctx):
    print("[case] amr walls closed box totals", flush=True)
    m = _mesh("2d_3lev_ref")
    p = _problem(m, recon="ppm")
    fd = hydro.FusedDriver(gpu_ctx, p.bs, p.cons, "ppm", "hllc", p.eos, p.cfl, amr_mesh=m)
    fd.initial_dt()
    t0 = _totals(p.bs, p.cons)
    for _ in range(20):
        fd.step()
    torch.cuda.synchronize()
    t1 = _totals(p.bs, fd.u0.cpu().numpy())
    assert abs(t1[0] - t0[0]) <= 1e-13 * t0[0] and abs(t1[4] - t0[4]) <= 1e-13 * t0[4]


def _gravity_pin(p):
    nrbin, rmax, G = p.monopole
    lo, hi = p.meta["box"]
    mesh = {"nx1": nrbin}  # nrbin's default is nx1: stand in the bin count
    for d in range(3):
        mesh[f"x{d + 1}min"], mesh[f"x{d + 1}max"] = lo[d], hi[d]
    return {"gravity": {"type": "monopole", "gravity_constant": G}, "parthenon/mesh": mesh}


def _plain_smr(ctx, p):
    m = p.meta["mesh"]
    md = hydro.MeshData(ctx, p.bs, p.cons)
    plan = amr.AmrPlan(ctx, md.dbs, m, p.bs)
    pkg = hydro.initialize_hydro(_pin(p, m.ng))
    grav = hydro.initialize_gravity(_gravity_pin(p))
    return hydro.ApophisDriver(pkg, md, amr=plan, gravity=grav), plan


def _fused_smr(ctx, p):
    return hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl,
                             tlim=p.tlim, integrator=p.integrator, monopole=p.monopole,
                             amr_mesh=p.meta["mesh"])


def test_evrard_smr_steps_match_oracle(gpu_ctx):
    """monopole_smr_32 (32^3 base in 16^3 blocks, [-0.5, 0.5]^3 at level 2: 120 blocks
    over three levels, outflow): three steps in each form against the oracle, bitwise."""
    print("[case] evrard smr 3 steps plain/fused", flush=True)
    p = problems.evrard(smr_level=2)
    ref = ao.AmrState(p, p.meta["mesh"])
    drv, _ = _plain_smr(gpu_ctx, p)
    fd = _fused_smr(gpu_ctx, p)
    dt = drv.initialize()
    assert dt == ref.estimate_dt()
    fd.initial_dt()
    fd.set_dt(dt)
    for _ in range(3):
        est = drv.step(dt)
        fd.step(fixed_dt=True)
        assert est == ref.step(dt)
    torch.cuda.synchronize()
    mask = np.tile(p.bs.interior_mask(), 5)
    assert np.array_equal(drv.md.cons.cpu().numpy(), ref.u0)
    assert np.array_equal(fd.u0.cpu().numpy()[mask], ref.u0[mask])


def test_evrard_smr_to_tlim_meets_the_reference_checks(gpu_ctx):
    """The deck's own mesh run to t = 0.8 with the device dt policy in the fused form, the
    plain form's run equal to it step for step, and the reference's profile checks
    (evrards_collapse.py) on the result."""
    print("[case] evrard smr to t=0.8 fused and plain", flush=True)
    p = problems.evrard(smr_level=2)
    fd = _fused_smr(gpu_ctx, p)
    fd.initial_dt()
    n = 0
    while float(fd.time.item()) < p.tlim * (1 - 1e-14):
        fd.step()
        n += 1
    torch.cuda.synchronize()
    drv, _ = _plain_smr(gpu_ctx, p)
    t, n_plain = drv.evolve(tlim=p.tlim)
    torch.cuda.synchronize()
    assert n_plain == n and t == float(fd.time.item())
    mask = np.tile(p.bs.interior_mask(), 5)
    u = fd.u0.cpu().numpy()
    assert np.array_equal(u[mask], drv.md.cons.cpu().numpy()[mask])
    st = orc.State(dataclasses.replace(p, cons=u))  # primitives of the final state
    fails, rms = evrard_verdict(evrard_profiles(p.bs, st.prim), evrard_exact())
    print(n, rms, flush=True)
    assert not fails, fails


def test_gravity_smr_meets_the_reference_limits(gpu_ctx):
    """gravity.py's monopole_smr_16 (16^3 base in 8^3 blocks, [-0.25, 0.25]^3 at level 2):
    the field on the GPU equals the oracle's bitwise and meets REFINED_LIMITS (phi L2 5e-3,
    g_x L2 5e-2, phi max 5e-2, g_x max 0.2)."""
    print("[case] gravity smr limits", flush=True)
    p = problems.grav_sphere(res=16, block=8, smr_level=2)
    drv, _ = _plain_smr(gpu_ctx, p)
    drv.initialize()
    drv.mono.add_tasks(drv.parts, drv.md.grav, drv.md.phi)
    torch.cuda.synchronize()
    grav, phi = drv.md.grav.cpu().numpy(), drv.md.phi.cpu().numpy()
    ref = ao.AmrState(p, p.meta["mesh"])
    ref.gravity_field()
    mask3 = np.tile(p.bs.interior_mask(), 3)
    assert np.array_equal(grav[mask3], ref.grav[mask3])
    l2p, l2g, mxp, mxg = sphere_errors_max(p.bs, grav, phi)
    print(l2p, l2g, mxp, mxg, flush=True)
    assert l2p <= 5e-3 and l2g <= 5e-2 and mxp <= 5e-2 and mxg <= 2e-1
