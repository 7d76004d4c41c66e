"""GPU parity of monopole gravity (aph_gravity.hip; monopole_gravity.hpp and
gravity.cpp:374-410) against the oracle, through the C ABI:

* the solver's tables (m_enc, phi_rad) and its field (g, phi) bitwise, in 1D, 2D and 3D,
  whatever the partitioning (the bin sums are exact integer sums);
* plain (ApophisDriver with the gravity package) and fused (aph_plan_set_monopole) steps
  bitwise, Evrard's collapse to t = 0.8 in the fused form bitwise, a captured graph
  replayed equal to eager steps;
* the reference's own accuracy criteria on the GPU results (the gravity test's falling L2
  errors; Evrard's collapse profiles at 64^3 against the exact solution).
"""
import dataclasses

import numpy as np
import pytest
import torch

from apophis_amd import problems
from oracle import compare
from oracle import oracle as orc
from tests.monopole_checks import (evrard_exact, evrard_profiles, evrard_verdict,
                                   sphere_errors)

pytestmark = pytest.mark.gpu


def _hydro():
    from apophis_amd import hydro

    return hydro


def _gravity_pin(p):
    """The deck's gravity block and mesh box; rmax and nrbin from their defaults."""
    nrbin, rmax, G = p.monopole
    lo, hi = p.meta["box"]
    mesh = {"nx1": nrbin}  # nrbin's default is nx1: stand in the bin count
    for d in range(3):
        mesh[f"x{d + 1}min"], mesh[f"x{d + 1}max"] = lo[d], hi[d]
    return {"gravity": {"type": "monopole", "gravity_constant": G},
            "parthenon/mesh": mesh}


def _plain(ctx, p, pack_size=None):
    hydro = _hydro()
    pin = {"parthenon/time": {"cfl": p.cfl, "integrator": p.integrator},
           "parthenon/mesh": {"nghost": int(p.bs.blocks["ng"].min())},
           "hydro": {"reconstruction": p.recon, "riemann": p.riemann},
           "eos": {"gm1": float(p.eos["gm1"][0])}}
    pkg = hydro.initialize_hydro(pin)
    grav = hydro.initialize_gravity(_gravity_pin(p))
    assert (grav.param("nrbin"), grav.param("rmax")) == p.monopole[:2]
    md = hydro.MeshData(ctx, p.bs, p.cons)
    return hydro.ApophisDriver(pkg, md, pack_size=pack_size, gravity=grav)


def _fused(ctx, p):
    hydro = _hydro()
    return hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl,
                             tlim=p.tlim, integrator=p.integrator, monopole=p.monopole)


CASES = {
    "sphere3d": lambda: problems.grav_sphere(res=32, block=16),
    "sphere2d": lambda: problems.grav_sphere(res=64, block=16, ndim=2, recon="plm",
                                             riemann="hllc"),
    "sphere1d": lambda: problems.grav_sphere(res=128, block=32, ndim=1, recon="ppm",
                                             riemann="hllc"),
    "evrard": lambda: problems.evrard(),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("pack_size", [None, 3])
def test_tables_and_field_match_oracle(gpu_ctx, name, pack_size):
    p = CASES[name]()
    print(f"[case] monopole tables/field {name} pack {pack_size}", flush=True)
    ref = orc.State(p)
    nrbin, rmax, G = p.monopole
    mb, _ = orc.mono_bin_masses(p.bs, ref.prim, nrbin, rmax)
    m_enc, phi_rad = orc.mono_tables(nrbin, p.bs.ndim, rmax, G, mb)
    grav, phi = orc.mono_field(p.bs, nrbin, rmax, G, m_enc, phi_rad)
    drv = _plain(gpu_ctx, p, pack_size=pack_size)
    drv.initialize()
    drv.md.grav.fill_(np.nan)
    drv.md.phi.fill_(np.nan)
    drv.mono.add_tasks(drv.parts, drv.md.grav, drv.md.phi)
    torch.cuda.synchronize()
    got_enc, got_phi = drv.mono.tables()
    assert np.array_equal(got_enc, m_enc)
    assert np.array_equal(got_phi, phi_rad)
    g = drv.md.grav.cpu().numpy()
    ph = drv.md.phi.cpu().numpy()
    for b in range(p.bs.nblocks):
        assert np.array_equal(p.bs.interior(g, b, 3), p.bs.interior(grav, b, 3))
        assert np.array_equal(p.bs.interior(ph, b, 1), p.bs.interior(phi, b, 1))


@pytest.mark.parametrize("name", list(CASES))
def test_plain_steps_match_oracle(gpu_ctx, name):
    """Three steps with the driver's dt policy (a fixed dt from the initial state at rest
    would let Evrard's cloud collapse through its centre in one step)."""
    p = CASES[name]()
    print(f"[case] monopole plain 3 steps {name}", flush=True)
    ref = orc.State(p)
    t_ref, _ = ref.run(nsteps=3)
    drv = _plain(gpu_ctx, p)
    t, n = drv.evolve(nlim=3)
    torch.cuda.synchronize()
    assert n == 3 and t == t_ref
    assert compare.bitwise_equal_interior(drv.md.cons.cpu().numpy(), ref.u0, p.bs)


@pytest.mark.parametrize("name", ["sphere2d", "sphere1d", "evrard"])
def test_fused_steps_match_oracle(gpu_ctx, name):
    p = CASES[name]()
    print(f"[case] monopole fused 3 steps {name}", flush=True)
    ref = orc.State(p)
    t_ref, _ = ref.run(nsteps=3)
    fd = _fused(gpu_ctx, p)
    fd.initial_dt()
    for _ in range(3):
        fd.step()
    torch.cuda.synchronize()
    assert float(fd.time.item()) == t_ref
    assert compare.bitwise_equal_interior(fd.u0.cpu().numpy(), ref.u0, p.bs)


def test_fused_evrard_to_tlim_matches_oracle(gpu_ctx):
    """The reference's monopole_32 configuration run to t = 0.8 with the device-side dt
    policy: the same steps and the same final state as the oracle's run."""
    p = problems.evrard()
    print("[case] monopole fused evrard to t=0.8", flush=True)
    ref = orc.State(p)
    t_ref, n_ref = ref.run(tlim=p.tlim)
    fd = _fused(gpu_ctx, p)
    fd.initial_dt()
    n = 0
    while float(fd.time.item()) < p.tlim * (1 - 1e-14):
        fd.step()
        n += 1
    torch.cuda.synchronize()
    assert n == n_ref
    assert float(fd.time.item()) == t_ref
    assert compare.bitwise_equal_interior(fd.u0.cpu().numpy(), ref.u0, p.bs)


def test_fused_graph_replay_equals_eager(gpu_ctx):
    p = problems.evrard()
    print("[case] monopole graph replay vs eager", flush=True)
    hydro = _hydro()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ctx = hydro._lib.Context(s.cuda_stream)
        a = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl,
                              tlim=p.tlim, monopole=p.monopole)
        b = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl,
                              tlim=p.tlim, monopole=p.monopole)
        a.initial_dt()
        b.initial_dt()
        b.capture(3)
        for r in range(3):
            for _ in range(3):
                a.step()
            b.replay()
            s.synchronize()
            assert np.array_equal(a.u0.cpu().numpy(), b.u0.cpu().numpy()), r
            assert a.dt.item() == b.dt.item()


def test_field_meets_the_reference_accuracy_criteria(gpu_ctx):
    """gravity.py's monopole configurations: L2 errors of phi and g_x against the analytic
    sphere fall from 16^3 to 32^3 to 64^3."""
    errs = []
    for res in (16, 32, 64):
        p = problems.grav_sphere(res=res, block=16 if res > 16 else 16)
        drv = _plain(gpu_ctx, p)
        drv.initialize()
        drv.mono.add_tasks(drv.parts, drv.md.grav, drv.md.phi)
        torch.cuda.synchronize()
        errs.append(sphere_errors(p.bs, drv.md.grav.cpu().numpy(), drv.md.phi.cpu().numpy()))
    errs = np.array(errs)
    print(errs)
    assert np.all(np.diff(errs, axis=0) < 0), errs


def test_evrard_64_meets_the_reference_checks(gpu_ctx):
    p = problems.evrard(res=64, block=32)
    print("[case] monopole evrard 64^3 profiles", flush=True)
    fd = _fused(gpu_ctx, p)
    fd.initial_dt()
    while float(fd.time.item()) < p.tlim * (1 - 1e-14):
        fd.step()
    torch.cuda.synchronize()
    from oracle import oracle as O  # the primitives of the final state, on the CPU
    st = O.State(dataclasses.replace(p, cons=fd.u0.cpu().numpy()))
    fails, rms = evrard_verdict(evrard_profiles(p.bs, st.prim), evrard_exact())
    print(rms)
    assert not fails, fails


def test_adjoint_refuses_monopole(gpu_ctx):
    hydro = _hydro()
    from apophis_amd import adjoint

    p = problems.evrard()
    fd = _fused(gpu_ctx, p)
    dt = float(fd.initial_dt().item())
    st = adjoint.Stepper(fd, dt)
    lam = torch.zeros_like(fd.u0)
    with pytest.raises(hydro._lib.HydroError, match="monopole"):
        st.adjoint_step_(fd.u0.clone(), lam)


def test_evrard_history_output_matches_oracle_sums(gpu_ctx, tmp_path):
    """The deck's history output (inputs/evrards_collapse.in:94-97: hst every 0.05) from
    the plain driver: a row at t = 0 and whenever t passes a multiple of 0.05, and its
    total_mass / total_energy equal to the oracle's sums of the same states."""
    from apophis_amd import output

    p = problems.evrard()
    print("[case] monopole evrard hst", flush=True)
    hst = output.HstOutput(str(tmp_path), dt=0.05, block_number=1)
    drv = _plain(gpu_ctx, p)
    t, n = drv.evolve(tlim=p.tlim, hst=hst)
    torch.cuda.synchronize()
    rows = np.genfromtxt(tmp_path / "parthenon.out1.hst")
    ref = orc.State(p)
    want, tt, k = [], 0.0, 0
    want.append((0.0, 0, orc.global_sum(p.bs, ref.u0, 0), orc.global_sum(p.bs, ref.u0, 4)))
    nxt = 0.05
    dt_est, dt = ref.estimate_dt(), np.inf
    while tt < p.tlim * (1 - 1e-14):
        dt = min(2.0 * dt, dt_est)
        if p.tlim - tt < dt:
            dt = p.tlim - tt
        dt_est = ref.step(dt)
        tt += dt
        k += 1
        if tt >= nxt:
            want.append((tt, k, orc.global_sum(p.bs, ref.u0, 0), orc.global_sum(p.bs, ref.u0, 4)))
            nxt += 0.05
    want = np.array(want)
    assert rows.shape == (len(want), 6)
    np.testing.assert_array_equal(rows[:, 2], want[:, 1])
    np.testing.assert_array_equal(rows[:, 3], p.bs.nblocks)
    # the file keeps 7 significant digits
    np.testing.assert_allclose(rows[:, 0], want[:, 0], rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(rows[:, 4:], want[:, 2:], rtol=1e-6)
