"""Benchmark of the hydro flux path on one MI355X.

Default workload (N=1): 3D blast wave, 256^3 cells as 512 meshblocks of 32^3, FP64,
gamma=5/3, PPM + HLLC, RK2, CFL 0.3, periodic.  A step is one full RK2 step (2 stages)
of the batched fused form replayed from a captured hipGraph with device-resident dt.

Printed (one JSON line): zone-cycles/s (interior cells x RK2 steps / s), the roofline
of the dominant kernel (the fused stage kernel, timed live with HIP events on its own
stream), the plain form's throughput on the same workload, the uneven-block (batched
vs per-block plain) comparison, and the oracle's CPU throughput on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def bytes_per_zone_cycle(nv, ndim, g, n):
    """Algorithmic bytes of one fused RK2 zone-cycle: 8*NV*(2f+3), f = 1 + 2*d*g/n."""
    f = 1.0 + 2.0 * ndim * g / n
    return 8.0 * nv * (2.0 * f + 3.0)


def stage_bytes(bs, stencil_g):
    """Algorithmic bytes of one stage launch averaged over the two stages
    (stage 1: f+1 state reads/writes per cell, stage 2: f+2)."""
    tot = 0.0
    for b in range(bs.nblocks):
        n = int(bs.blocks["nx"][b][0])
        cells = int(np.prod(bs.blocks["nx"][b][: bs.ndim]))
        tot += cells * bytes_per_zone_cycle(bs.nvar, bs.ndim, stencil_g, n) / 2.0
    return tot


def measured_traffic(name="k_fused_blast256_ppm_hllc.md"):
    """HBM bytes per launch of the fused kernel from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by a 1 GiB copy's reading), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def profiled_valu_per_cell(name="k_fused_blast256_ppm_hllc.md"):
    """FP64 VALU lane-instructions per cell per stage of the fused kernel, from the
    committed rocprofv3 SQ_INSTS_VALU summary (wave instructions x 64 / cells), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("| VALU lane-instructions per cell per stage |"):
                return float(line.split("|")[2])
    except (OSError, ValueError):
        pass
    return None


# MI355X FP64 vector peak: 78.6 TFLOP/s counting an FMA as 2 -> 39.3e12 lane-instr/s
# (256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz)
VALU_PEAK_LANE_INSTR = 39.3e12


def event_time_stage(fd, torch, reps=10):
    """Average duration (ms) of one fused stage launch, HIP events on fd's stream."""
    s = torch.cuda.current_stream()
    times = []
    for r in range(reps):
        for stage, (a, b, c) in ((1, (fd.u0, None, fd.u1)), (2, (fd.u1, fd.u0, fd.u0))):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(s)
            fd.stage(stage, a, b, c)
            e1.record(s)
            e1.synchronize()
            times.append(e0.elapsed_time(e1))
    return float(np.mean(times)), float(np.min(times))


def _cpu_info():
    """(model name, physical cores) of the host from lscpu, or (None, None)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return None, None
    d = {}
    for line in out.splitlines():
        k, _, v = line.partition(":")
        d[k.strip()] = v.strip()
    try:
        phys = int(d.get("Core(s) per socket", "0")) * int(d.get("Socket(s)", "0")) or None
    except ValueError:
        phys = None
    return d.get("Model name"), phys


def _time_oracle(p, threads, max_steps, budget_s):
    """Oracle RK2 steps of problem p at its initial CFL dt on `threads` OpenMP threads:
    one untimed warm-up step, then steps until max_steps or budget_s seconds."""
    from oracle import oracle as orc

    orc.set_threads(threads)
    st = orc.State(p)
    dt = st.estimate_dt()
    st.step(dt)
    n = 0
    t0 = time.perf_counter()
    while n < max_steps and time.perf_counter() - t0 < budget_s:
        st.step(dt)
        n += 1
    el = time.perf_counter() - t0
    return p.bs.interior_cells() * n / el, n, el


def cpu_baseline():
    """The oracle -- the C restatement of the reference's flux path, built with gcc -O3
    -ffp-contract=off and OpenMP over blocks -- on the headline workload itself (256^3
    blast, 512 blocks of 32^3, PPM+HLLC), with every thread of this GPU's CPU share (the
    affinity set, capped by the OMP_NUM_THREADS the machine sets), and single-threaded on
    a 128^3 sample of the same blocks (64 of them), the reference's default of Kokkos
    Serial per rank."""
    from apophis_amd import problems

    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(aff, int(env)) if env.isdigit() and int(env) > 0 else aff
    model, phys = _cpu_info()
    v_all, n_all, el_all = _time_oracle(problems.blast3d(nblk=8, n=32), threads, 6, 12.0)
    v_one, n_one, el_one = _time_oracle(problems.blast3d(nblk=4, n=32), 1, 4, 12.0)
    return {"value": v_all, "unit": "zone-cycles/s", "cores": threads, "kind": "port",
            "sample": f"oracle rk2 steps of the 256^3 blast itself (512 blocks of 32^3, "
                      f"PPM+HLLC), {n_all} steps after 1 warm-up, {el_all:.1f} s, "
                      f"{threads} OpenMP threads",
            "cpu_model": model, "host_physical_cores": phys, "affinity_cpus": aff,
            "single_thread": {"value": v_one, "unit": "zone-cycles/s", "cores": 1,
                              "sample": f"128^3 blast (64 blocks of 32^3), {n_one} steps "
                                        f"after 1 warm-up, {el_one:.1f} s"}}


def adjoint_leg(ctx, hydro, problems, torch):
    """Kelvin-Helmholtz 512^2 (16x16 blocks of 32^2, PLM+HLLC, one passive scalar): time
    of a frozen-dt forward RK2 step vs one adjoint step (both stage adjoints, given the
    step's input and stage-1 states, as the rollout's backward pass runs it; also without
    the stage-1 state, which is then recomputed), and (after an untimed one-segment
    warm-up) of a whole 200-step rollout with
    checkpoints every 20 steps, forward and backward (segment recomputation included)."""
    from apophis_amd import adjoint

    p = problems.kh2d(nblk=16, n=32)
    fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl)
    dt = float(fd.initial_dt().item())  # CFL 0.3 estimate of the initial state, frozen
    st = adjoint.Stepper(fd, dt)
    u = torch.from_numpy(p.cons.copy()).cuda()
    n = 20
    states = []
    for _ in range(3):
        st.step_(u)
    for _ in range(n):  # the steps' input states, untimed
        states.append(u.clone())
        st.step_(u)
    stage1 = []
    for k in range(n):
        st.stage1_(states[k])
        stage1.append(st.stage1().clone())
    lam = torch.ones_like(u) * adjoint.interior_mask_tensor(p.bs, u.device)
    # one untimed pass of each timed loop first (scratch, clocks, caches), then the timed
    # passes: forward steps alone, and adjoint steps given the stage-1 states
    for _pass in range(2):  # the second pass's times are kept
        w = states[0].clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            st.step_(w)
        torch.cuda.synchronize()
        t_fwd = (time.perf_counter() - t0) / n
        lam_w = lam.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(n - 1, -1, -1):
            st.adjoint_step_(states[k], lam_w, stage1=stage1[k])
        torch.cuda.synchronize()
        t_adj = (time.perf_counter() - t0) / n
    lam_w = lam.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(n - 1, -1, -1):
        st.adjoint_step_(states[k], lam_w)
    torch.cuda.synchronize()
    t_adj_rc = (time.perf_counter() - t0) / n
    # warm-up: one checkpoint segment forward and back (the allocator's blocks for the
    # segment's states), untimed
    amps_w = torch.from_numpy(p.meta["amps"]).cuda().requires_grad_(True)
    adjoint.scalar_variance(adjoint.rollout(adjoint.kh_state(p.bs, amps_w), st, 20, 20),
                            p.bs).backward()
    amps = torch.from_numpy(p.meta["amps"]).cuda().requires_grad_(True)
    torch.cuda.synchronize()
    u0 = adjoint.kh_state(p.bs, amps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    uN = adjoint.rollout(u0, st, 200, 20)
    loss = adjoint.scalar_variance(uN, p.bs)
    torch.cuda.synchronize()
    t_roll_f = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t_roll_b = time.perf_counter() - t0
    return {"workload": "KH 512^2 (256 blocks of 32^2), PLM+HLLC + scalar, frozen dt", "dt": dt,
            "forward_step_ms": t_fwd * 1e3, "adjoint_step_ms": t_adj * 1e3,
            "adjoint_over_forward": t_adj / t_fwd,
            "adjoint_step_recompute_ms": t_adj_rc * 1e3,
            "rollout_200_forward_s": t_roll_f, "rollout_200_backward_s": t_roll_b,
            "rollout_backward_over_forward": t_roll_b / t_roll_f,
            "finite_grad": bool(torch.isfinite(amps.grad).all().item())}


def rt_batch_leg(ctx, hydro, problems, torch, nsteps=300, every=20):
    """Batched Rayleigh-Taylor with adjoint: 64 independent instances (grids drawn from
    128x384 / 256x768 / 512x1536, blocks of 32^2), PPM+HLLC, gravity, reflecting walls,
    all instances in one fused launch per stage.  After an untimed one-segment warm-up,
    a 300-step rollout at the frozen initial CFL dt with checkpoints every 20 steps, then
    the gradient of the summed
    mixing-layer width with respect to the 64x6 mode amplitudes."""
    from apophis_amd import adjoint

    p = problems.rt_batch()
    cells = p.bs.interior_cells()
    fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl, g=p.g)
    dt = float(fd.initial_dt().item())
    fd.autotune(reps=2)
    st = adjoint.Stepper(fd, dt)
    # warm-up: one checkpoint segment forward and back (the adjoint's scratch and the
    # allocator's blocks for the segment's states), untimed
    a_w = torch.from_numpy(p.meta["amps"]).cuda().requires_grad_(True)
    adjoint.mixing_width(adjoint.rollout(adjoint.rt_state(p.bs, a_w), st, every, every),
                         p.bs).backward()
    del a_w
    a = torch.from_numpy(p.meta["amps"]).cuda().requires_grad_(True)
    u0 = adjoint.rt_state(p.bs, a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    uN = adjoint.rollout(u0, st, nsteps, every)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss = adjoint.mixing_width(uN, p.bs)
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    g = a.grad
    return {"workload": "64 RT instances (0.5x1.5, grids 128x384/256x768/512x1536, blocks "
                        "of 32^2), PPM+HLLC, g=-0.1, frozen dt, checkpoint every 20",
            "instances": int(p.meta["amps"].shape[0]), "blocks": p.bs.nblocks, "cells": cells,
            "steps": nsteps, "dt": dt,
            "tiles": {"x".join(map(str, k)): "x".join(map(str, v))
                      for k, v in fd.tiles().items()},
            "forward_s": t1 - t0, "forward_zone_cycles_per_s": cells * nsteps / (t1 - t0),
            "backward_s": t2 - t1, "backward_over_forward": (t2 - t1) / (t1 - t0),
            "loss": float(loss.item()), "grad_finite": bool(torch.isfinite(g).all().item()),
            "grad_max": float(g.abs().max().item())}


def composition_leg(ctx, hydro, problems, torch, nsteps=20):
    """The reference deck inputs/kh-comp-advection.in (ONe composition + one level set:
    NV = 14, PLM + HLLC, blocks of 64 x 128) at its own 256 x 128 mesh and widened to
    4096 x 2048 (1024 blocks): zone-cycles/s of the batched fused form, graph-replayed."""
    out = {}
    for name, nx in (("deck_256x128", (256, 128)), ("wide_4096x2048", (4096, 2048))):
        p = problems.kh_comp(nx=nx)
        cells = p.bs.interior_cells()
        fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl)
        fd.initial_dt()
        fd.autotune(reps=3)
        for _ in range(2):
            fd.step()
        fd.capture(nsteps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fd.replay()
        torch.cuda.synchronize()
        out[name] = {"cells": cells, "blocks": p.bs.nblocks, "nvar": p.bs.nvar,
                     "zone_cycles_per_s": cells * nsteps / (time.perf_counter() - t0),
                     "tiles": {"x".join(map(str, k)): "x".join(map(str, v))
                               for k, v in fd.tiles().items()}}
        del fd
    return out


def monopole_leg(ctx, hydro, problems, torch, nsteps=20, res=256):
    """Evrard's collapse with monopole self-gravity (inputs/evrards_collapse.in's physics:
    PLM + HLLC, RK2, outflow, G = 1, nrbin = 4 res) at 256^3 in blocks of 32^3, batched fused
    form, graph-replayed: zone-cycles/s with the solver before every stage, and the same
    steps with gravity off, whose difference is the solver's cost per stage."""
    import dataclasses

    p = problems.evrard(res=res, block=32)
    cells = p.bs.interior_cells()
    out = {"cells": cells, "blocks": p.bs.nblocks, "nrbin": p.monopole[0]}
    for name, mono in (("monopole", p.monopole), ("no_gravity", None)):
        fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl,
                               tlim=p.tlim, monopole=mono)
        fd.initial_dt()
        for _ in range(2):
            fd.step()
        fd.capture(nsteps)
        fd.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fd.replay()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[name] = {"zone_cycles_per_s": cells * nsteps / el,
                     "ms_per_step": el / nsteps * 1e3}
        del fd
    out["solver_ms_per_stage"] = (out["monopole"]["ms_per_step"] -
                                  out["no_gravity"]["ms_per_step"]) / 2
    torch.cuda.empty_cache()
    return out


def evrard_smr_leg(ctx, hydro, problems, torch, solver="monopole"):
    """The deck inputs/evrards_collapse.in on a static-refinement mesh ([-0.5, 0.5]^3 at
    level 2, outflow) to t = 0.8, batched fused form with the device dt policy: steps, wall
    time, zone-cycles/s, and the reference's profile checks on the result.  solver
    "monopole": the reference test's monopole_smr_32 (32^3 base in 16^3 blocks: 120 blocks
    over three levels); "poisson": its smr_16 (16^3 base in 8^3 blocks, the deck's
    MGBiCGSTAB problem with flux correction, solved by BiCGSTAB with a multigrid V-cycle
    on the composite operator), with the time of one solve of the initial state and its
    iterations."""
    from tests.monopole_checks import evrard_exact, evrard_profiles, evrard_verdict

    if solver == "monopole":
        p = problems.evrard(smr_level=2)
        grav = {"monopole": p.monopole}
    else:
        p = problems.evrard(res=16, block=8, smr_level=2, solver="poisson")
        grav = {"poisson": p.poisson}
    cells = p.bs.interior_cells()
    extra = {}
    if solver == "poisson":  # one solve of the initial state, timed on its own
        md = hydro.MeshData(ctx, p.bs, p.cons)
        ps = hydro.PoissonGravitySolver(ctx, md.dbs, *p.poisson, mesh=p.meta["mesh"])
        hydro.cons_to_prim(md, p.eos)
        g = torch.zeros(p.bs.field_size(3), dtype=torch.float64, device=md.cons.device)
        ps.solve(md, g)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            ps.solve(md, g)
        torch.cuda.synchronize()
        _, iters, res = ps.info()
        extra = {"solve_ms": (time.perf_counter() - t0) / 5 * 1e3, "solve_iterations": iters,
                 "solve_rel_residual": res}
        del ps, md
    fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl, tlim=p.tlim,
                           amr_mesh=p.meta["mesh"], **grav)
    fd.initial_dt()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n = 0
    while float(fd.time.item()) < p.tlim * (1 - 1e-14):
        fd.step()
        n += 1
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    md = hydro.MeshData(ctx, p.bs, fd.u0.cpu().numpy())  # primitives of the final state
    hydro.cons_to_prim(md, p.eos)
    torch.cuda.synchronize()
    fails, rms = evrard_verdict(evrard_profiles(p.bs, md.prim.cpu().numpy()), evrard_exact())
    return {"blocks": p.bs.nblocks, "cells": cells, "levels": [int(v) for v in np.bincount(
                p.meta["mesh"].level)], "steps": n, "wall_s": el,
            "zone_cycles_per_s": cells * n / el, "profile_rms": rms,
            "reference_checks_pass": not fails, **extra}


def multilevel_adjoint_leg(ctx, hydro, problems, torch, reps=5):
    """The adjoint on the two-level blast (2D PPM+HLLC, 48^2 base blocks of 32^2 with the
    central 16^2 refined: 3.1 M cells): one forward step vs one adjoint step (stage 1
    stored, and recomputed) at a frozen dt."""
    from apophis_amd import adjoint

    p, m = problems.blast2d_two_level()
    fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl, amr_mesh=m)
    dt = float(fd.initial_dt().item())
    st = adjoint.Stepper(fd, dt)
    u = torch.from_numpy(p.cons.copy()).cuda()
    lam = torch.ones_like(u) * adjoint.interior_mask_tensor(p.bs, u.device)
    st.step_(u)
    st.stage1_(u)
    s1 = st.stage1().clone()
    st.adjoint_step_(u, lam, stage1=s1)
    torch.cuda.synchronize()

    def timed(fn):
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    f = timed(lambda: st.step_(u))
    st.stage1_(u)
    s1 = st.stage1().clone()
    a1 = timed(lambda: st.adjoint_step_(u, lam, stage1=s1))
    a2 = timed(lambda: st.adjoint_step_(u, lam))
    out = {"cells": p.bs.interior_cells(), "blocks": p.bs.nblocks, "forward_step_ms": f,
           "adjoint_step_stored_ms": a1, "adjoint_step_recompute_ms": a2,
           "stored_over_forward": a1 / f, "recompute_over_forward": a2 / f,
           "finite": bool(torch.isfinite(lam).all().item())}
    del fd, st
    torch.cuda.empty_cache()
    return out


def two_level_leg(ctx, hydro, problems, torch, nsteps=20):
    """A two-level mesh (2D blast, 48^2 base blocks of 32^2 with the central 16^2 refined:
    3.1 M cells, 3072 blocks): zone-cycles/s of the batched form (cross-level ghosts, fine
    blocks first, flux correction from the interface buffer; graph-replayed) and of the
    plain task form (aph_amr_fill / aph_amr_flux_correct between the tasks)."""
    from apophis_amd import amr
    t_plan = time.perf_counter()
    p, m = problems.blast2d_two_level()
    cells = p.bs.interior_cells()
    fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl, amr_mesh=m)
    t_plan = time.perf_counter() - t_plan
    fd.initial_dt()
    for _ in range(2):
        fd.step()
    fd.capture(nsteps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fd.replay()
    torch.cuda.synchronize()
    vf = cells * nsteps / (time.perf_counter() - t0)
    del fd
    pkg = hydro.initialize_hydro({
        "parthenon/time": {"cfl": p.cfl, "integrator": "rk2"},
        "parthenon/mesh": {"nghost": int(p.bs.blocks["ng"][0])},
        "hydro": {"reconstruction": p.recon, "riemann": p.riemann},
        "eos": {"gm1": float(p.eos["gm1"][0])}})
    md = hydro.MeshData(ctx, p.bs, p.cons)
    plan = amr.AmrPlan(ctx, md.dbs, m, p.bs)
    drv = hydro.ApophisDriver(pkg, md, amr=plan)
    dt = drv.initialize()
    dt = drv.step(dt)
    torch.cuda.synchronize()
    npl = 5
    t0 = time.perf_counter()
    for _ in range(npl):
        dt = drv.step(dt)
    torch.cuda.synchronize()
    vp = cells * npl / (time.perf_counter() - t0)
    return {"cells": cells, "blocks": p.bs.nblocks, "fine_blocks": int((m.level == 1).sum()),
            "ghost_ops": len(plan.ops), "flux_fixes": len(plan.fixes),
            "plan_build_s": round(t_plan, 2), "fused_zone_cycles_per_s": vf,
            "plain_zone_cycles_per_s": vp, "fused_over_plain": vf / vp}


def blast_deck_leg(ctx, hydro, problems, torch, ncycles=300):
    """The reference's adaptive deck (inputs/blast.in: 128^2 of 16^2 blocks, three levels
    on the pressure-gradient criterion, PLM + HLLC, vl2, cfl 0.4, derefine_count 10):
    the adaptive start, then `ncycles` cycles of step + device tags + regrid in each form.
    Reports zone-cycles/s of the whole loop (host regrids included) and of the steps alone
    on the final mesh, the regrids and the host seconds they took."""
    from apophis_amd import adaptive
    pkg = hydro.initialize_hydro(problems.blast_deck_pin())
    t0 = time.perf_counter()
    m, bs, cons = adaptive.initial_mesh(
        problems.blast_deck_base_mesh(), problems.blast_pgen(),
        lambda mm, bb, cc: adaptive.device_tags(ctx, pkg, mm, bb, cc))
    out = {"start_s": round(time.perf_counter() - t0, 3), "start_blocks": bs.nblocks,
           "start_levels": [int(v) for v in np.bincount(m.level)], "cycles": ncycles}
    for form in ("fused", "plain"):
        drv = adaptive.AdaptiveDriver(ctx, pkg, m, bs, cons, form=form)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        drv.evolve(ncycles=ncycles)
        w = time.perf_counter() - t0
        st = drv.stats
        r = {"zone_cycles_per_s": st["zone_cycles"] / w, "wall_s": round(w, 3),
             "regrids": st["regrids"], "regrid_host_s": round(st["regrid_s"], 3),
             "sim_time": drv.time, "end_blocks": drv.bs.nblocks,
             "end_levels": [int(v) for v in np.bincount(drv.mesh.level)]}
        if form == "fused":  # the steps alone, on the final mesh
            fd = drv.fd
            fd.set_dt(drv.dt)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(50):
                fd.step(fixed_dt=True)
            torch.cuda.synchronize()
            r["steps_only_zone_cycles_per_s"] = drv.bs.interior_cells() * 50 / (
                time.perf_counter() - t1)
        out[form] = r
        del drv
    # the whole deck to its tlim (1e-2), fused form: cycles, regrids, the final mesh and
    # the drift of the conserved totals (periodic: exact up to round-off)
    def totals(bs, u):
        t = np.zeros(5)
        for b in range(bs.nblocks):
            t += bs.interior(u, b, bs.nvar).reshape(bs.nvar, -1)[:5].sum(axis=1) * \
                bs.blocks["vol"][b]
        return t
    t_start = totals(bs, cons)
    drv = adaptive.AdaptiveDriver(ctx, pkg, m, bs, cons, form="fused",
                                  tlim=problems.BLAST_DECK["tlim"])
    t0 = time.perf_counter()
    n = drv.evolve()
    w = time.perf_counter() - t0
    t_end = totals(drv.bs, drv.state().cpu().numpy())
    out["to_tlim"] = {"tlim": drv.time, "cycles": n, "wall_s": round(w, 2),
                      "zone_cycles_per_s": drv.stats["zone_cycles"] / w,
                      "regrids": drv.stats["regrids"], "regrid_host_s": round(drv.stats["regrid_s"], 2),
                      "end_blocks": drv.bs.nblocks,
                      "end_levels": [int(v) for v in np.bincount(drv.mesh.level)],
                      "mass_rel_drift": float(abs(t_end[0] - t_start[0]) / abs(t_start[0])),
                      "energy_rel_drift": float(abs(t_end[4] - t_start[4]) / abs(t_start[4]))}
    del drv
    return out


def profile_run(args, ctx, hydro, problems, torch):
    """Eager launches for rocprofv3: 3 RK2 steps of the chosen form on the headline
    workload (or the uneven block set), then a 1 GiB device-to-device copy repeated 3
    times, whose FETCH_SIZE / WRITE_SIZE calibrate the counters (2^30 bytes each way)."""
    if args.profile_run in ("adjoint", "rt_adjoint"):
        from apophis_amd import adjoint

        if args.profile_run == "adjoint":
            p = problems.kh2d(nblk=16, n=32)
        else:  # batched RT instances (8 by default), PPM+HLLC + gravity
            p = problems.rt_batch(ninst=args.rt_instances)
        fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl, g=p.g)
        st = adjoint.Stepper(fd, 1e-4)
        u = torch.from_numpy(p.cons.copy()).cuda()
        lam = torch.ones_like(u) * adjoint.interior_mask_tensor(p.bs, u.device)
        for _ in range(3):
            st.step_(u)
            st.adjoint_step_(u, lam)
        torch.cuda.synchronize()
        print(f"profile-run {args.profile_run}: done", flush=True)
        return
    if args.profile_run == "amr":
        # multilevel fused form: 3 RK2 steps of the two-level blast (3072 blocks of 32^2),
        # then 3 of the blast deck's three-level start mesh (148 blocks of 16^2, vl2)
        from apophis_amd import adaptive
        p, m = problems.blast2d_two_level()
        fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl, amr_mesh=m)
        fd.initial_dt()
        for _ in range(3):
            fd.step()
        torch.cuda.synchronize()
        del fd
        pkg = hydro.initialize_hydro(problems.blast_deck_pin())
        m, bs, cons = adaptive.initial_mesh(
            problems.blast_deck_base_mesh(), problems.blast_pgen(),
            lambda mm, bb, cc: adaptive.device_tags(ctx, pkg, mm, bb, cc))
        drv = adaptive.AdaptiveDriver(ctx, pkg, m, bs, cons, form="fused")
        for _ in range(3):
            drv.fd.step()
        torch.cuda.synchronize()
        print(f"profile-run {args.profile_run}: done", flush=True)
        return
    if args.profile_run == "graph":
        # the timed configuration: the headline workload's RK2 steps replayed from a
        # captured hipGraph (5 steps per graph, replayed three times: 30 stage launches, the
        # first 10 a warm-up), with
        # the tile the autotuner picks for 32^3 blocks set directly, so no tuning
        # launches enter the profile
        p = problems.blast3d(nblk=args.nblk, n=32)
        fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl)
        fd.set_tile(32, 8, 8, 8)
        fd.initial_dt()
        fd.capture(5)
        for _ in range(3):  # the first replay warms clocks and caches (read apart)
            fd.replay()
        torch.cuda.synchronize()
        _calibration_copies(torch)
        print(f"profile-run {args.profile_run}: done", flush=True)
        return
    if args.profile_run == "monopole":
        # Evrard 256^3 (512 blocks of 32^3) with monopole gravity: 3 RK2 steps, each stage
        # preceded by the solver's zero / bound / accumulate / tables / field kernels
        p = problems.evrard(res=256, block=32)
        fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl,
                               tlim=p.tlim, monopole=p.monopole)
        fd.initial_dt()
        for _ in range(3):
            fd.step()
        torch.cuda.synchronize()
        _calibration_copies(torch)
        print(f"profile-run {args.profile_run}: done", flush=True)
        return
    if args.profile_run == "poisson_smr":
        # Evrard's smr_16 (120 blocks of 8^3 over levels 1 and 2) with the deck's Poisson
        # problem on the multilevel mesh: 3 RK2 steps in the fused form, each stage preceded
        # by the solver's moments, right-hand side, BiCGSTAB iterations and field
        p = problems.evrard(res=16, block=8, smr_level=2, solver="poisson")
        fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl,
                               tlim=p.tlim, poisson=p.poisson, amr_mesh=p.meta["mesh"])
        fd.initial_dt()
        for _ in range(3):
            fd.step()
        torch.cuda.synchronize()
        print(f"profile-run {args.profile_run}: done", flush=True)
        return
    if args.profile_run == "occupancy":
        # the headline workload (256^3 blast, 8x8x8 tiles, 64 lanes) with more LDS reserved
        # per workgroup, so fewer workgroups are resident per CU: the time per stage launch
        # at each occupancy, from 10 graph-replayed RK2 steps (after one warm-up replay)
        p = problems.blast3d(nblk=args.nblk, n=32)
        cells = p.bs.interior_cells()
        fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl)
        fd.set_tile(32, 8, 8, 8)
        fd.set_workgroup(64)
        fd.initial_dt()
        u_init = fd.u0.clone()
        for reserve in (0, 32 << 10, 40 << 10, 64 << 10, 96 << 10):
            fd.u0.copy_(u_init)
            fd.time.zero_()
            fd.initial_dt()
            fd.set_lds_reserve(reserve)
            fd.capture(10)
            fd.replay()
            torch.cuda.synchronize()
            fd.u0.copy_(u_init)
            fd.time.zero_()
            fd.initial_dt()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fd.replay()
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            wg_per_cu = min(8, (160 << 10) // max(reserve, 20 << 10))
            print(f"occupancy: LDS {max(reserve, 20 << 10) >> 10} KiB per workgroup, "
                  f"{wg_per_cu} workgroups per CU ({wg_per_cu / 4:g} waves per SIMD): "
                  f"{el / 20 * 1e3:.3f} ms per stage, {cells * 10 / el:.4e} zone-cycles/s",
                  flush=True)
        print(f"profile-run {args.profile_run}: done", flush=True)
        return
    if args.profile_run == "autotune":
        autotune_tables(ctx, hydro, problems, torch)
        print(f"profile-run {args.profile_run}: done", flush=True)
        return
    if args.profile_run == "uneven":
        p = problems.uneven_blocks(nblocks=args.uneven_blocks)
    else:
        p = problems.blast3d(nblk=args.nblk, n=32)
    if args.profile_run == "plain":
        pkg = hydro.initialize_hydro({
            "parthenon/time": {"cfl": p.cfl, "integrator": "rk2"},
            "parthenon/mesh": {"nghost": 3},
            "hydro": {"reconstruction": p.recon, "riemann": p.riemann},
            "eos": {"gm1": float(p.eos["gm1"][0])}})
        drv = hydro.ApophisDriver(pkg, hydro.MeshData(ctx, p.bs, p.cons))
        dt = drv.initialize()
        for _ in range(3):
            dt = drv.step(dt)
    else:
        fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl)
        fd.initial_dt()
        for _ in range(3):
            fd.step()
    torch.cuda.synchronize()
    _calibration_copies(torch)
    print(f"profile-run {args.profile_run}: done", flush=True)


def _launch_desc(fd):
    """{block shape: "wg<lanes> seg <x>/<y>/<z>"} of a fused driver's plan."""
    return {"x".join(map(str, k)): f"wg{v[0]} seg {v[1]}/{v[2]}/{v[3]}"
            for k, v in fd.launch().items()}


def autotune_tables(ctx, hydro, problems, torch):
    """The autotuner's timing tables for the 256^3 blast cut into blocks of 8^3, 16^3 and
    32^3 (32768, 4096 and 512 blocks) and for the uneven 8^3/16^3/32^3 pack, each followed
    by the chosen launch's zone-cycles/s over 10 graph-replayed RK2 steps."""
    cases = [(f"blast3d 256^3, blocks of {n}^3", lambda n=n: problems.blast3d(nblk=256 // n,
                                                                               n=n))
             for n in (8, 16, 32)]
    cases.append(("uneven 8^3/16^3/32^3 pack, 4096 blocks",
                  lambda: problems.uneven_blocks(nblocks=4096)))
    for name, make in cases:
        p = make()
        cells = p.bs.interior_cells()
        fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl)
        fd.initial_dt()
        t0 = time.perf_counter()
        report = fd.autotune(reps=5)
        t_tune = time.perf_counter() - t0
        for _ in range(2):
            fd.step()
        fd.capture(10)
        torch.cuda.synchronize()
        fd.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fd.replay()
        torch.cuda.synchronize()
        zc = cells * 10 / (time.perf_counter() - t0)
        print(f"## {name}: {p.bs.nblocks} blocks, {cells} cells; tuning {t_tune:.1f} s; "
              f"chosen launch {_launch_desc(fd)}: {zc:.4e} zone-cycles/s", flush=True)
        print(report, flush=True)
        del fd, p
        torch.cuda.empty_cache()


def _calibration_copies(torch):
    """Three 1 GiB device-to-device copies (2^30 B read + 2^30 B written each), whose
    FETCH_SIZE / WRITE_SIZE readings calibrate the counters."""
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nblk", type=int, default=8, help="blocks per axis (8 -> 256^3)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip plain/uneven legs")
    ap.add_argument("--uneven-blocks", type=int, default=4096)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the KH adjoint leg")
    ap.add_argument("--no-rt", action="store_true", help="skip the batched RT adjoint leg")
    ap.add_argument("--rt-instances", type=int, default=8,
                    help="RT instances of --profile-run rt_adjoint")
    ap.add_argument("--profile-run", choices=["fused", "graph", "plain", "uneven", "autotune",
                                              "occupancy", "monopole", "adjoint",
                                              "rt_adjoint", "amr", "poisson_smr"],
                    default=None,
                    help="short eager run for rocprofv3 (no timing, no JSON)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design, not built)", file=sys.stderr)
        sys.exit(2)

    import torch

    from apophis_amd import hydro, problems

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx = hydro._lib.Context(stream.cuda_stream)
    if args.profile_run:
        profile_run(args, ctx, hydro, problems, torch)
        return

    # ------------------------------------------------------------ headline: fused
    t_setup = time.perf_counter()
    p = problems.blast3d(nblk=args.nblk, n=32)
    cells = p.bs.interior_cells()
    fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl)
    fd.initial_dt()
    t_tune = time.perf_counter()
    report = fd.autotune(reps=3)
    tiles = {"x".join(map(str, k)): "x".join(map(str, v)) for k, v in fd.tiles().items()}
    launch = _launch_desc(fd)
    print(f"# autotune {time.perf_counter() - t_tune:.1f}s: {tiles} {launch}", file=sys.stderr,
          flush=True)
    for _ in range(args.warmup):
        fd.step()
    fd.capture(args.steps)
    torch.cuda.synchronize()
    print(f"# setup {time.perf_counter() - t_setup:.1f}s, {cells} cells, "
          f"{p.bs.nblocks} blocks", file=sys.stderr, flush=True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    fd.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = cells * args.steps / el
    u = fd.u0.cpu().numpy()
    finite = bool(np.isfinite(u[np.tile(p.bs.interior_mask(), p.bs.nvar)]).all())
    # roofline of the fused stage kernel: HIP events on the launch stream around the timed
    # graph replay; a step is two stage launches plus two one-thread kernels, so the
    # replay time over the stage launches bounds the kernel's average duration from above
    avg_ms = e0.elapsed_time(e1) / (2 * args.steps)
    eager_ms, _ = event_time_stage(fd, torch)
    sb = stage_bytes(p.bs, 3)
    achieved = sb / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(),
            "kernel": "k_fused<PPM,HLLC,NV=5,3D>", "avg_launch_ms": round(avg_ms, 4),
            "avg_launch_ms_source": "graph replay events / stage launches",
            "eager_launch_ms": round(eager_ms, 4),
            "algorithmic_bytes_per_launch": sb}
    vpc = profiled_valu_per_cell()
    # the bound that applies: FP64 VALU issue, the profiled instructions per cell at the
    # event-timed launch duration
    valu_roof = None if vpc is None else {
        "bound": "valu_fp64", "unit": "lane-instr/s",
        "achieved": vpc * cells / (avg_ms * 1e-3), "peak": VALU_PEAK_LANE_INSTR,
        "frac": round(vpc * cells / (avg_ms * 1e-3) / VALU_PEAK_LANE_INSTR, 3),
        "lane_instr_per_cell_stage": vpc}
    print(f"# fused: {value:.4e} zone-cycles/s, stage {avg_ms:.3f} ms, "
          f"{achieved:.0f} GB/s", file=sys.stderr, flush=True)
    del fd
    torch.cuda.empty_cache()

    extras = {}
    if not args.no_extras:
        # plain form on the same workload (reference task structure, HBM round trips)
        pkg = hydro.initialize_hydro({
            "parthenon/time": {"cfl": p.cfl, "integrator": "rk2"},
            "parthenon/mesh": {"nghost": 3},
            "hydro": {"reconstruction": "ppm", "riemann": "hllc"},
            "eos": {"gm1": float(p.eos["gm1"][0])}})
        md = hydro.MeshData(ctx, p.bs, p.cons)
        drv = hydro.ApophisDriver(pkg, md)
        dt = drv.initialize()
        drv.step(dt)
        torch.cuda.synchronize()
        ns = max(3, args.steps // 10)
        t0 = time.perf_counter()
        for _ in range(ns):
            dt = drv.step(dt)
        torch.cuda.synchronize()
        extras["plain_form"] = {"value": cells * ns / (time.perf_counter() - t0),
                                "steps": ns, "note": "per-step dt read back to host"}
        print(f"# plain: {extras['plain_form']['value']:.4e}", file=sys.stderr, flush=True)
        del drv, md
        torch.cuda.empty_cache()
        # uneven AMR-like blocks: batched fused vs per-block plain launches
        q = problems.uneven_blocks(nblocks=args.uneven_blocks)
        qc = q.bs.interior_cells()
        fq = hydro.FusedDriver(ctx, q.bs, q.cons, q.recon, q.riemann, q.eos, q.cfl)
        fq.initial_dt()
        fq.autotune(reps=3)
        uneven_tiles = {"x".join(map(str, k)): "x".join(map(str, v))
                        for k, v in fq.tiles().items()}
        uneven_launch = _launch_desc(fq)
        for _ in range(2):
            fq.step()
        nq = 20
        fq.capture(nq)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fq.replay()
        torch.cuda.synchronize()
        vb = qc * nq / (time.perf_counter() - t0)
        del fq
        pkg = hydro.initialize_hydro({
            "parthenon/time": {"cfl": q.cfl, "integrator": "rk2"},
            "parthenon/mesh": {"nghost": 4},
            "hydro": {"reconstruction": "ppm", "riemann": "hllc"},
            "eos": {"gm1": float(q.eos["gm1"][0])}})
        mq = hydro.MeshData(ctx, q.bs, q.cons)
        dq = hydro.ApophisDriver(pkg, mq, pack_size=1)
        dt = dq.initialize()
        dq.step(dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        npl = 2
        for _ in range(npl):
            dt = dq.step(dt)
        torch.cuda.synchronize()
        vp = qc * npl / (time.perf_counter() - t0)
        del dq, mq
        # the plain form with one pack per block size (the reference's packs need blocks
        # of one shape): the same blocks ordered by size, three partitions
        qs = problems.uneven_blocks(nblocks=args.uneven_blocks, by_size=True)
        ed = np.asarray(qs.meta["edges"])
        cuts = [0] + [int(i) + 1 for i in np.flatnonzero(np.diff(ed))] + [len(ed)]
        runs = [(a, b - a) for a, b in zip(cuts, cuts[1:])]
        ms = hydro.MeshData(ctx, qs.bs, qs.cons)
        ds = hydro.ApophisDriver(pkg, ms, partitions=runs)
        dt = ds.initialize()
        ds.step(dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(npl):
            dt = ds.step(dt)
        torch.cuda.synchronize()
        vs = qc * npl / (time.perf_counter() - t0)
        extras["uneven_blocks"] = {"blocks": q.bs.nblocks, "cells": qc, "tiles": uneven_tiles,
                                   "launch": uneven_launch,
                                   "batched_value": vb, "plain_per_block_value": vp,
                                   "batched_over_plain": vb / vp,
                                   "plain_size_class_packs": {str(int(ed[a])): n
                                                              for a, n in runs},
                                   "plain_per_size_class_value": vs,
                                   "batched_over_plain_per_size_class": vb / vs}
        print(f"# uneven: batched {vb:.4e}, plain per-block {vp:.4e}, plain per size class "
              f"{vs:.4e}", file=sys.stderr, flush=True)
        del ds, ms

    if not args.no_extras:
        extras["evrard_monopole"] = monopole_leg(ctx, hydro, problems, torch)
        print(f"# monopole: {extras['evrard_monopole']}", file=sys.stderr, flush=True)
        extras["two_level_blast2d"] = two_level_leg(ctx, hydro, problems, torch)
        print(f"# two-level: {extras['two_level_blast2d']}", file=sys.stderr, flush=True)
        extras["kh_comp_ONe_lset"] = composition_leg(ctx, hydro, problems, torch)
        print(f"# composition: {extras['kh_comp_ONe_lset']}", file=sys.stderr, flush=True)
        extras["blast_deck_adaptive"] = blast_deck_leg(ctx, hydro, problems, torch)
        print(f"# blast deck: {extras['blast_deck_adaptive']}", file=sys.stderr, flush=True)
        extras["evrard_smr_monopole"] = evrard_smr_leg(ctx, hydro, problems, torch)
        print(f"# evrard smr: {extras['evrard_smr_monopole']}", file=sys.stderr, flush=True)
        extras["evrard_smr16_poisson"] = evrard_smr_leg(ctx, hydro, problems, torch, "poisson")
        print(f"# evrard smr poisson: {extras['evrard_smr16_poisson']}", file=sys.stderr,
              flush=True)
    if not args.no_adjoint:
        extras["adjoint_kh"] = adjoint_leg(ctx, hydro, problems, torch)
        print(f"# adjoint: {extras['adjoint_kh']}", file=sys.stderr, flush=True)
        extras["adjoint_two_level_blast2d"] = multilevel_adjoint_leg(ctx, hydro, problems,
                                                                     torch)
        print(f"# multilevel adjoint: {extras['adjoint_two_level_blast2d']}", file=sys.stderr,
              flush=True)
    if not args.no_rt:
        extras["rt_batch_adjoint"] = rt_batch_leg(ctx, hydro, problems, torch)
        print(f"# rt batch: {extras['rt_batch_adjoint']}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    cpu = None if args.no_cpu else cpu_baseline()
    out = {
        "metric": "zone-cycles/s + % HBM roofline on 1 GPU (batched vs plain, FP64); "
                  "adjoint/forward time",
        "value": value, "unit": "zone-cycles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded blast wave)",
        "config": {"workload": f"blast3d {args.nblk * 32}^3 periodic, {p.bs.nblocks} blocks "
                               "of 32^3, PPM+HLLC, RK2, CFL 0.3, gamma 5/3",
                   "form": "batched fused, hipGraph-replayed, device dt",
                   "tiles": tiles, "launch": launch, "finite": finite},
        "roofline": roof,
        "valu_roofline": valu_roof,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
