"""Multilevel meshes (two and three levels) on the GPU through the C ABI: aph_amr_fill / aph_amr_flux_correct
against the oracle's whole-level restatement (oracle/amr_oracle.py) bitwise, whole plain
steps with the flux correction against the oracle's step bitwise, conservation, and the
error paths (a corrupt plan, the fused form on a multilevel set)."""
import numpy as np
import pytest
import torch

from apophis_amd import amr, hydro
from oracle import amr_oracle as ao
from oracle import oracle as orc
from tests.test_amr import MESHES, _mesh, _problem, _totals

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(MESHES))
def test_amr_fill_and_flux_correct_match_oracle(gpu_ctx, name):
    print(f"[case] amr fill {name}", flush=True)
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
    torch.cuda.synchronize()
    assert np.array_equal(md.cons.cpu().numpy(), want)
    fx = [np.random.default_rng(9 + d).normal(size=bs.field_size(5)) for d in range(3)]
    for d in range(3):
        if md.flux[d] is not None:
            md.flux[d].copy_(torch.from_numpy(fx[d]))
    plan.flux_correct(md.flux)
    ao.flux_correct(m, bs, fx)
    torch.cuda.synchronize()
    for d in range(m.ndim):
        assert np.array_equal(md.flux[d].cpu().numpy(), fx[d]), d


@pytest.mark.parametrize("name,recon,riemann", [("2d_L", "plm", "hllc"),
                                                ("2d_ppm", "ppm", "hllc"),
                                                ("3d", "plm", "hlle"),
                                                ("2d_3lev", "plm", "hllc"),
                                                ("2d_3lev_ppm", "ppm", "hllc"),
                                                ("3d_3lev", "plm", "hlle")])
def test_plain_amr_steps_match_oracle(gpu_ctx, name, recon, riemann):
    print(f"[case] amr steps {name} {recon} {riemann}", flush=True)
    m = _mesh(name)
    p = _problem(m, recon=recon)
    p.riemann = riemann
    ref = ao.AmrState(p, m)
    pkg = hydro.initialize_hydro({
        "parthenon/time": {"cfl": p.cfl, "integrator": "rk2"},
        "parthenon/mesh": {"nghost": m.ng},
        "hydro": {"reconstruction": recon, "riemann": riemann},
        "eos": {"gm1": float(p.eos["gm1"][0])}})
    md = hydro.MeshData(gpu_ctx, p.bs, p.cons)
    plan = amr.AmrPlan(gpu_ctx, md.dbs, m, p.bs)
    drv = hydro.ApophisDriver(pkg, md, amr=plan)
    dt = drv.initialize()
    assert dt == ref.estimate_dt()
    t0 = _totals(p.bs, md.cons.cpu().numpy())
    for _ in range(2):
        est = drv.step(dt)
        est_ref = ref.step(dt)
        assert est == est_ref
    torch.cuda.synchronize()
    got = md.cons.cpu().numpy()
    assert np.array_equal(got, ref.u0)
    t1 = _totals(p.bs, got)
    assert np.all(np.abs(t1 - t0) <= 1e-13 * (np.abs(t0) + 1.0))


def test_amr_error_paths(gpu_ctx):
    m = _mesh("2d_L")
    bs = m.block_set()
    md = hydro.MeshData(gpu_ctx, bs, np.ones(bs.field_size(5)))
    ops, fixes, refs = amr.amr_plan(m, bs)
    bad = ops.copy()
    bad["dst"][0] = refs["offset"][0]  # an interior cell
    import ctypes
    from apophis_amd import _lib
    h = ctypes.c_void_p()
    vp = ctypes.c_void_p
    rc = _lib.lib().aph_amr_plan_create(gpu_ctx.h, md.dbs, bad.ctypes.data_as(vp), len(bad),
                                        fixes.ctypes.data_as(vp), len(fixes),
                                        refs.ctypes.data_as(vp), len(refs), ctypes.byref(h))
    assert rc != 0 and not h.value
    bad = ops.copy()
    bad["val"]["count"][0, 0] = 9
    rc = _lib.lib().aph_amr_plan_create(gpu_ctx.h, md.dbs, bad.ctypes.data_as(vp), len(bad),
                                        fixes.ctypes.data_as(vp), len(fixes),
                                        refs.ctypes.data_as(vp), len(refs), ctypes.byref(h))
    assert rc != 0
    # a face corrected twice, and a corrected face that is another fix's source
    def create(fx, rf):
        return _lib.lib().aph_amr_plan_create(gpu_ctx.h, md.dbs, ops.ctypes.data_as(vp),
                                              len(ops), fx.ctypes.data_as(vp), len(fx),
                                              rf.ctypes.data_as(vp), len(rf), ctypes.byref(h))
    assert create(np.concatenate([fixes, fixes[:1]]), refs) != 0
    assert "twice" in _lib.lib().aph_last_error(gpu_ctx.h).decode()
    j = int(np.flatnonzero(fixes["dir"] == fixes["dir"][0])[1])
    rf = np.concatenate([refs, np.array([(fixes["dst"][0], fixes["dst_stride"][0])],
                                        dtype=refs.dtype)])
    bad = fixes.copy()
    bad["val"]["first"][j], bad["val"]["count"][j] = len(refs), 1
    assert create(bad, rf) != 0
    assert "both corrected and a source" in _lib.lib().aph_last_error(gpu_ctx.h).decode()
    p = _problem(m)
    fd = hydro.FusedDriver(gpu_ctx, p.bs, p.cons, "plm", "hllc", p.eos, p.cfl)
    with pytest.raises(hydro._lib.HydroError, match="aph_plan_set_amr"):
        fd.step(fixed_dt=True)
    # an interface table whose fine side does not sit one level below its coarse region
    import ctypes as ct
    m3 = _mesh("2d_3lev")
    p3 = _problem(m3)
    f3 = hydro.FusedDriver(gpu_ctx, p3.bs, p3.cons, "plm", "hllc", p3.eos, p3.cfl)
    plan3 = amr.AmrPlan(gpu_ctx, f3.device_bs.h, m3, p3.bs)
    faces, level, nbuf = amr.fused_faces(m3, p3.bs)
    args = lambda fc, lv, n: (gpu_ctx.h, f3.plan, plan3.h, fc.ctypes.data_as(ct.c_void_p),
                              lv.ctypes.data_as(ct.c_void_p), ct.c_int64(n))
    assert _lib.lib().aph_plan_set_amr(*args(faces, level, nbuf)) == 0
    bad = level.copy()
    bad[bad == 2] = 0
    assert _lib.lib().aph_plan_set_amr(*args(faces, bad, nbuf)) != 0
    assert "one level coarser" in _lib.lib().aph_last_error(gpu_ctx.h).decode()
    bad = faces.copy()
    r1 = np.flatnonzero(bad["role"] == 1)[0]
    bad["qa"][r1] = bad["na"][r1] + 2  # past the end of its coarse region
    assert _lib.lib().aph_plan_set_amr(*args(bad, level, nbuf)) != 0
    assert _lib.lib().aph_plan_set_amr(*args(faces, level, nbuf - 1)) != 0


@pytest.mark.parametrize("name,recon,riemann", [("2d_L", "plm", "hllc"),
                                                ("2d_ppm", "ppm", "hllc"),
                                                ("3d", "plm", "hlle"),
                                                ("2d_3lev", "plm", "hllc"),
                                                ("2d_3lev_ppm", "ppm", "hllc"),
                                                ("3d_3lev", "plm", "hlle")])
def test_fused_amr_steps_match_oracle(gpu_ctx, name, recon, riemann):
    """The batched form on a multilevel mesh: cross-level ghosts from aph_amr_fill, one
    launch per level, the finest first, coarse faces corrected from the interface buffer
    -- the same bits as the plain form's step with aph_amr_flux_correct."""
    print(f"[case] fused amr steps {name} {recon} {riemann}", flush=True)
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
    got = fd.u0.cpu().numpy()
    mask = np.tile(p.bs.interior_mask(), 5)
    assert np.array_equal(got[mask], ref.u0[mask])


def test_fused_amr_wait_that_runs_out_is_reported(gpu_ctx):
    """A multilevel wait that gives up poisons the stage output, dt and time with NaN, and
    aph_plan_status reports it once, then clears it.  The failure is forced from the host
    (a negative wait limit: every wait gives up without polling), so every kernel still
    exits normally.  With the default limit restored the steps equal the oracle's again."""
    print("[case] fused amr wait reported", flush=True)
    m = _mesh("2d_L")
    p = _problem(m, recon="plm")
    p.riemann = "hllc"
    ref = ao.AmrState(p, m)
    dt = ref.estimate_dt()
    fd = hydro.FusedDriver(gpu_ctx, p.bs, p.cons, "plm", "hllc", p.eos, p.cfl, amr_mesh=m)
    fd.status()  # nothing to report yet
    fd.set_dt(dt)
    fd.set_wait_limit(-1)
    fd.step(fixed_dt=False)
    torch.cuda.synchronize()
    assert np.isnan(fd.dt.item()) and np.isnan(fd.time.item())
    assert np.isnan(fd.u0[0].item())
    with pytest.raises(hydro._lib.HydroError, match="ran out"):
        fd.status()
    fd.status()  # reported once, then cleared
    fd.set_wait_limit(0)
    fd.u0.copy_(torch.from_numpy(p.cons))
    fd.time.zero_()
    fd.set_dt(dt)
    for _ in range(2):
        fd.step(fixed_dt=True)
        ref.step(dt)
    fd.status()
    got = fd.u0.cpu().numpy()
    mask = np.tile(p.bs.interior_mask(), 5)
    assert np.array_equal(got[mask], ref.u0[mask])


@pytest.mark.parametrize("ndim", [2, 3])
def test_refinement_tags_match_oracle(gpu_ctx, ndim):
    """aph_refinement_tag on an evolved blast state (prim with ghosts from the plain form)
    against the oracle's restatement of the three criteria: the same reduced values,
    bitwise, and the same tags, with thresholds that make blocks refine, stay and
    derefine."""
    from apophis_amd import problems
    print(f"[case] refinement tags {ndim}d", flush=True)
    if ndim == 3:
        p = problems.blast3d(nblk=4, n=8, recon="plm")
    else:
        p, _ = problems.blast2d_two_level(nbase=8, n=16, nref=0, recon="plm")
    ref = orc.State(p)
    dt = ref.estimate_dt()
    for _ in range(6):
        dt = ref.step(dt)
    pin = {"parthenon/time": {"cfl": 0.3, "integrator": "rk2"},
           "parthenon/mesh": {"nghost": int(p.bs.blocks["ng"].min())},
           "hydro": {"reconstruction": p.recon, "riemann": p.riemann},
           "eos": {"gm1": float(p.eos["gm1"][0])}}
    md = hydro.MeshData(gpu_ctx, p.bs, ref.u0)
    hydro.fill_ghosts(md)
    hydro.cons_to_prim(md, p.eos)
    vals0 = {c: ao.refinement_tag(p.bs, ref.prim, c, 1.0, 2.0)[1] for c in (1, 2, 3)}
    q = lambda c, f: float(np.quantile(vals0[c], f))
    for refine in ({"type": "pressure_gradient", "threshold_pressure_gradient": q(1, 0.7)},
                   {"type": "xyvelocity_gradient",
                    "threshold_xyvelocity_gradient": q(2, 0.7)},
                   {"type": "maxdensity", "maxdensity_deref_below": q(3, 0.3),
                    "maxdensity_refine_above": q(3, 0.7)}):
        pkg = hydro.initialize_hydro(dict(pin, refinement=refine))
        tags, vals = hydro.refinement_tag(md, pkg)
        crit, a, b = pkg.param("refinement/criterion")
        want_t, want_v = ao.refinement_tag(p.bs, ref.prim, crit, a, b)
        torch.cuda.synchronize()
        assert np.array_equal(vals.cpu().numpy(), want_v), refine["type"]
        assert np.array_equal(tags.cpu().numpy(), want_t), refine["type"]
        assert len(set(want_t.tolist())) >= 2, (refine["type"], want_t)


def test_adaptive_cycle_matches_oracle(gpu_ctx):
    """Three cycles of: two plain steps on the two-level mesh, aph_refinement_tag on the
    device, regrid on the host, new plans -- the same bits as the oracle's cycle."""
    from apophis_amd import problems
    print("[case] adaptive cycle", flush=True)
    p, m = problems.blast2d_two_level(nbase=6, n=8, nref=2, recon="plm")
    ref_p, ref_m = p, m
    ref = ao.AmrState(ref_p, ref_m)
    pin = {"parthenon/time": {"cfl": p.cfl, "integrator": "rk2"},
           "parthenon/mesh": {"nghost": m.ng},
           "hydro": {"reconstruction": p.recon, "riemann": p.riemann},
           "eos": {"gm1": float(p.eos["gm1"][0])},
           "refinement": {"type": "pressure_gradient", "threshold_pressure_gradient": 0.5}}
    pkg = hydro.initialize_hydro(pin)
    cons = p.cons
    dt = ref.estimate_dt()
    for cycle in range(3):
        md = hydro.MeshData(gpu_ctx, p.bs, cons)
        plan = amr.AmrPlan(gpu_ctx, md.dbs, m, p.bs)
        drv = hydro.ApophisDriver(pkg, md, amr=plan)
        assert drv.initialize() == ref.estimate_dt()
        for _ in range(2):
            est = drv.step(dt)
            assert est == ref.step(dt)
            dt = min(dt, est)
        tags, _ = hydro.refinement_tag(md, pkg)
        want, _ = ao.refinement_tag(ref_p.bs, ref.prim, 1, 0.5, 0.0)
        got_u = md.cons.cpu().numpy()
        assert np.array_equal(got_u, ref.u0)
        assert np.array_equal(tags.cpu().numpy(), want)
        m, nbs, cons = amr.regrid(m, p.bs, got_u, tags.cpu().numpy())
        p = problems.Problem("adaptive", nbs, cons, p.eos, p.recon, p.riemann, p.cfl)
        ref_m, rbs, rc = amr.regrid(ref_m, ref_p.bs, ref.u0, want)
        ref_p = problems.Problem("adaptive", rbs, rc, p.eos, p.recon, p.riemann, p.cfl)
        ref = ao.AmrState(ref_p, ref_m)
        del drv, plan, md
    assert (m.level == 1).sum() > 0


@pytest.mark.parametrize("form", ["plain", "fused"])
def test_three_level_steps_with_a_passive_scalar(gpu_ctx, form):
    """NV = 6 (one passive scalar) on the three-level mesh: plain and fused steps equal
    the oracle's bitwise, and the scalar total is conserved across the level jumps."""
    print(f"[case] 3-level scalar {form}", flush=True)
    m = _mesh("2d_3lev")
    bs = m.block_set(nscalars=1)
    base = _problem(m)
    u = np.zeros(bs.field_size(6))
    for b in range(bs.nblocks):
        v = bs.interior(u, b, 6)
        v[:5] = base.bs.interior(base.cons, b, 5)
        x, y, _ = bs.cell_centers(b)
        v[5] = v[0] * (0.5 + 0.4 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y))
    from apophis_amd import problems
    p = problems.Problem("amr_scalar", bs, u, base.eos, "plm", "hllc", 0.3)
    ref = ao.AmrState(p, m)
    dt = ref.estimate_dt()
    tot = lambda arr: sum(bs.interior(arr, b, 6)[5].sum() * bs.blocks["vol"][b]
                          for b in range(bs.nblocks))
    s0 = tot(u)
    if form == "plain":
        pkg = hydro.initialize_hydro({
            "parthenon/time": {"cfl": p.cfl, "integrator": "rk2"},
            "parthenon/mesh": {"nghost": m.ng},
            "hydro": {"reconstruction": "plm", "riemann": "hllc", "nscalars": 1},
            "eos": {"gm1": float(p.eos["gm1"][0])}})
        md = hydro.MeshData(gpu_ctx, bs, u)
        drv = hydro.ApophisDriver(pkg, md, amr=amr.AmrPlan(gpu_ctx, md.dbs, m, bs))
        drv.initialize()
        for _ in range(2):
            drv.step(dt)
            ref.step(dt)
        got = md.cons.cpu().numpy()
    else:
        fd = hydro.FusedDriver(gpu_ctx, bs, u, "plm", "hllc", p.eos, p.cfl, amr_mesh=m)
        fd.set_dt(dt)
        for _ in range(2):
            fd.step(fixed_dt=True)
            ref.step(dt)
        got = fd.u0.cpu().numpy()
    torch.cuda.synchronize()
    mask = np.tile(bs.interior_mask(), 6)
    assert np.array_equal(got[mask], ref.u0[mask])
    assert abs(tot(ref.u0) - s0) <= 1e-13 * abs(s0)


@pytest.mark.parametrize("ndim", [2, 3])
def test_fused_multilevel_with_several_tiles_per_face(gpu_ctx, ndim):
    """Three levels with blocks larger than the tile (8-cell tiles: 4 tiles per block face),
    so every interface region is written by several fine tiles and read by several coarse
    tiles inside the one launch: the fused steps equal the plain form's bitwise."""
    print(f"[case] multilevel tiles {ndim}d", flush=True)
    if ndim == 2:
        nodes = [(0, (x, y)) for x in (1, 2) for y in (1, 2)] + [(1, (3, 3)), (1, (4, 3))]
        m = amr.AmrMesh(2, (4, 4), (32, 32), 2, (0, 0, 0), (1, 1, 1), nodes, max_level=2)
        tile = (32, 8, 8, 1)
    else:
        nodes = [(0, (x, y, z)) for x in (1, 2) for y in (1, 2) for z in (0, 1)] + \
            [(1, (3, 3, 1))]
        m = amr.AmrMesh(3, (4, 4, 2), (16, 16, 16), 2, (0, 0, 0), (1, 1, 1), nodes,
                        max_level=2)
        tile = (16, 8, 8, 8)
    p = _problem(m)
    ref = ao.AmrState(p, m) if ndim == 2 else None
    pkg = hydro.initialize_hydro({
        "parthenon/time": {"cfl": p.cfl, "integrator": "rk2"},
        "parthenon/mesh": {"nghost": m.ng},
        "hydro": {"reconstruction": "plm", "riemann": "hllc"},
        "eos": {"gm1": float(p.eos["gm1"][0])}})
    md = hydro.MeshData(gpu_ctx, p.bs, p.cons)
    drv = hydro.ApophisDriver(pkg, md, amr=amr.AmrPlan(gpu_ctx, md.dbs, m, p.bs))
    dt = drv.initialize()
    fd = hydro.FusedDriver(gpu_ctx, p.bs, p.cons, "plm", "hllc", p.eos, p.cfl, amr_mesh=m)
    fd.set_tile(*tile)
    assert fd.tiles()[tuple(int(v) for v in p.bs.blocks["nx"][0])] == tile[1:]
    fd.set_dt(dt)
    for _ in range(2):
        drv.step(dt)
        fd.step(fixed_dt=True)
        if ref is not None:
            ref.step(dt)
    torch.cuda.synchronize()
    mask = np.tile(p.bs.interior_mask(), 5)
    plain = md.cons.cpu().numpy()
    assert np.array_equal(fd.u0.cpu().numpy()[mask], plain[mask])
    assert np.isfinite(plain[mask]).all()
    if ref is not None:
        assert np.array_equal(plain, ref.u0)


def test_autotuned_multilevel_plan_matches_plain(gpu_ctx):
    """aph_plan_autotune on a multilevel plan (every candidate tile rebuilds the item list
    and the interface counts) leaves a plan whose steps equal the plain form's bitwise."""
    print("[case] multilevel autotune", flush=True)
    nodes = [(0, (x, y)) for x in (1, 2) for y in (1, 2)] + [(1, (3, 3))]
    m = amr.AmrMesh(2, (4, 4), (32, 32), 2, (0, 0, 0), (1, 1, 1), nodes, max_level=2)
    p = _problem(m)
    pkg = hydro.initialize_hydro({
        "parthenon/time": {"cfl": p.cfl, "integrator": "rk2"},
        "parthenon/mesh": {"nghost": m.ng},
        "hydro": {"reconstruction": "plm", "riemann": "hllc"},
        "eos": {"gm1": float(p.eos["gm1"][0])}})
    md = hydro.MeshData(gpu_ctx, p.bs, p.cons)
    drv = hydro.ApophisDriver(pkg, md, amr=amr.AmrPlan(gpu_ctx, md.dbs, m, p.bs))
    dt = drv.initialize()
    fd = hydro.FusedDriver(gpu_ctx, p.bs, p.cons, "plm", "hllc", p.eos, p.cfl, amr_mesh=m)
    fd.set_dt(dt)
    assert fd.autotune(reps=2)  # a non-empty timing report
    fd.u0.copy_(torch.from_numpy(p.cons))  # the cross-level fill wrote u0's ghost cells
    for _ in range(2):
        drv.step(dt)
        fd.step(fixed_dt=True)
    torch.cuda.synchronize()
    mask = np.tile(p.bs.interior_mask(), 5)
    assert np.array_equal(fd.u0.cpu().numpy()[mask], md.cons.cpu().numpy()[mask])


def test_multilevel_graph_replays_match_eager(gpu_ctx):
    """A captured multilevel RK2 step sequence (cross-level fill, epoch bump, one launch
    over every level, timeout check) replayed three times, with eager steps and
    synchronisations between the replays: every replay equals the eager steps bitwise."""
    print("[case] multilevel graph replays", flush=True)
    from apophis_amd import _lib
    m = _mesh("2d_3lev")
    p = _problem(m)
    nsteps = 3
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        ctx = _lib.Context(st.cuda_stream)
        fd = hydro.FusedDriver(ctx, p.bs, p.cons, "plm", "hllc", p.eos, p.cfl, amr_mesh=m)
        dt = float(fd.initial_dt().item())
        u_init = torch.from_numpy(p.cons.copy()).cuda()
        mask = torch.from_numpy(np.tile(p.bs.interior_mask(), 5)).cuda()

        def eager():
            fd.u0.copy_(u_init)
            fd.set_dt(dt)
            for _ in range(nsteps):
                fd.step(fixed_dt=True)
            st.synchronize()
            return fd.u0[mask].clone()

        want = eager()
        fd.u0.copy_(u_init)
        fd.set_dt(dt)
        fd.capture(nsteps, fixed_dt=True)
        for rep in range(3):
            fd.u0.copy_(u_init)
            fd.set_dt(dt)
            st.synchronize()
            fd.replay()
            st.synchronize()
            got = fd.u0[mask].clone()
            assert torch.equal(got, want), rep
            assert torch.isfinite(got).all()
            assert torch.equal(eager(), want), rep  # eager work between the replays
        st.synchronize()
