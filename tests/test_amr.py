"""Multilevel meshes (2:1 jumps, two and three levels): the host plan (apophis_amd/amr.py) and the oracle's
restatement (oracle/amr_oracle.py) of Parthenon's multilevel exchange and flux correction.

The reference delegates both to Parthenon, whose source is not in the reference tree, so
what is pinned here is what the mathematics fixes:
* the plan, evaluated op by op, equals the oracle's whole-level evaluation bitwise;
* restriction and minmod prolongation reproduce a linear field exactly (to round-off) in
  every cross-level ghost cell whose stencil does not wrap around the periodic domain;
* a uniform state is a fixed point of a step, bitwise;
* with flux correction the totals of every conserved variable are kept to round-off
  across the level jump; without it they drift;
* every op writes a ghost cell and reads interior cells only (what the library checks
  before it accepts a plan), so the fill kernel has no read/write overlap.
"""
import numpy as np
import pytest

from apophis_amd import amr, problems
from apophis_amd.mesh import lattice_mesh, make_eos, BlockSet
from oracle import amr_oracle as ao
from oracle import oracle as orc

MESHES = {
    "2d_L": dict(ndim=2, nbase=(4, 4), block_nx=(8, 8), ng=2, refined=[(1, 1), (2, 1), (1, 2)]),
    "2d_ppm": dict(ndim=2, nbase=(4, 3), block_nx=(8, 8), ng=3, refined=[(1, 1), (2, 1)]),
    "3d": dict(ndim=3, nbase=(3, 3, 2), block_nx=(6, 6, 6), ng=2, refined=[(1, 1, 0)]),
    # three levels: a level-2 region inside a level-1 region (refined as (level, loc)),
    # with blocks that hold faces to both other levels
    "2d_3lev": dict(ndim=2, nbase=(4, 4), block_nx=(8, 8), ng=2,
                    nodes=[(0, (1, 1)), (0, (2, 1)), (0, (1, 2)), (0, (2, 2)),
                           (1, (3, 3)), (1, (4, 4))]),
    "2d_3lev_ppm": dict(ndim=2, nbase=(4, 4), block_nx=(6, 6), ng=3,
                        nodes=[(0, (1, 1)), (0, (2, 1)), (0, (1, 2)), (0, (2, 2)),
                               (1, (3, 4))]),
    "3d_3lev": dict(ndim=3, nbase=(4, 4, 2), block_nx=(4, 4, 4), ng=2,
                    nodes=[(0, (x, y, z)) for x in (1, 2) for y in (1, 2) for z in (0, 1)]
                    + [(1, (3, 3, 1))]),
}


def _mesh(name):
    k = MESHES[name]
    if "nodes" in k:
        return amr.AmrMesh(k["ndim"], k["nbase"], k["block_nx"], k["ng"], (0.0, 0.0, 0.0),
                           (1.0, 1.0, 1.0), k["nodes"], max_level=2)
    return amr.TwoLevelMesh(k["ndim"], k["nbase"], k["block_nx"], k["ng"], (0.0, 0.0, 0.0),
                            (1.0, 1.0, 1.0), k["refined"])


def _eval_plan(ops, refs, nd, u, nv):
    def val(v, comp):
        r = refs[v["first"]: v["first"] + v["count"]]
        s = u[r["offset"][0] + comp * r["stride"][0]]
        for k in range(1, len(r)):
            s = s + u[r["offset"][k] + comp * r["stride"][k]]
        return s if len(r) == 1 else s * (1.0 / len(r))

    sg = lambda x: -1.0 if x < 0 else 1.0
    out = []
    for op in ops:
        for c in range(nv):
            v0 = val(op["val"][0], c)
            o = v0
            if op["kind"] == amr.AMR_PROLONG:
                for d in range(nd):
                    a = v0 - val(op["val"][1 + 2 * d], c)
                    b = val(op["val"][2 + 2 * d], c) - v0
                    o = o + (op["sign"][d] * 0.25) * ((0.5 * (sg(a) + sg(b))) * min(abs(a), abs(b)))
            out.append((op["dst"] + c * op["dst_stride"], o))
    for i, o in out:
        u[i] = o


@pytest.mark.parametrize("name", list(MESHES))
def test_plan_equals_oracle_fill(name):
    m = _mesh(name)
    bs = m.block_set()
    u = np.random.default_rng(1).normal(size=bs.field_size(5))
    a, b = u.copy(), u.copy()
    orc.fill_ghosts(bs, a)
    ao.amr_fill(m, bs, a)
    orc.fill_ghosts(bs, b)
    ops, fixes, refs = amr.amr_plan(m, bs)
    _eval_plan(ops, refs, m.ndim, b, 5)
    assert np.array_equal(a, b)
    # every ghost cell is set: by the exchange or by one op, never both
    mark = np.zeros(bs.field_size(5), bool)
    mark[ops["dst"]] = True
    assert len(np.unique(ops["dst"])) == len(ops)
    # flux fixes: the plan's means against the oracle's whole-level face arrays
    fx = [np.random.default_rng(2 + d).normal(size=bs.field_size(5)) for d in range(3)]
    want = [f.copy() for f in fx]
    ao.flux_correct(m, bs, want)
    for fx_ in fixes:
        d = int(fx_["dir"])
        r = refs[fx_["val"]["first"]: fx_["val"]["first"] + fx_["val"]["count"]]
        for c in range(5):
            s = fx[d][r["offset"][0] + c * r["stride"][0]]
            for k in range(1, len(r)):
                s = s + fx[d][r["offset"][k] + c * r["stride"][k]]
            fx[d][fx_["dst"] + c * fx_["dst_stride"]] = s * (1.0 / len(r))
    for d in range(3):
        assert np.array_equal(fx[d], want[d]), d


@pytest.mark.parametrize("name", list(MESHES))
def test_linear_field_exact_across_levels(name):
    m = _mesh(name)
    bs = m.block_set()
    nd = m.ndim
    coef = np.array([[1.0, 0.3, -0.7, 0.2], [2.0, -1.1, 0.4, 0.9], [0.5, 0.0, 0.0, 1.3],
                     [-1.0, 0.25, 0.5, -0.5], [3.0, 1.0, 1.0, 1.0]])
    u = np.zeros(bs.field_size(5))
    f = lambda c, x, y, z: c[0] + c[1] * x + c[2] * y + (c[3] * z if nd == 3 else 0.0)
    for b in range(bs.nblocks):
        x, y, z = bs.cell_centers(b)
        v = bs.interior(u, b, 5)
        for n in range(5):
            v[n] = f(coef[n], x, y, z)
    orc.fill_ghosts(bs, u)
    ao.amr_fill(m, bs, u)
    ops, _, _ = amr.amr_plan(m, bs)
    dst = set(int(v) for v in ops["dst"])
    checked = 0
    ext = bs.extents()
    for b in range(bs.nblocks):
        r = bs.blocks[b]
        view = bs.block_view(u, b, 5)
        g = int(r["ng"])
        E = ext[b]
        for k in range(int(E[2])):
            for j in range(int(E[1])):
                for i in range(int(E[0])):
                    loc = (i, j, k)
                    off = 5 * int(r["cell_base"]) + (k * int(E[1]) + j) * int(E[0]) + i
                    if off not in dst:
                        continue
                    xc = [r["xmin"][d] + (loc[d] - g + 0.5) * r["dx"][d] if d < nd else 0.5
                          for d in range(3)]
                    # keep stencils (two coarse cells either side) off the periodic seam
                    h = 2.5 * 2 * max(r["dx"][:nd])
                    if any(xc[d] < h or xc[d] > 1.0 - h for d in range(nd)):
                        continue
                    for n in range(5):
                        want = f(coef[n], *xc)
                        assert view[n, k, j, i] == pytest.approx(want, rel=1e-13, abs=1e-13)
                    checked += 1
    assert checked > 20


def _problem(m, recon="plm", riemann="hllc", seed=3, amp=0.1, uniform=False):
    bs = m.block_set()
    u = np.zeros(bs.field_size(5))
    rng = np.random.default_rng(seed)
    kx, ky = 2 * np.pi, 4 * np.pi
    for b in range(bs.nblocks):
        x, y, z = bs.cell_centers(b)
        v = bs.interior(u, b, 5)
        rho = 1.0 if uniform else 1.0 + amp * np.sin(kx * x) * np.cos(ky * y)
        vx = 0.3 if uniform else 0.3 + amp * np.cos(kx * x + 1.0)
        vy = -0.2
        p = 1.0 if uniform else 1.0 + amp * np.sin(ky * y + 0.5)
        v[0] = rho
        v[1] = rho * vx
        v[2] = rho * vy
        v[3] = 0.0
        v[4] = p / 0.4 + 0.5 * rho * (vx * vx + vy * vy)
    return problems.Problem("amr", bs, u, make_eos(0.4), recon, riemann, 0.3)


def test_uniform_state_is_a_fixed_point():
    m = _mesh("2d_L")
    p = _problem(m, uniform=True)
    st = ao.AmrState(p, m)
    before = st.u0.copy()
    dt = st.estimate_dt()
    st.step(dt)
    mask = np.tile(p.bs.interior_mask(), 5)
    assert np.array_equal(st.u0[mask], before[mask])


def _totals(bs, u):
    t = np.zeros(5)
    for b in range(bs.nblocks):
        t += bs.interior(u, b, 5).reshape(5, -1).sum(axis=1) * bs.blocks["vol"][b]
    return t


@pytest.mark.parametrize("name,recon", [("2d_L", "plm"), ("2d_ppm", "ppm"), ("3d", "plm"),
                                        ("2d_3lev", "plm"), ("2d_3lev_ppm", "ppm"),
                                        ("3d_3lev", "plm")])
def test_flux_correction_conserves_across_levels(name, recon):
    m = _mesh(name)
    p = _problem(m, recon=recon)
    st = ao.AmrState(p, m)
    t0 = _totals(p.bs, st.u0)
    dt = st.estimate_dt()
    for _ in range(3):
        st.step(dt)
    t1 = _totals(p.bs, st.u0)
    scale = np.abs(t0) + 1.0
    assert np.all(np.abs(t1 - t0) <= 1e-13 * scale), t1 - t0
    nc = ao.AmrState(_problem(m, recon=recon), m, flux_correction=False)
    for _ in range(3):
        nc.step(dt)
    assert np.max(np.abs(_totals(p.bs, nc.u0) - t0) / scale) > 1e-9


def _is_interior(bs, off):
    b = int(np.searchsorted(bs.nvar * bs.blocks["cell_base"], off, side="right")) - 1
    c = int(off) - bs.nvar * int(bs.blocks["cell_base"][b])
    N = bs.extents()[b]
    idx = (c % N[0], (c // N[0]) % N[1], c // (N[0] * N[1]))
    g = int(bs.blocks["ng"][b])
    nx = bs.blocks["nx"][b]
    return all(g <= idx[d] < g + int(nx[d]) for d in range(bs.ndim))


@pytest.mark.parametrize("name", list(MESHES))
def test_ops_write_ghosts_and_read_interiors(name):
    m = _mesh(name)
    bs = m.block_set()
    ops, fixes, refs = amr.amr_plan(m, bs)
    assert not any(_is_interior(bs, o) for o in ops["dst"])
    used = set()
    for op in ops:
        nval = 1 if op["kind"] == amr.AMR_MEAN else 1 + 2 * m.ndim
        for k in range(nval):
            v = op["val"][k]
            used.update(range(int(v["first"]), int(v["first"]) + int(v["count"])))
    assert all(_is_interior(bs, refs["offset"][i]) for i in used)
    # a fix writes a face of a level-L block and reads faces of level-(L+1) blocks only
    blk = lambda off: int(np.searchsorted(bs.nvar * bs.blocks["cell_base"], off,
                                          side="right")) - 1
    for f in fixes:
        r = refs[f["val"]["first"]: f["val"]["first"] + f["val"]["count"]]
        assert all(m.level[blk(o)] == m.level[blk(f["dst"])] + 1 for o in r["offset"])
    assert len(fixes) > 0 and len(np.unique(fixes["dst"] * 3 + fixes["dir"])) == len(fixes)


@pytest.mark.parametrize("name", list(MESHES))
def test_fused_interface_table_covers_every_slot_once(name):
    """The fused form's interface table (amr.fused_faces), read with the kernel's slot
    formula: every fine face cell on a face to a coarser leaf writes one slot of the
    region owned by the coarse face across it, at its own child position, and every
    slot of every region is written exactly once."""
    m = _mesh(name)
    bs = m.block_set()
    nd, nx, nv = m.ndim, m.nx, bs.nvar
    faces, level, total = amr.fused_faces(m, bs)
    nch = 2 ** (nd - 1)
    owner = {}  # slot -> (coarse block, face, coarse global transverse cell (A, B), child)
    trans = lambda d: ((1 if d == 0 else 0), (1 if d == 2 else 2))
    for b in range(bs.nblocks):
        for f in range(6):
            af = faces[b * 6 + f]
            d = f // 2
            is_amr = d < nd and bs.blocks["bc"][b][f] == amr.BC_AMR
            assert (af["role"] != 0) == is_amr
            if af["role"] != 2:
                continue
            A, B = trans(d)
            for cb in range(nx[B] if B < nd else 1):
                for ca in range(nx[A]):
                    for k in range(nch):
                        slot = ((cb * af["na"] + ca) * nch + k) * nv + af["base"]
                        assert slot not in owner
                        ga = int(m.loc[b][A]) * nx[A] + ca
                        gb = int(m.loc[b][B]) * nx[B] + cb if B < nd else 0
                        owner[slot] = (b, f, (ga, gb), k)
    assert len(owner) * nv == total
    written = set()
    for b in range(bs.nblocks):
        for f in range(6):
            af = faces[b * 6 + f]
            if af["role"] != 1:
                continue
            d = f // 2
            A, B = trans(d)
            for cb in range(nx[B] if B < nd else 1):
                for ca in range(nx[A]):
                    fa, fb = af["qa"] + ca, af["qb"] + cb
                    child = (fa & 1) * 2 + (fb & 1) if nd == 3 else (fa & 1)
                    slot = (((fb >> 1) * af["na"] + (fa >> 1)) * nch + child) * nv + af["base"]
                    assert slot in owner and slot not in written
                    written.add(slot)
                    cblk, cf, (ga, gb), k = owner[slot]
                    # one level coarser, the opposite face of the same axis, above this cell
                    assert level[cblk] == level[b] - 1 and cf == (f ^ 1)
                    assert k == child
                    assert (int(m.loc[b][A]) * nx[A] + ca) // 2 == ga
                    if B < nd:
                        assert (int(m.loc[b][B]) * nx[B] + cb) // 2 == gb
    assert written == set(owner)


def test_unbalanced_or_orphan_trees_are_refused():
    base = dict(ndim=2, nbase=(4, 4), block_nx=(8, 8), ng=2, xmin=(0, 0, 0), xmax=(1, 1, 1))
    with pytest.raises(ValueError, match="balance"):
        # a level-2 region next to a level-0 leaf
        amr.AmrMesh(**base, refined=[(0, (1, 1)), (1, (2, 2))])
    with pytest.raises(ValueError, match="parent"):
        amr.AmrMesh(**base, refined=[(1, (2, 2))])
    with pytest.raises(ValueError, match="max_level"):
        amr.AmrMesh(**base, refined=[(0, (1, 1))], max_level=0)


def _deck_mesh():
    from apophis_amd import adaptive, hydro
    from tests.test_adaptive import oracle_tagger
    pkg = hydro.initialize_hydro(problems.blast_deck_pin())
    m, bs, _ = adaptive.initial_mesh(problems.blast_deck_base_mesh(), problems.blast_pgen(),
                                     oracle_tagger(pkg))
    return m, bs


@pytest.mark.parametrize("name", list(MESHES) + ["blast_deck"])
def test_native_tree_builds_the_same_lists(name):
    """aph_amr_build (the library's tree, host only) gives the op lists of the Python
    statement, byte for byte."""
    if name == "blast_deck":
        m, bs = _deck_mesh()
    else:
        m = _mesh(name)
        bs = m.block_set()
    want = amr.amr_plan(m, bs)
    # no length hint (a counting build first), a hint too short (a second build at the
    # reported lengths) and the hint the last build left (one build)
    shape = (m.ndim, bs.nvar, tuple(int(v) for v in m.nx), int(bs.blocks["ng"][0]), m.bc)
    for hint in (None, [0, 0, 0], "kept"):
        if hint is None:
            amr._PLAN_HINT.pop(shape, None)
        elif hint != "kept":
            amr._PLAN_HINT[shape] = hint
        got = amr.native_plan(m, bs)
        for w, g in zip(want, got):
            assert w.dtype == g.dtype and len(w) == len(g)
            assert w.tobytes() == g.tobytes()
        assert amr._PLAN_HINT[shape] == [len(w) for w in want]


def test_native_tree_refuses_bad_trees():
    from apophis_amd import _lib
    m = _mesh("2d_3lev")
    bs = m.block_set()
    # a level-2 block moved next to level-0 leaves: not balanced
    bad = amr.AmrMesh.__new__(amr.AmrMesh)
    bad.__dict__.update(m.__dict__)
    bad.loc = m.loc.copy()
    bad.level = m.level.copy()
    b2 = int(np.flatnonzero(m.level == 2)[0])
    b0 = int(np.flatnonzero(m.level == 0)[0])
    bad.level[b0], bad.loc[b0] = 2, (0, 0, 0)  # inside a level-0 leaf's region
    with pytest.raises(_lib.HydroError, match="aph_amr_build"):
        amr.native_plan(bad, bs)
    bad.level, bad.loc = m.level.copy(), m.loc.copy()
    bad.loc[b2] = bad.loc[b2 + 1]  # two blocks at one node
    with pytest.raises(_lib.HydroError, match="same node"):
        amr.native_plan(bad, bs)
