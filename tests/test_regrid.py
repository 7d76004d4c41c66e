"""Regridding a two-level mesh from refinement tags (apophis_amd/amr.py regrid): the mesh
follows the tags (refine level-0 blocks tagged 1, merge level-1 sibling groups all tagged
-1, nothing else), conserved totals are kept to round-off, refining and merging back
returns the coarse data, a linear field is prolongated exactly, and the regridded state
steps on (conservation across a regrid in a short adaptive run of the oracle)."""
import numpy as np
import pytest

from apophis_amd import amr, problems
from oracle import amr_oracle as ao


def _tot(bs, u):
    return sum(bs.interior(u, b, bs.nvar).reshape(bs.nvar, -1).sum(axis=1) * bs.blocks["vol"][b]
               for b in range(bs.nblocks))


@pytest.mark.parametrize("nd", [2, 3])
def test_regrid_structure_conservation_and_round_trip(nd):
    if nd == 2:
        p, m = problems.blast2d_two_level(nbase=8, n=8, nref=2, recon="plm")
    else:
        m = amr.TwoLevelMesh(3, (3, 3, 2), (6, 6, 6), 2, (0, 0, 0), (1, 1, 1), [(1, 1, 0)])
        bs = m.block_set()
        u = np.zeros(bs.field_size(5))
        rng = np.random.default_rng(4)
        for b in range(bs.nblocks):
            v = bs.interior(u, b, 5)
            v[:] = 1.0 + 0.1 * rng.random(v.shape)
        p = problems.Problem("r", bs, u, None, "plm", "hllc", 0.3)
    bs = p.bs
    tags = np.zeros(bs.nblocks, np.int32)
    lev0 = [b for b in range(bs.nblocks) if m.level[b] == 0]
    tags[lev0[:2]] = 1
    tags[lev0[2:4]] = -1       # derefine on level 0: no effect
    fine = [b for b in range(bs.nblocks) if m.level[b] == 1]
    tags[fine[:1]] = -1        # one sibling only: the group stays
    nm, nbs, nc = amr.regrid(m, bs, p.cons, tags)
    assert (nm.level == 1).sum() == (m.level == 1).sum() + 2 * 2 ** nd
    np.testing.assert_allclose(_tot(nbs, nc), _tot(bs, p.cons), rtol=1e-14)
    # merge the new groups again: the original interior data comes back
    new_parents = {tuple(int(v) for v in m.loc[b][:nd]) for b in lev0[:2]}
    t2 = np.zeros(nbs.nblocks, np.int32)
    for b in range(nbs.nblocks):
        if nm.level[b] == 1 and tuple(int(v) // 2 for v in nm.loc[b][:nd]) in new_parents:
            t2[b] = -1
    m3, bs3, c3 = amr.regrid(nm, nbs, nc, t2)
    assert bs3.nblocks == bs.nblocks
    mask = np.tile(bs.interior_mask(), bs.nvar)
    np.testing.assert_allclose(c3[mask], p.cons[mask], rtol=2e-15, atol=0)


def test_prolongation_of_a_linear_field_is_exact():
    m = amr.TwoLevelMesh(2, (4, 4), (8, 8), 2, (0, 0, 0), (1, 1, 1), [])
    bs = m.block_set()
    u = np.zeros(bs.field_size(5))
    for b in range(bs.nblocks):
        x, y, _ = bs.cell_centers(b)
        v = bs.interior(u, b, 5)
        for n in range(5):
            v[n] = 1.0 + (n + 1) * 0.1 * x - 0.3 * y
    tags = np.zeros(bs.nblocks, np.int32)
    tags[5] = 1  # an inner block: its stencil does not wrap
    nm, nbs, nc = amr.regrid(m, bs, u, tags)
    for b in range(nbs.nblocks):
        if nm.level[b] != 1:
            continue
        x, y, _ = nbs.cell_centers(b)
        v = nbs.interior(nc, b, 5)
        for n in range(5):
            np.testing.assert_allclose(v[n], 1.0 + (n + 1) * 0.1 * x - 0.3 * y, rtol=1e-14)


def test_adaptive_run_conserves_across_regrids():
    p, m = problems.blast2d_two_level(nbase=6, n=8, nref=2, recon="plm")
    st = ao.AmrState(p, m)
    t0 = _tot(p.bs, st.u0)
    dt = st.estimate_dt()
    for cycle in range(3):
        for _ in range(2):
            dt = st.step(dt)
        tags, _ = ao.refinement_tag(p.bs, st.prim, 1, 0.5, 0.0)
        m, nbs, nc = amr.regrid(m, p.bs, st.u0, tags)
        p = problems.Problem("adaptive", nbs, nc, p.eos, p.recon, p.riemann, p.cfl)
        st = ao.AmrState(p, m)
        dt = min(dt, st.estimate_dt())
    np.testing.assert_allclose(_tot(p.bs, st.u0), t0, rtol=1e-13,
                               atol=1e-13 * np.abs(t0).max())
    assert (m.level == 1).sum() > 0
