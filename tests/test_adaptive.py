"""The adaptive loop's host logic (apophis_amd/adaptive.py) and the blast deck
(inputs/blast.in: three levels on the pressure-gradient criterion): the derefinement
delay, the adaptive start that refines around the blast until it settles (tags from the
oracle's restatement of the criterion), and the deck's problem generator."""
import numpy as np
import pytest

from apophis_amd import adaptive, amr, hydro, problems
from oracle import amr_oracle as ao
from oracle import oracle as orc


def oracle_tagger(pkg):
    crit, a, b = pkg.param("refinement/criterion")
    eos = pkg.param("eos")

    def tag(mesh, bs, cons):
        u = cons.copy()
        orc.fill_ghosts(bs, u)
        ao.amr_fill(mesh, bs, u)
        prim = np.zeros_like(u)
        gam = np.zeros(bs.field_size(2))
        orc.cons_to_prim(bs, eos, u, prim, gam)
        return ao.refinement_tag(bs, prim, crit, a, b)[0]

    return tag


def test_derefine_counter_delays_and_resets():
    m = amr.AmrMesh(2, (2, 2), (4, 4), 2, (0, 0, 0), (1, 1, 1), [(0, (0, 0))], max_level=1)
    c = adaptive.DerefineCounter(3)
    n = len(m.level)
    tags = np.full(n, -1, np.int32)
    for _ in range(2):
        assert not (c.flags(m, tags) == -1).any()  # counting 1, 2
    assert (c.flags(m, tags) == -1).all()           # the third in a row
    tags[0] = 0
    f = c.flags(m, tags)
    assert f[0] == 0 and (f[1:] == -1).all()
    tags[0] = -1
    assert c.flags(m, tags)[0] == 0                 # block 0 starts again
    tags[1] = 1
    assert c.flags(m, tags)[1] == 1


def test_derefine_counter_carries_counts_by_leaf_across_meshes():
    m = amr.AmrMesh(2, (2, 2), (4, 4), 2, (0, 0, 0), (1, 1, 1), [(0, (0, 0))], max_level=1)
    c = adaptive.DerefineCounter(2)
    c.flags(m, np.full(len(m.level), -1, np.int32))
    c.flags(m, np.full(len(m.level), -1, np.int32))      # every leaf counted twice
    # a mesh with (0, (1, 1)) refined too: the old leaves keep their counts, new ones start at 0
    m2 = amr.AmrMesh(2, (2, 2), (4, 4), 2, (0, 0, 0), (1, 1, 1),
                     [(0, (0, 0)), (0, (1, 1))], max_level=1)
    f = c.flags(m2, np.full(len(m2.level), -1, np.int32))
    old = {(int(l), tuple(int(v) for v in x)) for l, x in zip(m.level, m.loc)}
    for b in range(len(m2.level)):
        key = (int(m2.level[b]), tuple(int(v) for v in m2.loc[b]))
        assert f[b] == (-1 if key in old else 0), key
    assert all(v == (2 if k in old else 1) for k, v in c.seen.items())


def test_blast_deck_pgen_restates_blast_cpp():
    """blast.cpp:16-71: inside rin (= radius_outer - radius_inner) the (drat damb,
    prat pamb) state, outside radius_outer the ambient state, between them a logarithmic
    ramp; energy = pres / (hydro gamma - 1), no velocity."""
    deck = dict(radius_inner=0.05)
    m = problems.blast_deck_base_mesh(dict(deck, nx=(32, 32), block_nx=(8, 8)))
    bs = m.block_set()
    u = problems.blast_pgen(dict(deck, drat=2.0))(m, bs)
    k = dict(problems.BLAST_DECK, **deck)
    gm1 = k["gamma"] - 1.0
    seen = set()
    for b in range(bs.nblocks):
        x, y, _ = bs.cell_centers(b)
        v = bs.interior(u, b, 5)
        r = np.sqrt(x * x + y * y)
        assert np.all(v[1:4] == 0.0)
        inner, outer = r < 0.05, r >= 0.1
        assert np.all(v[0][inner] == 2.0) and np.allclose(v[4][inner], 1.6e5 / gm1)
        assert np.all(v[0][outer] == 1.0) and np.allclose(v[4][outer], 1e-3 / gm1)
        ramp = ~inner & ~outer
        f = (r[ramp] - 0.05) / 0.05
        np.testing.assert_allclose(v[0][ramp], 2.0 ** (1 - f), rtol=1e-14)
        np.testing.assert_allclose(v[4][ramp] * gm1, 1.6e5 ** (1 - f) * 1e-3 ** f, rtol=1e-12)
        seen |= {"in" if inner.any() else "", "ramp" if ramp.any() else ""}
    assert {"in", "ramp"} <= seen


def test_adaptive_start_refines_around_the_blast():
    """The deck's adaptive start: three levels, the finest level on the blocks the
    pressure jump crosses, a balanced tree, and a settled mesh (no block below the top
    level is still tagged to refine)."""
    pin = problems.blast_deck_pin()
    pkg = hydro.initialize_hydro(pin)
    tag = oracle_tagger(pkg)
    m0 = problems.blast_deck_base_mesh()
    m, bs, cons = adaptive.initial_mesh(m0, problems.blast_pgen(), tag)
    assert m.nlevels == 3 and m.max_level == 2
    tags = tag(m, bs, cons)
    assert not np.any((tags == 1) & (m.level < m.max_level))
    # every block the r = 0.1 circle crosses is at level 2
    for b in range(bs.nblocks):
        x, y, _ = bs.cell_centers(b)
        r = np.sqrt(x * x + y * y)
        if r.min() < 0.1 <= r.max():
            assert m.level[b] == 2, (b, m.loc[b])
    # the state is generated again on the final mesh, not prolongated onto it
    np.testing.assert_array_equal(cons, problems.blast_pgen()(m, bs))


@pytest.mark.parametrize("nd", [2, 3])
def test_no_change_check_never_skips_a_regrid(nd):
    """The driver skips the regrid when no leaf is flagged 1 and no sibling group is
    flagged -1 whole; on random flags the library's regrid then returns the same leaves,
    and the check passes every flag set that does change them."""
    rng = np.random.default_rng(7 + nd)
    base = (2, 2, 2)[:nd] + (1,) * (3 - nd)
    refined = [(0, (0, 0, 0)[:nd] + (0,) * (3 - nd)), (0, (1, 1, 0) if nd == 2 else (1, 1, 1))]
    m = amr.AmrMesh(nd, base[:nd], (4,) * nd, 1, (0, 0, 0), (1, 1, 1),
                    [(l, tuple(c[:nd])) for l, c in refined], max_level=2)
    skipped = changed = 0
    for trial in range(300):
        p = rng.choice([0.0, 0.02, 0.3])
        flags = np.where(rng.random(len(m.level)) < 0.85, -1, 0).astype(np.int32)
        flags[rng.random(len(m.level)) < p] = 1
        if trial % 3 == 0:  # break up sibling groups
            flags[rng.random(len(m.level)) < 0.5] = 0
        flags[(flags < 0) & (m.level == 0)] = 0
        flags[(flags > 0) & (m.level >= m.max_level)] = 0
        lv, lc = amr.regrid_leaves(m, flags)
        same = np.array_equal(lv, m.level) and np.array_equal(lc, m.loc)
        if not adaptive._may_change(m, flags):
            assert same
            skipped += 1
        changed += not same
    assert skipped > 20 and changed > 20
