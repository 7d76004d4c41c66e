"""The hydro package's refinement criteria (hydro.cpp:303-333, src/refinement/): option
parsing, and the oracle's restatement (oracle/amr_oracle.refinement_tag) pinned by closed
forms and by the cells at the edges of each criterion's index range:
* a uniform state has zero gradients: tag derefine for both gradient criteria;
* p = p0 + a x gives eps = |a| dx / p at every cell, so the block's value is
  |a| dx / min p over the range, which reaches one ghost cell past the interior;
* v1 = b y gives vg = |b| dy;
* max density reads i up to one past the interior (other.cpp:25) but no ghost row in j
  and no ghost column below.
"""
import numpy as np
import pytest

from apophis_amd import hydro
from apophis_amd.mesh import BlockSet, lattice_mesh, make_eos
from oracle import amr_oracle as ao
from oracle import oracle as orc


def _pin(refine):
    return {"parthenon/time": {"cfl": 0.3, "integrator": "rk2"},
            "parthenon/mesh": {"nghost": 2},
            "hydro": {"reconstruction": "plm", "riemann": "hllc"},
            "refinement": refine}


def test_refinement_options():
    pkg = hydro.initialize_hydro(_pin({"type": "pressure_gradient",
                                       "threshold_pressure_gradient": 0.2}))
    assert pkg.param("refinement/criterion") == (1, 0.2, 0.0)
    pkg = hydro.initialize_hydro(_pin({"type": "maxdensity", "maxdensity_deref_below": 0.5,
                                       "maxdensity_refine_above": 2.0}))
    assert pkg.param("refinement/criterion") == (3, 0.5, 2.0)
    for bad in ({"type": "pressure_gradient"}, {"type": "xyvelocity_gradient"},
                {"type": "maxdensity", "maxdensity_deref_below": 2.0,
                 "maxdensity_refine_above": 1.0},
                {"type": "maxdensity", "maxdensity_refine_above": 1.0}):
        with pytest.raises(hydro.ApophisError):
            hydro.initialize_hydro(_pin(bad))
    assert hydro.initialize_hydro(_pin({"type": "unset"})).param("refinement/type") == "unset"


def _bs(ndim=2, n=8, ng=2, nblk=2):
    nb = (nblk,) * ndim + (1,) * (3 - ndim)
    bnx = (n,) * ndim + (1,) * (3 - ndim)
    m = lattice_mesh(ndim, nb, bnx, ng, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), ["periodic"] * 6)
    return BlockSet(ndim, [m])


def _prim_from(bs, fn):
    """prim over every cell incl. ghosts from fn(x, y, z) -> (rho, v1, v2, v3, p)."""
    prim = np.zeros(bs.field_size(5))
    ext = bs.extents()
    for b in range(bs.nblocks):
        r = bs.blocks[b]
        g = int(r["ng"])
        view = bs.block_view(prim, b, 5)
        c = []
        for d in range(3):
            E = int(ext[b][d])
            off = g if d < bs.ndim else 0
            c.append(r["xmin"][d] + (np.arange(E) - off + 0.5) * r["dx"][d])
        z, y, x = np.meshgrid(c[2], c[1], c[0], indexing="ij")
        for v, f in enumerate(fn(x, y, z)):
            view[v] = f
    return prim


@pytest.mark.parametrize("ndim", [2, 3])
def test_uniform_state_derefines(ndim):
    bs = _bs(ndim)
    prim = _prim_from(bs, lambda x, y, z: (1.0, 0.3, -0.2, 0.1, 2.0))
    for crit in (1, 2):
        tags, vals = ao.refinement_tag(bs, prim, crit, 0.1, 0.0)
        assert (tags == -1).all() and (vals == 0.0).all()
    tags, vals = ao.refinement_tag(bs, prim, 3, 0.5, 2.0)
    assert (tags == 0).all() and (vals == 1.0).all()


def test_pressure_gradient_closed_form():
    bs = _bs(2)
    a = 0.7
    prim = _prim_from(bs, lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 3.0 + a * x))
    tags, vals = ao.refinement_tag(bs, prim, 1, 0.01, 0.0)
    dx = bs.blocks["dx"][0][0]
    for b in range(bs.nblocks):
        x0 = bs.blocks["xmin"][b][0]
        pmin = 3.0 + a * (x0 - 0.5 * dx)   # the ghost column below the block
        assert vals[b] == pytest.approx(a * dx / pmin, rel=1e-13)
        assert tags[b] == 1


def test_velocity_gradient_closed_form():
    bs = _bs(2)
    b_ = -1.3
    prim = _prim_from(bs, lambda x, y, z: (1.0, b_ * y, 0.0, 0.0, 1.0))
    tags, vals = ao.refinement_tag(bs, prim, 2, 1.0, 0.0)
    dy = bs.blocks["dx"][0][1]
    np.testing.assert_allclose(vals, abs(b_) * dy, rtol=1e-13)
    assert (tags == -1).all()  # |b| dy < threshold / 2


def test_max_density_range_edges():
    bs = _bs(2, n=8, ng=2, nblk=1)
    prim = _prim_from(bs, lambda x, y, z: (1.0 + 0 * x, 0.0, 0.0, 0.0, 1.0))
    view = bs.block_view(prim, 0, 5)
    g, n = 2, 8
    view[0, 0, g + 3, g + n] = 5.0      # first ghost column above: inside the range
    assert ao.refinement_tag(bs, prim, 3, 0.5, 2.0)[1][0] == 5.0
    view[0, 0, g + 3, g + n] = 1.0
    view[0, 0, g + 3, g - 1] = 5.0      # ghost column below: outside
    view[0, 0, g + n, g + 3] = 5.0      # ghost row above: outside
    tags, vals = ao.refinement_tag(bs, prim, 3, 0.5, 2.0)
    assert vals[0] == 1.0 and tags[0] == 0
