"""Pins of the multilevel Poisson restatement (oracle/poisson_amr_oracle.py: the discrete
problem src/gravity/gsolvers/poisson_gravity.hpp defines on the static-refinement decks'
meshes), and the library's host assembly of the same operator (aph_poisson_tree_matrix,
the matrix the device solver iterates with) against it.

* On a single-level mesh the operator is the lattice restatement's 7-point Laplacian
  (oracle/poisson_oracle.py, itself pinned by the exact DST solve): the same right-hand
  side, the same solution and the same field.
* Away from the domain faces (more than four of a cell's own widths: a ghost prolongated
  from a coarse cell on a domain face takes that face's image, minus the coarse cell, in
  its slope) a linear potential is annihilated on multilevel meshes:
  Linear prolongation, restriction by the mean and the flux correction are exact for it,
  while a slip at a level jump (a wrong sign, index or weight) is not; the Constant
  prolongation leaves a first-order error there, which the same check sees.
* With flux correction the fluxes through block faces balance: the volume-weighted
  column sums of the operator vanish for every cell off the domain faces; without it
  they do not.
* The reference's SMR accuracy limits (gravity.py's REFINED_LIMITS) on its poisson_smr_16
  mesh, with Linear prolongation (the deck's Quadratic form is not built).
* The library's operator and neighbour rows equal the restatement's, bitwise, on meshes of
  two and three levels, with fine blocks on the domain faces beside coarse ones, for
  both prolongations, with and without flux correction.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from apophis_amd import _lib, amr, problems
from oracle import oracle as O
from oracle import poisson_amr_oracle as PA
from oracle import poisson_oracle as P
from tests.monopole_checks import sphere_errors_max

BC = ["outflow"] * 6


def _mesh(name):
    if name == "grav16":
        return problems.grav_sphere(res=16, block=8, smr_level=2, solver="poisson").meta["mesh"]
    if name == "three":  # levels 0, 1, 2
        ref = amr.static_refinement(3, (4, 4, 4), (4, 4, 4), (0, 0, 0), (1, 1, 1),
                                    [((0.3, 0.3, 0.3), (0.4, 0.4, 0.4), 2)], BC)
        return amr.AmrMesh(3, (4, 4, 4), (4, 4, 4), 1, (0, 0, 0), (1, 1, 1), ref,
                           max_level=2, bc=BC)
    if name == "wallfine":  # fine blocks on the domain faces next to coarse leaves
        ref = amr.static_refinement(3, (3, 2, 2), (4, 4, 4), (0, 0, 0), (1.5, 1, 1),
                                    [((0.0, 0.0, 0.0), (0.2, 0.2, 0.2), 1)], BC)
        return amr.AmrMesh(3, (3, 2, 2), (4, 4, 4), 2, (0, 0, 0), (1.5, 1, 1), ref,
                           max_level=1, bc=BC)
    if name == "six":  # 6^3 blocks: no halving below the base level, a large coarsest level
        ref = amr.static_refinement(3, (3, 2, 2), (6, 6, 6), (0, 0, 0), (1.5, 1, 1),
                                    [((0.6, 0.4, 0.4), (0.8, 0.6, 0.6), 1)], BC)
        return amr.AmrMesh(3, (3, 2, 2), (6, 6, 6), 1, (0, 0, 0), (1.5, 1, 1), ref,
                           max_level=1, bc=BC)
    if name == "flat":
        return amr.AmrMesh(3, (2, 3, 2), (8, 4, 6), 1, (-0.5, -0.4, -0.6), (0.5, 0.8, 0.6),
                           set(), max_level=0, bc=BC)
    raise KeyError(name)


def _to_pack(pr, vec, comps=1):
    """Leaf vectors (comps, n) into a pack-layout field."""
    bs = pr.bs
    vec = np.atleast_2d(vec)
    out = np.zeros(bs.field_size(comps))
    nx = pr.nx
    for b in range(bs.nblocks):
        v = bs.interior(out, b, comps)
        for c in range(comps):
            v[c] = vec[c, b * pr.ncb:(b + 1) * pr.ncb].reshape(nx[2], nx[1], nx[0])
    return out


def _prim(p):
    prim = np.zeros_like(p.cons)
    gam = np.zeros(p.bs.field_size(2))
    u = p.cons.copy()
    O.fill_ghosts(p.bs, u)
    O.cons_to_prim(p.bs, p.eos, u, prim, gam)
    return prim


def test_single_level_equals_the_lattice_restatement():
    m = _mesh("flat")
    bs = m.block_set()
    pr = PA.Problem(m, bs, alpha=0.0)
    rng = np.random.default_rng(11)
    prim = np.zeros(bs.field_size(bs.nvar))
    for b in range(bs.nblocks):
        bs.interior(prim, b, bs.nvar)[0] = rng.uniform(0.5, 2.0, size=bs.interior(
            prim, b, bs.nvar)[0].shape)
    rho = pr.leaf_values(prim, bs.nvar, 0)
    coef = pr.mp_coeffs(rho, 0.7)
    # the two compute cell centres differently (per block vs over the lattice): a few ulp
    np.testing.assert_allclose(coef, P.mp_coeffs(bs, prim, 0.7), rtol=0,
                               atol=1e-14 * abs(coef).max())
    b_ml = pr.rhs(rho, 0.7, coef)
    b_lat = P.rhs(bs, prim, 0.7, coef)
    np.testing.assert_allclose(P.gather(bs, _to_pack(pr, b_ml), 1, 0), b_lat,
                               rtol=1e-14, atol=1e-14 * abs(b_lat).max())
    phi = pr.solve(b_ml)
    lo, dx, shape, _ = P.lattice(bs)
    phi_lat = P.solve(b_lat, dx)
    np.testing.assert_allclose(P.gather(bs, _to_pack(pr, phi), 1, 0), phi_lat, rtol=0,
                               atol=1e-11 * abs(phi_lat).max())
    g_lat, _ = P.field(bs, phi_lat, coef)
    g = pr.field(phi, coef)
    for d in range(3):
        np.testing.assert_allclose(P.gather(bs, _to_pack(pr, g[d]), 1, 0), g_lat[d], rtol=0,
                                   atol=1e-10 * abs(g_lat).max())


def _off_wall(pr):
    keep = np.ones(pr.n, dtype=bool)
    for cells, *_ in pr._walls():
        keep[cells] = False
    return keep


def _far_from_walls(pr, m, widths=4):
    cen = pr.centres()
    dx = np.repeat(pr.bs.blocks["dx"].astype(float), pr.ncb, axis=0)
    keep = np.ones(pr.n, dtype=bool)
    for d in range(3):
        gap = np.minimum(cen[d] - m.xmin[d], m.xmax[d] - cen[d])
        keep &= gap > widths * dx[:, d]
    return keep


@pytest.mark.parametrize("name", ["grav16", "three"])
def test_linear_potentials_are_annihilated_off_the_domain_faces(name):
    m = _mesh(name)
    bs = m.block_set()
    x, y, z = PA.Problem.centres(type("S", (), {"bs": bs})())
    lin = 0.3 + 1.7 * x - 0.9 * y + 2.3 * z
    for fc in (False, True):
        pr = PA.Problem(m, bs, flux_correct=fc)
        keep = _far_from_walls(pr, m)  # the Constant check below: level jumps among them
        scale = abs(pr.A).max() * abs(lin).max()
        r = abs(pr.A @ lin)[keep].max()
        assert r <= 1e-13 * scale, (fc, r, scale)
        const = PA.Problem(m, bs, flux_correct=fc, prolongation="Constant")
        assert abs(const.A @ np.ones(pr.n))[keep].max() <= 1e-13 * abs(const.A).max()
        assert abs(const.A @ lin)[keep].max() >= 1e-3 * scale  # first order at the jumps


@pytest.mark.parametrize("name", ["grav16", "three", "wallfine"])
def test_flux_correction_balances_the_block_faces(name):
    m = _mesh(name)
    bs = m.block_set()
    vol = np.repeat(bs.blocks["vol"].astype(float), int(np.prod(m.nx)))
    for fc in (True, False):
        pr = PA.Problem(m, bs, flux_correct=fc)
        colsum = np.asarray((sp.diags(vol) @ pr.A).sum(axis=0)).ravel()
        keep = _off_wall(pr)
        worst = abs(colsum[keep]).max() / (abs(pr.A).max() * vol.max())
        if fc:
            assert worst <= 1e-13, worst
        else:
            assert worst >= 1e-2, worst


def test_smr_accuracy_limits_of_the_reference():
    """gravity.py's poisson_smr_16: 16^3 base in 8^3 blocks, [-0.25, 0.25]^3 at level 2,
    the sphere at (0.01, 0, 0), flux_correct false; REFINED_LIMITS (phi L2 5e-3, g_x L2
    5e-2, phi max 5e-2, g_x max 0.2)."""
    p = problems.grav_sphere(res=16, block=8, smr_level=2, solver="poisson")
    pr = PA.Problem(p.meta["mesh"], p.bs, flux_correct=False)
    coef, phi, g = pr.gravity(_prim(p), 1.0)
    b = pr.rhs(pr.leaf_values(_prim(p), p.bs.nvar, 0), 1.0, coef)
    assert np.linalg.norm(pr.A @ phi - b) <= 1e-12 * np.linalg.norm(b)
    l2p, l2g, mxp, mxg = sphere_errors_max(p.bs, _to_pack(pr, g, 3), _to_pack(pr, phi),
                                           center=p.meta["center"])
    print(l2p, l2g, mxp, mxg)
    assert l2p <= 5e-3 and l2g <= 5e-2 and mxp <= 5e-2 and mxg <= 2e-1


@pytest.mark.skipif(not _lib.available(), reason="native library not built")
@pytest.mark.parametrize("name", ["grav16", "three", "wallfine", "flat", "six"])
def test_library_assembly_equals_the_restatement(name):
    m = _mesh(name)
    bs = m.block_set()
    for fc in (False, True):
        for prol, code in (("Linear", 0), ("Constant", 1)):
            pr = PA.Problem(m, bs, alpha=0.25, flux_correct=fc, prolongation=prol)
            rp, col, val, nbp, nbc, nbv = amr.poisson_matrix(m, bs, fc, code, 0.25)
            A = sp.csr_matrix((val, col, rp), shape=pr.A.shape)
            assert abs(A - pr.A).max() == 0.0, (fc, prol)
            NB = sp.csr_matrix((nbv, nbc, nbp), shape=(6 * pr.n, pr.n))
            grad = pr._grad_rows()
            for d in range(3):
                for s in range(2):
                    part = NB[(2 * d + s) * pr.n:(2 * d + s + 1) * pr.n]
                    assert abs(part - grad[d][s]).max() == 0.0, (fc, prol, d, s)


@pytest.mark.skipif(not _lib.available(), reason="native library not built")
def test_library_assembly_refusals():
    m = _mesh("three")
    bs = m.block_set()
    with pytest.raises(_lib.HydroError, match="quadratic form is not built"):
        amr.poisson_matrix(m, bs, False, 2)
    bad = bs.blocks.copy()
    bad["xmin"][3][0] += 0.01
    bs2 = m.block_set()
    bs2.blocks = bad
    with pytest.raises(_lib.HydroError, match="spacing or corner"):
        amr.poisson_matrix(m, bs2, False, 0)
