"""Pins of the Poisson-gravity restatement (oracle/poisson_oracle.py, restating
src/gravity/gsolvers/poisson_gravity.hpp and gravity.hpp's MultipolePotential).

* The DST solve is exact: the operator applied by its own stencil (ghost = minus the
  interior value) reproduces the right-hand side to rounding.
* The multipole coefficients and potential: for a compact cloud of masses the expansion
  matches the direct sum -G sum m / |x - x_i| at distance to third order in size/distance;
  dropping the quadrupole (or scaling it wrongly) leaves a second-order error, 10x larger.
* The reference's own accuracy test (gravity.py, the poisson configurations: sphere offset
  by 0.01 along x, 16^3, 32^3, 64^3): L2 errors of phi and g_x fall monotonically.
* Evrard's collapse with the deck's Poisson solver (evrards_collapse.py's poisson_32) passes
  every one of the reference's profile checks.
"""
import numpy as np
import pytest

from apophis_amd import problems
from oracle import oracle as O
from oracle import poisson_oracle as P
from tests.monopole_checks import (analytic_g, analytic_phi, evrard_exact, evrard_profiles,
                                   evrard_verdict)


def test_dst_solve_is_exact():
    rng = np.random.default_rng(7)
    b = rng.normal(size=(12, 10, 16))
    dx = (0.1, 0.07, 0.05)
    for alpha in (0.0, 3.0):
        phi = P.solve(b, dx, alpha)
        np.testing.assert_allclose(P.laplacian(phi, dx, alpha), b, rtol=0, atol=1e-11 * abs(b).max())


def test_multipole_expansion_matches_direct_sum():
    rng = np.random.default_rng(3)
    G = 0.7
    pts = rng.uniform(-0.05, 0.05, size=(40, 3))
    m = rng.uniform(0.5, 1.5, size=40)
    x, y, z = pts.T
    r2 = x * x + y * y + z * z
    mom = [m, m * y, m * z, m * x, m * x * y, m * y * z, m * (3 * z * z - r2), m * z * x,
           m * 0.5 * (x * x - y * y)]
    c = np.array([np.sum(v) * P.SCALE[n] * G for n, v in enumerate(mom)])
    far = rng.normal(size=(50, 3))
    far /= np.linalg.norm(far, axis=1)[:, None]
    direct = np.array([-G * np.sum(m / np.linalg.norm(q - pts, axis=1)) for q in far])
    expan = P.multipole_potential(c, far[:, 0], far[:, 1], far[:, 2])
    err = np.max(np.abs(expan - direct) / np.abs(direct))
    c_mono_dip = c.copy()
    c_mono_dip[4:] = 0.0
    err_md = np.max(np.abs(P.multipole_potential(c_mono_dip, far[:, 0], far[:, 1], far[:, 2])
                           - direct) / np.abs(direct))
    assert err < 1e-4 and err_md > 10 * err, (err, err_md)


def _errors(p, phi, g):
    Z, Y, X = P.centres(p.bs)
    c = p.meta["center"]
    dx = X - c[0]
    r = np.sqrt(dx ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
    e_phi = phi - analytic_phi(r)
    gx = np.where(r > 0, analytic_g(r) * dx / np.where(r > 0, r, 1.0), 0.0)
    return float(np.sqrt(np.mean(e_phi ** 2))), float(np.sqrt(np.mean((g[0] - gx) ** 2)))


def test_accuracy_test_of_the_reference():
    errs = []
    for res in (16, 32, 64):
        p = problems.grav_sphere(res=res, block=16, solver="poisson")
        st = O.State(p)
        _, phi, g = P.gravity(p.bs, st.prim, 1.0)
        errs.append(_errors(p, phi, g))
    errs = np.array(errs)
    print(errs)
    assert np.all(np.diff(errs, axis=0) < 0), errs
    # second order: each halving of dx cuts both errors by about 4
    assert np.all(errs[:-1] / errs[1:] > 3.5), errs


def test_evrard_poisson_passes_the_reference_checks():
    p = problems.evrard(solver="poisson")
    st = O.State(p)
    t, n = st.run(tlim=p.tlim)
    assert t == pytest.approx(0.8)
    fails, rms = evrard_verdict(evrard_profiles(p.bs, st.prim), evrard_exact())
    print(n, rms)
    assert not fails, fails
