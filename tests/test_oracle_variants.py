"""Oracle checks of the method variants: the LHLLC and "none" Riemann solvers
(hydro_lhllc.hpp, hydro_none.hpp) and the rk1 / rk3 / vl2 integrators with the upstream
low-storage coefficients and vl2's first-order predictor (hydro.cpp:114-139).

* LHLLC equals HLLC bitwise when max(|v_L|, |v_R|) = 1 (phi = 1), and its middle pressure
  follows phi = chi (2 - chi) with chi the velocity magnitude: for a symmetric collision
  the contact is at rest exactly and the momentum flux is p + rho u c q(p*) in closed form;
* "none" leaves the state unchanged;
* on first-order upwind advection (a linear operator A) one step of each integrator is a
  polynomial in c A (rk1: degree 1, rk2 and vl2: degree 2, rk3: degree 3);
* vl2's predictor is DC: a step equals the composition of the task functions with DC
  fluxes and beta = 1/2 in stage 1 and the PLM fluxes in stage 2.
"""
import numpy as np
import pytest

from apophis_amd import problems
from apophis_amd.mesh import BlockSet, lattice_mesh, make_eos
from oracle import oracle as orc

rng = np.random.default_rng(31)


def _states(rng):
    wl = np.array([rng.uniform(0.2, 2), rng.normal(0, 0.3), rng.normal(0, 0.3),
                   rng.normal(0, 0.3), rng.uniform(0.2, 2)])
    wr = np.array([rng.uniform(0.2, 2), 0.0, 0.0, 0.0, rng.uniform(0.2, 2)])
    return wl, wr


def test_lhllc_is_hllc_at_unit_velocity_magnitude():
    for _ in range(500):
        wl, wr = _states(rng)
        wl[1:4] *= 0.99 / max(np.sqrt((wl[1:4] ** 2).sum()), 1.0)  # |v_L| < 1
        axis = rng.integers(0, 3)
        wr[1 + axis] = rng.choice([-1.0, 1.0])  # |v_R| = 1 exactly
        g = [rng.choice([1.4, 5.0 / 3.0])] * 2
        a = orc.riemann("lhllc", wl, wr, g, g)
        b = orc.riemann("hllc", wl, wr, g, g)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (wl, wr)


@pytest.mark.parametrize("u", [0.05, 0.3, 0.5, 1.0, 1.5])
def test_lhllc_symmetric_collision_closed_form(u):
    """wl = (rho, u, 0, 0, p), wr = (rho, -u, 0, 0, p): s* = 0 exactly, so the flux is
    (0, cp, 0, 0, 0) with cp = p + rho u c q_L and q_L = sqrt(1 + (g+1)/(2g) (p*/p - 1)),
    p* = p + rho u c phi, phi = u (2 - u) for LHLLC and 1 for HLLC."""
    rho, p, g = 1.3, 0.7, 1.4
    c = np.sqrt(g * p / rho)  # the C_LIGHT form agrees to ~1e-16
    wl = np.array([rho, u, 0.0, 0.0, p])
    wr = np.array([rho, -u, 0.0, 0.0, p])
    for kind, phi in (("lhllc", u * (2.0 - u)), ("hllc", 1.0)):
        f = orc.riemann(kind, wl, wr, [g, g], [g, g])
        pstar = p + rho * u * c * phi
        q = np.sqrt(1.0 + (g + 1.0) / (2.0 * g) * (pstar / p - 1.0))
        assert f[0] == 0.0 and f[2] == 0.0 and f[3] == 0.0 and f[4] == 0.0
        assert f[1] == pytest.approx(p + rho * u * c * q, rel=1e-12), (kind, u)
    if u != 1.0:  # phi != 1: LHLLC really differs from HLLC
        assert orc.riemann("lhllc", wl, wr, [g, g], [g, g])[1] != \
            orc.riemann("hllc", wl, wr, [g, g], [g, g])[1]


def _advection(nscalars=1, n=32, recon="dc", integrator="rk2", riemann="hllc"):
    m = lattice_mesh(1, (1, 1, 1), (n, 1, 1), 2, (0, 0, 0), (1, 1, 1), ["periodic"] * 6)
    bs = BlockSet(1, [m], nscalars=nscalars)
    u = np.zeros(bs.field_size(bs.nvar))
    x, _, _ = bs.cell_centers(0)
    s = 1.0 + 0.5 * np.sin(2 * np.pi * x[0, 0]) ** 3
    vel, rho, pres, gm1 = 0.75, 1.0, 1.0, 0.4
    v = bs.interior(u, 0, bs.nvar)
    v[0], v[1], v[4], v[5] = rho, rho * vel, pres / gm1 + 0.5 * rho * vel ** 2, rho * s
    p = problems.Problem("adv", bs, u, make_eos(gm1), recon, riemann, 0.3,
                         integrator=integrator)
    return p, s, vel


@pytest.mark.parametrize("integrator", ["rk1", "rk2", "rk3", "vl2"])
def test_integrator_closed_form_linear_advection(integrator):
    """First-order upwind advection of a passive scalar in a uniform flow is linear
    (s <- s + c A s per forward Euler stage, c = v dt / dx), so one step is the
    integrator's stability polynomial of c A applied to s."""
    p, s, vel = _advection(integrator=integrator)
    n = s.size
    st = orc.State(p)
    dt = 0.4 / n
    st.step(dt)
    got = p.bs.interior(st.u0, 0, p.bs.nvar)[5, 0, 0]
    c = vel * dt * n
    A = lambda y: -(y - np.roll(y, 1))
    terms = [s, c * A(s), 0.5 * c * c * A(A(s)), c ** 3 / 6.0 * A(A(A(s)))]
    degree = {"rk1": 1, "rk2": 2, "vl2": 2, "rk3": 3}[integrator]
    expect = sum(terms[:degree + 1])
    np.testing.assert_allclose(got, expect, rtol=2e-14)
    # one degree more or less is far off
    assert np.max(np.abs(got - sum(terms[:degree]))) > 1e-4
    if degree < 3:
        assert np.max(np.abs(got - sum(terms[:degree + 2]))) > 1e-6


def test_integrator_coefficients():
    """The stage tables against the schemes they encode: stage k of a low-storage
    step is u <- gam0 u + gam1 u_n + beta dt L(u); for L = lambda (a scalar ODE) the
    composition must equal the scheme's amplification factor (Taylor polynomial for
    rk1/rk2/rk3, the midpoint rule for vl2)."""
    z = 0.37
    for name, poly in (("rk1", 1 + z), ("rk2", 1 + z + z * z / 2),
                       ("rk3", 1 + z + z * z / 2 + z ** 3 / 6), ("vl2", 1 + z + z * z / 2)):
        g0, g1, be = orc.integrator_stages(name)
        u, un = 1.0, 1.0
        for k in range(len(g0)):
            u = g0[k] * u + g1[k] * un + be[k] * z * u
        assert u == pytest.approx(poly, rel=1e-15), name


def test_vl2_predictor_is_dc():
    p, s, vel = _advection(recon="plm", integrator="vl2")
    st = orc.State(p)
    dt = 0.4 / s.size
    u_start = st.u0.copy()
    st.step(dt)
    # the same step composed from the task functions
    bs, eos = p.bs, p.eos
    u0, u1 = u_start.copy(), u_start.copy()
    prim, gam = np.zeros_like(u0), np.zeros(bs.field_size(2))
    fx = [np.zeros_like(u0) for _ in range(3)]
    orc.fill_ghosts(bs, u0)
    orc.cons_to_prim(bs, eos, u0, prim, gam)
    for recon, beta in (("dc", 0.5), ("plm", 1.0)):
        orc.calculate_fluxes(bs, recon, "hllc", prim, gam, *fx)
        orc.update(bs, u0, u1, *fx, 0.0, 1.0, beta * dt)
        orc.fill_ghosts(bs, u0)
        orc.cons_to_prim(bs, eos, u0, prim, gam)
    mask = np.tile(bs.interior_mask(), bs.nvar)
    assert np.array_equal(st.u0[mask], u0[mask])
    # a PLM predictor gives a different state
    p2, _, _ = _advection(recon="plm", integrator="rk2")
    st2 = orc.State(p2)
    st2.step(dt)
    assert np.max(np.abs(st2.u0[mask] - st.u0[mask])) > 1e-8


def test_none_solver_leaves_the_state_unchanged():
    p = problems.blast3d(nblk=1, n=8, recon="dc", riemann="none")
    st = orc.State(p)
    before = st.u0.copy()
    st.step(1e-3)
    mask = np.tile(p.bs.interior_mask(), p.bs.nvar)
    assert np.array_equal(st.u0[mask], before[mask])


def test_lhllc_unclipped_phi_quirk():
    """phi = chi (2 - chi) is not clipped (hydro_lhllc.hpp:71-79): above |v| = 2 it turns
    negative and flips the PVRS term, so a supersonic flow at |v| = 5 moving left is no
    longer upwinded, while HLLC upwinds it."""
    g = 1.4
    m = np.array([1, -1, 1, 1, 1])
    wl = np.array([1.0, 5.0, 0.3, -0.2, 1.0]) * m
    wr = np.array([0.5, 4.0, 0.1, 0.0, 0.4]) * m
    gam = [g, g]
    E = wr[4] / (g - 1) + 0.5 * wr[0] * (wr[1:4] ** 2).sum()
    upwind = np.array([wr[0] * wr[1], wr[0] * wr[1] ** 2 + wr[4], wr[0] * wr[1] * wr[2],
                       wr[0] * wr[1] * wr[3], (E + wr[4]) * wr[1]])
    np.testing.assert_allclose(orc.riemann("hllc", wl, wr, gam, gam), upwind, rtol=1e-13)
    assert np.max(np.abs(orc.riemann("lhllc", wl, wr, gam, gam) - upwind)) > 1e-3
