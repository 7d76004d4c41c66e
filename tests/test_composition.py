"""The reference's composition scalars (hydro.cpp:141-202): option parsing, the
kh-comp-advection deck's initial state (kh.cpp:268-313, iprob 6) and what the oracle
must satisfy with 7 or 9 passive scalars (NV = 12 / 14).

* The pgen draws from std::mt19937 through std::uniform_real_distribution; the restatement
  in problems.std_uniform_m05_05 and the draw order of problems.kh_comp are checked
  bitwise against a program built with the system's C++ library (tests/native/
  std_rng_check.cpp), which is what the reference links.
* Passive scalars do not feed back: the hydro fields of a step with 9 scalars are the bits
  of the same step with none.
* A scalar equal to c * density stays c * density (its primitive is constant, so every
  reconstruction is exact and its flux is c times the mass flux), to round-off.
* Each scalar's total is conserved on a periodic mesh (telescoping fluxes).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from apophis_amd import hydro, problems
from apophis_amd.mesh import BlockSet, lattice_mesh
from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def rng_prog(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("rng") / "std_rng_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off",
                    os.path.join(HERE, "native", "std_rng_check.cpp"), "-o", exe], check=True)
    return exe


def _hex(lines):
    return np.array([[float.fromhex(t) for t in l.split()] for l in lines if l.strip()])


def test_std_uniform_matches_libstdcxx(rng_prog):
    for seed in (0, 1, 3, 12345):
        out = subprocess.run([rng_prog, "draws", str(seed), "3000"], check=True,
                             capture_output=True, text=True).stdout.splitlines()
        want = _hex(out)[:, 0]
        got = problems.std_uniform_m05_05(seed, 3000)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), seed


@pytest.mark.parametrize("nlset", [0, 1])
def test_kh_comp_initial_state_matches_pgen_loop(rng_prog, nlset):
    p = problems.kh_comp(nlset=nlset)
    bs = p.bs
    assert bs.nvar == 12 + 2 * nlset and bs.nblocks == 4
    for b in (0, 3):
        r = bs.blocks[b]
        nx, ny = int(r["nx"][0]), int(r["nx"][1])
        args = [rng_prog, "khcomp", str(b), str(nx), str(ny), repr(float(r["xmin"][0])),
                repr(float(r["xmin"][1])), repr(float(r["dx"][0])), repr(float(r["dx"][1])),
                str(nlset)]
        out = subprocess.run(args, check=True, capture_output=True, text=True).stdout
        want = _hex(out.splitlines())  # (ny * nx, nv), loop order j, i
        got = bs.interior(p.cons, b, bs.nvar)[:, 0].reshape(bs.nvar, -1).T
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), b


def _pin(composition="ONe", nlset=1, eos="ideal"):
    return {"parthenon/time": {"cfl": 0.4, "integrator": "rk2"},
            "parthenon/mesh": {"nghost": 2},
            "hydro": {"reconstruction": "plm", "riemann": "hllc",
                      "composition": composition, "nlset": nlset},
            "eos": {"type": eos, "gm1": 0.66666666667}}


def test_initialize_hydro_composition_options():
    pkg = hydro.initialize_hydro(_pin())
    assert pkg.param("nscalars") == 9 and pkg.param("ncomp") == 6 and pkg.param("nlset") == 1
    assert pkg.param("comp_abar") == [4.0, 12.0, 16.0, 20.0, 30.0, 56.0]
    assert pkg.param("cons_labels")[5:] == [
        "scalar_density_alpha", "scalar_density_C12", "scalar_density_O16",
        "scalar_density_Ne20", "scalar_density_IME", "scalar_density_IGE",
        "scalar_density_ye", "scalar_density_lset0", "scalar_density_xfuel0"]
    assert hydro.initialize_hydro(_pin(nlset=0)).param("nscalars") == 7
    # an unknown composition adds nothing (hydro.cpp:154 only matches "ONe")
    assert hydro.initialize_hydro(_pin(composition="unset", nlset=0)).param("nscalars") == 0
    with pytest.raises(hydro.ApophisError, match="Levelset"):
        hydro.initialize_hydro(_pin(composition="unset", nlset=1))
    with pytest.raises(hydro.ApophisError, match="Helmholtz"):
        hydro.initialize_hydro(_pin(eos="helm"))


def _hydro_only(p):
    """The same mesh and hydro state with no scalars."""
    r = p.bs.blocks
    nbx = int(round((r["xmin"][:, 0].max() - r["xmin"][:, 0].min()) / r["dx"][0, 0] /
                    r["nx"][0, 0])) + 1
    nby = p.bs.nblocks // nbx
    m = lattice_mesh(2, (nbx, nby, 1), (int(r["nx"][0, 0]), int(r["nx"][0, 1]), 1),
                     int(r["ng"][0]), (-1.0, -0.5, -0.5), (1.0, 0.5, 0.5), ["periodic"] * 6)
    bs = BlockSet(2, [m], nscalars=0)
    cons = np.zeros(bs.field_size(5))
    for b in range(bs.nblocks):
        bs.interior(cons, b, 5)[:] = p.bs.interior(p.cons, b, p.bs.nvar)[:5]
    return problems.Problem("kh_hydro", bs, cons, p.eos, p.recon, p.riemann, p.cfl)


@pytest.mark.parametrize("recon,riemann", [("plm", "hllc"), ("ppm", "hlle")])
def test_scalars_do_not_feed_back(recon, riemann):
    p = problems.kh_comp(nx=(64, 32), block=(16, 16), recon=recon, riemann=riemann)
    q = _hydro_only(p)
    a, b = orc.State(p), orc.State(q)
    dt = a.estimate_dt()
    assert dt == b.estimate_dt()
    for _ in range(2):
        a.step(dt)
        b.step(dt)
    for k in range(p.bs.nblocks):
        ha = p.bs.interior(a.u0, k, p.bs.nvar)[:5]
        hb = q.bs.interior(b.u0, k, 5)
        assert np.array_equal(ha, hb)


def test_scalar_proportional_to_density_stays_so():
    p = problems.kh_comp(nx=(64, 32), block=(16, 16))
    for b in range(p.bs.nblocks):
        v = p.bs.interior(p.cons, b, p.bs.nvar)
        v[5 + 4] = 0.3 * v[0]   # IME = 0.3 rho
        v[5 + 8] = 0.0          # xfuel0 = 0 rho
    st = orc.State(p)
    dt = st.estimate_dt()
    for _ in range(3):
        st.step(dt)
    for b in range(p.bs.nblocks):
        v = p.bs.interior(st.u0, b, p.bs.nvar)
        np.testing.assert_allclose(v[5 + 4] / v[0], 0.3, rtol=2e-15, atol=0)
        assert not v[5 + 8].any()


def test_scalar_totals_conserved():
    p = problems.kh_comp(nx=(64, 32), block=(16, 16))
    st = orc.State(p)
    tot = lambda u: np.array([sum(p.bs.interior(u, b, p.bs.nvar)[n].sum()
                                  for b in range(p.bs.nblocks)) for n in range(p.bs.nvar)])
    t0 = tot(st.u0)
    dt = st.estimate_dt()
    for _ in range(3):
        st.step(dt)
    t1 = tot(st.u0)
    for n in (0, 5 + 2, 5 + 3, 5 + 6, 5 + 7):  # mass, O16, Ne20, Ye, lset0
        assert abs(t1[n] - t0[n]) <= 1e-12 * np.abs(t0[n]).max() * 2048, n
