"""CPU checks of the hand-written adjoint point functions (apophis_amd/csrc/aph_adjoint.hpp,
host build in tests/native) against central finite differences of their forward functions:
for random inputs x, directions d and cotangents v,  v . (f(x+hd) - f(x-hd)) / 2h  must equal
(J^T v) . d.  Inputs where a branch boundary lies within the FD stencil (one-sided differences
disagree) are skipped; the test asserts that most samples are checked.  The host build of the
forward functions must also match the oracle bitwise."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "native", "_build", "libadj_check.so")
_dp = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("tests/native not built (run __graft_entry__.build())")
    return ctypes.CDLL(LIB)


def _p(a):
    return a.ctypes.data_as(_dp)


def _call(fn, nout, *args):
    out = np.zeros(nout)
    fn(*[_p(np.ascontiguousarray(a, dtype=float)) if isinstance(a, np.ndarray) else a
         for a in args], _p(out))
    return out


def fd_check(f, adj, x, nout, rng, h_rel=1e-6, n_dirs=3, tol=1e-6):
    """Returns True if checked (and asserts), False if skipped for a nearby kink."""
    v = rng.normal(size=nout)
    xb = adj(x, v)
    for _ in range(n_dirs):
        d = rng.normal(size=x.size) * np.maximum(np.abs(x), 1e-3)
        h = h_rel
        fp, fm, f0 = f(x + h * d), f(x - h * d), f(x)
        c = v @ (fp - fm) / (2 * h)
        fwd = v @ (fp - f0) / h
        bwd = v @ (f0 - fm) / h
        scale = max(abs(c), np.abs(v).max() * np.abs(f0).max() * 1e-8, 1e-300)
        if abs(fwd - bwd) > 1e-4 * scale:
            return False
        a = xb @ d
        assert abs(a - c) <= tol * max(abs(c), abs(a), 1e-8 * scale), (x, d, a, c)
    return True


def test_host_forward_matches_oracle(lib):
    rng = np.random.default_rng(0)
    for _ in range(500):
        q = rng.normal(size=5)
        got = _call(lib.chk_ppm, 2, q)
        assert tuple(got) == orc.ppm(*q)
        got = _call(lib.chk_plm, 2, q[:3])
        assert tuple(got) == orc.plm(*q[:3])


@pytest.mark.parametrize("kind", ["ppm", "plm"])
def test_recon_adjoint_matches_fd(lib, kind):
    rng = np.random.default_rng(1)
    n = 5 if kind == "ppm" else 3
    f = lambda x: _call(getattr(lib, f"chk_{kind}"), 2, x)
    adj = lambda x, v: _call(getattr(lib, f"chk_{kind}_adj"), n, x, v)
    checked = 0
    for trial in range(3000):
        if trial % 3 == 0:  # smooth data through different limiter branches
            c = rng.normal(size=3)
            t = np.arange(n) - n // 2
            x = c[0] + c[1] * t + c[2] * t * t + 1e-3 * rng.normal(size=n)
        else:
            x = rng.normal(size=n)
        checked += fd_check(f, adj, x, 2, rng)
    assert checked > 2000


def _states(rng):
    wl = np.array([rng.uniform(0.3, 3), rng.normal(0, 0.8), rng.normal(), rng.normal(),
                   rng.uniform(0.2, 3)])
    wr = np.array([rng.uniform(0.3, 3), rng.normal(0, 0.8), rng.normal(), rng.normal(),
                   rng.uniform(0.2, 3)])
    g = rng.choice([1.4, 5.0 / 3.0])
    gam = g * (1 + 1e-14 * rng.normal(size=4))
    return np.concatenate([wl, wr, gam[:2], gam[2:]])


@pytest.mark.parametrize("kind,code", [("hllc", 3), ("hlle", 2), ("lhllc", 4)])
def test_riemann_adjoint_matches_fd(lib, kind, code):
    rng = np.random.default_rng(2)
    f = lambda x: _call(lib.chk_riemann, 5, code, x)
    adj = lambda x, v: _call(lib.chk_riemann_adj, 14, code, x, v)
    checked = 0
    for _ in range(1500):
        x = _states(rng)
        checked += fd_check(f, adj, x, 5, rng, tol=2e-6)
    assert checked > 1200


def test_cons_to_prim_adjoint_matches_fd(lib):
    rng = np.random.default_rng(3)
    eos = np.array([2.0 / 3.0, 1.5, 3.4e-18, 3.4e-18])
    f = lambda u: _call(lib.chk_c2p, 8, u, eos)
    adj = lambda u, v: _call(lib.chk_c2p_adj, 6, u, eos, v)
    checked = 0
    for _ in range(1000):
        rho = rng.uniform(0.3, 3)
        vel = rng.normal(size=3)
        p = rng.uniform(0.1, 3)
        u = np.array([rho, rho * vel[0], rho * vel[1], rho * vel[2],
                      p / eos[0] + 0.5 * rho * (vel ** 2).sum(), rho * rng.uniform()])
        checked += fd_check(f, adj, u, 8, rng)
    assert checked > 900



def test_face_shared_ppm_adjoint_matches_ppm_adj(lib):
    """The split reverse (ppm_face_tape / ppm_cell_adj / ppm_face_adj, as the batched
    adjoint's reconstruction kernel composes them) against the one-cell ppm_adj: the same
    branches, so the same cotangents up to rounding; the taped faces are ppm_face bitwise."""
    rng = np.random.default_rng(5)
    worst = 0.0
    for trial in range(4000):
        if trial % 2 == 0:
            c = rng.normal(size=3)
            t = np.arange(5) - 2
            q = c[0] + c[1] * t + c[2] * t * t + 1e-3 * rng.normal(size=5)
        else:
            q = rng.normal(size=5)
        ob = rng.normal(size=2)
        a = _call(lib.chk_ppm_adj, 5, q, ob)
        b = _call(lib.chk_ppm_split_adj, 7, q, ob)
        shared = _call(lib.chk_ppm_shared, 2, q)
        lo = _call(lib.chk_ppm_face, 1, q[:4], ctypes.c_long(1), 0)
        assert b[5] == lo[0]
        assert tuple(shared) == orc.ppm(*q)
        scale = max(np.abs(a).max(), np.abs(ob).max(), 1e-300)
        worst = max(worst, np.abs(a - b[:5]).max() / scale)
    assert worst < 1e-13, worst
