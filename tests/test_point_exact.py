"""CPU checks of the exact reformulations of the PPM arithmetic used by the device march
(apophis_amd/csrc/aph_point.hpp, host build in tests/native):

* div6_fast(x) (one multiply, two fma) equals x / 6.0 bitwise for x = +-0 and
  2^-1019 <= |x| <= 2^1020;
* ppm_face<true> (div6_fast) equals ppm_face<false> (true divisions) bitwise on windows
  whose values are +-0 or 2^-900 <= |q| < 2^1000;
* ppm_cell(q, ppm_face(q[0..3]), ppm_face(q[1..4])) equals the oracle's ppm(q)
  (ppm_simple.hpp:31-161) bitwise, i.e. the limited face value is the same whichever
  adjacent cell computes it.
"""
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


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def test_div6_fast_range(lib):
    rng = np.random.default_rng(11)
    n = 4_000_000
    sig = rng.integers(0, 2 ** 52, size=n, dtype=np.uint64)
    exp = rng.integers(1023 - 1019, 1023 + 1021, size=n, dtype=np.uint64)  # 2^-1019..2^1020
    sign = rng.integers(0, 2, size=n, dtype=np.uint64) << np.uint64(63)
    x = (sign | (exp << np.uint64(52)) | sig).view(np.float64)
    x = x[np.abs(x) <= 2.0 ** 1020]
    x = np.concatenate([x, [0.0, -0.0, 6.0, -6.0, 3.0, 2.0 ** -1019, 2.0 ** 1020,
                            6.0 * 2.0 ** -1019, np.nextafter(2.0 ** -1019, 1.0)],
                        rng.integers(-10 ** 6, 10 ** 6, size=100_000) * 1.0,  # exact quotients
                        rng.normal(size=100_000)])
    out = np.empty_like(x)
    lib.chk_div6_fast(_p(x), ctypes.c_long(x.size), _p(out))
    ref = x / 6.0
    bad = _bits(out) != _bits(ref)
    assert not bad.any(), (x[bad][:4], out[bad][:4], ref[bad][:4])


def _plain_windows(rng, n):
    """Windows of four values, each +-0 or 2^-900 <= |q| < 2^1000, drawn to reach the
    limiter's branches and the range ends."""
    kind = rng.integers(0, 5, size=n)
    q = rng.normal(size=(n, 4))
    smooth = kind == 0  # near-quadratic data: limiter mostly inactive
    t = np.arange(4) - 1.5
    c = rng.normal(size=(n, 3))
    q[smooth] = (c[:, :1] + c[:, 1:2] * t + c[:, 2:3] * t * t)[smooth] \
        + 1e-7 * rng.normal(size=(n, 4))[smooth]
    q[kind == 1] = np.where(rng.random((n, 4)) < 0.4, 0.0, q)[kind == 1]  # zeros
    lo = kind == 2  # near the lower end of the range
    q[lo] = rng.choice([-1.0, 1.0], size=(n, 4))[lo] * 2.0 ** rng.uniform(-900, -880, (n, 4))[lo]
    hi = kind == 3  # near the upper end
    q[hi] = rng.choice([-1.0, 1.0], size=(n, 4))[hi] * 2.0 ** rng.uniform(980, 999.9, (n, 4))[hi]
    mixed = kind == 4  # wildly mixed scales
    q[mixed] = (rng.choice([-1.0, 1.0], size=(n, 4)) * rng.uniform(1.0, 2.0, (n, 4))
                * 2.0 ** rng.integers(-900, 999, (n, 4)).astype(float))[mixed]
    return q


def test_fast_face_matches_exact_on_plain_windows(lib):
    rng = np.random.default_rng(12)
    n = 2_000_000
    q = np.ascontiguousarray(_plain_windows(rng, n))
    a = np.abs(q)
    assert ((a == 0) | ((a >= 2.0 ** -900) & (a < 2.0 ** 1000))).all()
    fast = np.empty(n)
    slow = np.empty(n)
    lib.chk_ppm_face(_p(q), ctypes.c_long(n), 1, _p(fast))
    lib.chk_ppm_face(_p(q), ctypes.c_long(n), 0, _p(slow))
    bad = _bits(fast) != _bits(slow)
    assert not bad.any(), (q[bad][:3], fast[bad][:3], slow[bad][:3])


def test_face_shared_ppm_is_bitwise_ppm(lib):
    rng = np.random.default_rng(9)
    for trial in range(20000):
        if trial % 2:
            t = np.arange(5) - 2.0
            c = rng.normal(size=3)
            q = c[0] + c[1] * t + c[2] * t * t + 1e-6 * rng.normal(size=5)
        else:
            q = rng.normal(size=5) * rng.choice([1e-8, 1.0, 1e8])
        out = np.zeros(2)
        lib.chk_ppm_shared(_p(np.ascontiguousarray(q)), _p(out))
        assert tuple(out) == orc.ppm(*q)


def _pencils(rng):
    """Pencils of the kinds the march meets: smooth, noisy, plateaus and exact extrema,
    jumps, zeros and sign changes, and values a few ulps apart (the adiabatic-index
    fields), at several scales."""
    n = 40
    x = np.arange(n) / n
    kinds = [
        np.sin(2 * np.pi * x * rng.uniform(0.5, 3)) * rng.uniform(0.1, 10),
        1.0 + 0.01 * rng.uniform(-1, 1, n),
        np.where(x < rng.uniform(0.2, 0.8), 1.0, 0.125),
        np.repeat(rng.normal(size=n // 4), 4),
        rng.normal(size=n) * (rng.random(n) < 0.5),
        (5.0 / 3.0) * (1 + rng.integers(-4, 5, n) * 2.0 ** -52),
        np.round(rng.normal(size=n) * 4) / 4,
        rng.normal(size=n) * rng.choice([1e-8, 1.0, 1e8]),
    ]
    return kinds


@pytest.mark.parametrize("fast", [0, 1])
def test_ppm_march_sharing_matches_oracle_bitwise(lib, fast):
    """The device march's form of PPM -- each face's limited value formed once for both
    cells beside it, each cell's second difference formed once and carried to the next
    step, the limiter-only arithmetic formed only where its branch is taken -- equals the
    oracle's ppm (ppm_simple.hpp:31-161) cell by cell, bit for bit."""
    rng = np.random.default_rng(29)
    for _ in range(300):
        for q in _pencils(rng):
            q = np.ascontiguousarray(q, dtype=np.float64)
            out = np.zeros(2 * (q.size - 4))
            lib.chk_ppm_march(_p(q), ctypes.c_long(q.size), fast, _p(out))
            want = np.array([orc.ppm(*q[i - 2:i + 3]) for i in range(2, q.size - 2)]).ravel()
            assert _bits(out).tolist() == _bits(want).tolist()


def _face_states(rng):
    """Left/right primitive states and adiabatic indices covering both signs of the
    contact speed, one- and two-sided shocks (pstar above one or both pressures),
    rarefactions, and velocity magnitudes on both sides of 1 and 2 (LHLLC's phi)."""
    wl = np.array([rng.uniform(0.1, 3), rng.normal(0, 2), rng.normal(), rng.normal(),
                   rng.uniform(0.05, 5)])
    wr = np.array([rng.uniform(0.1, 3), rng.normal(0, 2), rng.normal(), rng.normal(),
                   rng.uniform(0.05, 5)])
    if rng.random() < 0.3:  # low-Mach regime
        wl[1:4] *= 0.01
        wr[1:4] *= 0.01
    g = rng.choice([1.4, 5.0 / 3.0])
    gam = g * (1 + 1e-14 * rng.normal(size=4))
    return wl, wr, gam[:2].copy(), gam[2:].copy()


@pytest.mark.parametrize("kind,code", [("hllc", 3), ("hlle", 2), ("lhllc", 4)])
def test_riemann_matches_oracle_bitwise(lib, kind, code):
    """The device Riemann solvers (operands selected before dividing, shared quotients)
    against the oracle's restatement of hydro_hllc.hpp / hydro_hlle.hpp."""
    rng = np.random.default_rng(13)
    for _ in range(20000):
        wl, wr, gl, gr = _face_states(rng)
        x = np.ascontiguousarray(np.concatenate([wl, wr, gl, gr]))
        out = np.zeros(5)
        lib.chk_riemann(code, _p(x), _p(out))
        ref = orc.riemann(kind, wl, wr, gl, gr)
        assert (_bits(out) == _bits(ref)).all(), (x, out, ref)


def test_floor_cons_matches_cons_to_prim(lib):
    """floor_cons_pt leaves the conserved state exactly as cons_to_prim_pt does, with and
    without the density and pressure floors firing."""
    rng = np.random.default_rng(14)
    eos = np.array([2.0 / 3.0, 1.5, 1e-3, 1e-2])
    fired = 0
    for _ in range(20000):
        rho = rng.choice([rng.uniform(1e-4, 2e-3), rng.uniform(0.1, 3)])
        v = rng.normal(size=3)
        p = rng.choice([rng.uniform(-1e-2, 2e-2), rng.uniform(0.1, 3)])
        u = np.array([rho, rho * v[0], rho * v[1], rho * v[2],
                      p / eos[0] + 0.5 * rho * (v ** 2).sum()])
        out = np.zeros(10)
        lib.chk_floor(_p(u), _p(eos), _p(out))
        assert (_bits(out[:5]) == _bits(out[5:])).all(), (u, out)
        fired += out[4] != u[4]
    assert fired > 1000  # the pressure-floor branch was exercised


def test_ppm_limiter_is_discontinuous_plm_is_lipschitz():
    """Along a line in input space the reference PPM (ppm_simple.hpp) jumps by ~1 where a
    limiter branch switches -- the jump does not shrink with the sampling step -- while
    PLM's van Leer limiter is Lipschitz.  Central finite differences of a PPM rollout
    therefore do not converge to the (exact, active-branch) adjoint gradient when a jump
    lies within +-h; gradients of PPM runs are checked against autograd instead."""
    q0 = np.array([0.3, -1.2, 0.7, 0.9, -0.4])
    d = np.array([1.0, 0.2, -0.5, 0.3, 0.8])
    jumps = []
    for n in (2001, 20001, 200001):
        t = np.linspace(-3.0, 3.0, n)
        v = np.array([orc.ppm(*(q0 + s * d)) for s in t])
        jumps.append(np.abs(np.diff(v, axis=0)).max())
    assert min(jumps) > 0.9 and max(jumps) - min(jumps) < 0.01, jumps
    t = np.linspace(-3.0, 3.0, 200001)
    v = np.array([orc.plm(*(q0[:3] + s * d[:3])) for s in t])
    assert np.abs(np.diff(v, axis=0)).max() / (t[1] - t[0]) < 10.0
