"""ctypes front end of the CPU oracle (oracle/hydro_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libhydro_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def build(force=False):
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_plm.argtypes = [ctypes.c_double] * 3 + [_dp, _dp]
        L.orc_ppm.argtypes = [ctypes.c_double] * 5 + [_dp, _dp]
        for f in (L.orc_hlle, L.orc_hllc, L.orc_lhllc):
            f.argtypes = [_dp] * 5
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_cons_to_prim.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                                       _dp, _dp, _dp]
        L.orc_calculate_fluxes.argtypes = [_vp] + [ctypes.c_int] * 5 + [_dp] * 5
        L.orc_calculate_fluxes.restype = ctypes.c_int
        L.orc_update.argtypes = [_vp] + [ctypes.c_int] * 3 + [_dp] * 5 + [ctypes.c_double] * 3
        L.orc_apply_gravity.argtypes = [_vp] + [ctypes.c_int] * 3 + [_dp, _dp, _dp,
                                                                     ctypes.c_double]
        L.orc_fill_ghosts.argtypes = [_vp] + [ctypes.c_int] * 3 + [_dp]
        L.orc_estimate_timestep.argtypes = [_vp] + [ctypes.c_int] * 3 + [_vp, _dp, _dp,
                                                                         ctypes.c_double]
        L.orc_estimate_timestep.restype = ctypes.c_double
        L.orc_global_sum.argtypes = [_vp] + [ctypes.c_int] * 3 + [_dp, ctypes.c_int]
        L.orc_global_sum.restype = ctypes.c_double
        L.orc_rk2_step.argtypes = ([_vp] + [ctypes.c_int] * 5 + [_vp, _dp, ctypes.c_double,
                                                                 ctypes.c_double] + [_dp] * 7)
        L.orc_rk2_step.restype = ctypes.c_double
        L.orc_step.argtypes = ([_vp] + [ctypes.c_int] * 6 + [_vp, _dp, ctypes.c_double,
                                                             ctypes.c_double] + [_dp] * 7)
        L.orc_step.restype = ctypes.c_double
        L.orc_integrator_stages.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_integrator_stages.restype = ctypes.c_int
        L.orc_mono_scale.argtypes = [ctypes.c_int, ctypes.c_long]
        L.orc_mono_scale.restype = ctypes.c_int
        L.orc_mono_bin_masses.argtypes = ([_vp] + [ctypes.c_int] * 3 +
                                          [_dp, ctypes.c_int, ctypes.c_double, _dp])
        L.orc_mono_bin_masses.restype = ctypes.c_int
        L.orc_mono_tables.argtypes = ([ctypes.c_int] * 2 + [ctypes.c_double] * 2 +
                                      [_dp] * 3)
        L.orc_mono_point.argtypes = ([ctypes.c_int] * 2 + [ctypes.c_double] * 2 + [_dp] * 2 +
                                     [ctypes.c_double] * 3 + [_dp] * 2)
        L.orc_mono_field.argtypes = ([_vp] + [ctypes.c_int] * 3 + [ctypes.c_double] * 2 +
                                     [_dp] * 4)
        L.orc_apply_gravity_field.argtypes = ([_vp] + [ctypes.c_int] * 3 + [_dp] * 3 +
                                              [ctypes.c_double])
        L.orc_step_monopole.argtypes = ([_vp] + [ctypes.c_int] * 6 + [_vp, ctypes.c_int] +
                                        [ctypes.c_double] * 4 + [_dp] * 10)
        L.orc_step_monopole.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _bp(bs):
    return bs.blocks.ctypes.data_as(_vp)


# ------------------------------------------------------------ point functions
def plm(qm1, q0, qp1):
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().orc_plm(qm1, q0, qp1, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value  # (ql_ip1, qr_i)


def ppm(qm2, qm1, q0, qp1, qp2):
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().orc_ppm(qm2, qm1, q0, qp1, qp2, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value  # (ql_ip1, qr_i)


def riemann(kind, wl, wr, ifl, ifr):
    wl, wr = np.ascontiguousarray(wl, float), np.ascontiguousarray(wr, float)
    ifl, ifr = np.ascontiguousarray(ifl, float), np.ascontiguousarray(ifr, float)
    out = np.zeros(5)
    f = {"hllc": lib().orc_hllc, "lhllc": lib().orc_lhllc, "hlle": lib().orc_hlle}[kind]
    f(_p(wl), _p(wr), _p(ifl), _p(ifr), _p(out))
    return out


def integrator_stages(name):
    """(gam0, gam1, beta) of the named integrator's stages."""
    g0, g1, be = np.zeros(3), np.zeros(3), np.zeros(3)
    n = lib().orc_integrator_stages(INTEGRATOR[name], _p(g0), _p(g1), _p(be))
    return g0[:n], g1[:n], be[:n]


# ------------------------------------------------------------ mesh functions
RECON = {"dc": 1, "plm": 2, "ppm": 3}
RIEMANN = {"none": 1, "hlle": 2, "hllc": 3, "lhllc": 4}
INTEGRATOR = {"rk1": 1, "rk2": 2, "rk3": 3, "vl2": 4}


def set_threads(n):
    lib().orc_set_threads(int(n))


class State:
    """Oracle-side state of a problem: u0, u1, prim, gamma and flux scratch."""

    def __init__(self, prob):
        self.prob = prob
        bs = prob.bs
        self.bs = bs
        nv = bs.nvar
        self.u0 = prob.cons.copy()
        self.u1 = np.zeros_like(self.u0)
        self.prim = np.zeros_like(self.u0)
        self.gamma = np.zeros(bs.field_size(2))
        self.fx = [np.zeros_like(self.u0) for _ in range(3)]
        self.g = np.ascontiguousarray(prob.g, dtype=np.float64)
        self.mono = getattr(prob, "monopole", None)  # (nrbin, rmax, G) or None
        self.pois = getattr(prob, "poisson", None)  # (G, alpha, tol, max_it) or None
        if self.pois is not None:
            self.grav = np.zeros(bs.field_size(3))
        if self.mono is not None:
            self.grav = np.zeros(bs.field_size(3))
            self.m_enc = np.zeros(int(self.mono[0]) + 1)
            self.phi_rad = np.zeros(int(self.mono[0]) + 1)
        fill_ghosts(bs, self.u0)
        cons_to_prim(bs, prob.eos, self.u0, self.prim, self.gamma)

    def estimate_dt(self):
        return estimate_timestep(self.bs, self.prob.eos, self.prim, self.u0, self.prob.cfl)

    def step(self, dt):
        p = self.prob
        need = {"dc": 1, "plm": 2, "ppm": 3}[p.recon]
        if int(self.bs.blocks["ng"].min()) < need:
            raise ValueError("Not enough ghost zones for reconstruction")
        if self.pois is not None:
            return self._step_poisson(dt)
        if self.mono is not None:
            nrbin, rmax, G = self.mono
            return lib().orc_step_monopole(
                _bp(self.bs), self.bs.nblocks, self.bs.ndim, self.bs.nvar,
                INTEGRATOR[p.integrator], RECON[p.recon], RIEMANN[p.riemann],
                p.eos.ctypes.data_as(_vp), int(nrbin), float(rmax), float(G), p.cfl, dt,
                _p(self.u0), _p(self.u1), _p(self.prim), _p(self.gamma), _p(self.fx[0]),
                _p(self.fx[1]), _p(self.fx[2]), _p(self.grav), _p(self.m_enc),
                _p(self.phi_rad))
        return lib().orc_step(_bp(self.bs), self.bs.nblocks, self.bs.ndim, self.bs.nvar,
                              INTEGRATOR[p.integrator], RECON[p.recon], RIEMANN[p.riemann],
                              p.eos.ctypes.data_as(_vp), _p(self.g), p.cfl, dt,
                              _p(self.u0), _p(self.u1), _p(self.prim), _p(self.gamma),
                              _p(self.fx[0]), _p(self.fx[1]), _p(self.fx[2]))

    def _step_poisson(self, dt):
        """One step with Poisson gravity (the exact discrete solution of each stage's
        problem, oracle/poisson_oracle.py) between the flux update and the ghost fill, as
        orc_step_monopole orders the tasks."""
        from . import poisson_oracle as P

        p, bs = self.prob, self.bs
        G, alpha = self.pois[0], self.pois[1]
        g0, g1, be = integrator_stages(p.integrator)
        self.u1[:] = self.u0
        for s in range(len(g0)):
            rec = "dc" if (s == 0 and p.integrator == "vl2") else p.recon
            _, _, g = P.gravity(bs, self.prim, G, alpha)
            for d in range(3):
                P.scatter(bs, g[d], self.grav, 3, d)
            calculate_fluxes(bs, rec, p.riemann, self.prim, self.gamma, *self.fx)
            update(bs, self.u0, self.u1, *self.fx, g0[s], g1[s], be[s] * dt)
            apply_gravity_field(bs, self.u0, self.prim, self.grav, be[s] * dt)
            fill_ghosts(bs, self.u0)
            cons_to_prim(bs, p.eos, self.u0, self.prim, self.gamma)
        return self.estimate_dt()

    def run(self, nsteps=None, tlim=None, fixed_dt=None):
        """Evolve with the upstream dt policy (<= 2x growth, clip to tlim)."""
        t = 0.0
        dt_est = self.estimate_dt()
        dt = np.inf
        n = 0
        while True:
            if nsteps is not None and n >= nsteps:
                break
            if tlim is not None and t >= tlim * (1 - 1e-14):
                break
            if fixed_dt is not None:
                dt = fixed_dt
            else:
                dt = min(2.0 * dt, dt_est)
                if tlim is not None and t < tlim and tlim - t < dt:
                    dt = tlim - t
            dt_est = self.step(dt)
            t += dt
            n += 1
        return t, n


def cons_to_prim(bs, eos, cons, prim, gamma):
    lib().orc_cons_to_prim(_bp(bs), bs.nblocks, bs.ndim, bs.nvar, eos.ctypes.data_as(_vp),
                           _p(cons), _p(prim), _p(gamma))


def calculate_fluxes(bs, recon, riemann_, prim, gamma, fx1, fx2, fx3):
    rc = lib().orc_calculate_fluxes(_bp(bs), bs.nblocks, bs.ndim, bs.nvar, RECON[recon],
                                    RIEMANN[riemann_], _p(prim), _p(gamma), _p(fx1),
                                    _p(fx2), _p(fx3))
    if rc == -2:
        raise ValueError("Not enough ghost zones for reconstruction")
    if rc != 0:
        raise ValueError("unsupported reconstruction/Riemann combination")


def update(bs, u0, u1, fx1, fx2, fx3, gam0, gam1, beta_dt):
    lib().orc_update(_bp(bs), bs.nblocks, bs.ndim, bs.nvar, _p(u0), _p(u1), _p(fx1),
                     _p(fx2), _p(fx3), gam0, gam1, beta_dt)


def apply_gravity(bs, cons, prim, g, dt):
    g = np.ascontiguousarray(g, dtype=np.float64)
    lib().orc_apply_gravity(_bp(bs), bs.nblocks, bs.ndim, bs.nvar, _p(cons), _p(prim),
                            _p(g), dt)


def fill_ghosts(bs, cons):
    lib().orc_fill_ghosts(_bp(bs), bs.nblocks, bs.ndim, bs.nvar, _p(cons))


def estimate_timestep(bs, eos, prim, cons, cfl):
    return lib().orc_estimate_timestep(_bp(bs), bs.nblocks, bs.ndim, bs.nvar,
                                       eos.ctypes.data_as(_vp), _p(prim), _p(cons), cfl)


def global_sum(bs, cons, var):
    return lib().orc_global_sum(_bp(bs), bs.nblocks, bs.ndim, bs.nvar, _p(cons), var)


# ------------------------------------------------------------ monopole gravity
def mono_scale(emax, ncells):
    return lib().orc_mono_scale(int(emax), int(ncells))


def mono_bin_masses(bs, prim, nrbin, rmax):
    """(bin masses, fixed-point scale s) of AccumulateMass over every block."""
    mbin = np.zeros(int(nrbin))
    s = lib().orc_mono_bin_masses(_bp(bs), bs.nblocks, bs.ndim, bs.nvar, _p(prim),
                                  int(nrbin), float(rmax), _p(mbin))
    return mbin, s


def mono_tables(nrbin, ndim, rmax, G, mbin):
    m_enc, phi_rad = np.zeros(int(nrbin) + 1), np.zeros(int(nrbin) + 1)
    lib().orc_mono_tables(int(nrbin), int(ndim), float(rmax), float(G),
                          _p(np.ascontiguousarray(mbin, dtype=np.float64)), _p(m_enc),
                          _p(phi_rad))
    return m_enc, phi_rad


def mono_point(nrbin, ndim, rmax, G, m_enc, phi_rad, x, y, z):
    g, phi = np.zeros(3), np.zeros(1)
    lib().orc_mono_point(int(nrbin), int(ndim), float(rmax), float(G), _p(m_enc),
                         _p(phi_rad), float(x), float(y), float(z), _p(g), _p(phi))
    return g, float(phi[0])


def mono_field(bs, nrbin, rmax, G, m_enc, phi_rad):
    """(gravity field, 3 components; phi, 1 component), interior cells, pack layout."""
    grav, phi = np.zeros(bs.field_size(3)), np.zeros(bs.field_size(1))
    lib().orc_mono_field(_bp(bs), bs.nblocks, bs.ndim, int(nrbin), float(rmax), float(G),
                         _p(m_enc), _p(phi_rad), _p(grav), _p(phi))
    return grav, phi


def apply_gravity_field(bs, cons, prim, grav, dt):
    lib().orc_apply_gravity_field(_bp(bs), bs.nblocks, bs.ndim, bs.nvar, _p(cons),
                                  _p(prim), _p(grav), dt)
