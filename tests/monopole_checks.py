"""Shared checks of monopole gravity results against the reference's own regression
criteria (used by the oracle pins and the GPU tests).

* The gravity accuracy test (tst/regression/test_suites/gravity/gravity.py:52-98,
  186-200): the analytic potential and radial acceleration of the sphere
  rho = rho0 (1 - r^2/r0^2)^2 (Mandal, Mukherjee & Mignone 2023, sec. 4.1); the L2 norms
  of phi - phi_exact and g_x - g_x,exact over the interior cells.
* Evrard's collapse (tst/regression/test_suites/evrards_collapse/evrards_collapse.py:
  515-528, 556-727): volume-weighted radial profiles of rho, v_r and P/rho^gamma in 80
  geometric bins, their RMS against the exact profiles (tests/golden/evrard_exact.txt),
  and the collapse signatures.
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def analytic_phi(r, rho0=1.0, r0=0.25, G=1.0):
    r = np.asarray(r, dtype=float)
    mtot = 32.0 * np.pi * rho0 * r0 ** 3 / 105.0
    out = -G * mtot / np.where(r > 0.0, r, 1.0)
    q = r[r <= r0]
    out[r <= r0] = (-2.0 / 3.0 * np.pi * G * rho0 * r0 ** 2 +
                    4.0 * np.pi * G * rho0 * (q ** 2 / 6.0 - q ** 4 / (10.0 * r0 ** 2) +
                                              q ** 6 / (42.0 * r0 ** 4)))
    return out


def analytic_g(r, rho0=1.0, r0=0.25, G=1.0):
    r = np.asarray(r, dtype=float)
    mtot = 32.0 * np.pi * rho0 * r0 ** 3 / 105.0
    out = -G * mtot / np.where(r > 0.0, r * r, 1.0)
    q = r[r <= r0]
    out[r <= r0] = -4.0 * np.pi * G * rho0 * (q / 3.0 - 2.0 / 5.0 * q ** 3 / r0 ** 2 +
                                              q ** 5 / (7.0 * r0 ** 4))
    return out


def interior(bs, field, nv, comp):
    """Concatenated interior values of component comp of a pack-layout field."""
    return np.concatenate([bs.interior(field, b, nv)[comp].ravel()
                           for b in range(bs.nblocks)])


def centres(bs):
    xs, ys, zs = [], [], []
    for b in range(bs.nblocks):
        x, y, z = bs.cell_centers(b)
        xs.append(x.ravel())
        ys.append(y.ravel())
        zs.append(z.ravel())
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(zs)


def sphere_errors(bs, grav, phi, rho0=1.0, r0=0.25, G=1.0):
    """(phi L2, g_x L2) of a gravity field against the sphere's analytic solution."""
    return sphere_errors_max(bs, grav, phi, rho0, r0, G)[:2]


def sphere_errors_max(bs, grav, phi, rho0=1.0, r0=0.25, G=1.0, center=(0.0, 0.0, 0.0)):
    """(phi L2, g_x L2, phi max, g_x max) of a gravity field against the sphere's analytic
    solution about `center` (gravity.py's get_errors: the Poisson runs' sphere sits at
    X0_OFF = 0.01): gravity.py's l2_norm (the mean over every cell of every block,
    unweighted) and max |error| (its REFINED_LIMITS quantities, gravity.py:48-53,
    169-200, 348-372)."""
    x, y, z = centres(bs)
    x, y, z = x - center[0], y - center[1], z - center[2]
    r = np.sqrt(x * x + y * y + z * z)
    ph = interior(bs, phi, 1, 0) - analytic_phi(r, rho0, r0, G)
    gr = analytic_g(r, rho0, r0, G)
    gx_exact = np.where(r > 0.0, gr * x / np.where(r > 0.0, r, 1.0), 0.0)
    gx = interior(bs, grav, 3, 0) - gx_exact
    return (float(np.sqrt(np.mean(ph ** 2))), float(np.sqrt(np.mean(gx ** 2))),
            float(np.max(np.abs(ph))), float(np.max(np.abs(gx))))


def evrard_exact(gamma=5.0 / 3.0):
    d = np.loadtxt(os.path.join(GOLDEN, "evrard_exact.txt"), comments="#")
    return {"r": d[:, 0], "rho": d[:, 1], "vr": d[:, 2], "entropy": d[:, 3] / d[:, 1] ** gamma}


def evrard_profiles(bs, prim, nbins=80, gamma=5.0 / 3.0, radius=1.0):
    x, y, z = centres(bs)
    rho, v1, v2, v3, p = (interior(bs, prim, bs.nvar, c) for c in range(5))
    r = np.sqrt(x * x + y * y + z * z)
    vr = np.where(r > 0.0, (v1 * x + v2 * y + v3 * z) / np.where(r > 0.0, r, 1.0), 0.0)
    vol = np.concatenate([np.full(int(np.prod(bs.blocks["nx"][b])), bs.blocks["vol"][b])
                          for b in range(bs.nblocks)])
    rr = r / radius
    ok = (rr > 0.0) & np.isfinite(rr) & (vol > 0.0)
    edges = np.geomspace(max(rr[ok].min(), 1e-3), min(rr[ok].max(), 2.0), nbins + 1)
    which = np.digitize(rr, edges) - 1
    out = {"r": np.sqrt(edges[:-1] * edges[1:])}
    for name, val in (("rho", rho), ("vr", vr), ("entropy", p / rho ** gamma)):
        good = ok & (which >= 0) & (which < nbins) & np.isfinite(val)
        if name != "vr":
            good &= val > 0.0
        ws = np.bincount(which[good], weights=val[good] * vol[good], minlength=nbins)
        w = np.bincount(which[good], weights=vol[good], minlength=nbins)
        prof = np.full(nbins, np.nan)
        prof[w > 0] = ws[w > 0] / w[w > 0]
        out[name] = prof
    return out


def evrard_verdict(prof, exact):
    """The reference's checks; returns (failures, RMS per quantity)."""
    limits = {"rho": 100.0, "vr": 1.5, "entropy": 1000.0}
    fails, rms = [], {}
    for q in ("rho", "vr", "entropy"):
        y = prof[q]
        fin = np.isfinite(y)
        if np.count_nonzero(fin) < 20:
            fails.append(f"{q}: {np.count_nonzero(fin)} finite bins")
        rng = fin & (prof["r"] >= exact["r"][0]) & (prof["r"] <= exact["r"][-1])
        rms[q] = float(np.sqrt(np.mean((y[rng] - np.interp(prof["r"][rng], exact["r"],
                                                             exact[q])) ** 2)))
        if not np.isfinite(rms[q]) or rms[q] > limits[q]:
            fails.append(f"{q}: rms {rms[q]}")
    r, vr = prof["r"], prof["vr"]
    infall = (r > 0.05) & (r < 0.6) & np.isfinite(vr)
    if not infall.any() or np.nanmin(vr[infall]) > -0.2:
        fails.append("no inward collapse")
    rebound = (r > 0.8) & (r < 1.6) & np.isfinite(vr)
    if not rebound.any() or np.nanmax(vr[rebound]) < 0.02:
        fails.append("no outer positive v_r")
    return fails, rms
