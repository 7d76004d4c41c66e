"""Output formats of the reference used for cross-validation.

* linear_wave_errors / write_linear_wave_errors: the L1 / max error table that
  linear_wave.cpp's UserWorkAfterLoop computes (linear_wave.cpp:164-301) and the exact
  text layout of its ``linearwave-errors.dat``, which the reference's convergence test
  reads (convergence.py:120-132).
* HistoryFile: the history sums the hydro package registers (GlobalQuantHst total_mass,
  total_energy, hydro.cpp:336-346) written per step as a whitespace table.  The column
  layout of Parthenon's own .hst writer is upstream and not restated bit for bit.
"""
from __future__ import annotations

import os

import numpy as np


def linear_wave_errors(bs, cons, exact, domain_volume, nhydro=5):
    """l1[v] = sum |U - U_exact| V / domain volume, max[v] = max |U - U_exact|, the RMS of
    the l1 errors and the largest max/l1 ratio, over the interior of every block.
    exact: per block arrays (nhydro, nz, ny, nx) of the analytic state."""
    l1 = np.zeros(nhydro)
    mx = np.zeros(nhydro)
    for b in range(bs.nblocks):
        u = bs.interior(cons, b, bs.nvar)[:nhydro]
        d = np.abs(u - exact[b][:nhydro])
        vol = float(bs.blocks["vol"][b])
        l1 += (d * vol).reshape(nhydro, -1).sum(axis=1)
        mx = np.maximum(mx, d.reshape(nhydro, -1).max(axis=1))
    l1 = l1 / domain_volume
    rms = float(np.sqrt((l1 ** 2).sum()))
    max_over_l1 = float(np.max(mx / l1))
    return {"l1": l1, "max": mx, "rms": rms, "max_over_l1": max_over_l1}


def write_linear_wave_errors(path, nx, ncycle, err):
    """Append one row to `path` in linearwave-errors.dat's layout (header when new)."""
    new = not os.path.exists(path)
    with open(path, "a") as f:
        if new:
            f.write("# Nx1  Nx2  Nx3  Ncycle  ")
            f.write("RMS-L1-Error  d_L1  M1_L1  M2_L1  M3_L1  E_L1 ")
            f.write("  Largest-Max/L1  d_max  M1_max  M2_max  M3_max  E_max ")
            f.write("\n")
        l1, m = err["l1"], err["max"]
        f.write("%d  %d" % (nx[0], nx[1]))
        f.write("  %d  %d" % (nx[2], ncycle))
        f.write("  %e  %e" % (err["rms"], l1[0]))
        f.write("  %e  %e  %e" % (l1[1], l1[2], l1[3]))
        f.write("  %e" % l1[4])
        f.write("  %e  %e  " % (err["max_over_l1"], m[0]))
        f.write("%e  %e  %e" % (m[1], m[2], m[3]))
        f.write("  %e" % m[4])
        f.write("\n")


class HistoryFile:
    """time, dt and the registered history sums, one row per call of write()."""

    def __init__(self, path, names=("total_mass", "total_energy")):
        self.path = path
        self.names = tuple(names)
        with open(path, "w") as f:
            cols = ["time", "dt"] + list(self.names)
            f.write("# History data\n")
            f.write("# " + "  ".join(f"[{i + 1}]={c}" for i, c in enumerate(cols)) + "\n")

    def write(self, time, dt, values):
        with open(self.path, "a") as f:
            f.write("  ".join("%.14e" % v for v in [time, dt] + list(values)) + "\n")
