"""Output formats of the reference used for cross-validation.

* linear_wave_errors / write_linear_wave_errors: the L1 / max error table that
  linear_wave.cpp's UserWorkAfterLoop computes (linear_wave.cpp:164-301) and the exact
  text layout of its ``linearwave-errors.dat``, which the reference's convergence test
  reads (convergence.py:120-132).
* HstOutput: the history output (<parthenon/outputN> file_type = hst, e.g.
  inputs/evrards_collapse.in:94-97) of the sums the hydro package registers
  (GlobalQuantHst total_mass, total_energy, hydro.cpp:336-346).  The file layout is
  Parthenon's history writer (upstream outputs/history.cpp, not vendored in the reference):
  "<basename>.out<N>.hst", a "#  History data" line, one header line of "[i]=label"
  columns (time, dt, cycle, nbtotal, then the registered labels), and one row per output
  time with each real printed by the output's data_format (default "%14.6e") after a
  blank.  Restated from the upstream writer's published source; parity unpinned (the
  reference holds no .hst file or reader to check it against).
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


class HstOutput:
    """One history output block: writes a row whenever the simulation time reaches the
    next output time (the first at the initial time), then advances it by dt.

    basename / block_number / dt / data_format are the output block's file_basename
    ("parthenon"), its number N (outputN), dt and data_format."""

    def __init__(self, directory, dt, block_number=0, basename="parthenon",
                 labels=("total_mass", "total_energy"), data_format="%14.6e"):
        self.path = os.path.join(directory, f"{basename}.out{int(block_number)}.hst")
        self.dt = float(dt)
        self.labels = tuple(labels)
        self.fmt = " " + data_format  # a blank separates the columns
        self.next_time = None
        self.file_number = 0

    def due(self, time):
        return self.next_time is None or time >= self.next_time

    def write(self, time, dt, ncycle, nbtotal, values):
        """Append one row (and the header before the first)."""
        with open(self.path, "a") as f:
            if self.file_number == 0:
                f.write("#  History data\n")
                cols = ["time", "dt", "cycle", "nbtotal"]
                f.write("# [1]=time     [2]=dt       [3]=cycle    [4]=nbtotal  ")
                for i, lbl in enumerate(self.labels):
                    f.write("[%d]=%-8s" % (len(cols) + 1 + i, lbl))
                f.write("\n")
            f.write(self.fmt % time)
            f.write(self.fmt % dt)
            f.write("  %d " % ncycle)
            f.write("  %d " % nbtotal)
            for v in values:
                f.write(self.fmt % v)
            f.write("\n")
        self.file_number += 1
        self.next_time = (time if self.next_time is None else self.next_time) + self.dt

    def maybe_write(self, time, dt, ncycle, nbtotal, values_fn):
        """write() if an output is due at `time`; values_fn() gives the sums."""
        if self.due(time):
            self.write(time, dt, ncycle, nbtotal, values_fn())
            return True
        return False
