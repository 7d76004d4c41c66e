"""Parity metric of the hydro flux path (test infrastructure).

A GPU state passes against the oracle when, on every interior cell and variable,
    |U_gpu - U_ref| <= tol * max(|U_ref|, s_v)
where s_v is the largest |U_ref| of variable v over the cell's block (so momentum
components that pass through zero are compared on the block's scale).
"""
from __future__ import annotations

import numpy as np


def worst_rel(got, ref, bs, nv=None):
    """Largest |got-ref| / max(|ref|, s_v) over interior cells (0.0 if identical)."""
    nv = nv or bs.nvar
    worst = 0.0
    for b in range(bs.nblocks):
        g = bs.interior(got, b, nv)
        r = bs.interior(ref, b, nv)
        for v in range(nv):
            rv = r[v]
            s = np.max(np.abs(rv)) if rv.size else 0.0
            scale = np.maximum(np.abs(rv), s)
            d = np.abs(g[v] - rv)
            with np.errstate(divide="ignore", invalid="ignore"):
                q = np.where(d == 0.0, 0.0, d / np.where(scale > 0, scale, 1.0))
            if not np.all(np.isfinite(g[v])):
                return float("inf")
            worst = max(worst, float(q.max()) if q.size else 0.0)
    return worst


def bitwise_equal_interior(got, ref, bs, nv=None):
    nv = nv or bs.nvar
    for b in range(bs.nblocks):
        if not np.array_equal(bs.interior(got, b, nv), bs.interior(ref, b, nv)):
            return False
    return True
