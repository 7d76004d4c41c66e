"""CPU restatement of the reference's Poisson self-gravity (src/gravity/gsolvers/
poisson_gravity.hpp) on a uniform 3D lattice mesh with physical boundaries on every face.

TEST INFRASTRUCTURE: imported only by tests/ as the checker.

* mp_coeffs: ComputeMPCoeff + ScaleMPCoeff (poisson_gravity.hpp:430-500): the nine
  multipole moments of rho V about the origin, each scaled by its constant times G.
* multipole_potential: MultipolePotential (gravity.hpp:45-58).
* rhs: ComputeRhs (:502-589): 4 pi G rho per interior cell, minus 2 phi_b A_f / (dx V) on
  every face of the domain, phi_b the multipole potential at the face centre.
* solve: the discrete problem PoissonEquation defines (:31-388) with D = 1 and
  diagonal_alpha: the 7-point Laplacian over cell centres, minus alpha, with phi = 0 at
  the domain faces (the ghost cell holds minus the interior value, SetZeroDirichlet-
  Boundary), solved exactly by the type-II discrete sine transform, which diagonalises
  that operator (a cell-centred Dirichlet problem's eigenvectors are
  sin(pi (k+1) (i+1/2) / N)).  The reference reaches the same solution iteratively
  (Parthenon's MG / BiCGSTAB) to its residual tolerance.
* field: ComputeGravityVector (:591-625) with the ghost values MultipoleGravityBC writes
  (gravity.cpp:37-104: 2 phi_0(face) - the mirrored interior value): g = -grad phi by
  centred differences.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.fft

SCALE = (-1.0, -1.0, -1.0, -1.0, -3.0, -3.0, -0.25, -3.0, -3.0)  # :479-495


def lattice(bs):
    """(origin, dx, shape (nz, ny, nx), per-block offsets (bz, by, bx)) of a uniform 3D
    lattice mesh (every block the same size)."""
    b0 = bs.blocks[0]
    dx = np.array(b0["dx"], dtype=float)
    nxb = np.array(b0["nx"], dtype=int)
    lo = np.min(np.stack([bs.blocks[b]["xmin"] for b in range(bs.nblocks)]), axis=0)
    hi = np.max(np.stack([bs.blocks[b]["xmin"] + bs.blocks[b]["nx"] * bs.blocks[b]["dx"]
                          for b in range(bs.nblocks)]), axis=0)
    n = np.rint((hi - lo) / dx).astype(int)
    offs = []
    for b in range(bs.nblocks):
        o = np.rint((bs.blocks[b]["xmin"] - lo) / dx).astype(int)
        offs.append(o)
    assert np.prod(n) == bs.nblocks * np.prod(nxb), "blocks must tile a box"
    return lo, dx, (n[2], n[1], n[0]), offs


def gather(bs, field, nv, comp):
    lo, dx, shape, offs = lattice(bs)
    out = np.empty(shape)
    for b in range(bs.nblocks):
        o = offs[b]
        nx = bs.blocks[b]["nx"]
        out[o[2]:o[2] + nx[2], o[1]:o[1] + nx[1], o[0]:o[0] + nx[0]] = \
            bs.interior(field, b, nv)[comp]
    return out


def scatter(bs, arr, field, nv, comp):
    lo, dx, shape, offs = lattice(bs)
    for b in range(bs.nblocks):
        o = offs[b]
        nx = bs.blocks[b]["nx"]
        bs.interior(field, b, nv)[comp] = arr[o[2]:o[2] + nx[2], o[1]:o[1] + nx[1],
                                              o[0]:o[0] + nx[0]]


def centres(bs):
    lo, dx, shape, _ = lattice(bs)
    z = lo[2] + (np.arange(shape[0]) + 0.5) * dx[2]
    y = lo[1] + (np.arange(shape[1]) + 0.5) * dx[1]
    x = lo[0] + (np.arange(shape[2]) + 0.5) * dx[0]
    return np.meshgrid(z, y, x, indexing="ij")  # (Z, Y, X)


def mp_coeffs(bs, prim, G):
    """The nine scaled multipole coefficients (each moment summed exactly, math.fsum)."""
    rho = gather(bs, prim, bs.nvar, 0)
    Z, Y, X = centres(bs)
    vol = float(bs.blocks[0]["vol"])
    s = rho * vol
    r2 = X * X + Y * Y + Z * Z
    mom = [s, s * Y, s * Z, s * X, s * X * Y, s * Y * Z, s * (3.0 * Z * Z - r2), s * Z * X,
           s * 0.5 * (X * X - Y * Y)]
    return np.array([math.fsum(m.ravel()) * SCALE[n] * G for n, m in enumerate(mom)])


def multipole_potential(c, x, y, z):
    x, y, z = np.broadcast_arrays(np.asarray(x, float), np.asarray(y, float),
                                  np.asarray(z, float))
    r2 = x * x + y * y + z * z
    out = np.zeros_like(r2)
    nz = r2 != 0.0
    r = np.sqrt(r2[nz])
    r3 = r * r2[nz]
    r5 = r3 * r2[nz]
    xx, yy, zz = x[nz], y[nz], z[nz]
    out[nz] = (c[0] / r + (c[1] * yy + c[2] * zz + c[3] * xx) / r3 +
               (c[4] * xx * yy + c[5] * yy * zz + c[6] * (3.0 * zz * zz - r2[nz]) +
                c[7] * zz * xx + c[8] * 0.5 * (xx * xx - yy * yy)) / r5)
    return out


def boundary_potentials(bs, c):
    """phi_0 at the centre of every domain face cell: dict face -> 2D array."""
    lo, dx, shape, _ = lattice(bs)
    hi = lo + dx * np.array([shape[2], shape[1], shape[0]])
    Z, Y, X = centres(bs)
    return {
        "x-": multipole_potential(c, lo[0], Y[:, :, 0], Z[:, :, 0]),
        "x+": multipole_potential(c, hi[0], Y[:, :, -1], Z[:, :, -1]),
        "y-": multipole_potential(c, X[:, 0, :], lo[1], Z[:, 0, :]),
        "y+": multipole_potential(c, X[:, -1, :], hi[1], Z[:, -1, :]),
        "z-": multipole_potential(c, X[0], Y[0], lo[2]),
        "z+": multipole_potential(c, X[-1], Y[-1], hi[2]),
    }


def rhs(bs, prim, G, c):
    lo, dx, shape, _ = lattice(bs)
    b = 4.0 * np.pi * G * gather(bs, prim, bs.nvar, 0)
    pb = boundary_potentials(bs, c)
    # A_f / (dxc V) = 1 / dx^2 on a uniform lattice
    b[:, :, 0] -= 2.0 * pb["x-"] / (dx[0] * dx[0])
    b[:, :, -1] -= 2.0 * pb["x+"] / (dx[0] * dx[0])
    b[:, 0, :] -= 2.0 * pb["y-"] / (dx[1] * dx[1])
    b[:, -1, :] -= 2.0 * pb["y+"] / (dx[1] * dx[1])
    b[0] -= 2.0 * pb["z-"] / (dx[2] * dx[2])
    b[-1] -= 2.0 * pb["z+"] / (dx[2] * dx[2])
    return b


def laplacian(phi, dx, alpha=0.0):
    """The operator applied directly: sum_d (phi_- - 2 phi + phi_+) / dx_d^2 - alpha phi,
    with the ghost value -phi at every domain face."""
    out = -alpha * phi
    for ax, d in ((2, dx[0]), (1, dx[1]), (0, dx[2])):
        p = np.moveaxis(phi, ax, 0)
        ext = np.concatenate([-p[:1], p, -p[-1:]], axis=0)
        lap = (ext[:-2] - 2.0 * ext[1:-1] + ext[2:]) / (d * d)
        out = out + np.moveaxis(lap, 0, ax)
    return out


def solve(b, dx, alpha=0.0):
    """phi with laplacian(phi) = b, exactly (to rounding), by DST-II."""
    nz, ny, nx = b.shape
    lam = -alpha
    lz = -4.0 * np.sin(np.pi * (np.arange(nz) + 1) / (2 * nz)) ** 2 / (dx[2] * dx[2])
    ly = -4.0 * np.sin(np.pi * (np.arange(ny) + 1) / (2 * ny)) ** 2 / (dx[1] * dx[1])
    lx = -4.0 * np.sin(np.pi * (np.arange(nx) + 1) / (2 * nx)) ** 2 / (dx[0] * dx[0])
    lam = lam + lz[:, None, None] + ly[None, :, None] + lx[None, None, :]
    bh = scipy.fft.dstn(b, type=2, norm="ortho")
    return scipy.fft.idstn(bh / lam, type=2, norm="ortho")


def field(bs, phi, c):
    """(g as (3, nz, ny, nx), phi) with the multipole ghost values."""
    lo, dx, shape, _ = lattice(bs)
    pb = boundary_potentials(bs, c)
    g = np.zeros((3,) + phi.shape)
    for comp, ax, (m, p) in ((0, 2, ("x-", "x+")), (1, 1, ("y-", "y+")),
                             (2, 0, ("z-", "z+"))):
        q = np.moveaxis(phi, ax, 0)
        lo_g = 2.0 * np.expand_dims(pb[m], 0) - q[:1]
        hi_g = 2.0 * np.expand_dims(pb[p], 0) - q[-1:]
        ext = np.concatenate([lo_g, q, hi_g], axis=0)
        d = [dx[0], dx[1], dx[2]][comp]
        grad = (ext[2:] - ext[:-2]) / (2.0 * d)
        g[comp] = -np.moveaxis(grad, 0, ax)
    return g, phi


def gravity(bs, prim, G, alpha=0.0):
    """The whole pass: (coefficients, phi, g) as global arrays."""
    c = mp_coeffs(bs, prim, G)
    lo, dx, shape, _ = lattice(bs)
    phi = solve(rhs(bs, prim, G, c), dx, alpha)
    g, _ = field(bs, phi, c)
    return c, phi, g
