"""Seeded problem generators for the hydro flux path's workloads.

Each generator returns a :class:`Problem`: the block set, the conserved state in
pack layout (interior filled; ghosts are filled by the caller's first boundary
exchange), and the run parameters.  Random fields come from a counter-based
SplitMix64 stream keyed by (seed, counter), so the same inputs are produced on any
host and in any order.

Formulas follow the reference's problem generators where one exists:
``src/pgen/sod.cpp:11-46`` (shock tube), ``src/pgen/blast.cpp:16-71`` (blast),
``src/pgen/kh.cpp:236-262`` (iprob 5 shear layer), ``src/pgen/rt.cpp:18-111``
(Rayleigh-Taylor), ``src/pgen/linear_wave.cpp`` (sound wave).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .mesh import BlockSet, lattice_mesh, make_eos

RECON = {"dc": 1, "plm": 2, "ppm": 3}
RIEMANN = {"none": 1, "hlle": 2, "hllc": 3, "lhllc": 4}
NEED_GHOST = {"dc": 1, "plm": 2, "ppm": 3}  # hydro.cpp:56-67


@dataclasses.dataclass
class Problem:
    name: str
    bs: BlockSet
    cons: np.ndarray  # flat pack layout, nvar components
    eos: np.ndarray  # EOS_DTYPE record
    recon: str
    riemann: str
    cfl: float
    tlim: float = np.inf
    nsteps: int = 0
    g: tuple = (0.0, 0.0, 0.0)
    meta: dict = dataclasses.field(default_factory=dict)
    integrator: str = "rk2"  # rk1 / rk2 / rk3 / vl2 (main.hpp Integrator)
    monopole: tuple = None  # monopole gravity (nrbin, rmax, gravity_constant), or None
    # Poisson gravity (gravity_constant, diagonal_alpha, tolerance, max_iterations), or None
    poisson: tuple = None


# ---------------------------------------------------------------- RNG
_M64 = (1 << 64) - 1


def splitmix64(x):
    """SplitMix64 finaliser on a uint64 array (wrapping arithmetic)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, counter):
    """U[0,1) doubles for (seed, counter) pairs; counter may be an array."""
    c = np.asarray(counter, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = splitmix64(np.uint64(seed & _M64)) ^ c
    z = splitmix64(key)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _set_interior(bs, cons, b, rho, m1, m2, m3, en, scalars=()):
    v = bs.interior(cons, b, bs.nvar)
    v[0] = rho
    v[1] = m1
    v[2] = m2
    v[3] = m3
    v[4] = en
    for n, s in enumerate(scalars):
        v[5 + n] = s


# ---------------------------------------------------------------- Sod
def sod(ncells=256, nblock_cells=None, xmin=0.0, xmax=1.0, x_discont=0.5,
        left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1), gamma=1.4, gm1=0.4,
        recon="plm", riemann="hlle", cfl=0.4, tlim=0.2, ng=None):
    """1D shock tube (sod.cpp:11-46); outflow at both ends.

    left/right = (rho, u, p).  The energy uses ``p / (gamma - 1.0)`` like the
    reference pgen while the EOS uses ``gm1`` (eos/gm1), as in inputs/sod.in."""
    nbx = ncells // (nblock_cells or ncells)
    ng = ng if ng is not None else NEED_GHOST[recon]
    m = lattice_mesh(1, (nbx, 1, 1), (ncells // nbx, 1, 1), ng, (xmin, -0.5, -0.5),
                     (xmax, 0.5, 0.5), ["outflow"] * 6)
    bs = BlockSet(1, [m])
    cons = np.zeros(bs.field_size(bs.nvar))
    for b in range(bs.nblocks):
        x, _, _ = bs.cell_centers(b)
        lft = x < x_discont
        rl, ul, pl = left
        rr, ur, pr = right
        rho = np.where(lft, rl, rr)
        mom = np.where(lft, rl * ul, rr * ur)
        en = np.where(lft, 0.5 * rl * ul * ul + pl / (gamma - 1.0),
                      0.5 * rr * ur * ur + pr / (gamma - 1.0))
        _set_interior(bs, cons, b, rho, mom, 0.0, 0.0, en)
    return Problem("sod", bs, cons, make_eos(gm1), recon, riemann, cfl, tlim=tlim)


# ---------------------------------------------------------------- blast
def blast3d(nblk=8, n=32, seed=20240601, recon="ppm", riemann="hllc", cfl=0.3,
            nsteps=100, ng=3, noise=0.01, p_amb=0.1, p_in=10.0, radius=0.1):
    """3D blast wave on [-1/2,1/2]^3, nblk^3 blocks of n^3, periodic (blast.cpp:16-71
    with radius_inner = radius_outer, i.e. no ramp), plus +-noise uniform density noise
    keyed by the global cell index."""
    gm1 = 2.0 / 3.0
    m = lattice_mesh(3, (nblk,) * 3, (n,) * 3, ng, (-0.5,) * 3, (0.5,) * 3,
                     ["periodic"] * 6)
    bs = BlockSet(3, [m])
    cons = np.zeros(bs.field_size(bs.nvar))
    N = nblk * n
    for b in range(bs.nblocks):
        x, y, z = bs.cell_centers(b)
        r = bs.blocks[b]
        lx, ly, lz = (int(round((r["xmin"][d] + 0.5) / (n * r["dx"][d]))) for d in range(3))
        gi = lx * n + np.arange(n)
        gj = ly * n + np.arange(n)
        gk = lz * n + np.arange(n)
        gid = (gk[:, None, None] * N + gj[None, :, None]) * N + gi[None, None, :]
        u = uniform(seed, gid)
        rho = 1.0 * (1.0 + noise * (2.0 * u - 1.0))
        rad = np.sqrt(x * x + y * y + z * z)
        pres = np.where(rad < radius, p_in, p_amb)
        _set_interior(bs, cons, b, rho, 0.0, 0.0, 0.0, pres / gm1)
    return Problem("blast3d", bs, cons, make_eos(gm1), recon, riemann, cfl,
                   nsteps=nsteps, meta={"seed": seed, "N": N})


# ---------------------------------------------------------------- monopole gravity
def _default_rmax(xmin, xmax):
    """gravity/rmax's default (gravity.cpp:262-272): the distance from the origin to the
    farthest corner of the mesh box."""
    edge = [max(abs(a), abs(b)) for a, b in zip(xmin, xmax)]
    return float(np.sqrt(edge[0] * edge[0] + edge[1] * edge[1] + edge[2] * edge[2]))


def _static_mesh(ndim, nb, block, ng, xmin, xmax, region, level):
    """An outflow-bounded multilevel mesh (amr.AmrMesh) of nb^ndim base blocks of block^ndim
    cells with the static refinement block (region, level) of a deck."""
    from .amr import AmrMesh, static_refinement

    nbase = tuple(nb if d < ndim else 1 for d in range(3))
    bnx = tuple(block if d < ndim else 1 for d in range(3))
    bc = ["outflow"] * 6
    lo, hi = (-region,) * 3, (region,) * 3
    ref = static_refinement(ndim, nbase[:ndim], bnx[:ndim], xmin, xmax, [(lo, hi, level)], bc)
    return AmrMesh(ndim, nbase[:ndim], bnx[:ndim], ng, xmin, xmax, ref, max_level=level,
                   bc=bc)


def grav_sphere(res=32, block=None, ndim=3, rho0=1.0, r0=0.25, nrbin=None, G=1.0,
                gm1=0.666666667, recon="dc", riemann="hlle", cfl=0.4, center=None,
                solver="monopole", tol=1e-8, smr_level=0):
    """The gravity accuracy test (grav_test.cpp:15-55, pid 1; inputs/grav_test.in as
    tst/.../gravity/gravity.py runs it): rho = rho0 (1 - r^2/r0^2)^2 inside r0 and 0
    outside (ConsToPrim's floor then applies), at rest, E = 1, on [-1/2, 1/2]^ndim with
    outflow boundaries.  solver "monopole": centred at the origin, nrbin = 4 res;
    "poisson": centred at (0.01, 0, 0), the residual tolerance `tol`, flux_correct false
    (the deck's).  smr_level > 0: the deck's static refinement block ([-0.25, 0.25]^3 at
    that level, gravity.py's SMR cases with res 16 and 8^3 blocks) on a multilevel mesh,
    kept in meta["mesh"]; Poisson's fine/coarse prolongation there is "Linear" (the
    deck's boundary_prolongation = Quadratic is not built)."""
    if center is None:
        center = (0.01, 0.0, 0.0) if solver == "poisson" else (0.0, 0.0, 0.0)
    block = block or res
    nb = res // block
    nblocks = tuple(nb if d < ndim else 1 for d in range(3))
    bnx = tuple(block if d < ndim else 1 for d in range(3))
    xmin, xmax = (-0.5,) * 3, (0.5,) * 3
    amesh = None
    if smr_level:
        amesh = _static_mesh(ndim, nb, block, NEED_GHOST[recon], xmin, xmax, 0.25, smr_level)
        bs = amesh.block_set()
    else:
        m = lattice_mesh(ndim, nblocks, bnx, NEED_GHOST[recon], xmin, xmax, ["outflow"] * 6)
        bs = BlockSet(ndim, [m])
    cons = np.zeros(bs.field_size(bs.nvar))
    for b in range(bs.nblocks):
        x, y, z = bs.cell_centers(b)
        r = np.sqrt((x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2)
        rho = np.where(r <= r0, rho0 * (1.0 - r * r * (1.0 / (r0 * r0))) ** 2, 0.0)
        _set_interior(bs, cons, b, rho, 0.0, 0.0, 0.0, 1.0)
    # the mesh box of the deck on every axis (gravity.cpp reads x2/x3 extents too)
    rmax = _default_rmax(xmin, xmax)
    grav = {"monopole": (nrbin or 4 * res, rmax, G)} if solver == "monopole" else \
        {"poisson": (G, 0.0, tol, 10000, False, "Linear")}
    return Problem("grav_sphere", bs, cons, make_eos(gm1), recon, riemann, cfl,
                   integrator="rk1", **grav,
                   meta={"rho0": rho0, "r0": r0, "res": res, "box": (xmin, xmax),
                         "center": center, "mesh": amesh})


def evrard(res=32, block=16, mass=1.0, radius=1.0, u_therm=0.05, nrbin=None, G=1.0,
           gm1=0.666666667, recon="plm", riemann="hllc", cfl=0.4, tlim=0.8,
           solver="monopole", tol=1e-8, smr_level=0):
    """Evrard's collapse (evrards_collapse.cpp:15-49, inputs/evrards_collapse.in) with the
    monopole solver as tst/.../evrards_collapse/evrards_collapse.py runs it: rho =
    M / (2 pi R^2 r) inside R and 1e-5 outside, at rest, E = u_therm rho, on [-1, 1]^3
    (res^3 cells in blocks of block^3), outflow, PLM + HLLC, RK2, CFL 0.4, to t = 0.8;
    solver "monopole" (nrbin = 4 res) or "poisson" (the deck's own: residual tolerance
    `tol`, flux_correct true, Linear prolongation).  smr_level > 0: the deck's static
    refinement block ([-0.5, 0.5]^3 at that level; the test's monopole_smr_32 is res 32,
    block 16, level 2, its smr_16 the Poisson solver at res 16, block 8) on a multilevel
    mesh, kept in meta["mesh"]."""
    nb = res // block
    xmin, xmax = (-1.0,) * 3, (1.0,) * 3
    amesh = None
    if smr_level:
        amesh = _static_mesh(3, nb, block, NEED_GHOST[recon], xmin, xmax, 0.5, smr_level)
        bs = amesh.block_set()
    else:
        m = lattice_mesh(3, (nb,) * 3, (block,) * 3, NEED_GHOST[recon], xmin, xmax,
                         ["outflow"] * 6)
        bs = BlockSet(3, [m])
    cons = np.zeros(bs.field_size(bs.nvar))
    for b in range(bs.nblocks):
        x, y, z = bs.cell_centers(b)
        r = np.sqrt(x * x + y * y + z * z)
        rho = np.where(r < radius, mass / (2.0 * np.pi * (radius * radius) * r), 1e-5)
        _set_interior(bs, cons, b, rho, 0.0, 0.0, 0.0, u_therm * rho)
    grav = {"monopole": (nrbin or 4 * res, _default_rmax(xmin, xmax), G)} \
        if solver == "monopole" else {"poisson": (G, 0.0, tol, 10000, True, "Linear")}
    return Problem("evrard", bs, cons, make_eos(gm1), recon, riemann, cfl, tlim=tlim,
                   **grav, meta={"res": res, "radius": radius, "box": (xmin, xmax),
                                 "mesh": amesh})


# ---------------------------------------------------------------- uneven blocks
def _fourier_field(seed, ctr0, x, y, z, ndim):
    """Sum of 4 seeded Fourier modes with wave-vector components in {-1, 0, 1} on the
    unit box (one wavelength per block at most, so an 8-cell block resolves it),
    normalised to [-1, 1]."""
    r = uniform(seed, ctr0 + np.arange(32, dtype=np.uint64))
    amps = r[0:4] + 0.1
    amps = amps / amps.sum()
    s = np.zeros_like(x)
    for mdx in range(4):
        kv = np.floor(r[4 + 3 * mdx: 7 + 3 * mdx] * 3.0) - 1.0  # integers in [-1, 1]
        kv[ndim:] = 0.0
        if not kv.any():
            kv[0] = 1.0
        ph = 2.0 * np.pi * r[16 + mdx]
        s = s + amps[mdx] * np.sin(2.0 * np.pi * (kv[0] * x + kv[1] * y + kv[2] * z) + ph)
    return s


def uneven_blocks(nblocks=4096, sizes=(8, 16, 32), probs=(0.5, 0.35, 0.15), ng=4,
                  seed=7, recon="ppm", riemann="hllc", cfl=0.3, nsteps=50,
                  jump_frac=0.3, ndim=3, by_size=False):
    """Independent cubic blocks of uneven edge (uneven AMR-like block set).

    Builder-defined rules (the reference has no uneven packs): every block spans the
    unit cube (dx = 1/n) and is periodic onto itself, so it needs no neighbour; the
    global dt is the minimum over all blocks.  Each block holds a smooth field from 4
    seeded Fourier modes (rho in [0.5,2], p in [0.1,1], |u| <= 1) and, in jump_frac of
    the blocks, planar discontinuities at a random orientation (density x4, pressure x10
    inside a slab bounded by two parallel planes, see below).

    by_size: the same blocks (data keyed by the block's draw index, kept as mesh_id) in
    the order of their edge, so that blocks of one size form one contiguous range, as the
    reference's packs of equal-shape blocks need."""
    gm1 = 2.0 / 3.0
    r = uniform(seed, np.arange(nblocks, dtype=np.uint64))
    cum = np.cumsum(probs)
    which = np.searchsorted(cum, r * cum[-1], side="right")
    edges = np.asarray(sizes)[np.minimum(which, len(sizes) - 1)]
    order = np.argsort(edges, kind="stable") if by_size else np.arange(nblocks)
    meshes = []
    for b in order:
        n = int(edges[b])
        nx = (n, n if ndim > 1 else 1, n if ndim > 2 else 1)
        meshes.append(lattice_mesh(ndim, (1, 1, 1), nx, ng, (0.0, 0.0, 0.0),
                                   (1.0, 1.0, 1.0), ["periodic"] * 6, mesh_id=int(b)))
    bs = BlockSet(ndim, meshes)
    cons = np.zeros(bs.field_size(bs.nvar))
    for slot, b in enumerate(order):
        x, y, z = bs.cell_centers(slot)
        base = np.uint64(1_000_000 + 256 * int(b))
        rho = 1.25 + 0.75 * _fourier_field(seed, base, x, y, z, ndim)
        prs = 0.55 + 0.45 * _fourier_field(seed, base + np.uint64(40), x, y, z, ndim)
        vel = [_fourier_field(seed, base + np.uint64(80 + 40 * d), x, y, z, ndim) / np.sqrt(3.0)
               for d in range(3)]
        rj = uniform(seed, base + np.uint64(250 + np.arange(5, dtype=np.uint64)))
        if rj[0] < jump_frac:
            # a slab between two parallel planes with an integer normal (13 possible
            # orientations), so the discontinuity is compatible with the block's own
            # periodicity; the inner side of the slab gets density x4, pressure x10
            nrm = np.floor(rj[1:4] * 3.0) - 1.0
            nrm[ndim:] = 0.0
            if not nrm.any():
                nrm[0] = 1.0
            phase = (x * nrm[0] + y * nrm[1] + z * nrm[2] + rj[4]) % 1.0
            side = phase < 0.5
            rho = np.where(side, rho * 4.0, rho)
            prs = np.where(side, prs * 10.0, prs)
        ek = 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
        _set_interior(bs, cons, slot, rho, rho * vel[0], rho * vel[1], rho * vel[2],
                      prs / gm1 + ek)
    return Problem("uneven", bs, cons, make_eos(gm1), recon, riemann, cfl,
                   nsteps=nsteps, meta={"edges": edges[order], "seed": seed})


# ---------------------------------------------------------------- KH
def kh2d(nblk=16, n=32, seed=11, amps=None, a=0.01, sigma2=0.01, drat=2.0,
         vflow=0.5, p0=2.5, recon="plm", riemann="hllc", cfl=0.3, nsteps=200):
    """2D Kelvin-Helmholtz on [0,1]x[-1/2,1/2] periodic (kh.cpp:236-262 iprob 5 form):
    w = (tanh((|y|-1/4)/a)+1)/2, rho = w + (1-w)*drat, v_x = vflow*(w - (1-w)),
    v_y = sum_k a_k sin(2 pi k x) exp(-(|y|-1/4)^2/sigma2), p = p0, and one passive
    scalar s = 1-w marking the inner layer (conserved rho*s)."""
    gm1 = 2.0 / 3.0
    if amps is None:
        amps = 0.01 * uniform(seed, np.arange(8, dtype=np.uint64))
    amps = np.asarray(amps, dtype=np.float64)
    m = lattice_mesh(2, (nblk, nblk, 1), (n, n, 1), NEED_GHOST[recon], (0.0, -0.5, -0.5),
                     (1.0, 0.5, 0.5), ["periodic"] * 6)
    bs = BlockSet(2, [m], nscalars=1)
    cons = np.zeros(bs.field_size(bs.nvar))
    for b in range(bs.nblocks):
        x, y, _ = bs.cell_centers(b)
        w = (np.tanh((np.abs(y) - 0.25) / a) + 1.0) * 0.5
        rho = w + (1.0 - w) * drat
        vx = vflow * w - (1.0 - w) * vflow
        env = np.exp(-((np.abs(y) - 0.25) ** 2) / sigma2)
        vy = np.zeros_like(x)
        for k in range(len(amps)):
            vy = vy + amps[k] * np.sin(2.0 * np.pi * (k + 1) * x) * env
        en = p0 / gm1 + 0.5 * rho * (vx * vx + vy * vy)
        _set_interior(bs, cons, b, rho, rho * vx, rho * vy, 0.0, en, [rho * (1.0 - w)])
    return Problem("kh2d", bs, cons, make_eos(gm1), recon, riemann, cfl, nsteps=nsteps,
                   meta={"amps": amps})


# ---------------------------------------------------------------- RT
def rt2d(nx=128, ny=384, nblk=(4, 12), seed=5, amps=None, gamma=1.4, g=-0.1,
         recon="ppm", riemann="hllc", cfl=0.3, nsteps=300, mesh_id=0, width=0.0):
    """2D Rayleigh-Taylor on [-1/4,1/4]x[-3/4,3/4]: rho=2 above y=0, 1 below,
    hydrostatic p = 2.5 + g*rho*y, 6-mode interface perturbation of v_y, periodic in
    x, reflecting in y, constant gravity g in y (rt.cpp:18-111, gravity.cpp:374-410).
    width > 0 replaces the sharp interface by rho = 1.5 + 0.5 tanh(y/width) with the
    hydrostatic p = 2.5 + g (1.5 y + 0.5 width log cosh(y/width))."""
    gm1 = gamma - 1.0
    if amps is None:
        amps = 0.01 * uniform(seed, np.arange(6, dtype=np.uint64))
    amps = np.asarray(amps, dtype=np.float64)
    m = lattice_mesh(2, (nblk[0], nblk[1], 1), (nx // nblk[0], ny // nblk[1], 1),
                     NEED_GHOST[recon], (-0.25, -0.75, -0.5), (0.25, 0.75, 0.5),
                     ["periodic", "periodic", "reflect", "reflect", "periodic", "periodic"],
                     mesh_id=mesh_id)
    bs = BlockSet(2, [m])
    cons = np.zeros(bs.field_size(bs.nvar))
    for b in range(bs.nblocks):
        x, y, _ = bs.cell_centers(b)
        rho, vy, p = _rt_fields(x, y, amps, g, width)
        _set_interior(bs, cons, b, rho, 0.0, rho * vy, 0.0, p / gm1 + 0.5 * rho * vy * vy)
    return Problem("rt2d", bs, cons, make_eos(gm1), recon, riemann, cfl, nsteps=nsteps,
                   g=(0.0, g, 0.0), meta={"amps": amps})


RT_GRIDS = ((128, 384), (256, 768), (512, 1536))


def rt_batch(ninst=64, seed=11, grids=RT_GRIDS, block=32, gamma=1.4, g=-0.1, recon="ppm",
             riemann="hllc", cfl=0.3, nsteps=300, width=0.0):
    """Batched Rayleigh-Taylor: `ninst` independent rt2d instances (same domain, EOS and
    gravity) in one block set, instance i = mesh_id i.  Each instance draws its grid
    uniformly from `grids` and six mode amplitudes U(0, 0.01) from the SplitMix64 stream
    (seed, 7*i .. 7*i+6); every mesh is split into blocks of block x block cells.  The
    state of instance i equals rt2d(nx, ny, ..., amps=amps[i], mesh_id=i) bitwise."""
    gm1 = gamma - 1.0
    r = uniform(seed, np.arange(7 * ninst, dtype=np.uint64)).reshape(ninst, 7)
    shapes, amps, meshes = [], [], []
    for i in range(ninst):
        nx, ny = grids[min(int(r[i, 0] * len(grids)), len(grids) - 1)]
        shapes.append((nx, ny))
        amps.append(0.01 * r[i, 1:7])
        meshes.append(lattice_mesh(2, (nx // block, ny // block, 1), (block, block, 1),
                                   NEED_GHOST[recon], (-0.25, -0.75, -0.5), (0.25, 0.75, 0.5),
                                   ["periodic", "periodic", "reflect", "reflect", "periodic",
                                    "periodic"], mesh_id=i))
    amps = np.array(amps)
    bs = BlockSet(2, meshes)
    cons = np.zeros(bs.field_size(bs.nvar))
    x, y, _ = bs.cell_centers_all()
    ab = amps[bs.blocks["mesh_id"].astype(np.int64)].T[:, :, None, None, None]  # (6, nb, 1,1,1)
    rho, vy, p = _rt_fields(x, y, ab, g, width)
    fields = [rho, 0.0 * rho, rho * vy, 0.0 * rho, p / gm1 + 0.5 * rho * vy * vy]
    idx = bs.interior_index()
    nb = bs.nblocks
    for v, f in enumerate(fields):
        cons[idx[:, v, :]] = np.reshape(f, (nb, -1))
    return Problem("rt_batch", bs, cons, make_eos(gm1), recon, riemann, cfl, nsteps=nsteps,
                   g=(0.0, g, 0.0), meta={"amps": amps, "grids": shapes, "width": width})


def _rt_fields(x, y, amps, g, width=0.0):
    """rho, v_y, p of the rt2d initial state at cell centres (x, y)."""
    if width > 0.0:
        rho = 1.5 + 0.5 * np.tanh(y / width)
        p = 2.5 + g * (1.5 * y + 0.5 * width * np.log(np.cosh(y / width)))
    else:
        rho = np.where(y < 0.0, 1.0, 2.0)
        p = 2.5 + g * rho * y
    vy = np.zeros_like(x)
    for k in range(len(amps)):
        vy = vy + amps[k] * np.cos(2.0 * np.pi * (k + 1) * x / 0.5) * np.exp(-(y / 0.05) ** 2)
    return rho, vy, p


# ---------------------------------------------------------------- linear wave
def linear_wave(ndim=1, res=32, amp=1.0e-6, recon="plm", riemann="hllc", cfl=0.3,
                gamma=5.0 / 3.0, ncycles=1.0):
    """Sound wave (wave_flag 0) along x1 on a periodic [0,1] (x [0,1] ...) box of
    res^ndim cells in one block, following linear_wave.cpp's hydro eigenvector with
    d0 = 1, p0 = 1/gamma (sound speed 1).  Returns the problem; meta['exact'] holds
    the initial interior state, which the wave returns to after whole periods."""
    gm1 = gamma - 1.0
    nx = (res, res if ndim > 1 else 1, res if ndim > 2 else 1)
    ng = NEED_GHOST[recon]
    m = lattice_mesh(ndim, (1, 1, 1), nx, ng, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0),
                     ["periodic"] * 6)
    bs = BlockSet(ndim, [m])
    cons = np.zeros(bs.field_size(bs.nvar))
    d0, p0 = 1.0, 1.0 / gamma
    c0 = np.sqrt(gamma * p0 / d0)
    h0 = (p0 / gm1 + p0) / d0
    # right eigenvector of the sound wave moving in -x (Roe form, u0 = 0)
    rem = np.array([1.0, -c0, 0.0, 0.0, h0 - 0.0 * c0])
    x, _, _ = bs.cell_centers(0)
    # cell-average of sin(2 pi x) over each cell
    dx = bs.blocks[0]["dx"][0]
    sn = (np.cos(2 * np.pi * (x - dx / 2)) - np.cos(2 * np.pi * (x + dx / 2))) / (2 * np.pi * dx)
    e0 = p0 / gm1
    st = [d0 + amp * rem[0] * sn, amp * rem[1] * sn, np.zeros_like(x), np.zeros_like(x),
          e0 + amp * rem[4] * sn]
    _set_interior(bs, cons, 0, *st)
    period = 1.0 / c0
    return Problem("linear_wave", bs, cons, make_eos(gm1), recon, riemann, cfl,
                   tlim=ncycles * period,
                   meta={"exact": np.stack(st), "amp": amp, "period": period})


def linear_wave3d(res=64, recon="plm", riemann="hllc", integrator="vl2", amp=1.0e-6,
                  wave_flag=0, gamma=1.666666666666667, eos_gm1=0.66666666667, cfl=0.3,
                  tlim_periods=2.0, block=None):
    """The reference's 3D linear-wave convergence problem (inputs/linear_wave3d.in with
    the overrides of convergence.py:55-80): a (2 res) x res x res periodic mesh on
    [0,3]x[0,1.5]x[0,1.5], a hydro wave along the grid diagonal (linear_wave.cpp:42-147,
    304-338), amplitude 1e-6, run for tlim_periods wave periods.  The problem generator
    uses hydro/gamma = 1.666666666666667 and the EOS eos/gm1 = 0.66666666667, as the deck
    does.  block = meshblock extent (default: the whole mesh, as convergence.py sets for
    one rank); meta['exact'] holds the analytic state the errors are measured against."""
    nx = (2 * res, res, res)
    block = tuple(block) if block is not None else nx
    ng = 3 if recon == "ppm" else 2
    size = (3.0, 1.5, 1.5)
    gm1 = gamma - 1.0
    ang_3 = np.arctan(size[0] / size[1])
    sin_a3, cos_a3 = np.sin(ang_3), np.cos(ang_3)
    ang_2 = np.arctan(0.5 * (size[0] * cos_a3 + size[1] * sin_a3) / size[2])
    sin_a2, cos_a2 = np.sin(ang_2), np.cos(ang_2)
    x1 = size[0] * cos_a2 * cos_a3
    x2 = size[1] * cos_a2 * sin_a3
    x3 = size[2] * sin_a2
    lam = x1
    if ang_3 != 0.0:
        lam = min(lam, x2)
    if ang_2 != 0.0:
        lam = min(lam, x3)
    k_par = 2.0 * np.pi / lam
    d0, u0 = 1.0, 0.0
    p0 = 1.0 / gamma
    h0 = ((p0 / gm1 + 0.5 * d0 * (u0 * u0)) + p0) / d0
    a = np.sqrt(gm1 * max(h0 - 0.5 * u0 * u0, 1.0e-20))
    ev = [u0 - a, u0, u0, u0, u0 + a]
    rem = {0: [1.0, u0 - a, 0.0, 0.0, h0 - u0 * a], 4: [1.0, u0 + a, 0.0, 0.0, h0 + u0 * a]}
    if wave_flag not in rem:
        raise ValueError("sound waves only (wave_flag 0 or 4)")
    r = rem[wave_flag]
    nblk = tuple(n // b for n, b in zip(nx, block))
    m = lattice_mesh(3, nblk, block, ng, (0.0, 0.0, 0.0), size, ["periodic"] * 6)
    bs = BlockSet(3, [m])
    cons = np.zeros(bs.field_size(bs.nvar))
    exact = []
    for b in range(bs.nblocks):
        xc, yc, zc = bs.cell_centers(b)
        x = cos_a2 * (xc * cos_a3 + yc * sin_a3) + zc * sin_a2
        sn = np.sin(k_par * x)
        d1 = d0 + amp * sn * r[0]
        mx = d0 * u0 + amp * sn * r[1]
        my = amp * sn * r[2]
        mz = amp * sn * r[3]
        m1 = mx * cos_a2 * cos_a3 - my * sin_a3 - mz * sin_a2 * cos_a3
        m2 = mx * cos_a2 * sin_a3 + my * cos_a3 - mz * sin_a2 * sin_a3
        m3 = mx * sin_a2 + mz * cos_a2
        e0 = p0 / gm1 + 0.5 * d0 * u0 * u0 + amp * sn * r[4]
        _set_interior(bs, cons, b, d1, m1, m2, m3, e0)
        exact.append(np.stack([d1, m1, m2, m3, e0]))
    tlim = lam / abs(ev[wave_flag]) * tlim_periods
    return Problem("linear_wave3d", bs, cons, make_eos(eos_gm1), recon, riemann, cfl,
                   tlim=tlim, integrator=integrator,
                   meta={"exact": exact, "nx": nx, "domain_volume": size[0] * size[1] * size[2],
                         "lambda": lam})


# ---------------------------------------------------------------- KH with composition
def std_uniform_m05_05(seed, n):
    """The first n values of std::uniform_real_distribution<double>(-0.5, 0.5) driven by
    std::mt19937(seed), as libstdc++ computes them (the reference's pgen, kh.cpp:38-39):
    generate_canonical<double, 53> takes two 32-bit words w1, w2 and forms
    (w1 + w2 * 2^32) / 2^64 (one rounding, in the sum; at 1.0 it steps down to the largest
    double below 1), then the distribution returns (b - a) * that + a.  numpy's legacy
    RandomState seeds MT19937 with the same init_genrand and randint over the full
    uint32 range returns its raw words in order."""
    words = np.random.RandomState(seed & 0xFFFFFFFF).randint(
        0, 2 ** 32, size=2 * n, dtype=np.uint64).astype(np.float64)
    c = (words[0::2] + words[1::2] * 4294967296.0) / 18446744073709551616.0
    c = np.where(c >= 1.0, np.nextafter(1.0, 0.0), c)
    return 1.0 * c + -0.5


def kh_comp(nx=(256, 128), block=(64, 128), vflow=1.0, drat=0.1, amp=0.01, ofrac=0.65,
            ye=0.4983, pamp=2.5, gamma=1.666666666666667, eos_gm1=0.66666666667,
            nlset=1, recon="plm", riemann="hllc", cfl=0.4, tlim=1.0, nsteps=100):
    """The reference deck inputs/kh-comp-advection.in: Kelvin-Helmholtz slip surfaces with
    random perturbations and the ONe composition (kh.cpp:268-313, iprob = 6) on
    [-1,1]x[-1/2,1/2], periodic, nghost 2 (3 for PPM), PLM + HLLC, cfl 0.4, ideal gas with
    gm1 = 0.66666666667 (the pgen's pressure uses hydro/gamma).  Scalars (hydro.cpp:
    187-198): 6 species densities (alpha, C12, O16, Ne20, IME, IGE), Ye, then per level
    set (lset, xfuel); the pgen sets O16, Ne20, Ye and lset0 and leaves the rest 0.
    Block b draws its perturbations from std::mt19937(b) (Parthenon's gid; the blocks of
    one level are numbered along x first) in the pgen's loop order k, j, i."""
    gm1 = gamma - 1.0
    nbx, nby = nx[0] // block[0], nx[1] // block[1]
    m = lattice_mesh(2, (nbx, nby, 1), (block[0], block[1], 1), max(2, NEED_GHOST[recon]),
                     (-1.0, -0.5, -0.5), (1.0, 0.5, 0.5), ["periodic"] * 6)
    nsc = 7 + 2 * nlset
    bs = BlockSet(2, [m], nscalars=nsc)
    cons = np.zeros(bs.field_size(bs.nvar))
    rho = 1.0
    for b in range(bs.nblocks):
        _, y, _ = bs.cell_centers(b)
        inner = np.abs(y) < 0.25
        ncell = y.size
        # draws per cell in loop order: IM1, IM2, and for the inner layer IM1, IM2 again
        per = np.where(inner.ravel(), 4, 2)
        r = std_uniform_m05_05(b, int(per.sum()))
        first = np.concatenate([[0], np.cumsum(per)[:-1]])
        r1, r2 = r[first], r[first + 1]
        r3 = np.where(inner.ravel(), r[np.minimum(first + 2, r.size - 1)], 0.0)
        r4 = np.where(inner.ravel(), r[np.minimum(first + 3, r.size - 1)], 0.0)
        shp = y.shape
        d = np.full(ncell, rho)
        m1 = vflow + amp * r1
        m2 = amp * r2
        o16 = np.full(ncell, ofrac * rho)
        ne20 = np.full(ncell, (1.0 - ofrac) * rho)
        yes = np.full(ncell, ye * rho)
        ii = inner.ravel()
        d[ii] = drat * rho
        m1[ii] = -drat * (vflow + amp * r3[ii]) * rho
        m2[ii] = drat * amp * r4[ii] * rho
        o16[ii] = 1.0 * drat * rho
        ne20[ii] = 0.0 * drat * rho
        yes[ii] = 0.5 * drat * rho
        lset = (0.25 - np.abs(y.ravel())) * d
        en = pamp / gm1 + 0.5 * (m1 * m1 + m2 * m2) / d
        z = np.zeros(ncell)
        sc = [z, z, o16, ne20, z, z, yes]
        for i in range(nlset):
            sc += [lset if i == 0 else z, z]
        _set_interior(bs, cons, b, d.reshape(shp), m1.reshape(shp), m2.reshape(shp), 0.0,
                      en.reshape(shp), [s.reshape(shp) for s in sc])
    return Problem("kh_comp", bs, cons, make_eos(eos_gm1), recon, riemann, cfl, tlim=tlim,
                   nsteps=nsteps, meta={"nlset": nlset, "composition": "ONe"})


# ---------------------------------------------------------------- two-level blast
def blast2d_two_level(nbase=48, n=32, nref=16, seed=20240601, recon="ppm", riemann="hllc",
                      cfl=0.3, noise=0.01, p_amb=0.1, p_in=10.0, radius=0.1):
    """2D blast wave (blast.cpp:16-71, no ramp) on [-1/2,1/2]^2, periodic, on a two-level
    mesh: a base of nbase^2 blocks of n^2 whose central nref^2 base blocks are refined once
    (each replaced by 4 blocks of n^2 at half the spacing).  Density noise is keyed by the
    block and the cell's index in it.  Returns (Problem, amr.TwoLevelMesh)."""
    from .amr import TwoLevelMesh

    gm1 = 2.0 / 3.0
    lo = (nbase - nref) // 2
    refined = [(x, y) for x in range(lo, lo + nref) for y in range(lo, lo + nref)]
    m = TwoLevelMesh(2, (nbase, nbase), (n, n), NEED_GHOST[recon], (-0.5, -0.5, -0.5),
                     (0.5, 0.5, 0.5), refined)
    bs = m.block_set()
    cons = np.zeros(bs.field_size(bs.nvar))
    for b in range(bs.nblocks):
        x, y, _ = bs.cell_centers(b)
        key = (np.uint64(b) << np.uint64(20)) + np.arange(x.size, dtype=np.uint64)
        u = uniform(seed, key).reshape(x.shape)
        rho = 1.0 * (1.0 + noise * (2.0 * u - 1.0))
        pres = np.where(np.sqrt(x * x + y * y) < radius, p_in, p_amb)
        _set_interior(bs, cons, b, rho, 0.0, 0.0, 0.0, pres / gm1)
    p = Problem("blast2d_two_level", bs, cons, make_eos(gm1), recon, riemann, cfl,
                meta={"nbase": nbase, "nref": nref})
    return p, m


# ---------------------------------------------------------------- the adaptive blast deck
# inputs/blast.in: a 128^2 mesh of 16^2 blocks on [-1/2, 1/2]^2, periodic, three levels
# (refinement = adaptive, numlevel = 3) on the pressure-gradient criterion (threshold 0.1),
# PLM + HLLC, vl2, cfl 0.4, tlim 1e-2; <eos> is not set, so the EOS takes the default
# gm1 = 0.6666667 (hydro.cpp:262) while the problem generator uses hydro/gamma.
BLAST_DECK = dict(nx=(128, 128), block_nx=(16, 16), numlevel=3, xmin=(-0.5, -0.5, -0.5),
                  xmax=(0.5, 0.5, 0.5), recon="plm", riemann="hllc", integrator="vl2",
                  cfl=0.4, tlim=1.0e-2, gamma=1.666666666666667, eos_gm1=0.6666667,
                  threshold=0.1, pamb=1.0e-3, prat=1.6e8, radius_outer=0.1,
                  radius_inner=0.0, damb=1.0, drat=1.0, x0=(0.0, 0.0, 0.0))


def blast_deck_pin(deck=None):
    """The deck's parameter input for hydro.initialize_hydro."""
    k = dict(BLAST_DECK, **(deck or {}))
    return {"parthenon/time": {"cfl": k["cfl"], "integrator": k["integrator"]},
            "parthenon/mesh": {"nghost": NEED_GHOST[k["recon"]]},
            "hydro": {"reconstruction": k["recon"], "riemann": k["riemann"],
                      "gamma": k["gamma"]},
            "eos": {"gm1": k["eos_gm1"]},
            "refinement": {"type": "pressure_gradient",
                           "threshold_pressure_gradient": k["threshold"]}}


def blast_pgen(deck=None):
    """The blast problem generator (blast.cpp:16-71) as a function (mesh, bs) -> cons:
    uniform (damb, pamb) outside radius_outer, (drat damb, prat pamb) inside
    rin = radius_outer - radius_inner, and between them the logarithmic ramp."""
    k = dict(BLAST_DECK, **(deck or {}))
    rout = k["radius_outer"]
    rin = rout - k["radius_inner"]
    da, pa, drat, prat = k["damb"], k["pamb"], k["drat"], k["prat"]
    gm1 = k["gamma"] - 1.0

    def pgen(mesh, bs):
        cons = np.zeros(bs.field_size(bs.nvar))
        for b in range(bs.nblocks):
            x, y, z = bs.cell_centers(b)
            if bs.ndim < 3:
                z = np.zeros_like(x)
            rad = np.sqrt((x - k["x0"][0]) ** 2 + (y - k["x0"][1]) ** 2 + (z - k["x0"][2]) ** 2)
            den = np.full(x.shape, da)
            pres = np.full(x.shape, pa)
            inner = rad < rin
            den[inner], pres[inner] = drat * da, prat * pa
            ramp = (rad < rout) & ~inner
            if ramp.any():
                f = (rad[ramp] - rin) / (rout - rin)
                den[ramp] = np.exp((1.0 - f) * np.log(drat * da) + f * np.log(da))
                pres[ramp] = np.exp((1.0 - f) * np.log(prat * pa) + f * np.log(pa))
            _set_interior(bs, cons, b, den, 0.0, 0.0, 0.0, pres / gm1)
        return cons

    return pgen


def blast_deck_base_mesh(deck=None):
    """The deck's root-level mesh (no block refined yet) as an amr.AmrMesh."""
    from .amr import AmrMesh

    k = dict(BLAST_DECK, **(deck or {}))
    nd = len(k["nx"])
    nbase = tuple(k["nx"][d] // k["block_nx"][d] for d in range(nd))
    return AmrMesh(nd, nbase, k["block_nx"], NEED_GHOST[k["recon"]], k["xmin"], k["xmax"],
                   [], max_level=k["numlevel"] - 1)
