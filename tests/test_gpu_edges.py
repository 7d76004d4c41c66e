"""Edge cases through the C ABI: an empty block set (every entry point is a no-op, the dt
estimate is unconstrained and sums are zero), blocks at the minimum size each stencil
allows (plain and fused against the oracle bitwise), and one large block that is its own
periodic neighbour on every axis (64-bit offsets, one block feeding all its ghosts)."""
import numpy as np
import pytest
import torch

from apophis_amd import hydro, problems
from apophis_amd.mesh import BlockSet, lattice_mesh, make_eos
from oracle import compare
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

DBL_MAX = np.finfo(np.float64).max


def _pin(p):
    return {"parthenon/time": {"cfl": p.cfl, "integrator": p.integrator},
            "parthenon/mesh": {"nghost": int(p.bs.blocks["ng"].min()) if p.bs.nblocks else 3},
            "hydro": {"reconstruction": p.recon, "riemann": p.riemann},
            "eos": {"gm1": float(p.eos["gm1"][0])}}


def test_empty_block_set(gpu_ctx):
    print("[case] empty block set", flush=True)
    bs = BlockSet(3, [])
    assert bs.nblocks == 0 and bs.field_size(5) == 0
    p = problems.Problem("empty", bs, np.zeros(0), make_eos(2.0 / 3.0), "ppm", "hllc", 0.3)
    pkg = hydro.initialize_hydro(_pin(p))
    md = hydro.MeshData(gpu_ctx, bs, p.cons)
    drv = hydro.ApophisDriver(pkg, md)
    assert drv.initialize() == DBL_MAX
    assert drv.step(1e-3) == DBL_MAX
    assert float(hydro.global_quant_hst(md, 0).item()) == 0.0
    fd = hydro.FusedDriver(gpu_ctx, bs, p.cons, "ppm", "hllc", p.eos, p.cfl)
    fd.initial_dt()
    torch.cuda.synchronize()
    assert float(fd.dt.item()) == DBL_MAX
    fd.set_dt(1e-3)
    fd.step(fixed_dt=True)
    torch.cuda.synchronize()


def _lattice(ndim, n, ng, nblk, recon, riemann, seed=3):
    nb = (nblk,) * ndim + (1,) * (3 - ndim)
    bnx = (n,) * ndim + (1,) * (3 - ndim)
    m = lattice_mesh(ndim, nb, bnx, ng, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), ["periodic"] * 6)
    bs = BlockSet(ndim, [m])
    u = np.zeros(bs.field_size(5))
    rng = np.random.default_rng(seed)
    for b in range(bs.nblocks):
        x, y, z = bs.cell_centers(b)
        v = bs.interior(u, b, 5)
        rho = 1.0 + 0.2 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + 0.01 * rng.random(x.shape)
        vx = 0.3 * np.cos(2 * np.pi * y)
        vy = -0.2 * np.sin(2 * np.pi * x)
        pres = 1.0 + 0.1 * np.cos(2 * np.pi * (x + y + z))
        v[0], v[1], v[2], v[3] = rho, rho * vx, rho * vy, 0.0
        v[4] = pres / 0.4 + 0.5 * rho * (vx * vx + vy * vy)
    return problems.Problem("lattice", bs, u, make_eos(0.4), recon, riemann, 0.3)


@pytest.mark.parametrize("ndim,n,recon", [(2, 2, "plm"), (3, 2, "plm"), (2, 3, "ppm"),
                                          (3, 3, "ppm")])
def test_minimum_block_sizes_match_oracle(gpu_ctx, ndim, n, recon):
    """Blocks of n = the stencil width (the fused form reads up to n cells of a neighbour)."""
    print(f"[case] minimum blocks {ndim}d n={n} {recon}", flush=True)
    ng = 2 if recon == "plm" else 3
    p = _lattice(ndim, n, ng, 4, recon, "hllc")
    ref = orc.State(p)
    dt = ref.estimate_dt()
    md = hydro.MeshData(gpu_ctx, p.bs, p.cons)
    drv = hydro.ApophisDriver(hydro.initialize_hydro(_pin(p)), md)
    assert drv.initialize() == dt
    drv.step(dt)
    fd = hydro.FusedDriver(gpu_ctx, p.bs, p.cons, recon, "hllc", p.eos, p.cfl)
    fd.set_dt(dt)
    fd.step(fixed_dt=True)
    ref.step(dt)
    torch.cuda.synchronize()
    assert compare.bitwise_equal_interior(md.cons.cpu().numpy(), ref.u0, p.bs)
    assert compare.bitwise_equal_interior(fd.u0.cpu().numpy(), ref.u0, p.bs)


def test_single_large_self_periodic_block(gpu_ctx):
    """One 96^3 block, periodic onto itself on every axis, PPM: fused and plain against
    the oracle (one step)."""
    print("[case] one 96^3 block", flush=True)
    p = _lattice(3, 96, 3, 1, "ppm", "hllc")
    assert p.bs.nblocks == 1 and (p.bs.blocks["nbr"][0][[12, 14, 10, 16, 4, 22]] == 0).all()
    ref = orc.State(p)
    dt = ref.estimate_dt()
    fd = hydro.FusedDriver(gpu_ctx, p.bs, p.cons, "ppm", "hllc", p.eos, p.cfl)
    fd.set_dt(dt)
    fd.step(fixed_dt=True)
    ref.step(dt)
    torch.cuda.synchronize()
    assert compare.bitwise_equal_interior(fd.u0.cpu().numpy(), ref.u0, p.bs)
