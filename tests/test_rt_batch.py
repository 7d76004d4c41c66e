"""Batched Rayleigh-Taylor workload (CPU): the batch generator, its differentiable twin
and the mixing-layer width loss.

* every instance of problems.rt_batch equals problems.rt2d with that instance's grid and
  amplitudes, bitwise;
* adjoint.rt_state (torch) reproduces the generator's state;
* adjoint.mixing_width on a linear heavy-fraction ramp of N cells across a layer of
  width h gives the midpoint sum of 4 t (1 - t) exactly:  W = h (2/3 + 1/(3 N^2)), and
  W = 0 on the unmixed state.
"""
import numpy as np
import torch

from apophis_amd import adjoint, problems

GRIDS = ((32, 96), (64, 192))


def test_instances_equal_rt2d_bitwise():
    p = problems.rt_batch(ninst=4, grids=GRIDS, block=16, seed=3)
    assert len(set(p.meta["grids"])) == 2  # both grid sizes drawn
    off = 0
    for i, (nx, ny) in enumerate(p.meta["grids"]):
        q = problems.rt2d(nx=nx, ny=ny, nblk=(nx // 16, ny // 16), amps=p.meta["amps"][i],
                          mesh_id=i)
        for b in range(q.bs.nblocks):
            assert int(p.bs.blocks["mesh_id"][off + b]) == i
            a = p.bs.interior(p.cons, off + b, 5)
            c = q.bs.interior(q.cons, b, 5)
            assert np.array_equal(a.view(np.uint64), c.view(np.uint64)), (i, b)
        off += q.bs.nblocks
    assert off == p.bs.nblocks
    assert (p.meta["amps"] >= 0).all() and (p.meta["amps"] < 0.01).all()


def test_rt_state_matches_generator():
    for width in (0.0, 0.05):
        p = problems.rt_batch(ninst=3, grids=GRIDS, block=16, seed=4, width=width)
        u = adjoint.rt_state(p.bs, torch.from_numpy(p.meta["amps"]), width=width)
        np.testing.assert_allclose(u.numpy(), p.cons, rtol=1e-14, atol=1e-15)


def test_smooth_interface_is_hydrostatic():
    """dp/dy = g rho for the smooth-interface state (central differences, O(dy^2))."""
    p = problems.rt2d(nx=16, ny=1536, nblk=(1, 16), width=0.05, amps=np.zeros(6))
    _, y, _ = p.bs.cell_centers_all()
    yy = y[:, 0, :, 0].reshape(-1)
    rho = p.bs.interior_index()[:, 0, :]
    r = p.cons[rho].reshape(p.bs.nblocks, -1, 16)[:, :, 0].reshape(-1)
    en = p.cons[p.bs.interior_index()[:, 4, :]].reshape(p.bs.nblocks, -1, 16)[:, :, 0]
    mom = p.cons[p.bs.interior_index()[:, 2, :]].reshape(p.bs.nblocks, -1, 16)[:, :, 0]
    pres = 0.4 * (en.reshape(-1) - 0.5 * mom.reshape(-1) ** 2 / r)
    dy = yy[1] - yy[0]
    dpdy = (pres[2:] - pres[:-2]) / (2 * dy)
    np.testing.assert_allclose(dpdy, -0.1 * r[1:-1], rtol=1e-4)


def test_mixing_width_closed_form():
    p = problems.rt_batch(ninst=2, grids=GRIDS, block=16, seed=5)
    u = torch.from_numpy(p.cons.copy())
    w0 = adjoint.mixing_width(u, p.bs, per_instance=True)
    assert torch.equal(w0, torch.zeros(2, dtype=torch.float64))  # sharp interface
    idx = p.bs.interior_index()[:, 0, :]
    _, y, _ = p.bs.cell_centers_all()
    dy = p.bs.blocks["dx"][:, 1][:, None, None, None]
    ncell = 8  # ramp of 8 cells centred on y = 0 in every instance
    h = ncell * dy
    t = np.clip((y + h / 2) / h, 0.0, 1.0)
    rho = (1.0 + t).reshape(p.bs.nblocks, -1)
    u[torch.from_numpy(idx.reshape(-1))] = torch.from_numpy(rho.reshape(-1))
    w = adjoint.mixing_width(u, p.bs, per_instance=True).numpy()
    mid = p.bs.blocks["mesh_id"]
    for i in range(2):
        hi = float(h[mid == i].ravel()[0])
        expect = hi * (2.0 / 3.0 + 1.0 / (3.0 * ncell ** 2))
        assert abs(w[i] - expect) <= 1e-13 * expect, (i, w[i], expect)
    assert abs(float(adjoint.mixing_width(u, p.bs)) - w.sum()) <= 1e-15 * w.sum()
