"""GPU checks of the tile autotuner (aph_plan_autotune / aph_plan_get_tile): it times
every block shape of a plan, its choices divide the blocks, tiles set by hand are kept,
and -- since a tile only decides which lane computes which cell -- a step taken with the
tuned tiles equals a step with the default tiles bitwise."""
import numpy as np
import pytest
import torch

from apophis_amd import problems

pytestmark = pytest.mark.gpu


def _driver(ctx, p):
    from apophis_amd import hydro

    return hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl)


def _steps(fd, n):
    fd.initial_dt()
    for _ in range(n):
        fd.step()
    torch.cuda.synchronize()
    return fd.u0.cpu().numpy()


@pytest.mark.parametrize("case", ["uneven3d", "kh2d", "sod1d"])
def test_autotune_choices_and_bitwise_steps(gpu_ctx, case):
    print(f"[case] autotune {case}", flush=True)
    if case == "uneven3d":
        p = problems.uneven_blocks(nblocks=24, seed=5)
    elif case == "kh2d":
        p = problems.kh2d(nblk=4, n=32)
    else:
        p = problems.sod()
    ref = _steps(_driver(gpu_ctx, p), 3)
    fd = _driver(gpu_ctx, p)
    fd.initial_dt()
    report = fd.autotune(reps=2)
    shapes = {tuple(int(v) for v in p.bs.blocks["nx"][b]) for b in range(p.bs.nblocks)}
    assert report.count("chosen") == len(shapes), report
    tiles = fd.tiles()
    assert set(tiles) == shapes
    for shape, t in tiles.items():
        assert all(s % d == 0 for s, d in zip(shape, t)), (shape, t)
    fd.u0.copy_(torch.from_numpy(p.cons))
    fd.time.zero_()
    got = _steps(fd, 3)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_autotune_keeps_user_tile(gpu_ctx):
    p = problems.uneven_blocks(nblocks=16, seed=6, sizes=(8, 16), probs=(0.5, 0.5))
    print("[case] autotune keeps user tile", flush=True)
    fd = _driver(gpu_ctx, p)
    fd.initial_dt()
    fd.set_tile(16, 16, 4, 4)
    report = fd.autotune(reps=1)
    assert "not tuned" in report
    b16 = [b for b in range(p.bs.nblocks) if int(p.bs.blocks["nx"][b][0]) == 16][0]
    assert fd.tiles()[tuple(int(v) for v in p.bs.blocks["nx"][b16])] == (16, 4, 4)
