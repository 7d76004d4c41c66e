"""GPU checks of the launch autotuner (aph_plan_autotune / aph_plan_get_tile /
aph_plan_get_launch / aph_plan_set_workgroup): it times every block shape of a plan for
each workgroup size, its choices divide the blocks, tiles set by hand are kept, and --
since tile, segment length and workgroup size only decide which lane computes which cell
-- a step taken with the tuned launch equals a step with the default launch bitwise."""
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
    assert report.count("chosen: workgroup") == 1, report
    for wg in (64, 128, 256):
        assert f"workgroup {wg}, block" in report, report
    tiles = fd.tiles()
    assert set(tiles) == shapes
    launch = fd.launch()
    assert set(launch) == shapes
    assert len({v[0] for v in launch.values()}) == 1  # one workgroup size per plan
    assert launch[next(iter(shapes))][0] in (64, 128, 256)
    for shape, t in tiles.items():
        assert all(s % d == 0 for s, d in zip(shape, t)), (shape, t)
        assert all(d % g == 0 for d, g in zip(t, launch[shape][1:])), (t, launch[shape])
        assert f"  block {shape[0]}x{shape[1]}x{shape[2]}: tile {t[0]}x{t[1]}x{t[2]}" in report
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


@pytest.mark.parametrize("lanes", [128, 256])
@pytest.mark.parametrize("case", ["uneven3d", "kh2d", "sod1d"])
def test_workgroup_size_bitwise(gpu_ctx, case, lanes):
    """Steps with 128- and 256-lane workgroups (tiles of 64 cells leave lanes idle) equal
    the 64-lane steps bit for bit, final dt included."""
    print(f"[case] workgroup {lanes} {case}", flush=True)
    if case == "uneven3d":
        p = problems.uneven_blocks(nblocks=24, seed=5)
    elif case == "kh2d":
        p = problems.kh2d(nblk=4, n=32)
    else:
        p = problems.sod()
    a = _driver(gpu_ctx, p)
    ref = _steps(a, 3)
    fd = _driver(gpu_ctx, p)
    fd.set_workgroup(lanes)
    assert all(v[0] == lanes for v in fd.launch().values())
    got = _steps(fd, 3)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    assert fd.dt.item() == a.dt.item()


def test_set_workgroup_refuses_other_sizes(gpu_ctx):
    from apophis_amd import hydro

    fd = _driver(gpu_ctx, problems.sod())
    with pytest.raises(hydro._lib.HydroError):
        fd.set_workgroup(96)


def test_lds_reserve_changes_no_bit(gpu_ctx):
    """A larger LDS reservation only lowers the resident workgroups per CU."""
    p = problems.uneven_blocks(nblocks=24, seed=5)
    print("[case] lds reserve", flush=True)
    ref = _steps(_driver(gpu_ctx, p), 3)
    fd = _driver(gpu_ctx, p)
    fd.set_lds_reserve(96 << 10)
    got = _steps(fd, 3)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
