"""aph_global_sum_noalloc -- the history sum from the library's fixed device scratch, for
callers that cannot allocate -- against aph_global_sum through the C ABI: the same bits
on two calls in a row (the second on a smaller block set, so partial sums left by the
first call must not leak into it), and the refusal of a block set beyond its scratch."""
import ctypes

import numpy as np
import pytest
import torch

from apophis_amd import _lib, hydro, problems
from apophis_amd.mesh import BlockSet, lattice_mesh

pytestmark = pytest.mark.gpu


def _sums(ctx, dbs, cons, nvar, fn):
    out = torch.full((nvar,), -1.0, dtype=torch.float64, device="cuda")
    for v in range(nvar):
        ctx.check(fn(ctx.h, dbs.h, ctypes.c_void_p(cons.data_ptr()), v,
                     ctypes.c_void_p(out.data_ptr() + 8 * v)))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_sum_equals_allocating_sum_twice(gpu_ctx):
    L = _lib.lib()
    cases = [("uneven 96 blocks", problems.uneven_blocks(nblocks=96, seed=3)),
             ("blast 8 blocks", problems.blast3d(nblk=2, n=8))]
    for name, p in cases:
        print(f"[case] noalloc sum: {name}", flush=True)
        dbs = hydro.DeviceBlockSet(gpu_ctx, p.bs)
        cons = torch.from_numpy(p.cons).cuda()
        want = _sums(gpu_ctx, dbs, cons, p.bs.nvar, L.aph_global_sum)
        got = _sums(gpu_ctx, dbs, cons, p.bs.nvar, L.aph_global_sum_noalloc)
        assert got.tobytes() == want.tobytes(), name
        # a closed form of the sum the blocks hold: the interior mass
        mass = sum(float(p.bs.interior(p.cons, b, p.bs.nvar)[0].sum()) *
                   float(p.bs.blocks["vol"][b]) for b in range(p.bs.nblocks))
        assert got[0] == pytest.approx(mass, rel=1e-12)


def test_noalloc_sum_refuses_more_blocks_than_its_scratch(gpu_ctx):
    print("[case] noalloc sum: 65537 blocks refused", flush=True)
    L = _lib.lib()
    one = lattice_mesh(1, (1, 1, 1), (4, 1, 1), 2, (0, 0, 0), (1, 1, 1), ["outflow"] * 6)
    many = np.repeat(one, 65537)
    bs = BlockSet(1, [many])
    dbs = hydro.DeviceBlockSet(gpu_ctx, bs)
    tiny = torch.zeros(8, dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = L.aph_global_sum_noalloc(gpu_ctx.h, dbs.h, ctypes.c_void_p(tiny.data_ptr()), 0,
                                  ctypes.c_void_p(out.data_ptr()))
    assert rc == -1  # APH_E_INVALID
    assert b"65536" in L.aph_last_error(gpu_ctx.h)
    # the entry point still works afterwards
    p = problems.blast3d(nblk=2, n=8)
    d2 = hydro.DeviceBlockSet(gpu_ctx, p.bs)
    cons = torch.from_numpy(p.cons).cuda()
    assert (_sums(gpu_ctx, d2, cons, 5, L.aph_global_sum_noalloc).tobytes() ==
            _sums(gpu_ctx, d2, cons, 5, L.aph_global_sum).tobytes())
