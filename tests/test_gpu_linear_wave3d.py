"""The reference's linear-wave convergence test (tst/regression/test_suites/convergence/
convergence.py) run through the batched fused form on the GPU: 4 method configurations
(rk2/vl2 x HLLE/HLLC, PLM) x 3 resolutions of the 3D diagonal sound wave, errors written
to linearwave-errors.dat in the reference's layout and checked the way convergence.py
checks them (row count, RMS-L1 of the last row <= 4.63e-8, convergence.py:120-132); plus
the fused state against the oracle, bitwise, after a full run at 32x16x16."""
import numpy as np
import pytest
import torch

from apophis_amd import output, problems
from oracle import compare
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

METHODS = [("rk2", "plm", "hlle"), ("rk2", "plm", "hllc"), ("vl2", "plm", "hlle"),
           ("vl2", "plm", "hllc")]
RES = (16, 32, 64)


def _run_fused(ctx, p):
    from apophis_amd import hydro

    fd = hydro.FusedDriver(ctx, p.bs, p.cons, p.recon, p.riemann, p.eos, p.cfl, tlim=p.tlim,
                           integrator=p.integrator)
    fd.initial_dt()
    n = 0
    while float(fd.time.item()) < p.tlim * (1 - 1e-14):
        fd.step()
        n += 1
    torch.cuda.synchronize()
    return fd.u0.cpu().numpy(), n


def test_linear_wave_convergence_kat(gpu_ctx, tmp_path):
    path = tmp_path / "linearwave-errors.dat"
    rms = {}
    for integ, recon, rs in METHODS:
        for res in RES:
            print(f"[case] linear wave 3D {integ} {recon} {rs} res {res}", flush=True)
            p = problems.linear_wave3d(res=res, recon=recon, riemann=rs, integrator=integ)
            u, n = _run_fused(gpu_ctx, p)
            e = output.linear_wave_errors(p.bs, u, p.meta["exact"], p.meta["domain_volume"])
            output.write_linear_wave_errors(str(path), p.meta["nx"], n, e)
            rms[(integ, rs, res)] = e["rms"]
            print(f"  steps {n} rms-L1 {e['rms']:.4e}", flush=True)
    lines = open(path).read().splitlines()
    assert len(lines) == len(METHODS) * len(RES) + 1
    data = np.genfromtxt(path)
    assert data[-1, 4] <= 4.63e-08  # convergence.py:130, vl2 + PLM + HLLC at 128x64x64
    for integ, _, rs in METHODS:
        e = [rms[(integ, rs, r)] for r in RES]
        assert e[2] < e[1] < e[0]
        assert e[1] / e[2] > 2.5  # second order (convergence.py plots against it)


def test_fused_matches_oracle_after_full_run(gpu_ctx):
    print("[case] linear wave 3D fused vs oracle, 32x16x16, vl2 PLM HLLC", flush=True)
    p = problems.linear_wave3d(res=16)
    got, n = _run_fused(gpu_ctx, p)
    st = orc.State(p)
    t, n_ref = st.run(tlim=p.tlim)
    assert n == n_ref
    assert compare.bitwise_equal_interior(got, st.u0, p.bs)
