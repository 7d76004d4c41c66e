"""The PyTorch restatement of a multilevel step (tests/torch_amr.py), which the GPU tests of
the multilevel adjoint differentiate with autograd, against the oracle's multilevel step
(oracle/amr_oracle.py AmrState): the same state after two steps, bitwise, on periodic
meshes of two and three levels and on meshes with outflow and reflecting faces, with PLM
and PPM."""
import numpy as np
import pytest
import torch

from oracle import amr_oracle as ao
from tests import test_amr, test_amr_physical, torch_amr

CASES = [("2d_L", "plm", "hllc", test_amr), ("2d_ppm", "ppm", "hllc", test_amr),
         ("3d", "plm", "hlle", test_amr), ("2d_3lev", "plm", "hllc", test_amr),
         ("2d_out", "plm", "hllc", test_amr_physical),
         ("2d_ref", "ppm", "lhllc", test_amr_physical),
         ("3d_ref", "ppm", "hllc", test_amr_physical)]


@pytest.mark.parametrize("name,recon,riemann,mod", CASES,
                         ids=[f"{c[0]}-{c[1]}-{c[2]}" for c in CASES])
def test_restatement_steps_equal_oracle(name, recon, riemann, mod):
    m = mod._mesh(name)
    p = mod._problem(m, recon=recon)
    p.riemann = riemann
    ref = ao.AmrState(p, m)
    dt = ref.estimate_dt()
    eos = (float(p.eos["gm1"][0]), float(p.eos["dfloor"][0]), float(p.eos["pfloor"][0]))
    tm = torch_amr.Multilevel(m, p.bs, recon, riemann, eos)
    u = tm.interior(p.cons)
    with torch.no_grad():
        for _ in range(2):
            u = tm.step(u, dt)
            ref.step(dt)
    want = tm.interior(ref.u0)
    assert torch.equal(u, want)
