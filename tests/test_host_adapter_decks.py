"""The C++ host adapter's deck handling (apophis_amd/csrc/host): InitializeGravity's
PARTHENON_FAIL messages (gravity.cpp:220-370) and the runner's own refusals, which all
come before any device call, so this runs without a GPU.  The runs themselves are in
tests/test_gpu_host_adapter.py."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "apophis_amd", "_build", "apophis_run")
EVRARD = os.path.join(ROOT, "apophis_amd", "inputs", "evrards_collapse.in")


@pytest.mark.parametrize("overrides,message", [
    (["gravity/type=bogus"], "[Apophis]: Gravity not recognized. Exiting."),
    (["gravity/solver=CG"], "Unknown solver type CG."),
    (["gravity/boundary_prolongation=Cubic"],
     "Unknown prolongation method for gravity boundaries."),
    (["gravity/set_flux_boundary=true"], "set_flux_boundary = true is not built"),
    (["gravity/type=monopole", "gravity/rmax=-1"],
     "[Apophis]: gravity/rmax must be positive for monopole gravity."),
    (["gravity/type=constant"], "no problem generator here sets the field"),
    (["parthenon/mesh/refinement=static"], "one uniform lattice"),
    (["parthenon/meshblock/nx2=12"], "meshblock nx2 must divide mesh nx2"),
    (["parthenon/meshblock/nx1=0"], "mesh and meshblock nx1 must be positive"),
    (["parthenon/mesh/ix3_bc=bogus"], "unknown boundary condition bogus"),
    (["hydro/reconstruction=ppm"], "Not enough ghost zones for reconstruction"),
    (["job/problem_id=blast"], "implements problem_id = sod and evrards_collapse"),
])
def test_evrard_deck_refusals(overrides, message):
    r = subprocess.run([EXE, EVRARD, *overrides], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1, (r.stdout, r.stderr)
    assert message in r.stderr, r.stderr
