"""The reference's 3D linear-wave convergence problem (linear_wave.cpp, convergence.py)
on the oracle, and the error-file / history formats (apophis_amd/output.py).

* the analytic state is a sound wave of unit speed along the grid diagonal with
  lambda = 1 (linear_wave.cpp:76-117): the initial state has the eigenvector's structure;
* vl2 + PLM + HLLC: the RMS L1 error after two periods falls with resolution (the GPU
  test runs the reference's full 4 x 3 matrix and its 4.63e-8 threshold);
* linearwave-errors.dat has the reference's layout: convergence.py reads column 4 of
  the last row as the RMS error.
"""
import numpy as np

from apophis_amd import output, problems
from oracle import oracle as orc


def test_initial_state_is_a_diagonal_sound_wave():
    p = problems.linear_wave3d(res=8)
    assert abs(p.meta["lambda"] - 1.0) < 1e-14
    assert abs(p.tlim - 2.0) < 1e-14  # two periods at |ev| = 1
    u = p.bs.interior(p.cons, 0, 5)
    dr = u[0] - 1.0
    # momentum = -a * drho along the unit wave vector, a = 1
    k = np.array([np.cos(np.arctan(np.sqrt(0.8))) * np.cos(np.arctan(2.0)),
                  np.cos(np.arctan(np.sqrt(0.8))) * np.sin(np.arctan(2.0)),
                  np.sin(np.arctan(np.sqrt(0.8)))])
    for d in range(3):
        np.testing.assert_allclose(u[1 + d], -dr * k[d], rtol=1e-9, atol=1e-21)
    np.testing.assert_allclose(np.linalg.norm(k), 1.0, rtol=1e-15)


def test_oracle_converges_and_error_file_layout(tmp_path):
    path = tmp_path / "linearwave-errors.dat"
    errs = []
    for res in (16, 32):
        p = problems.linear_wave3d(res=res)
        st = orc.State(p)
        t, n = st.run(tlim=p.tlim)
        assert abs(t - p.tlim) < 1e-12
        e = output.linear_wave_errors(p.bs, st.u0, p.meta["exact"], p.meta["domain_volume"])
        output.write_linear_wave_errors(str(path), p.meta["nx"], n, e)
        errs.append(e["rms"])
    assert errs[1] < errs[0] / 2.5  # close to second order
    assert errs[1] < 1e-7
    lines = open(path).read().splitlines()
    assert lines[0] == ("# Nx1  Nx2  Nx3  Ncycle  RMS-L1-Error  d_L1  M1_L1  M2_L1  M3_L1  "
                        "E_L1   Largest-Max/L1  d_max  M1_max  M2_max  M3_max  E_max ")
    data = np.genfromtxt(path)
    assert data.shape == (2, 16)
    np.testing.assert_array_equal(data[:, 0], [32, 64])
    np.testing.assert_allclose(data[:, 4], errs, rtol=1e-6)  # %e keeps 7 digits


def test_hst_output_layout(tmp_path):
    """The history writer's layout (upstream Parthenon's, restated; parity unpinned):
    file name, the two header lines, blank-separated %14.6e reals, integer cycle and
    block count, and the output times every dt from the first call."""
    h = output.HstOutput(str(tmp_path), dt=0.05, block_number=1)
    t, n = 0.0, 0
    for dt in (0.02, 0.02, 0.02, 0.03, 0.03, 0.01):
        h.maybe_write(t, dt, n, 8, lambda: [1.0 + t, 2.5 - t])
        t += dt
        n += 1
    path = tmp_path / "parthenon.out1.hst"
    lines = open(path).read().splitlines()
    assert lines[0] == "#  History data"
    assert lines[1] == ("# [1]=time     [2]=dt       [3]=cycle    [4]=nbtotal  "
                        "[5]=total_mass[6]=total_energy")
    assert lines[2] == ("   0.000000e+00   2.000000e-02  0   8    1.000000e+00"
                        "   2.500000e+00")
    data = np.genfromtxt(path)
    assert data.shape == (3, 6)  # t = 0, 0.06 (>= 0.05), 0.12 (>= 0.10)
    np.testing.assert_allclose(data[:, 0], [0.0, 0.06, 0.12])
    np.testing.assert_array_equal(data[:, 2], [0, 3, 5])
