"""Residual-history files and system dumps (openaccel_amd.sysio) in the
reference's formats: file naming (<system>_NNN.out numbered after the
system's existing files, <name>_<call>_b.bin), the header's column names,
the tab-separated lines with 3 / 6 significant digits, raw FP64 rhs in node
order; the block-CRS matrix dump round-trips."""
import os

import numpy as np

from openaccel_amd import sysio


def test_history_naming_and_format(tmp_path):
    d = str(tmp_path / "0" / "residuals")
    h1 = sysio.ResidualHistory(d, "momentum equations", ["U-momentum", "V-momentum", "W-momentum"])
    h2 = sysio.ResidualHistory(d, "momentum equations", ["U-momentum", "V-momentum", "W-momentum"])
    h3 = sysio.ResidualHistory(d, "pressure correction", ["P-correction"])
    assert os.path.basename(h1.path) == "momentum_equations_001.out"
    assert os.path.basename(h2.path) == "momentum_equations_002.out"
    assert os.path.basename(h3.path) == "pressure_correction_001.out"
    h1.write(1, 1, 7, 0.0, [1.25e-2, 3.5e-3, 1e-9])
    h1.write(2, 2, 6, 0.0, [1.0e-2, 3.0e-3, 2e-9])
    text = open(h1.path).read().splitlines()
    assert text[0].startswith("# Accel solver timestamp: ")
    assert text[-3] == ("# global_iterations\tinner_iterations\tsolver_calls\tsolver_iterations\t"
                        "sim_time[s]\tU-momentum[control_residual]\tV-momentum[control_residual]\t"
                        "W-momentum[control_residual]")
    assert text[-2] == "1\t1\t1\t7\t0.000e+00\t1.250000e-02\t3.500000e-03\t1.000000e-09"
    names, rows = sysio.read_history(h1.path)
    assert names[5] == "U-momentum[control_residual]"
    np.testing.assert_array_equal(rows[:, 2], [1, 2])
    np.testing.assert_allclose(rows[1, 5:], [1.0e-2, 3.0e-3, 2e-9])


def test_system_dump_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    rp = np.array([0, 2, 5, 7], dtype=np.int32)
    cols = np.array([0, 1, 0, 1, 2, 1, 2], dtype=np.int32)
    vals = rng.normal(size=(7, 3, 3))
    b = rng.normal(size=9)
    a_path, b_path = sysio.write_system(str(tmp_path / "momentum_equations"), 12, rp, cols, vals, b)
    assert os.path.basename(b_path) == "momentum_equations_0012_b.bin"
    # the rhs file is raw FP64 in node order, nothing else
    assert os.path.getsize(b_path) == 9 * 8
    np.testing.assert_array_equal(np.fromfile(b_path, dtype=np.float64), b)
    rp2, cols2, vals2, b2 = sysio.read_system(a_path, b_path)
    assert np.array_equal(rp2, rp) and np.array_equal(cols2, cols)
    assert np.array_equal(vals2, vals) and np.array_equal(b2, b)
