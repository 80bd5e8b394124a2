"""On-disk outputs of a linear system, written from the device path.

* Residual history, as linearSystem<N>::initializeHistory_ / writeHistory_
  write it (src/equation/linearSystem.h:573-670): one file per system and
  run, <residual_dir>/<system name, blanks as '_'>_NNN.out with NNN one more
  than the files of that system already there; a header of '# ' comment
  lines ending with the column names global_iterations, inner_iterations,
  solver_calls, solver_iterations, sim_time[s] and <equation>[control_residual]
  per block component; then one tab-separated line per solve, the time with
  3 and the residuals with 6 significant digits in scientific notation.
  (The reference's default residual directory is
  <post-processing dir>/0/residuals, simulationIO.cpp:998.)
* System dump, as linearSystem<N>::writeLinearSystem_ (linearSystem.h:672-729):
  <name>_<call count, 4 digits>_b.bin holds the right-hand side, raw FP64 in
  node order ([node][N]), as writeVector_ writes it.  The matrix goes through
  liblinsolve's Matrix::writeMatrix, whose source the reference tree does not
  hold; here it is <name>_<call>_A.bin in a format of this module: int32 n,
  N, nnz, then row_ptr[n+1] and cols[nnz] (int32) and the block values
  [nnz][N][N] (FP64, row-major blocks) -- the block-CRS the reference's
  rowVals / rowCols expose.
"""
from __future__ import annotations

import datetime
import os

import numpy as np

COMMENT = "# "


def canonical_name(system_name: str) -> str:
    return system_name.replace(" ", "_")


class ResidualHistory:
    """Residual history file of one linear system (writeHistory_ format)."""

    def __init__(self, residual_dir: str, system_name: str, equation_names, info_lines=()):
        os.makedirs(residual_dir, exist_ok=True)
        base = canonical_name(system_name)
        prefix = os.path.join(residual_dir, base)
        nfiles = sum(1 for e in os.listdir(residual_dir)
                     if os.path.join(residual_dir, e).startswith(prefix))
        self.path = f"{prefix}_{nfiles + 1:03d}.out"
        self.equations = list(equation_names)
        self.calls = 0
        with open(self.path, "w") as fh:
            now = datetime.datetime.now().strftime("%c")
            fh.write(f"{COMMENT}Accel solver timestamp: {now}\n")
            fh.write(f"{COMMENT}Git revision: openaccel_amd\n")
            fh.write(f"{COMMENT}\n")
            for line in info_lines:
                fh.write(f"{COMMENT}{line}\n")
            fh.write(f"{COMMENT}\n")
            fh.write(f"{COMMENT}System residual history:\n")
            cols = ["global_iterations", "inner_iterations", "solver_calls", "solver_iterations",
                    "sim_time[s]"] + [f"{e}[control_residual]" for e in self.equations]
            fh.write(COMMENT + "\t".join(cols) + "\n")

    def write(self, global_iteration: int, inner_iteration: int, solver_iterations: int,
              sim_time: float, control_residual):
        """One line per solve (the solver call count advances by one)."""
        res = np.atleast_1d(np.asarray(control_residual, dtype=np.float64))
        assert res.size == len(self.equations)
        self.calls += 1
        parts = [str(int(global_iteration)), str(int(inner_iteration)), str(self.calls),
                 str(int(solver_iterations)), f"{sim_time:.3e}"] + [f"{r:.6e}" for r in res]
        with open(self.path, "a") as fh:
            fh.write("\t".join(parts) + "\n")


def read_history(path: str):
    """(column names, rows [n][5 + components]) of a residual history file."""
    names, rows = None, []
    with open(path) as fh:
        for line in fh:
            if line.startswith(COMMENT):
                body = line[len(COMMENT):].rstrip("\n")
                if body.startswith("global_iterations"):
                    names = body.split("\t")
                continue
            if line.strip():
                rows.append([float(v) for v in line.split("\t")])
    return names, np.array(rows)


def write_system(prefix: str, call_count: int, row_ptr, cols, vals, b):
    """Dump a block-CRS system: <prefix>_<call>_b.bin (raw FP64 rhs, as the
    reference's writeVector_) and <prefix>_<call>_A.bin (this module's
    block-CRS format); returns the two paths."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(cols, dtype=np.int32)
    v = np.ascontiguousarray(vals, dtype=np.float64)
    n, nnz = rp.size - 1, ci.size
    N = int(round(np.sqrt(v.size // max(nnz, 1)))) if nnz else 1
    assert v.size == nnz * N * N and np.asarray(b).size == n * N
    stem = f"{prefix}_{call_count:04d}"
    with open(stem + "_A.bin", "wb") as fh:
        np.array([n, N, nnz], dtype=np.int32).tofile(fh)
        rp.tofile(fh)
        ci.tofile(fh)
        v.tofile(fh)
    np.ascontiguousarray(b, dtype=np.float64).tofile(stem + "_b.bin")
    return stem + "_A.bin", stem + "_b.bin"


def read_system(a_path: str, b_path: str):
    """(row_ptr, cols, vals [nnz][N][N], b) of a dump written by write_system."""
    with open(a_path, "rb") as fh:
        n, N, nnz = (int(x) for x in np.fromfile(fh, dtype=np.int32, count=3))
        rp = np.fromfile(fh, dtype=np.int32, count=n + 1)
        ci = np.fromfile(fh, dtype=np.int32, count=nnz)
        vals = np.fromfile(fh, dtype=np.float64, count=nnz * N * N).reshape(nnz, N, N)
    b = np.fromfile(b_path, dtype=np.float64)
    assert b.size == n * N
    return rp, ci, vals, b


def write_flow_system(flow, N: int, prefix: str, call_count: int):
    """Dump the flow's current system N (3 momentum, 1 pressure correction or
    the last scalar system) from the device (oagpu_system_download)."""
    rp, cols, _ = flow.mesh.graph()
    vals, b = flow.system(N)
    return write_system(prefix, call_count, rp, cols, vals, b)
