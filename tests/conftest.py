import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def meshgen():
    from openaccel_amd import meshgen as mg
    return mg


def rel_err(a, b):
    """max |a - b| over the largest magnitude of b."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def row_rel_err(vals, ref, row_ptr):
    """Largest error of a CRS matrix row relative to that row's largest
    coefficient (vals / ref: [nnz][...] blocks in CRS order): a small
    boundary row is held to its own scale, not to the matrix's."""
    a = np.abs(np.asarray(vals, dtype=np.float64) - np.asarray(ref, dtype=np.float64))
    a = a.reshape(a.shape[0], -1).max(axis=1)
    r = np.abs(np.asarray(ref, dtype=np.float64)).reshape(a.shape[0], -1).max(axis=1)
    starts = np.asarray(row_ptr[:-1])
    err = np.maximum.reduceat(a, starts)
    scale = np.maximum(np.maximum.reduceat(r, starts), 1e-300)
    return float((err / scale).max())


def block_rel_err(a, b, k):
    """Largest error of a k-component block (a node's gradient, an element's
    ip mass flow rates) relative to that block's largest magnitude: a
    component that nearly cancels is held to its block's scale."""
    a = np.asarray(a, dtype=np.float64).reshape(-1, k)
    b = np.asarray(b, dtype=np.float64).reshape(-1, k)
    scale = np.maximum(np.abs(b).max(axis=1), 1e-300)
    return float((np.abs(a - b).max(axis=1) / scale).max())


def rhs_rel_err(b, b_ref, vals_ref, row_ptr, cols, x, block=False):
    """Error of a block-CRS system's rhs relative, row by row, to the size of
    the terms it sums: sum_j |A_ij| |x_j| + |b_i| (x: the field the system is
    assembled about).  Near convergence b is a small difference of large
    fluxes; this holds each row to the fluxes' scale.  block: a node's
    components share the largest of their scales (a velocity component that
    vanishes, as the spanwise one of a two-dimensional flow, still sums
    cross-component viscous terms of the others' size)."""
    vals_ref = np.asarray(vals_ref, dtype=np.float64)
    nnz = vals_ref.shape[0]
    N = int(round(np.sqrt(vals_ref.size // nnz)))
    A = np.abs(vals_ref.reshape(nnz, N, N))
    rp = np.asarray(row_ptr)
    nn = rp.size - 1
    xa = np.abs(np.asarray(x, dtype=np.float64).reshape(nn, N))
    terms = np.einsum("sij,sj->si", A, xa[np.asarray(cols)])          # [nnz][N]
    scale = np.add.reduceat(terms, rp[:-1], axis=0)                   # [nn][N]
    br = np.asarray(b_ref, dtype=np.float64).reshape(nn, N)
    scale = np.maximum(scale + np.abs(br), 1e-300)
    if block:
        scale = np.repeat(scale.max(axis=1, keepdims=True), N, axis=1)
    err = np.abs(np.asarray(b, dtype=np.float64).reshape(nn, N) - br)
    return float((err / scale).max())


HEX_LRSCV = np.array([[0, 1], [1, 2], [2, 3], [0, 3], [4, 5], [5, 6], [6, 7], [4, 7], [0, 4],
                      [1, 5], [2, 6], [3, 7]])


def mass_rel_err(b, b_ref, mdot, conn):
    """Error of a p' rhs (a node's net mass imbalance) relative to the mass
    flow through that node's control-volume faces, sum over its SCS ips of
    |mdot| (hex8 meshes): near convergence the imbalance is a small
    difference of those flows."""
    md = np.abs(np.asarray(mdot, dtype=np.float64)).reshape(-1, 12)
    c = np.asarray(conn)[:, :8]
    nn = np.asarray(b_ref).size
    scale = np.zeros(nn)
    for side in (0, 1):
        np.add.at(scale, c[:, HEX_LRSCV[:, side]].ravel(), md.ravel())
    br = np.asarray(b_ref, dtype=np.float64).ravel()
    scale = np.maximum(scale + np.abs(br), 1e-300)
    return float((np.abs(np.asarray(b, dtype=np.float64).ravel() - br) / scale).max())


def entry_rel_err(a, b, floor=1e-3):
    """Largest per-entry error relative to max(|b_i|, floor * max|b|): every
    entry above `floor` of the array's scale is held to its own magnitude."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    scale = np.maximum(np.abs(b), floor * max(np.abs(b).max(), 1e-300))
    return float((np.abs(a - b) / scale).max())
