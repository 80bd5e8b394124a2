"""Synthetic hex8 meshes for the lid-driven cavity workloads.

The reference reads Exodus meshes through STK (src/mesh/meshIO.cpp:24); the
workloads here are generated instead, with the reference's own conventions:

* hex8 node ordering of Hex8CVFEM.C:428-432 (s1 = x, s2 = y, s3 = z),
* side ordinals of Hex8CVFEM.C:419-426 (0: -y, 1: +x, 2: +y, 3: -x, 4: -z, 5: +z),
* boundary faces as (parent element, face ordinal, boundary id), the data a
  side bucket carries in navierStokesAssemblerElemBoundaryConditions.cpp:2775-2790.

Node ids are lexicographic (x fastest), element ids likewise.  Boundary id 0 is
the moving lid (y = 1), id 1 every other wall.
"""
from __future__ import annotations

import numpy as np

LID, WALL = 0, 1


def cavity_hex(n: int, length: float = 1.0, perturb: float = 0.0, seed: int = 0):
    """Uniform n^3 hex8 grid of the cube [0, length]^3.

    perturb > 0 moves every node by U(-perturb*h, perturb*h) per component,
    boundary nodes only tangentially (the component normal to each boundary
    plane they lie on is kept), seeded by `seed`.
    """
    h = length / n
    m = n + 1
    g = np.arange(m, dtype=np.float64) * h
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    if perturb > 0.0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-perturb * h, perturb * h, size=xyz.shape)
        ijk = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m),
                                   indexing="ij"), axis=-1).reshape(-1, 3)[:, ::-1]
        on_bnd = (ijk == 0) | (ijk == n)
        d[on_bnd] = 0.0
        xyz = xyz + d
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = (i + m * (j + m * k)).ravel().astype(np.int64)
    off = np.array([0, 1, 1 + m, m, m * m, 1 + m * m, 1 + m + m * m, m + m * m],
                   dtype=np.int64)
    conn = (base[:, None] + off[None, :]).astype(np.int32)
    eid = (i + n * (j + n * k)).ravel()
    ii, jj, kk = i.ravel(), j.ravel(), k.ravel()
    faces = []
    for sel, ordv, bc in ((jj == n - 1, 2, LID), (jj == 0, 0, WALL),
                          (ii == n - 1, 1, WALL), (ii == 0, 3, WALL),
                          (kk == 0, 4, WALL), (kk == n - 1, 5, WALL)):
        e = eid[sel]
        faces.append(np.stack([e, np.full_like(e, ordv), np.full_like(e, bc)], axis=1))
    bf = np.concatenate(faces).astype(np.int32)
    return {
        "xyz": np.ascontiguousarray(xyz),
        "conn": np.ascontiguousarray(conn),
        "bf_elem": np.ascontiguousarray(bf[:, 0]),
        "bf_ord": np.ascontiguousarray(bf[:, 1]),
        "bf_bc": np.ascontiguousarray(bf[:, 2]),
        "n": n,
    }


def initial_velocity(n_nodes: int, seed: int, sigma: float = 1.0e-3) -> np.ndarray:
    """Seeded N(0, sigma) initial velocity [n_nodes][3]."""
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(rng.normal(0.0, sigma, size=(n_nodes, 3)))
