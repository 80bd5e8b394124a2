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


def cube_rotations() -> np.ndarray:
    """The 24 proper rotations of the reference cube as permutations of the
    hex8 local nodes: row r maps new local node k to old local node
    perm[r][k] (node k of the relabelled element sits where node perm[r][k]
    of the original was)."""
    v = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1],
                  [1, 1, 1], [-1, 1, 1]])
    perms = []
    for axes in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        for signs in np.ndindex(2, 2, 2):
            R = np.zeros((3, 3), dtype=int)
            for i, (a, s) in enumerate(zip(axes, signs)):
                R[i, a] = 1 if s == 0 else -1
            if round(np.linalg.det(R)) != 1:
                continue
            w = v @ R.T
            perms.append([int(np.flatnonzero((v == w[k]).all(axis=1))[0]) for k in range(8)])
    return np.array(perms, dtype=np.int64)


def relabel_hex(mesh: dict, seed: int) -> dict:
    """The same hex8 mesh with each element's local node order changed by a
    random proper cube rotation (positive Jacobians kept; node ids, node
    coordinates and element ids unchanged), boundary side ordinals remapped
    to the same physical faces.  A node then is the same local node of
    several elements, as in meshes from unstructured meshers."""
    rots = cube_rotations()
    rng = np.random.default_rng(seed)
    conn = np.asarray(mesh["conn"])
    ne = conn.shape[0]
    pick = rng.integers(0, len(rots), size=ne)
    perm = rots[pick]                                   # [ne][8]
    new = np.take_along_axis(conn[:, :8], perm, axis=1)
    # side s of the original (local nodes S) is the side of the relabelled
    # element whose local nodes map onto S
    side_sets = [frozenset(s) for s in _HEX_SIDES.tolist()]
    side_map = np.zeros((len(rots), 6), dtype=np.int64)
    for r, p in enumerate(rots):
        for s_new in range(6):
            image = frozenset(int(p[k]) for k in _HEX_SIDES[s_new])
            side_map[r, side_sets.index(image)] = s_new
    out = dict(mesh)
    out["conn"] = np.ascontiguousarray(new.astype(np.int32))
    bfe = np.asarray(mesh["bf_elem"])
    out["bf_ord"] = np.ascontiguousarray(
        side_map[pick[bfe], np.asarray(mesh["bf_ord"])].astype(np.int32))
    return out


def repeated_local_nodes(conn: np.ndarray, n_nodes: int) -> int:
    """Number of (node, local node) pairs held by more than one hex8 element:
    the incidences the row kernels visit through their extra list."""
    c = np.asarray(conn)[:, :8]
    key = c.astype(np.int64) * 8 + np.arange(8)[None, :]
    counts = np.bincount(key.ravel(), minlength=8 * n_nodes)
    return int(np.maximum(counts - 1, 0).sum())


def initial_velocity(n_nodes: int, seed: int, sigma: float = 1.0e-3) -> np.ndarray:
    """Seeded N(0, sigma) initial velocity [n_nodes][3]."""
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(rng.normal(0.0, sigma, size=(n_nodes, 3)))


# boundary ids of the backward-facing-step channel
BFS_INLET, BFS_OUTLET, BFS_WALL, BFS_SIDE = 0, 1, 2, 3


def graded(a: float, b: float, n: int, ratio: float) -> np.ndarray:
    """n cells on [a, b], refined towards both ends: cell sizes grow by
    `ratio` from each end to the middle (n even; ratio 1: uniform)."""
    if ratio == 1.0 or n < 2:
        return np.linspace(a, b, n + 1)
    half = n // 2
    h = ratio ** np.arange(half)
    sizes = np.concatenate([h, [ratio ** half] if n % 2 else [], h[::-1]])
    x = np.concatenate([[0.0], np.cumsum(sizes)])
    return a + (b - a) * x / x[-1]


def bfs_hex(nx_up: int, nx_dn: int, ny_low: int, ny_high: int, nz: int, step: float = 1.0,
            length_up: float = 5.0, length_dn: float = 25.0, height: float = 3.0,
            width: float = 4.0, ratio: float = 1.1, ratio_z: float = 1.0):
    """Backward-facing step: a channel of height `height` and width `width`
    with a step of height `step` at x = 0; the inlet section x in
    [-length_up, 0] spans y in [step, height], the section behind the step
    x in [0, length_dn] the full height.  Structured hex8 blocks: nx_up and
    nx_dn cells in x (uniform), ny_low cells in y below the step edge and
    ny_high above it, nz cells in z, the y cells graded by `ratio` from each
    wall (y = 0, step, height), the z cells by ratio_z from each side.

    Boundary ids: BFS_INLET (x = -length_up), BFS_OUTLET (x = length_dn),
    BFS_WALL (floor, step face, roof), BFS_SIDE (z = 0 and z = width).
    Returns the mesh dict of cavity_hex plus "first_dy" (the wall cells'
    thickness)."""
    xs = np.concatenate([np.linspace(-length_up, 0.0, nx_up + 1),
                         np.linspace(0.0, length_dn, nx_dn + 1)[1:]])
    ys = np.concatenate([graded(0.0, step, ny_low, ratio),
                         graded(step, height, ny_high, ratio)[1:]])
    zs = graded(0.0, width, nz, ratio_z)
    NX, NY, NZ = nx_up + nx_dn, ny_low + ny_high, nz
    # cells of the box; the block below the step edge upstream is removed
    k, j, i = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    alive = ~((i < nx_up) & (j < ny_low))
    # node (i, j, k) is used iff one of its up to 8 cells is alive
    used = np.zeros((NZ + 1, NY + 1, NX + 1), dtype=bool)
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                used[dk:dk + NZ, dj:dj + NY, di:di + NX] |= alive
    nid = np.full(used.shape, -1, dtype=np.int64)
    nid[used] = np.arange(int(used.sum()))
    kk, jj, ii = np.nonzero(used)  # lexicographic, x fastest
    xyz = np.stack([xs[ii], ys[jj], zs[kk]], axis=1)
    ci, cj, ck = i[alive], j[alive], k[alive]
    corners = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))
    conn = np.stack([nid[ck + dk, cj + dj, ci + di] for di, dj, dk in corners], axis=1)
    eid = np.full(alive.shape, -1, dtype=np.int64)
    eid[alive] = np.arange(int(alive.sum()))
    pad = np.zeros((NZ + 2, NY + 2, NX + 2), dtype=bool)
    pad[1:-1, 1:-1, 1:-1] = alive
    faces = []
    # side ordinal -> neighbour offset (di, dj, dk)
    for ordv, (di, dj, dk) in ((0, (0, -1, 0)), (1, (1, 0, 0)), (2, (0, 1, 0)),
                               (3, (-1, 0, 0)), (4, (0, 0, -1)), (5, (0, 0, 1))):
        nb = pad[1 + dk:1 + dk + NZ, 1 + dj:1 + dj + NY, 1 + di:1 + di + NX]
        sel = alive & ~nb
        e = eid[sel]
        if ordv == 3:
            bc = np.where(i[sel] == 0, BFS_INLET, BFS_WALL)
        elif ordv == 1:
            bc = np.full(e.shape, BFS_OUTLET)
        elif ordv in (4, 5):
            bc = np.full(e.shape, BFS_SIDE)
        else:
            bc = np.full(e.shape, BFS_WALL)
        faces.append(np.stack([e, np.full_like(e, ordv), bc], axis=1))
    bf = np.concatenate(faces).astype(np.int32)
    return {
        "xyz": np.ascontiguousarray(xyz),
        "conn": np.ascontiguousarray(conn.astype(np.int32)),
        "bf_elem": np.ascontiguousarray(bf[:, 0]),
        "bf_ord": np.ascontiguousarray(bf[:, 1]),
        "bf_bc": np.ascontiguousarray(bf[:, 2]),
        "first_dy": float(min(ys[1] - ys[0], ys[ny_low + 1] - ys[ny_low])),
    }


def bfs_wall_distance(xyz: np.ndarray, step: float = 1.0, height: float = 3.0) -> np.ndarray:
    """Distance to the nearest no-slip wall of bfs_hex (floor, step face,
    roof; the sides are not walls)."""
    x, y = xyz[:, 0], xyz[:, 1]
    roof = height - y
    floor = np.where(x < 0.0, y - step, y)
    # the step face x = 0, y in [0, step], seen from downstream
    face = np.where(x >= 0.0, np.hypot(x, np.maximum(y - step, 0.0)), np.inf)
    return np.minimum(np.minimum(roof, floor), face)


# ------------------------------------------------------------------ mixed
# element topology codes (the C ABI's OAGPU_HEX8 / TET4 / WED6 / PYR5)
HEX8, TET4, WED6, PYR5 = 0, 1, 2, 3
DUCT_INLET, DUCT_OUTLET, DUCT_WALL = 0, 1, 2

# hex8 sides (outward, counter-clockwise seen from outside), as above
_HEX_SIDES = np.array([[0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [0, 4, 7, 3], [0, 3, 2, 1],
                       [4, 5, 6, 7]])
# cube corner offsets (di, dj, dk) of the hex8 node order
_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1],
                     [0, 1, 1]])
# side -> neighbour cube offset
# side-node ordinals of every topology (side order, outward), the tables of
# the reference's master elements (HexSCS, TetSCS, WedSCS, PyrSCS)
SIDES = {HEX8: [[0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [0, 4, 7, 3], [0, 3, 2, 1], [4, 5, 6, 7]],
         TET4: [[0, 1, 3], [1, 2, 3], [0, 3, 2], [0, 2, 1]],
         WED6: [[0, 1, 4, 3], [1, 2, 5, 4], [0, 3, 5, 2], [0, 2, 1], [3, 4, 5]],
         PYR5: [[0, 1, 4], [1, 2, 4], [2, 3, 4], [0, 4, 3], [0, 3, 2, 1]]}


def boundary_sides(conn: np.ndarray, etype=None):
    """Sides that belong to one element only: (element, side ordinal) pairs
    in element order, and the sides' node ids ([n][4], -1 padded)."""
    conn = np.asarray(conn)
    et = np.zeros(conn.shape[0], dtype=np.int32) if etype is None else np.asarray(etype)
    keys, elems, ords, nodes = [], [], [], []
    for t, sides in SIDES.items():
        sel = np.where(et == t)[0]
        if sel.size == 0:
            continue
        for o, sn in enumerate(sides):
            nd = conn[sel][:, sn]
            k = np.sort(nd, axis=1)
            if k.shape[1] == 3:
                k = np.concatenate([np.full((k.shape[0], 1), -1), k], axis=1)
                nd = np.concatenate([nd, np.full((nd.shape[0], 1), -1)], axis=1)
            keys.append(k)
            elems.append(sel)
            ords.append(np.full(sel.size, o))
            nodes.append(nd)
    keys, elems = np.concatenate(keys), np.concatenate(elems)
    ords, nodes = np.concatenate(ords), np.concatenate(nodes)
    _, inv, cnt = np.unique(keys, axis=0, return_inverse=True, return_counts=True)
    once = cnt[inv.ravel()] == 1
    order = np.lexsort((ords[once], elems[once]))
    return elems[once][order], ords[once][order], nodes[once][order]


_SIDE_NB = np.array([[0, -1, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, 0, -1], [0, 0, 1]])


def duct_mixed(m: int, layers=(187, 8, 25, 75, 25), length: float = 1.0, width: float = 0.2,
               jitter: float = 0.1, seed: int = 0, hex_ring: bool = False):
    """Mixed-topology duct [0, length] x [0, width]^2 of cubes (m across,
    sum(layers) along x with length / sum(layers) == width / m): along x the
    cube layers are hex8 (inlet side), pyramids (each cube 6 pyramids about
    a centre node), tets, wedges, tets and hex8 again (outlet side), layers =
    (hex total, pyramid, tet, wedge, tet).  The walls and the ends are sides
    of whatever elements touch them: quad sides of hex8, wedge6 and pyramid5,
    tri sides of tet4 (and of wedge6 where a wedge layer ends the duct).
    hex_ring: the outer ring of cubes is hex8 everywhere instead, so every
    boundary side is a quad side of a hex8.

    Tet cubes: each side split into two triangles (the diagonal through its
    lowest node id, the same from both cubes sharing it), each triangle
    joined to the cube's centre node; a side facing a hex8 / pyramid cube or
    a wedge cube's quad side stays a quad and makes a pyramid instead.
    Wedge cubes: two prisms along x, split by the diagonal that the lowest-id
    rule gives on their x sides.  Node ordering per topology: tet4 base
    (counter-clockwise seen from the apex) + apex, wedge6 bottom triangle +
    top triangle, pyramid5 base + apex (positive Jacobians).

    Interior grid nodes move by N(0, jitter*h) per component, clipped at
    +-2.5 jitter*h; boundary nodes stay; centre nodes sit at their cube's
    corner mean.  Where that leaves a tet or pyramid with a volume below
    h^3 / 100, the moves of its grid nodes are halved until none is left.
    Boundary ids DUCT_INLET (x = 0), DUCT_OUTLET (x = length),
    DUCT_WALL.  Returns xyz, conn [E][8] (-1 padded), etype, bf_elem,
    bf_ord, bf_bc."""
    H, P, T, W = 0, 3, 1, 2
    nh, npyr, nt1, nw, nt2 = layers
    nx, ny, nz = nh + npyr + nt1 + nw + nt2, m, m
    h = width / m
    assert abs(length / nx - h) < 1e-12 * h, "cubes: length / nx must equal width / m"
    nh1 = nh // 2
    xtype = np.array([H] * nh1 + [P] * npyr + [T] * nt1 + [W] * nw + [T] * nt2 + [H] * (nh - nh1))
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ctype = xtype[i]
    if hex_ring:
        ring = (j == 0) | (j == ny - 1) | (k == 0) | (k == nz - 1)
        ctype = np.where(ring, H, ctype)
    # grid nodes
    g1 = nx + 1
    gid = lambda ii, jj, kk: ii + g1 * (jj + (ny + 1) * kk)
    kk_, jj_, ii_ = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    xyz0 = np.stack([ii_.ravel() * h, jj_.ravel() * h, kk_.ravel() * h], axis=1)
    d = np.zeros_like(xyz0)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        d = np.clip(rng.normal(0.0, jitter * h, size=xyz0.shape), -2.5 * jitter * h, 2.5 * jitter * h)
        on_b = ((ii_ == 0) | (ii_ == nx) | (jj_ == 0) | (jj_ == ny) | (kk_ == 0) | (kk_ == nz)).ravel()
        d[on_b] = 0.0
    xyz = xyz0 + d
    ng = xyz.shape[0]
    ci, cj, ck = i.ravel(), j.ravel(), k.ravel()
    ct = ctype.ravel()
    corner = np.stack([gid(ci + a, cj + b, ck + c) for a, b, c in _CORNERS], axis=1)  # [ncube][8]
    # centre nodes of pyramid and tet cubes
    has_c = (ct == P) | (ct == T)
    cnode = np.full(ct.size, -1, dtype=np.int64)
    cnode[has_c] = ng + np.arange(int(has_c.sum()))
    xyz = np.concatenate([xyz, xyz[corner[has_c]].mean(axis=1)])
    elems, types = [], []

    def add(conn, t):
        c = np.full((conn.shape[0], 8), -1, dtype=np.int64)
        c[:, :conn.shape[1]] = conn
        elems.append(c)
        types.append(np.full(conn.shape[0], t, dtype=np.int32))

    # hex cubes (their element ids come first, in cube order)
    hsel = np.where(ct == H)[0]
    add(corner[hsel], HEX8)
    hex_eid = np.full(ct.size, -1, dtype=np.int64)
    hex_eid[hsel] = np.arange(hsel.size)
    # wedge cubes
    ws = np.where(ct == W)[0]
    cw = corner[ws]
    add(cw[:, [0, 3, 7, 1, 2, 6]], WED6)
    add(cw[:, [0, 7, 4, 1, 6, 5]], WED6)
    # neighbour types per side (-1 outside)
    padded = np.full((nz + 2, ny + 2, nx + 2), -1)
    padded[1:-1, 1:-1, 1:-1] = ctype
    nbt = np.stack([padded[1 + ck + dk, 1 + cj + dj, 1 + ci + di] for di, dj, dk in _SIDE_NB], axis=1)
    for cube_t in (P, T):
        sel = np.where(ct == cube_t)[0]
        for side in range(6):
            q = corner[sel][:, _HEX_SIDES[side]]          # outward quad
            apex = cnode[sel]
            if cube_t == T:
                nb = nbt[sel, side]
                tri = (nb == T) | (nb == -1) | ((nb == W) & ((side == 1) | (side == 3)))
            else:
                tri = np.zeros(sel.size, dtype=bool)
            pq = q[~tri]
            add(np.stack([pq[:, 3], pq[:, 2], pq[:, 1], pq[:, 0], apex[~tri]], axis=1), PYR5)
            tq, ta = q[tri], apex[tri]
            if tq.size:
                first = np.argmin(tq, axis=1) % 2 == 0      # diagonal q0-q2, else q1-q3
                t1 = np.where(first[:, None], tq[:, [0, 1, 2]], tq[:, [1, 2, 3]])
                t2 = np.where(first[:, None], tq[:, [0, 2, 3]], tq[:, [1, 3, 0]])
                for tr in (t1, t2):
                    add(np.stack([tr[:, 0], tr[:, 2], tr[:, 1], ta], axis=1), TET4)
    conn = np.concatenate(elems).astype(np.int32)
    etype = np.concatenate(types)

    def tet_vol(a, b, cc, dd):
        return np.einsum("ij,ij->i", np.cross(b - a, cc - a), dd - a) / 6.0

    for _ in range(20):   # halve the moves around too-thin tets and pyramids
        ct4, cp5 = conn[etype == TET4], conn[etype == PYR5]
        v = [tet_vol(*(xyz[ct4[:, k]] for k in range(4)))]
        q = [xyz[cp5[:, k]] for k in range(5)]
        v += [tet_vol(q[0], q[1], q[2], q[4]), tet_vol(q[0], q[2], q[3], q[4]),
              tet_vol(q[0], q[1], q[3], q[4]), tet_vol(q[1], q[2], q[3], q[4])]
        bad = [ct4[v[0] < 1e-2 * h ** 3]] + [cp5[vk < 0.5e-2 * h ** 3] for vk in v[1:]]
        nodes = np.unique(np.concatenate([b.ravel() for b in bad]))
        nodes = nodes[nodes < ng]
        if nodes.size == 0:
            break
        d[nodes] *= 0.5
        xyz[:ng] = xyz0 + d
        xyz[ng:] = xyz[corner[has_c]].mean(axis=1)
    # boundary sides: sides of one element only, typed by the plane they lie on
    be, bo, bn = boundary_sides(conn, etype)
    xs = np.where(bn >= 0, xyz[np.maximum(bn, 0), 0], np.nan)
    bc = np.full(be.shape, DUCT_WALL)
    bc[np.nanmax(np.abs(xs), axis=1) < 1e-9 * length] = DUCT_INLET
    bc[np.nanmin(xs, axis=1) > length * (1 - 1e-9)] = DUCT_OUTLET
    bf = np.stack([be, bo, bc], axis=1).astype(np.int32)
    return {
        "xyz": np.ascontiguousarray(xyz),
        "conn": np.ascontiguousarray(conn),
        "etype": np.ascontiguousarray(etype),
        "bf_elem": np.ascontiguousarray(bf[:, 0]),
        "bf_ord": np.ascontiguousarray(bf[:, 1]),
        "bf_bc": np.ascontiguousarray(bf[:, 2]),
        "h": h,
    }
