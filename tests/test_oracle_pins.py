"""Pins of the oracle's momentum and pressure-correction element terms that a
plausible mistake (a dropped or sign-flipped transpose term, a scaled
viscosity, a wrong shape-function derivative, a lost lumping) fails:

* orientation invariance: relabelling every hex8 by a random proper cube
  rotation moves no physical point, so every nodal result (matrices, rhs,
  du, divergence, gradients) is unchanged to rounding;
* rigid rotation U = w x X + c has grad U + grad U^T = 0, so the viscous
  stress vanishes at every SCS ip and the momentum rhs is zero at EVERY node
  of a mesh without boundary faces, boundary nodes included (their control
  volumes are open, so a missing or sign-flipped transpose term leaves a
  residual there);
* a symmetric strain U = S X leaves, at a node whose control volume is cut
  by the domain boundary, rhs = 2 mu S . sum of its interior SCS areas =
  -2 mu S . (its exposed outward area) -- a closed form in mu and h;
* one uniform element of side h: the momentum and p' local matrices
  restated from the trilinear shape functions at the SCS ips
  (s_d = 0, s_k = sigma_Lk / 4, or / 2 at the shifted ips), areas h^2/4,
  against the assembled full-stencil matrix, and the reduced stencil's
  lumping of the absent columns into the diagonal.
"""
import numpy as np
import pytest

from conftest import entry_rel_err, rel_err, row_rel_err

BC = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
SIGMA = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1],
                  [1, 1, 1], [-1, 1, 1]], dtype=float)
LRSCV = [(0, 1), (1, 2), (2, 3), (0, 3), (4, 5), (5, 6), (6, 7), (4, 7), (0, 4), (1, 5), (2, 6),
         (3, 7)]


def no_faces(mesh):
    m = dict(mesh)
    for k in ("bf_elem", "bf_ord", "bf_bc"):
        m[k] = np.zeros(0, dtype=np.int32)
    return m


def state(c, U, nn, ne):
    c.field("U")[:] = U
    c.field("p")[:] = 0.0
    c.initialize()
    c.field("mdot")[:] = 0.0
    c.field("div")[:] = 0.0
    c.field("gradP")[:] = 0.0
    c.field("du")[:] = 1.0e-3


# ------------------------------------------------------- orientation invariance
@pytest.mark.parametrize("reduced", [1, 0])
def test_orientation_invariance(oracle, meshgen, reduced):
    base = meshgen.cavity_hex(5, perturb=0.2, seed=6)
    rot = meshgen.relabel_hex(base, 31)
    assert meshgen.repeated_local_nodes(rot["conn"], base["xyz"].shape[0]) > 0
    rng = np.random.default_rng(4)
    nn = base["xyz"].shape[0]
    U = rng.normal(scale=0.3, size=(nn, 3))
    p = rng.normal(scale=0.1, size=nn)
    out = []
    for m in (base, rot):
        c = oracle.Case(m, oracle.default_params(reduced=reduced), BC)
        c.field("U")[:] = U
        c.field("p")[:] = p
        c.initialize()
        r = {"div": c.field("div").copy(), "gradU": c.field("gradU").copy(),
             "gradP": c.field("gradP").copy(), "vol": c.field("vol").copy()}
        c.assemble_momentum()
        c.compute_du()
        r.update(A3=c.vals(3).copy(), b3=c.rhs(3).copy(), du=c.field("du").copy())
        c.assemble_pcorr()
        r.update(A1=c.vals(1).copy(), b1=c.rhs(1).copy(), rp=c.row_ptr.copy())
        h = c.outer_iteration([20, 0.1, 200, 1e-8])
        r.update(hist=h, U1=c.field("U").copy(), p1=c.field("p").copy())
        out.append(r)
    a, b = out
    assert np.array_equal(a["rp"], b["rp"])
    assert row_rel_err(b["A3"], a["A3"], a["rp"]) < 1e-12
    assert row_rel_err(b["A1"], a["A1"], a["rp"]) < 1e-12
    for k in ("b3", "b1", "du", "gradU", "gradP", "vol"):
        assert entry_rel_err(b[k], a[k]) < 1e-12, k
    assert rel_err(b["div"], a["div"]) < 1e-12
    np.testing.assert_allclose(b["hist"][:4], a["hist"][:4], rtol=1e-9)
    assert rel_err(b["U1"], a["U1"]) < 1e-9 and rel_err(b["p1"], a["p1"]) < 1e-9


# ---------------------------------------------------------- rigid rotation / strain
def exposed_area(xyz, h, L=1.0):
    """Outward area of the boundary part of each node's control volume on a
    uniform grid of the cube [0, L]^3 with spacing h."""
    out = np.zeros_like(xyz)
    tol = 1e-9 * L
    for d in range(3):
        o1, o2 = [k for k in range(3) if k != d]
        for side, sgn in ((0.0, -1.0), (L, 1.0)):
            on = np.abs(xyz[:, d] - side) < tol

            def extent(k):
                edge = (np.abs(xyz[:, k]) < tol) | (np.abs(xyz[:, k] - L) < tol)
                return np.where(edge, h / 2, h)

            out[on, d] += sgn * (extent(o1) * extent(o2))[on]
    return out


@pytest.mark.parametrize("shifted", [0, 1])
def test_rigid_rotation_zero_viscous_residual(oracle, meshgen, shifted):
    for perturb in (0.0, 0.2):
        m = no_faces(meshgen.cavity_hex(4, perturb=perturb, seed=3))
        nn, ne = m["xyz"].shape[0], m["conn"].shape[0]
        w, c0 = np.array([0.7, -1.3, 0.4]), np.array([0.2, 0.1, -0.3])
        U = np.cross(w, m["xyz"]) + c0
        c = oracle.Case(m, oracle.default_params(mu=0.37, ugrad_shifted=shifted, p_ref_node=-1), BC)
        state(c, U, nn, ne)
        c.assemble_momentum()
        b = c.rhs(3).reshape(nn, 3)
        # scale: the same mesh under a unit symmetric strain
        S = np.array([[0.5, 0.2, 0.0], [0.2, -0.3, 0.4], [0.0, 0.4, -0.2]])
        c2 = oracle.Case(m, oracle.default_params(mu=0.37, ugrad_shifted=shifted, p_ref_node=-1), BC)
        state(c2, m["xyz"] @ S.T, nn, ne)
        c2.assemble_momentum()
        scale = np.abs(c2.rhs(3)).max()
        assert scale > 0
        assert np.abs(b).max() < 1e-13 * scale, perturb


@pytest.mark.parametrize("shifted", [0, 1])
def test_symmetric_strain_boundary_rhs_closed_form(oracle, meshgen, shifted):
    n, mu = 4, 0.37
    h = 1.0 / n
    m = no_faces(meshgen.cavity_hex(n))
    nn, ne = m["xyz"].shape[0], m["conn"].shape[0]
    S = np.array([[0.5, 0.2, 0.0], [0.2, -0.3, 0.4], [0.0, 0.4, -0.2]])
    c = oracle.Case(m, oracle.default_params(mu=mu, ugrad_shifted=shifted, p_ref_node=-1), BC)
    state(c, m["xyz"] @ S.T, nn, ne)
    c.assemble_momentum()
    b = c.rhs(3).reshape(nn, 3)
    want = -2.0 * mu * exposed_area(m["xyz"], h) @ S.T
    assert np.abs(b - want).max() < 1e-13 * np.abs(want).max()
    assert np.abs(want).max() > 0


# ------------------------------------------------------------- one element
def ip_points(shifted):
    """(L, R, axis, s) of the 12 SCS ips of the reference cube [-1/2, 1/2]^3."""
    out = []
    for L, R in LRSCV:
        d = int(np.flatnonzero(SIGMA[L] != SIGMA[R])[0])
        s = SIGMA[L] * (0.5 if shifted else 0.25)
        s[d] = 0.0
        out.append((L, R, d, s))
    return out


def dndx(s, h):
    """d N_b / d x at reference point s for the cube of side h: [8][3]."""
    g = np.zeros((8, 3))
    for b in range(8):
        f = 0.5 + SIGMA[b] * s
        for j in range(3):
            g[b, j] = SIGMA[b, j] * np.prod([f[k] for k in range(3) if k != j]) / h
    return g


def element_mom(h, mu, shifted):
    K = np.zeros((8, 3, 8, 3))
    for L, R, d, s in ip_points(shifted):
        A = np.zeros(3)
        A[d] = h * h / 4.0 * SIGMA[R, d]       # from L towards R
        G = dndx(s, h)
        for sgn, row in ((1.0, L), (-1.0, R)):
            for i in range(3):
                for b in range(8):
                    K[row, i, b, i] += sgn * -mu * (G[b] @ A)
                    for j in range(3):
                        K[row, i, b, j] += sgn * -mu * G[b, i] * A[j]
    return K


def element_pc(h, rho, d_u, shifted):
    K = np.zeros((8, 8))
    for L, R, d, s in ip_points(shifted):
        A = np.zeros(3)
        A[d] = h * h / 4.0 * SIGMA[R, d]       # from L towards R
        G = dndx(s, h)
        for sgn, row in ((1.0, L), (-1.0, R)):
            K[row] += sgn * -rho * (G @ (d_u * A))
    return K


def to_global(K, conn):
    """Local element matrix [8][N][8][N] in the element's node order -> node ids."""
    out = np.zeros_like(K)
    out[np.ix_(conn, range(K.shape[1]), conn, range(K.shape[3]))] = K
    return out


def dense_from_crs(vals, rp, cols, nn, N):
    M = np.zeros((nn, N, nn, N))
    for r in range(nn):
        for k in range(rp[r], rp[r + 1]):
            M[r, :, cols[k], :] = vals[k].reshape(N, N)
    return M


def lump(K, rp, cols, N):
    """The reduced stencil's row: columns absent from the graph go into the
    diagonal block."""
    nn = K.shape[0]
    out = np.zeros_like(K)
    for r in range(nn):
        present = set(cols[rp[r]:rp[r + 1]].tolist())
        for b in range(nn):
            if b in present:
                out[r, :, b, :] += K[r, :, b, :]
            else:
                out[r, :, r, :] += K[r, :, b, :]
    return out


@pytest.mark.parametrize("reduced", [0, 1])
@pytest.mark.parametrize("shifted", [0, 1])
def test_single_element_momentum(oracle, meshgen, reduced, shifted):
    h, mu, rho, dt, urf = 0.3, 0.37, 1.7, 0.5, 0.8
    m = no_faces(meshgen.cavity_hex(1, length=h))
    c = oracle.Case(m, oracle.default_params(mu=mu, rho=rho, dt=dt, urf_U=urf, reduced=reduced,
                                             ugrad_shifted=shifted, p_ref_node=-1), BC)
    state(c, np.zeros((8, 3)), 8, 1)
    c.field("rho")[:] = rho
    c.assemble_momentum()
    got = dense_from_crs(c.vals(3), c.row_ptr, c.cols, 8, 3)
    K = to_global(element_mom(h, mu, shifted), m["conn"][0])
    for a in range(8):                          # pseudo-time term rho V / dt, V = h^3 / 8
        K[a, :, a, :] += np.eye(3) * rho * h ** 3 / 8 / dt
    if reduced:
        K = lump(K, c.row_ptr, c.cols, 3)
    for a in range(8):                          # implicit relaxation of the diagonal
        K[a, :, a, :] += np.diag(np.diag(K[a, :, a, :])) * (1.0 / urf - 1.0)
    assert np.abs(got - K).max() < 1e-14 * np.abs(K).max()
    if reduced:
        assert len(c.cols) == 8 * 4             # 3 edge neighbours + the node itself


@pytest.mark.parametrize("reduced", [0, 1])
@pytest.mark.parametrize("shifted", [0, 1])
def test_single_element_pcorr(oracle, meshgen, reduced, shifted):
    h, rho = 0.3, 1.7
    d_u = np.array([1.1e-3, 0.7e-3, 1.9e-3])
    m = no_faces(meshgen.cavity_hex(1, length=h))
    c = oracle.Case(m, oracle.default_params(rho=rho, reduced=reduced, pgrad_shifted=shifted,
                                             p_ref_node=-1), BC)
    state(c, np.zeros((8, 3)), 8, 1)
    c.field("rho")[:] = rho
    c.field("du")[:] = d_u
    c.assemble_pcorr()
    got = dense_from_crs(c.vals(1), c.row_ptr, c.cols, 8, 1)[:, 0, :, 0]
    K = to_global(element_pc(h, rho, d_u, shifted)[:, None, :, None], m["conn"][0])[:, 0, :, 0]
    if reduced:
        K = lump(K[:, None, :, None], c.row_ptr, c.cols, 1)[:, 0, :, 0]
    assert np.abs(got - K).max() < 1e-14 * np.abs(K).max()
    assert np.abs(K.sum(axis=1)).max() < 1e-15 * np.abs(K).max()
