"""Pins the oracle's symmetry-plane treatment with closed forms:

* assembled symmetry area: the symmetry-face area vectors summed onto their
  nearest nodes add up, plane by plane, to the plane's area vector, and a
  node inside a plane of a uniform grid carries -h^2 n (its dual face);
* gradient: for a field that is mirror-symmetric about the symmetry planes
  z = 0 and z = 1 (u, v independent of z, w linear in z), the corrected
  Green-Gauss gradient is exact at every node, planes and their edges
  included (dropping the 2 n n^T (n^T G n) term of the correction would flip
  dw/dz); for w = x the skipped symmetry-face terms leave dw/dz = +-1/2 on
  the plane edges of a uniform grid, a closed form of the open control
  volume;
* momentum: with every face a symmetry plane, each control volume is closed,
  so a linear velocity field (constant stress) leaves a zero residual at
  every node, boundary nodes included;
* projection and normal relaxation: the assembled rhs of a symmetry node has
  no normal component; the relaxed diagonal block D' satisfies D' t = D t for
  tangential t and D' n = (1 + 0.25) D n.
"""
import numpy as np
import pytest

BC3 = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
SYM = 3


def sym_mesh(meshgen, n, perturb=0.0, seed=0, all_sym=False):
    m = meshgen.cavity_hex(n, perturb=perturb, seed=seed)
    bc = m["bf_bc"].copy()
    zf = (m["bf_ord"] == 4) | (m["bf_ord"] == 5)
    bc[zf] = 2
    if all_sym:
        bc[:] = 2
    m["bf_bc"] = np.ascontiguousarray(bc)
    return m


def make(oracle, meshgen, n=5, perturb=0.0, seed=0, all_sym=False, **kw):
    m = sym_mesh(meshgen, n, perturb, seed, all_sym)
    c = oracle.Case(m, oracle.default_params(**kw), BC3)
    c.set_bc_types([0, 0, SYM])
    return c, m


def test_symmetry_area(oracle, meshgen):
    n = 4
    c, m = make(oracle, meshgen, n=n)
    x = m["xyz"]
    A = c.field("symA")
    for z0, nz in ((0.0, -1.0), (1.0, 1.0)):
        on = np.abs(x[:, 2] - z0) < 1e-12
        np.testing.assert_allclose(A[on].sum(axis=0), [0.0, 0.0, nz], atol=1e-14)
        inner = on & (x[:, 0] > 1e-9) & (x[:, 0] < 1 - 1e-9) & (x[:, 1] > 1e-9) & (x[:, 1] < 1 - 1e-9)
        np.testing.assert_allclose(A[inner], np.tile([0.0, 0.0, nz / n**2], (inner.sum(), 1)), atol=1e-15)
    off = (x[:, 2] > 1e-9) & (x[:, 2] < 1 - 1e-9)
    assert np.abs(A[off]).max() == 0.0


def test_unsupported_type_rejected(oracle, meshgen):
    m = sym_mesh(meshgen, 3)
    c = oracle.Case(m, oracle.default_params(), BC3)
    with pytest.raises(ValueError):
        c.set_bc_types([0, 7, SYM])


@pytest.mark.parametrize("n", [4, 5])
def test_gradient_mirror_symmetric_field_exact(oracle, meshgen, n):
    # uniform grid: Green-Gauss with trilinear ip interpolation is exact for
    # linear fields on affine elements only
    c, m = make(oracle, meshgen, n=n, grad_urf=1.0)
    x = m["xyz"]
    a, b, cc, d, e = 0.7, -0.4, 0.3, 1.1, -0.9
    U = np.stack([a * x[:, 0] + b * x[:, 1], cc * x[:, 0] + d * x[:, 1], e * x[:, 2]], axis=1)
    c.field("U")[:] = U
    c.field("gradU")[:] = 0.0
    c.gradient("U")
    G = c.field("gradU").reshape(-1, 3, 3)
    exact = np.array([[a, b, 0.0], [cc, d, 0.0], [0.0, 0.0, e]])
    np.testing.assert_allclose(G, np.broadcast_to(exact, G.shape), rtol=0, atol=1e-12)
    # pressure, mirror-symmetric: p = f x + g y
    c.field("p")[:] = 0.5 * x[:, 0] - 1.5 * x[:, 1]
    c.field("gradP")[:] = 0.0
    c.gradient("p")
    np.testing.assert_allclose(c.field("gradP"), np.tile([0.5, -1.5, 0.0], (c.nn, 1)), atol=1e-12)


def test_gradient_symmetry_face_terms_skipped(oracle, meshgen):
    """w = x is not mirror-symmetric: the symmetry planes' own face terms
    (skipped) would have closed the control volumes.  On a uniform grid with
    gradient relaxation 1, the open control volume of a plane node on the
    x = 0 edge of z = 0 misses 2 (h/4)(-h^2/4) out of V = h^3/4 (or 1 of
    V = h^3/8 at a corner), so dw/dz = +1/2 there, -1/2 on the x = 1 edge,
    signs swapped on z = 1, 0 inside the planes; the correction keeps
    n^T G n and removes dw/dx on the planes."""
    n = 4
    c, m = make(oracle, meshgen, n=n, grad_urf=1.0)
    x = m["xyz"]
    U = np.zeros((c.nn, 3))
    U[:, 2] = x[:, 0]
    c.field("U")[:] = U
    c.field("gradU")[:] = 0.0
    c.gradient("U")
    G = c.field("gradU").reshape(-1, 3, 3)
    for z0, sgn in ((0.0, 1.0), (1.0, -1.0)):
        on = np.abs(x[:, 2] - z0) < 1e-12
        want = np.zeros(c.nn)
        want[on & (x[:, 0] < 1e-12)] = 0.5 * sgn
        want[on & (x[:, 0] > 1 - 1e-12)] = -0.5 * sgn
        np.testing.assert_allclose(G[on, 2, 2], want[on], atol=1e-13)
        np.testing.assert_allclose(G[on, 2, 0], 0.0, atol=1e-13)
    off = (x[:, 2] > 1e-9) & (x[:, 2] < 1 - 1e-9)
    np.testing.assert_allclose(G[off, 2, 0], 1.0, atol=1e-12)


@pytest.mark.parametrize("perturb", [0.0, 0.15])
def test_momentum_all_symmetry_linear_field(oracle, meshgen, perturb):
    n = 4
    c, m = make(oracle, meshgen, n=n, perturb=perturb, seed=3, all_sym=True)
    G = np.array([[0.3, -0.2, 0.5], [0.1, 0.4, -0.3], [-0.2, 0.25, 0.1]])
    c.field("U")[:] = m["xyz"] @ G.T
    for f in ("mdot", "gradP", "div"):
        c.field(f)[:] = 0.0
    c.assemble_momentum()
    b = c.rhs(3).reshape(-1, 3)
    assert np.abs(b).max() < 1e-14


def test_projection_and_normal_relaxation(oracle, meshgen):
    n = 5
    c, m = make(oracle, meshgen, n=n, perturb=0.1, seed=4)
    rng = np.random.default_rng(5)
    c.field("U")[:] = rng.normal(scale=0.3, size=(c.nn, 3))
    c.initialize()
    c.field("mdot")[:] = rng.normal(scale=0.05, size=(c.ne, 12))
    c.field("gradP")[:] = rng.normal(scale=0.2, size=(c.nn, 3))
    c.mass_divergence()
    c.assemble_momentum()
    A = c.field("symA")
    on = np.linalg.norm(A, axis=1) > 0
    nrm = A[on] / np.linalg.norm(A[on], axis=1)[:, None]
    b = c.rhs(3).reshape(-1, 3)
    assert np.abs(np.einsum("ij,ij->i", b[on], nrm)).max() < 1e-15
    assert np.abs(b[on]).max() > 1e-6
    rp, dg = c.row_ptr, c.diag_off
    D0 = c.vals(3)[rp[:-1] + dg].copy()
    c.compute_du()
    c.normal_relaxation()
    D1 = c.vals(3)[rp[:-1] + dg].reshape(-1, 3, 3)
    D0 = D0.reshape(-1, 3, 3)
    np.testing.assert_array_equal(D1[~on], D0[~on])
    t = np.cross(nrm, [0.3, 0.5, 0.8])
    t /= np.linalg.norm(t, axis=1)[:, None]
    np.testing.assert_allclose(np.einsum("nij,nj->ni", D1[on], t), np.einsum("nij,nj->ni", D0[on], t),
                               rtol=0, atol=1e-14 * np.abs(D0).max())
    np.testing.assert_allclose(np.einsum("nij,nj->ni", D1[on], nrm),
                               1.25 * np.einsum("nij,nj->ni", D0[on], nrm),
                               rtol=0, atol=1e-14 * np.abs(D0).max())
