"""Pins the oracle's tet4 / wedge6 / pyramid5 master elements (oracle/topo.c):

* against the reference's own Fortran (MasterElementWork.F built from the
  reference tree into oracle/_ref by oracle/Makefile): SCS area vectors, SCV
  volumes and the gradient operator of random distorted elements, 1e-14
  relative;
* against closed forms: SCV volumes sum to the element volume (tet and
  pyramid by the triple product, wedge as a prism); every control volume of
  an element is closed (sum over its SCS, signed L->R, of the area vectors
  plus its share of the element faces is zero) -- checked through the
  element total: the sum of signed SCS areas around each node equals minus
  its exterior face-area share, which for the element's vertices sums to 0;
* derivatives reproduce linear functions in parametric space (partition of
  unity and d/ds of the vertex parametric coordinates);
* the gradient operator is exact for linear fields on the element.
"""
import numpy as np
import pytest

TOPOS = {"tet4": 1, "wed6": 2, "pyr5": 3}
REF_VERTS = {
    1: np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float),
    2: np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], float),
    3: np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0], [0, 0, 1]], float),
}
# parametric vertex coordinates (wedge: (r, s, xi), xi in [-1, 1])
PAR_VERTS = {
    1: REF_VERTS[1],
    2: np.array([[0, 0, -1], [1, 0, -1], [0, 1, -1], [0, 0, 1], [1, 0, 1], [0, 1, 1]], float),
    3: REF_VERTS[3],
}


def element(topo, rng, jitter=0.15):
    v = REF_VERTS[topo].copy()
    A = np.eye(3) + rng.normal(scale=0.2, size=(3, 3))
    if np.linalg.det(A) < 0.3:
        A = np.eye(3)
    return (v + rng.uniform(-jitter, jitter, v.shape) * 0.5) @ A.T + rng.normal(size=3)


def volume(topo, c):
    def tet(a, b, d, e):
        return abs(np.dot(b - a, np.cross(d - a, e - a))) / 6.0
    if topo == 1:
        return tet(*c)
    if topo == 3:
        m = c[:4].mean(axis=0)
        return sum(tet(c[i], c[(i + 1) % 4], m, c[4]) for i in range(4))
    return None


@pytest.mark.parametrize("name", list(TOPOS))
def test_against_reference_fortran(oracle, name):
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    topo = TOPOS[name]
    rng = np.random.default_rng(topo)
    for trial in range(20):
        c = element(topo, rng)
        for shifted in (0, 1):
            a, v, g, d = oracle.topo_geometry(topo, c, shifted)
            ips = np.ctypeslib.as_array(oracle.lib().or_topo_scs_ips(topo, shifted),
                                        shape=(a.shape[0] * 3,)).reshape(-1, 3)
            deriv = oracle.topo_deriv(topo, shifted, ips)
            ra, rv, rg, rd = oracle.ref_topo_geometry(topo, c, deriv)
            np.testing.assert_allclose(a, ra, rtol=0, atol=1e-14 * np.abs(ra).max())
            np.testing.assert_allclose(v, rv, rtol=0, atol=1e-14 * np.abs(rv).max())
            np.testing.assert_allclose(g, rg, rtol=0, atol=1e-13 * np.abs(rg).max())
            np.testing.assert_allclose(d, rd, rtol=1e-14)


@pytest.mark.parametrize("name", ["tet4", "pyr5"])
def test_scv_sum_is_element_volume(oracle, name):
    topo = TOPOS[name]
    rng = np.random.default_rng(10 + topo)
    for _ in range(10):
        c = element(topo, rng, jitter=0.0)
        a, v, g, d = oracle.topo_geometry(topo, c)
        assert v.sum() == pytest.approx(volume(topo, c), rel=1e-13)
        assert np.all(v > 0)


def test_wedge_prism_volume(oracle):
    # a right prism: base triangle area * height
    c = np.array([[0, 0, 0], [2, 0, 0], [0, 1.5, 0], [0, 0, 0.7], [2, 0, 0.7], [0, 1.5, 0.7]], float)
    a, v, g, d = oracle.topo_geometry(2, c)
    assert v.sum() == pytest.approx(0.5 * 2 * 1.5 * 0.7, rel=1e-14)


@pytest.mark.parametrize("name", list(TOPOS))
def test_scs_areas_close_with_faces(oracle, name):
    """Around an interior point of the element the SCS sum telescopes: the
    signed sum over all SCS of (L gets +A, R gets -A) is zero in total."""
    topo = TOPOS[name]
    c = element(topo, np.random.default_rng(20 + topo))
    a, v, g, d = oracle.topo_geometry(topo, c)
    lr = np.ctypeslib.as_array(oracle.lib().or_topo_lrscv(topo), shape=(a.shape[0] * 2,)).reshape(-1, 2)
    net = np.zeros((c.shape[0], 3))
    for ip, (L, R) in enumerate(lr):
        net[L] += a[ip]
        net[R] -= a[ip]
    np.testing.assert_allclose(net.sum(axis=0), 0.0, atol=1e-14 * np.abs(a).max())
    # each node's net outflow area is nonzero (an open sub-volume of the element)
    assert np.all(np.linalg.norm(net, axis=1) > 1e-6)


@pytest.mark.parametrize("name", list(TOPOS))
@pytest.mark.parametrize("shifted", [0, 1])
def test_derivatives_reproduce_linear_functions(oracle, name, shifted):
    topo = TOPOS[name]
    ips = np.ctypeslib.as_array(oracle.lib().or_topo_scs_ips(topo, shifted),
                                shape=(oracle.lib().or_topo_nscs(topo) * 3,)).reshape(-1, 3)
    d = oracle.topo_deriv(topo, shifted, ips)
    np.testing.assert_allclose(d.sum(axis=1), 0.0, atol=1e-14)
    P = PAR_VERTS[topo]
    for j in range(ips.shape[0]):
        J = d[j].T @ P
        if topo == 3 and shifted:
            # collapsed-hexahedron basis (Pyr5CVFEM.C shifted_pyr_derivative):
            # the vertex map is x = r (1 - t), y = s (1 - t), z = t
            r, s_, t = ips[j]
            want = np.array([[1 - t, 0, 0], [0, 1 - t, 0], [-r, -s_, 1]])
            np.testing.assert_allclose(J, want, atol=1e-14)
        else:
            np.testing.assert_allclose(J, np.eye(3), atol=1e-14)


@pytest.mark.parametrize("name", list(TOPOS))
def test_gradient_operator_exact_for_linear_fields(oracle, name):
    topo = TOPOS[name]
    c = element(topo, np.random.default_rng(30 + topo))
    a, v, g, d = oracle.topo_geometry(topo, c)
    coef = np.array([0.7, -1.3, 2.1])
    phi = c @ coef + 0.4
    grad = np.einsum("inj,n->ij", g, phi)
    np.testing.assert_allclose(grad, np.tile(coef, (g.shape[0], 1)), rtol=1e-13)
    assert np.all(d > 0)
