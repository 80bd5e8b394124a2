"""The oracle's boundary sides of every topology (topo.c), pinned against the
reference:

* the exposed-face ip locations equal the reference's intgExpFace_ tables
  entry by entry (Tet4CVFEM.C:384-403, Wed6CVFEM.C:458-480,
  Pyr5CVFEM.C:737-757; the tet's fourth side lists its second and third ip
  in element-node order, immaterial for the tet's constant gradient);
* tri sub-face area vectors equal the reference Fortran tri3d_scs_det
  (MasterElementWork.F:1589-1697, built from source under oracle/_ref);
* the sides close every element (their sub-face area vectors sum to zero)
  and point out of it, on distorted elements;
* the face gradient operator at every face ip (standard and shifted)
  differentiates a linear field exactly and, on the tet, equals the
  element's constant gradient; at a pyramid's apex the shifted operator is
  singular (det J = 0) and the gradient-operator fallback gives zero.
"""
import numpy as np
import pytest

S7, E11, F5, T29, S18 = 7 / 36, 11 / 18, 5 / 12, 29 / 36, 7 / 18

# intgExpFace_ of the reference, per side (face ips in order)
REF_IPS = {
    1: [[(S7, 0, S7), (E11, 0, S7), (S7, 0, E11)],
        [(E11, S7, S7), (S7, E11, S7), (S7, S7, E11)],
        [(0, S7, S7), (0, S7, E11), (0, E11, S7)],
        [(S7, S7, 0), (E11, S7, 0), (S7, E11, 0)]],
    2: [[(0.25, 0, -0.5), (0.75, 0, -0.5), (0.75, 0, 0.5), (0.25, 0, 0.5)],
        [(0.75, 0.25, -0.5), (0.25, 0.75, -0.5), (0.25, 0.75, 0.5), (0.75, 0.25, 0.5)],
        [(0, 0.25, -0.5), (0, 0.25, 0.5), (0, 0.75, 0.5), (0, 0.75, -0.5)],
        [(S7, S7, -1), (S7, E11, -1), (E11, S7, -1)],
        [(S7, S7, 1), (E11, S7, 1), (S7, E11, 1)]],
    3: [[(-F5, -T29, S7), (F5, -T29, S7), (0, -S18, E11)],
        [(T29, -F5, S7), (T29, F5, S7), (S18, 0, E11)],
        [(F5, T29, S7), (-F5, T29, S7), (0, S18, E11)],
        [(-T29, -F5, S7), (-S18, 0, E11), (-T29, F5, S7)],
        [(-0.5, -0.5, 0), (-0.5, 0.5, 0), (0.5, 0.5, 0), (0.5, -0.5, 0)]],
}
# sideNodeOrdinals_ of the reference
REF_SIDES = {
    0: [[0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [0, 4, 7, 3], [0, 3, 2, 1], [4, 5, 6, 7]],
    1: [[0, 1, 3], [1, 2, 3], [0, 3, 2], [0, 2, 1]],
    2: [[0, 1, 4, 3], [1, 2, 5, 4], [0, 3, 5, 2], [0, 2, 1], [3, 4, 5]],
    3: [[0, 1, 4], [1, 2, 4], [2, 3, 4], [0, 4, 3], [0, 3, 2, 1]],
}
# parametric vertex locations (the shifted face ips are these)
VPAR = {1: [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
        2: [(0, 0, -1), (1, 0, -1), (0, 1, -1), (0, 0, 1), (1, 0, 1), (0, 1, 1)],
        3: [(-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0), (0, 0, 1)]}
# physical reference elements (positive orientation)
PHYS = {0: [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)],
        1: [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
        2: [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)],
        3: [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0.5, 0.5, 1)]}


def distorted(topo, seed):
    rng = np.random.default_rng(seed)
    c = np.array(PHYS[topo], dtype=np.float64)
    return c + rng.uniform(-0.08, 0.08, size=c.shape)


@pytest.mark.parametrize("topo", [0, 1, 2, 3])
def test_side_tables(oracle, topo):
    assert oracle.topo_sides(topo) == REF_SIDES[topo]


@pytest.mark.parametrize("topo", [1, 2, 3])
def test_face_ips_match_reference(oracle, topo):
    for ord_, want in enumerate(REF_IPS[topo]):
        got = oracle.side_ips(topo, ord_, 0)
        want = np.array(want, dtype=np.float64)
        if topo == 1 and ord_ == 3:
            want = want[[0, 2, 1]]     # side order 0, 2, 1 (see the module docstring)
        np.testing.assert_allclose(got, want, atol=1e-15, err_msg=f"topo {topo} side {ord_}")
        sh = oracle.side_ips(topo, ord_, 1)
        np.testing.assert_array_equal(sh, np.array([VPAR[topo][n] for n in REF_SIDES[topo][ord_]],
                                                   dtype=np.float64))


def test_tri_areas_match_reference_fortran(oracle):
    rng = np.random.default_rng(11)
    for _ in range(20):
        p = rng.normal(size=(3, 3))
        np.testing.assert_allclose(oracle.tri_face_areav(p), oracle.ref_tri_areav(p),
                                   rtol=1e-14, atol=1e-15)
    # the three sub-faces partition the triangle: they sum to its area vector
    p = rng.normal(size=(3, 3))
    tri = 0.5 * np.cross(p[1] - p[0], p[2] - p[0])
    np.testing.assert_allclose(oracle.tri_face_areav(p).sum(axis=0), tri, rtol=1e-13)


@pytest.mark.parametrize("topo", [0, 1, 2, 3])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_sides_close_and_point_out(oracle, topo, seed):
    c = distorted(topo, seed)
    ctr = c.mean(axis=0)
    total = np.zeros(3)
    for nodes in REF_SIDES[topo]:
        p = c[nodes]
        av = oracle.tri_face_areav(p) if len(nodes) == 3 else oracle.quad_face_areav(p)
        a = av.sum(axis=0)
        assert np.dot(a, p.mean(axis=0) - ctr) > 0
        total += a
    assert np.abs(total).max() < 1e-14


@pytest.mark.parametrize("topo", [0, 1, 2, 3])
@pytest.mark.parametrize("shifted", [0, 1])
def test_face_gradop_linear_exact(oracle, topo, shifted):
    c = distorted(topo, 5)
    a = np.array([0.3, -1.2, 2.5])
    f = c @ a + 0.7
    for ord_ in range(len(REF_SIDES[topo])):
        g, det = oracle.side_gradop(topo, shifted, ord_, c)
        ok = np.ones(det.size, dtype=bool)
        if topo == 3 and shifted:
            # shifted_face_grad_op at the apex: the shifted pyramid derivatives
            # of the base nodes vanish there, det J = 0 exactly, and the
            # gradient operator's fallback (det -> 1) yields a zero operator
            apex = np.array([n == 4 for n in REF_SIDES[3][ord_]])
            assert np.all(det[apex] == 0.0) and np.all(g[apex] == 0.0)
            ok = ~apex
        assert np.all(det[ok] > 0)
        grad = np.einsum("knd,n->kd", g[ok], f)
        np.testing.assert_allclose(grad, np.broadcast_to(a, grad.shape), rtol=1e-12, atol=1e-12)
        if topo == 1:   # constant on the tet: the element's gradient
            _, _, ge, _ = oracle.topo_geometry(1, c)
            for k in range(g.shape[0]):
                np.testing.assert_allclose(g[k], ge[0], rtol=1e-13, atol=1e-14)
