"""Device master-element geometry (oagpu_element_geometry through the C ABI)
of hex8, tet4, wedge6 and pyramid5 batches against the oracle (itself pinned
to the reference's Fortran): SCS area vectors, SCV volumes, SCS gradient
operators and det J within 1e-13 relative, standard and shifted ips."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REF_VERTS = {
    0: np.array([[-.5, -.5, -.5], [.5, -.5, -.5], [.5, .5, -.5], [-.5, .5, -.5],
                 [-.5, -.5, .5], [.5, -.5, .5], [.5, .5, .5], [-.5, .5, .5]]),
    1: np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float),
    2: np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], float),
    3: np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0], [0, 0, 1]], float),
}


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def batch(topo, n, rng):
    v = REF_VERTS[topo]
    out = np.empty((n,) + v.shape)
    for e in range(n):
        A = np.eye(3) + rng.normal(scale=0.15, size=(3, 3))
        out[e] = (v + rng.uniform(-0.08, 0.08, v.shape)) @ A.T + rng.normal(size=3)
    return out


@pytest.mark.parametrize("topo", [0, 1, 2, 3])
@pytest.mark.parametrize("shifted", [0, 1])
def test_element_geometry_matches_oracle(oracle, oa, topo, shifted):
    print(f"[case] element geometry topo={topo} shifted={shifted}", flush=True)
    rng = np.random.default_rng(100 + 10 * topo + shifted)
    c = batch(topo, 257, rng)
    A, V, G, D = oa.element_geometry(topo, c, shifted)
    for e in range(c.shape[0]):
        a, v, g, d = oracle.topo_geometry(topo, c[e], shifted)
        np.testing.assert_allclose(A[e], a, rtol=0, atol=1e-13 * np.abs(a).max())
        np.testing.assert_allclose(V[e], v, rtol=1e-13)
        np.testing.assert_allclose(G[e], g, rtol=0, atol=1e-13 * np.abs(g).max())
        np.testing.assert_allclose(D[e], d, rtol=1e-13)


def test_element_geometry_rejects_inverted_elements(oa):
    c = REF_VERTS[1][None].copy()
    c[0, [1, 2]] = c[0, [2, 1]]           # swap two vertices: negative volume
    with pytest.raises(oa.OagpuError):
        oa.element_geometry(1, c)
