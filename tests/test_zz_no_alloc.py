"""The allocation-free geometry entry point (oagpu_element_geometry_noalloc,
device arrays in and out, its error word in the library's fixed scratch
array) against the allocating one (oagpu_element_geometry) through the C ABI:
bitwise equal outputs for every topology and both ip sets, on two calls in a
row with the second batch smaller, and the error word cleared between calls
(an inverted element reports OAGPU_EGEOM, the next good batch OAGPU_OK)."""
import ctypes as C

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
OAGPU_EGEOM = 5


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


def noalloc(oa, topo, coords, shifted):
    """Device buffers (hipMalloc through ctypes, test plumbing), the call
    through the C ABI; returns (status, areav, scv, dndx, detj) on the host."""
    from hipmem import DeviceArray, hip
    npe, ns = oa.TOPOLOGY_SIZES[topo]
    n = coords.shape[0]
    c = DeviceArray(coords.shape, host=coords)
    outs = [DeviceArray(s) for s in ((n, ns, 3), (n, npe), (n, ns, npe, 3), (n, ns))]
    assert hip().hipDeviceSynchronize() == 0
    rc = oa.load_library().oagpu_element_geometry_noalloc(topo, shifted, n, c.ptr,
                                                          *[o.ptr for o in outs], None)
    assert hip().hipDeviceSynchronize() == 0
    res = (rc, *[o.get() for o in outs])
    for o in (c, *outs):
        o.close()
    return res


@pytest.mark.parametrize("topo", [0, 1, 2, 3])
@pytest.mark.parametrize("shifted", [0, 1])
def test_noalloc_matches_allocating_entry_point(oa, topo, shifted):
    print(f"[case] noalloc geometry topo={topo} shifted={shifted}", flush=True)
    rng = np.random.default_rng(500 + 10 * topo + shifted)
    for n in (1000, 37):                      # two calls in a row, the second smaller
        c = batch(topo, n, rng)
        rc, A, V, G, D = noalloc(oa, topo, c, shifted)
        assert rc == 0, oa.load_library().oagpu_last_error()
        A0, V0, G0, D0 = oa.element_geometry(topo, c, shifted)
        for got, want in ((A, A0), (V, V0), (G, G0), (D, D0)):
            assert np.array_equal(got, want)


def test_noalloc_error_word_cleared_between_calls(oa):
    print("[case] noalloc geometry error word", flush=True)
    bad = REF_VERTS[1][None].copy()
    bad[0, [1, 2]] = bad[0, [2, 1]]           # swapped vertices: negative volume
    rc, *_ = noalloc(oa, 1, bad, 0)
    assert rc == OAGPU_EGEOM
    with pytest.raises(oa.OagpuError):
        oa.element_geometry(1, bad)
    good = batch(1, 5, np.random.default_rng(3))
    rc, A, V, G, D = noalloc(oa, 1, good, 0)
    assert rc == 0
    A0, V0, G0, D0 = oa.element_geometry(1, good, 0)
    assert np.array_equal(A, A0) and np.array_equal(D, D0)


def test_noalloc_rejects_bad_batches(oa):
    L = oa.load_library()
    assert L.oagpu_element_geometry_noalloc(7, 0, 1, C.c_void_p(8), None, None, None, None,
                                            None) != 0
    assert L.oagpu_element_geometry_noalloc(0, 0, -1, None, None, None, None, None, None) != 0
    assert L.oagpu_element_geometry_noalloc(0, 0, (1 << 30) + 1, C.c_void_p(8), None, None,
                                            None, None, None) != 0
    assert L.oagpu_element_geometry_noalloc(0, 0, 0, None, None, None, None, None, None) == 0
