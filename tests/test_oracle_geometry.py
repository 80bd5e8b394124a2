"""Pins the oracle's hex8/quad4 master-element restatement.

* against golden vectors produced by the reference's own Fortran
  (tests/golden/hex8_geometry.npz, written by make_geometry_golden.py from
  oracle/_ref, i.e. external/nalu/src/master_element/*.F compiled by flang);
* against that reference build live, on fresh random elements, when present;
* against closed forms: uniform spacing h gives SCS areas h^2/4 along the
  edge and SCV volumes h^3/8, SCV volumes sum to the element volume, shape
  functions form a partition of unity, dN/dx of a linear field is exact, the
  SCS of a node plus its exposed faces close its control volume.
"""
import os

import numpy as np
import pytest

from conftest import rel_err

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hex8_geometry.npz")
SCS = np.array([[0.00, -0.25, -0.25], [0.25, 0.00, -0.25], [0.00, 0.25, -0.25], [-0.25, 0.00, -0.25],
                [0.00, -0.25, 0.25], [0.25, 0.00, 0.25], [0.00, 0.25, 0.25], [-0.25, 0.00, 0.25],
                [-0.25, -0.25, 0.00], [0.25, -0.25, 0.00], [0.25, 0.25, 0.00], [-0.25, 0.25, 0.00]])
LR = np.array([[0, 1], [1, 2], [2, 3], [0, 3], [4, 5], [5, 6], [6, 7], [4, 7], [0, 4], [1, 5], [2, 6], [3, 7]])
SIDES = np.array([[0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [0, 4, 7, 3], [0, 3, 2, 1], [4, 5, 6, 7]])


def _geo(po, co):
    L = po.lib()
    co = np.ascontiguousarray(co)
    a = np.zeros((12, 3)); v = np.zeros(8)
    L.or_hex_scs_areav(po.dptr(co), po.dptr(a))
    L.or_hex_scv_vol(po.dptr(co), po.dptr(v))
    g0 = np.zeros((12, 8, 3)); d0 = np.zeros(12); g1 = np.zeros((12, 8, 3)); d1 = np.zeros(12)
    L.or_hex_scs_gradop(0, po.dptr(co), po.dptr(g0), po.dptr(d0))
    L.or_hex_scs_gradop(1, po.dptr(co), po.dptr(g1), po.dptr(d1))
    return a, v, g0, d0, g1, d1


def test_against_reference_golden_vectors(oracle):
    G = np.load(GOLD)
    L = oracle.lib()
    for k, co in enumerate(G["coords"]):
        a, v, g0, d0, g1, d1 = _geo(oracle, co)
        assert rel_err(a, G["scs_areav"][k]) < 1e-14
        assert rel_err(v, G["scv_vol"][k]) < 1e-14
        assert rel_err(g0, G["gradop_std"][k]) < 1e-13
        assert rel_err(d0, G["detj_std"][k]) < 1e-14
        assert rel_err(g1, G["gradop_shift"][k]) < 1e-13
        assert rel_err(d1, G["detj_shift"][k]) < 1e-14
        for f in range(6):
            gf = np.zeros((4, 8, 3)); df = np.zeros(4)
            L.or_hex_face_gradop(0, f, oracle.dptr(np.ascontiguousarray(co)), oracle.dptr(gf), oracle.dptr(df))
            assert rel_err(gf, G["face_gradop"][k, f]) < 1e-13
            L.or_hex_face_gradop(1, f, oracle.dptr(np.ascontiguousarray(co)), oracle.dptr(gf), oracle.dptr(df))
            assert rel_err(gf, G["face_gradop_shift"][k, f]) < 1e-13
            q = np.ascontiguousarray(co[SIDES[f]]); qa = np.zeros((4, 3))
            L.or_quad_face_areav(oracle.dptr(q), oracle.dptr(qa))
            assert rel_err(qa, G["quad_areav"][k, f]) < 1e-14
    sh = np.zeros((12, 8))
    L.or_hex_scs_shape(0, oracle.dptr(sh))
    assert np.array_equal(sh, G["shape_std"])
    L.or_hex_scs_shape(1, oracle.dptr(sh))
    assert np.array_equal(sh, G["shape_shift"])


def test_against_reference_build_live(oracle):
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(7)
    base = np.array([[-.5, -.5, -.5], [.5, -.5, -.5], [.5, .5, -.5], [-.5, .5, -.5],
                     [-.5, -.5, .5], [.5, -.5, .5], [.5, .5, .5], [-.5, .5, .5]])
    for _ in range(50):
        co = base * rng.uniform(0.01, 3.0, size=3) + rng.uniform(-0.12, 0.12, size=base.shape)
        a, v, g0, d0, g1, d1 = _geo(oracle, co)
        assert rel_err(a, oracle.ref_hex_scs_areav(co)) < 1e-14
        assert rel_err(v, oracle.ref_hex_scv_vol(co)) < 1e-14
        gr, dr = oracle.ref_hex_gradop(2 * SCS, co)
        assert rel_err(g1, gr) < 1e-13 and rel_err(d1, dr) < 1e-14


def test_uniform_element_closed_forms(oracle):
    h = 0.37
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float) * h + 2.0
    a, v, g0, d0, g1, d1 = _geo(oracle, base)
    np.testing.assert_allclose(v, h**3 / 8, rtol=1e-14)
    for ip, (l, r) in enumerate(LR):
        e = (base[r] - base[l]) / h           # unit edge direction L -> R
        np.testing.assert_allclose(a[ip], e * h * h / 4, atol=1e-15)
    # gradient of a linear field is exact at every ip
    c = np.array([0.3, -1.7, 2.2])
    phi = base @ c + 4.0
    for g in (g0, g1):
        np.testing.assert_allclose(np.einsum("inj,n->ij", g, phi), np.tile(c, (12, 1)), rtol=1e-12)


def test_partition_of_unity_and_volume(oracle):
    rng = np.random.default_rng(3)
    G = np.load(GOLD)
    L = oracle.lib()
    sh = np.zeros((12, 8))
    for s in (0, 1):
        L.or_hex_scs_shape(s, oracle.dptr(sh))
        np.testing.assert_allclose(sh.sum(axis=1), 1.0, rtol=0, atol=1e-15)
    for co in G["coords"]:
        a, v, g0, d0, g1, d1 = _geo(oracle, co)
        np.testing.assert_allclose(g0.sum(axis=1), 0.0, atol=1e-12 * np.abs(g0).max())
        # closed control volume: node SCS (signed) + its share of exposed faces = 0
        qa = {}
        for f in range(6):
            q = np.ascontiguousarray(co[SIDES[f]]); out = np.zeros((4, 3))
            L.or_quad_face_areav(oracle.dptr(q), oracle.dptr(out))
            qa[f] = out
        for node in range(8):
            s = np.zeros(3)
            for ip, (l, r) in enumerate(LR):
                if l == node: s += a[ip]
                if r == node: s -= a[ip]
            for f in range(6):
                for k in range(4):
                    if SIDES[f][k] == node:
                        s += qa[f][k]
            assert np.abs(s).max() < 1e-13 * np.abs(a).max()
    # SCV volumes sum to the element volume of a trilinear affine element
    M = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    co = base @ M.T
    _, v, *_ = _geo(oracle, co)
    np.testing.assert_allclose(v.sum(), abs(np.linalg.det(M)), rtol=1e-13)


def test_planted_geometry_mistake_is_caught(oracle):
    """A swapped L/R orientation (sign error) must fail the closed-CV check."""
    G = np.load(GOLD)
    a, *_ = _geo(oracle, G["coords"][10])
    bad = a.copy()
    bad[3] = -bad[3]
    assert rel_err(bad, G["scs_areav"][10]) > 1e-3
