"""Exodus II ingest (openaccel_amd.exodus) on the reference's own cavity mesh
(tests/golden/cavity_mesh.e, a byte copy of examples/cavity/mesh.e): sizes,
side sets on the planes they name with outward side ordinals, positive
element volumes, element orientations varied enough that nodes repeat as the
same local node of several elements; and, where the reference tree is
present, every classic-format example mesh it ships."""
import glob
import os

import numpy as np
import pytest

from openaccel_amd import exodus

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "cavity_mesh.e")
REF_EXAMPLES = "/root/reference/examples"
SIDES = np.array([[0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [0, 4, 7, 3], [0, 3, 2, 1], [4, 5, 6, 7]])


def hex_jacobian_centre(xyz, conn):
    e = xyz[conn[:, :8]]
    d1 = e[:, [1, 2, 5, 6]].sum(1) - e[:, [0, 3, 4, 7]].sum(1)
    d2 = e[:, [2, 3, 6, 7]].sum(1) - e[:, [0, 1, 4, 5]].sum(1)
    d3 = e[:, [4, 5, 6, 7]].sum(1) - e[:, [0, 1, 2, 3]].sum(1)
    return np.einsum("ij,ij->i", np.cross(d1, d2), d3) / 64.0


def test_cavity_fixture_sizes_and_sets(meshgen):
    m = exodus.read_exodus(FIXTURE)
    assert m["xyz"].shape == (882, 3) and m["conn"].shape == (400, 8)
    assert np.all(m["etype"] == exodus.HEX8)
    assert m["sideset_names"] == ["TOP", "SIDES", "FRONTANDBACK"]
    assert np.bincount(m["bf_bc"]).tolist() == [20, 60, 800]
    np.testing.assert_allclose(m["xyz"].min(0), [0, 0, 0], atol=1e-15)
    np.testing.assert_allclose(m["xyz"].max(0), [0.1, 0.1, 0.01], atol=1e-15)
    assert hex_jacobian_centre(m["xyz"], m["conn"]).min() > 0
    assert meshgen.repeated_local_nodes(m["conn"], 882) > 0
    assert exodus.nearest_node(m["xyz"], (0, 0, 0)) == int(np.argmin(np.linalg.norm(m["xyz"], axis=1)))


def test_cavity_side_sets_are_outward_faces_on_their_planes():
    m = exodus.read_exodus(FIXTURE)
    xyz, conn = m["xyz"], m["conn"]
    ctr = xyz[conn].mean(axis=1)
    planes = {0: [(1, 0.1)], 1: [(0, 0.0), (0, 0.1), (1, 0.0)], 2: [(2, 0.0), (2, 0.01)]}
    for f in range(m["bf_elem"].size):
        e, o, b = m["bf_elem"][f], m["bf_ord"][f], m["bf_bc"][f]
        q = xyz[conn[e][SIDES[o]]]
        on = [d for d, v in planes[b] if np.all(np.abs(q[:, d] - v) < 1e-8)]   # float32 coordinates
        assert len(on) == 1, (f, b)
        n = np.cross(q[2] - q[0], q[3] - q[1])            # right-handed: outward
        assert np.dot(n, q.mean(0) - ctr[e]) > 0
    # every exposed element side is in exactly one set
    key = lambda e, s: tuple(sorted(conn[e][SIDES[s]].tolist()))
    count = {}
    for e in range(conn.shape[0]):
        for s in range(6):
            count[key(e, s)] = count.get(key(e, s), 0) + 1
    exposed = {k for k, v in count.items() if v == 1}
    listed = [key(e, o) for e, o in zip(m["bf_elem"], m["bf_ord"])]
    assert len(listed) == len(set(listed)) and set(listed) == exposed


def test_fixture_is_the_reference_file():
    ref = os.path.join(REF_EXAMPLES, "cavity", "mesh.e")
    if not os.path.exists(ref):
        pytest.skip("reference tree not present")
    with open(ref, "rb") as a, open(FIXTURE, "rb") as b:
        assert a.read() == b.read()


def test_reference_example_meshes():
    paths = sorted(glob.glob(os.path.join(REF_EXAMPLES, "*", "mesh.e")))
    if not paths:
        pytest.skip("reference tree not present")
    read = 0
    for p in paths:
        with open(p, "rb") as fh:
            magic = fh.read(4)
        if magic[:3] != b"CDF":
            with pytest.raises(ValueError):
                exodus.read_exodus(p)
            continue
        m = exodus.read_exodus(p)
        nn = m["xyz"].shape[0]
        npe = np.array([8, 4, 6, 5])[m["etype"]]
        for k in range(8):
            used = k < npe
            assert np.all((m["conn"][used, k] >= 0) & (m["conn"][used, k] < nn))
            assert np.all(m["conn"][~used, k] == -1)
        hexes = m["etype"] == exodus.HEX8
        assert hex_jacobian_centre(m["xyz"], m["conn"][hexes]).min() > 0, p
        assert m["bf_elem"].size == 0 or m["bf_elem"].max() < m["conn"].shape[0]
        read += 1
    assert read >= 10


def test_rejects_non_netcdf(tmp_path):
    bad = tmp_path / "x.e"
    bad.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(ValueError):
        exodus.read_exodus(str(bad))
    bad.write_bytes(b"nope")
    with pytest.raises(ValueError):
        exodus.read_exodus(str(bad))
