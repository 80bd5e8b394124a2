"""The oracle's slip wall (OR_BC_SLIP), pinned by what the reference's
assembleElemTermsBoundarySlipWall_ (navierStokesAssemblerElemBoundaryConditions.cpp:
2376-2657) leaves: no face terms at all.  A channel with slip walls
assembles the momentum and p' matrices of the same channel with those sides
taken out of the side list, and the same rhs except at nodes that lie only
on slip sides, whose momentum rhs keeps the steady boundary-node factor
(navierStokesAssembler.cpp:289: slip walls are relaxed like no-slip ones).
"""
import numpy as np

INLET, OUTLET, SLIP = 1, 2, 4
BC = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3])


def channel(meshgen, drop_slip):
    m = meshgen.cavity_hex(5, perturb=0.15, seed=6)
    bc = np.full(m["bf_bc"].shape, 2, dtype=np.int32)      # id 2: slip walls
    bc[m["bf_ord"] == 3] = 0
    bc[m["bf_ord"] == 1] = 1
    m["bf_bc"] = bc
    if drop_slip:
        keep = bc != 2
        for k in ("bf_elem", "bf_ord", "bf_bc"):
            m[k] = np.ascontiguousarray(m[k][keep])
    return m


def assemble(oracle, meshgen, drop_slip):
    m = channel(meshgen, drop_slip)
    c = oracle.Case(m, oracle.default_params(), BC)
    c.set_bc_types([INLET, OUTLET, SLIP])
    rng = np.random.default_rng(3)
    nn = c.nn
    U = rng.normal(scale=0.3, size=(nn, 3))
    U[:, 0] += 1.0
    c.field("U")[:] = U
    c.field("p")[:] = rng.normal(scale=0.1, size=nn)
    c.initialize()
    c.field("gradP")[:] = rng.normal(scale=0.2, size=(nn, 3))
    c.field("gradU")[:] = rng.normal(size=(nn, 9))
    c.assemble_momentum()
    c.compute_du()
    out = {"A3": c.vals(3).copy(), "b3": c.rhs(3).copy()}
    c.assemble_pcorr()
    out.update(A1=c.vals(1).copy(), b1=c.rhs(1).copy())
    return out, m


def test_slip_wall_adds_no_face_terms(oracle, meshgen):
    slip, m = assemble(oracle, meshgen, False)
    bare, mb = assemble(oracle, meshgen, True)
    np.testing.assert_array_equal(slip["A3"], bare["A3"])
    np.testing.assert_array_equal(slip["A1"], bare["A1"])
    np.testing.assert_array_equal(slip["b1"], bare["b1"])
    # nodes on slip sides only: boundary nodes in one mesh, interior in the other
    nn = m["xyz"].shape[0]

    def on(mesh):
        s = np.zeros(nn, bool)
        sides = np.array([[0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [0, 4, 7, 3], [0, 3, 2, 1],
                          [4, 5, 6, 7]])
        s[mesh["conn"][mesh["bf_elem"]][np.arange(mesh["bf_elem"].size)[:, None],
                                         sides[mesh["bf_ord"]]].ravel()] = True
        return s
    only = on(m) & ~on(mb)
    assert only.sum() > 0
    b_s, b_b = slip["b3"].reshape(-1, 3), bare["b3"].reshape(-1, 3)
    np.testing.assert_array_equal(b_s[~only], b_b[~only])
    np.testing.assert_allclose(b_s[only], 0.75 * b_b[only], rtol=1e-15, atol=0)
