"""Pins the oracle's velocity-inlet / static-pressure-outlet treatment with
closed forms on a channel: the unit cube with a velocity inlet at x = 0
(id 0), a static-pressure outlet at x = 1 (id 1) and symmetry planes on the
four other faces (id 2).

* Uniform flow U = (U0, 0, 0), p = p_out is an exact steady solution: the
  side mass flows are rho U0 times the face-ip areas (in at the inlet, out
  at the outlet), the momentum and pressure-correction right-hand sides
  vanish at every node, and a full outer iteration leaves the state
  unchanged; likewise a uniform k = k_in for the scalar element and
  boundary terms;
* reversal flags (flowModel::updateFlowReversalFlag): an outlet face whose
  side flow sums to a net inflow is flagged and its rates zeroed; a face
  with net outflow keeps its positive rates, negative ones clipped to 0;
* with an outlet no pressure reference row is imposed (the reference node's
  row keeps its coefficients): the outlet's masked face columns make the p'
  system definite.
"""
import numpy as np
import pytest

U0, P_OUT, K_IN, W_IN = 1.3, 0.25, 2.0e-3, 5.0
INLET, OUTLET, SYM = 1, 2, 3


def channel(meshgen, n, perturb=0.0, seed=0):
    m = meshgen.cavity_hex(n, perturb=perturb, seed=seed)
    bc = np.full(m["bf_bc"].shape, 2, dtype=np.int32)
    bc[m["bf_ord"] == 3] = 0
    bc[m["bf_ord"] == 1] = 1
    m["bf_bc"] = bc
    return m


def make(oracle, meshgen, n=4, perturb=0.0, seed=0, **kw):
    m = channel(meshgen, n, perturb, seed)
    BC = np.array([[U0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    c = oracle.Case(m, oracle.default_params(**kw), BC)
    c.set_bc_types([INLET, OUTLET, SYM])
    c.set_bc_values(p=[0.0, P_OUT, 0.0], k=[K_IN, 0.0, 0.0], w=[W_IN, 0.0, 0.0])
    return c, m


def uniform_state(c):
    U = np.zeros((c.nn, 3))
    U[:, 0] = U0
    c.field("U")[:] = U
    c.field("p")[:] = P_OUT
    c.initialize()


@pytest.mark.parametrize("perturb", [0.0, 0.2])
def test_uniform_flow_is_steady(oracle, meshgen, perturb):
    c, m = make(oracle, meshgen, n=4, perturb=perturb, seed=2)
    uniform_state(c)
    md = c.field("msd")
    A = c.field("bareav")
    ords = m["bf_ord"]
    np.testing.assert_allclose(md[ords == 3], c.params.rho * U0 * A[ords == 3][:, :, 0], rtol=1e-14)
    np.testing.assert_allclose(md[ords == 1], c.params.rho * U0 * A[ords == 1][:, :, 0], rtol=1e-14)
    assert np.abs(md[(ords != 1) & (ords != 3)]).max() == 0.0
    c.assemble_momentum()
    scale = c.params.rho * U0 * U0 / 4**2
    assert np.abs(c.rhs(3)).max() < 1e-13 * scale
    c.compute_du()
    c.update_mdot()
    c.assemble_pcorr()
    assert np.abs(c.rhs(1)).max() < 1e-13 * scale
    U1 = c.field("U").copy()
    c.outer_iteration([20, 0.1, 200, 1e-6])
    np.testing.assert_allclose(c.field("U"), U1, rtol=0, atol=1e-12)
    np.testing.assert_allclose(c.field("p"), P_OUT, rtol=0, atol=1e-12)
    assert np.abs(c.field("div")).max() < 1e-13 * scale


def test_uniform_scalar_is_steady(oracle, meshgen):
    c, m = make(oracle, meshgen, n=4, perturb=0.15, seed=3)
    uniform_state(c)
    c.sst_enable(oracle.default_sst())
    c.field("k")[:] = K_IN
    c.field("omega")[:] = W_IN
    c.field("mut")[:] = 1e-3
    c.field("F1")[:] = 0.5
    for which in ("k", "omega"):
        c.assemble_scalar(which, terms=4)   # element and boundary terms only
        assert np.abs(c.rhs(1)).max() < 1e-15


def test_reversal_flags(oracle, meshgen):
    n = 4
    c, m = make(oracle, meshgen, n=n)
    uniform_state(c)
    ords = m["bf_ord"]
    out = np.where(ords == 1)[0]
    md = c.field("msd")
    rng = np.random.default_rng(5)
    vals = rng.normal(scale=0.05, size=(out.size, 4))
    vals[0] = [-0.2, 0.05, 0.01, -0.03]     # net inflow: flagged, zeroed
    vals[1] = [0.2, -0.05, 0.01, 0.03]      # net outflow: clipped only
    md[out] = vals
    c.update_reversal()
    rf = c.reversal()
    side = vals.sum(axis=1)
    flagged = side < -np.finfo(float).eps
    np.testing.assert_array_equal(rf[out], np.repeat(flagged[:, None].astype(np.int32), 4, axis=1))
    want = np.where(flagged[:, None], 0.0, np.maximum(vals, 0.0))
    np.testing.assert_array_equal(md[out], want)
    assert np.all(rf[ords != 1] == 0)
    # a flagged face is released once the velocity leaves through it and the
    # outlet pressure is not above the face's mean pressure
    c.field("p")[:] = P_OUT + 1.0
    c.update_reversal()
    assert np.all(c.reversal()[out] == 0)


def test_outlet_replaces_pressure_reference(oracle, meshgen):
    c, m = make(oracle, meshgen, n=4, perturb=0.1, seed=4)
    uniform_state(c)
    rng = np.random.default_rng(6)
    c.field("U")[:] += rng.normal(scale=0.05, size=(c.nn, 3))
    c.assemble_momentum()
    c.compute_du()
    c.update_mdot()
    c.assemble_pcorr()
    import scipy.sparse as sp
    A = sp.csr_matrix((c.vals(1).ravel().copy(), c.cols.copy(), c.row_ptr.copy()), shape=(c.nn, c.nn))
    r0 = c.params.p_ref_node
    assert A[r0].nnz > 1 and np.abs(A[r0].toarray()).sum() > abs(A[r0, r0])  # row kept
    ev = np.linalg.eigvalsh(0.5 * (A + A.T).toarray())
    assert ev.min() > 1e-6 * ev.max()   # positive definite without a reference row
