"""Pins the oracle's static-pressure opening (OR_BC_OPENING) on a channel:
the unit cube (perturbed interior nodes, flat faces) with a velocity inlet
at x = 0 (id 0), the boundary under test at x = 1 (id 1, outward normal +x)
and no-slip walls elsewhere (id 2).

What the reference's opening kernels leave, each checked against another
boundary type or a closed form:

* where every side flow leaves (msd > 0), an opening assembles exactly
  what a static-pressure outlet with no reversed ip assembles: momentum
  (implicit upwind + zero normal stress, navierStokesAssemblerElemBoundary
  Conditions.cpp:3505-3510 vs :1468-1775), p' (the outlet's terms at every
  ip, pressureCorrectionAssemblerElemBoundaryConditions.cpp:3026-3560) and
  k / omega (implicit upwind, phiAssemblerElemBoundaryConditions.hpp:
  1439-1760);
* the side mass flow: initial rates rho U_ip . A are not clipped
  (flowModel.cpp:4817-4992): in uniform inflow they equal a velocity
  inlet's rho U_bc . A; the update (:8094-8380) is the outlet's Rhie-Chow
  with the nodal pressure on the side's nodes, so with a uniform pressure
  p_s on the opening it equals an outlet's update with p_bc = p_s;
* where a side flow enters (msd <= 0), the momentum face terms are the
  explicit flux msd Iu d with Iu = U_ip . n / (d . n + SMALL) and the zero
  normal stress without its grad(U)^T terms: with uniform U = u0 the rhs
  change against a slip wall (no face terms) is 0.75 sum(msd) (u0 - (u0 .
  n / d . n) d) at each nearest node (the side rates also enter the node's
  mass divergence, whose term div U is on the rhs; 0.75: the boundary-node
  rhs factor), zero when d is along u0; on a
  flat face the off-diagonal 3 x 3 face blocks are s (I - n n^T), the
  outlet's tangential part with its grad(U)^T column removed;
* no reversal flags are ever set on an opening, and an opening, like an
  outlet, replaces the pressure reference row.
"""
import numpy as np
import pytest

INLET, OUTLET, WALL, SLIP, OPENING = 1, 2, 0, 4, 5
P_OUT = 0.2
SIDES = np.array([[0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [0, 4, 7, 3], [0, 3, 2, 1],
                  [4, 5, 6, 7]])


def channel(meshgen, n=5, seed=6):
    m = meshgen.cavity_hex(n, perturb=0.15, seed=seed)
    bc = np.full(m["bf_bc"].shape, 2, dtype=np.int32)
    bc[m["bf_ord"] == 3] = 0
    bc[m["bf_ord"] == 1] = 1
    m["bf_bc"] = bc
    return m


def make(oracle, meshgen, t1, bcU1=(0.0, 0.0, 0.0), p1=P_OUT, n=5, seed=6, u_in=1.0):
    m = channel(meshgen, n, seed)
    BC = np.array([[u_in, 0.0, 0.0], list(bcU1), [0.0] * 3])
    c = oracle.Case(m, oracle.default_params(), BC)
    c.set_bc_types([INLET, t1, WALL])
    c.set_bc_values(p=[0.0, p1, 0.0], k=[2e-3, 0.0, 0.0], w=[5.0, 0.0, 0.0])
    return c, m


def state(c, U, p, seed=3):
    c.field("U")[:] = U
    c.field("p")[:] = p
    c.initialize()
    rng = np.random.default_rng(seed)
    c.field("gradP")[:] = rng.normal(scale=0.2, size=(c.nn, 3))
    c.field("gradU")[:] = rng.normal(size=(c.nn, 9))


def noisy_u(nn, mean, seed=4, scale=0.1):
    rng = np.random.default_rng(seed)
    U = rng.normal(scale=scale, size=(nn, 3))
    return U + np.asarray(mean)[None, :]


def noisy_p(nn, seed=5):
    return P_OUT + np.random.default_rng(seed).normal(scale=0.1, size=nn)


def systems(oracle, c, du=None):
    c.assemble_momentum()
    c.compute_du()
    out = {"A3": c.vals(3).copy(), "b3": c.rhs(3).copy(), "du": c.field("du").copy()}
    if du is not None:
        c.field("du")[:] = du
    c.assemble_pcorr()
    out.update(A1=c.vals(1).copy(), b1=c.rhs(1).copy())
    c.sst_enable(oracle.default_sst())
    c.field("k")[:] = np.linspace(1e-3, 3e-3, c.nn)
    c.field("omega")[:] = np.linspace(4.0, 6.0, c.nn)
    c.field("mut")[:] = 1e-3
    c.field("F1")[:] = 0.5
    for which in ("k", "omega"):
        c.assemble_scalar(which, terms=4)
        out["A_" + which] = c.vals(1).copy()
        out["b_" + which] = c.rhs(1).copy()
    return out


def nearest_sum(m, msd, nn, ord_=1):
    """sum of the side rates whose face ip is nearest to each node"""
    s = np.zeros(nn)
    sel = np.where(m["bf_ord"] == ord_)[0]
    nodes = m["conn"][m["bf_elem"][sel][:, None], SIDES[m["bf_ord"][sel]]]
    np.add.at(s, nodes.ravel(), msd[sel].ravel())
    return s


def test_outflow_opening_is_an_outlet(oracle, meshgen):
    res = {}
    for t in (OUTLET, OPENING):
        c, m = make(oracle, meshgen, t)
        state(c, noisy_u(c.nn, [1.0, 0.0, 0.0]), noisy_p(c.nn))
        md = c.field("msd")[m["bf_ord"] == 1]
        assert md.min() > 0.0
        res[t] = systems(oracle, c)
    for k in res[OUTLET]:
        np.testing.assert_array_equal(res[OPENING][k], res[OUTLET][k], err_msg=k)


def test_pcorr_and_scalar_match_outlet_in_inflow(oracle, meshgen):
    """p' and k / omega take the outlet's terms at every ip whatever the
    flow direction (the outlet takes the opening's side rates, which its
    init would clip, and its du, which its momentum diagonal changes)"""
    res = {}
    for t in (OPENING, OUTLET):
        c, m = make(oracle, meshgen, t)
        state(c, noisy_u(c.nn, [0.4, 0.0, 0.0], scale=0.4), noisy_p(c.nn))
        if t == OPENING:
            msd = c.field("msd").copy()
            assert msd[m["bf_ord"] == 1].min() < 0.0 < msd[m["bf_ord"] == 1].max()
            res[t] = systems(oracle, c)
        else:
            c.field("msd")[:] = msd
            res[t] = systems(oracle, c, du=res[OPENING]["du"])
    for k in ("A1", "b1", "A_k", "b_k", "A_omega", "b_omega"):
        np.testing.assert_array_equal(res[OPENING][k], res[OUTLET][k], err_msg=k)


def test_init_rates_not_clipped(oracle, meshgen):
    u0 = np.array([-0.7, 0.2, 0.1])
    c, m = make(oracle, meshgen, OPENING)
    state(c, np.tile(u0, (c.nn, 1)), np.full(c.nn, P_OUT))
    ci, _ = make(oracle, meshgen, INLET, bcU1=u0)
    state(ci, np.tile(u0, (ci.nn, 1)), np.full(ci.nn, P_OUT))
    sel = m["bf_ord"] == 1
    a, b = c.field("msd")[sel], ci.field("msd")[sel]
    assert a.max() < 0.0
    np.testing.assert_allclose(a, b, rtol=1e-14, atol=0)


def test_update_uses_nodal_side_pressure(oracle, meshgen):
    m = channel(meshgen)
    nn = m["xyz"].shape[0]
    U = noisy_u(nn, [0.8, 0.0, 0.0])
    p = np.random.default_rng(9).normal(scale=0.1, size=nn)
    side = np.unique(m["conn"][m["bf_elem"][m["bf_ord"] == 1][:, None], SIDES[1]].ravel())
    p[side] = 0.35
    out = {}
    for t, p1 in ((OPENING, -3.0), (OUTLET, 0.35)):
        c, m = make(oracle, meshgen, t, p1=p1)
        state(c, U, p)
        c.assemble_momentum()
        c.compute_du()
        c.update_mdot()
        out[t] = c.field("msd")[m["bf_ord"] == 1].copy()
    np.testing.assert_array_equal(out[OPENING], out[OUTLET])


@pytest.mark.parametrize("direction", ["along_u", "slanted", "normal"])
def test_entrainment_momentum(oracle, meshgen, direction):
    u0 = np.array([-0.8, 0.3, -0.1])
    d = {"along_u": u0, "slanted": np.array([-1.0, 0.5, 0.2]),
         "normal": np.zeros(3)}[direction]
    res = {}
    for t in (OPENING, SLIP, OUTLET):
        c, m = make(oracle, meshgen, t, bcU1=d)
        state(c, np.tile(u0, (c.nn, 1)), np.full(c.nn, P_OUT))
        if t == OPENING:
            msd = c.field("msd").copy()
        c.assemble_momentum()
        res[t] = (c.vals(3).copy(), c.rhs(3).copy().reshape(-1, 3))
    nn = c.nn
    S = nearest_sum(m, msd, nn)
    dd = d / np.linalg.norm(d) if d.any() else np.array([-1.0, 0.0, 0.0])
    want = 0.75 * S[:, None] * (u0[None, :] - (u0[0] / dd[0]) * dd[None, :])
    got = res[OPENING][1] - res[SLIP][1]
    scale = 0.75 * np.abs(S).max() * np.abs(u0).max()
    assert scale > 0
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * scale)
    if direction != "along_u":
        assert np.abs(want).max() > 0.1 * scale
    # off-diagonal face blocks: s (I - n n^T), n = +x
    F = res[OPENING][0] - res[SLIP][0]
    G = res[OUTLET][0] - res[SLIP][0]
    rows = np.repeat(np.arange(nn), np.diff(c.row_ptr))
    off = c.cols != rows
    Fo, Go = F[off], G[off]
    big = np.abs(Fo).max()
    assert big > 0
    assert np.abs(Fo[:, :, 0]).max() < 1e-12 * big
    assert np.abs(Fo[:, 0, :]).max() < 1e-12 * big
    assert np.abs(Fo[:, 1, 2]).max() < 1e-12 * big and np.abs(Fo[:, 2, 1]).max() < 1e-12 * big
    np.testing.assert_allclose(Fo[:, 1, 1], Fo[:, 2, 2], rtol=0, atol=1e-12 * big)
    np.testing.assert_allclose(Fo[:, 1:, 1:], Go[:, 1:, 1:], rtol=0, atol=1e-12 * big)
    assert np.abs(Go[:, 1:, 0]).max() > 1e-3 * big   # the grad(U)^T column an outlet keeps


def test_no_reversal_flags_and_no_reference_row(oracle, meshgen):
    c, m = make(oracle, meshgen, OPENING)
    state(c, np.tile([-0.5, 0.0, 0.0], (c.nn, 1)), np.full(c.nn, P_OUT))
    c.update_reversal()
    assert np.all(c.reversal() == 0)
    c.assemble_momentum()
    c.compute_du()
    c.assemble_pcorr()
    r0 = c.params.p_ref_node
    lo, hi = c.row_ptr[r0], c.row_ptr[r0 + 1]
    assert np.count_nonzero(c.vals(1)[lo:hi]) > 1   # the reference row keeps its coefficients
