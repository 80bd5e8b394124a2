"""Pins the oracle's high-resolution advection (limited deferred correction)
with closed forms:

* blending factor (Barth-Jespersen, smooth form): for phi = x with its exact
  gradient on a uniform grid, beta = 1 wherever x has neighbours on both
  sides and beta = 0 on the planes x = 0 and x = 1 (extrema); a node that
  is a strict local maximum gets beta = 0;
* deferred correction: for a linear velocity field and beta = 1 the upwind
  value plus the correction is the field at the ip, so the change of the
  momentum rhs between beta = 1 and beta = 0 is -sum(+-mdot (U(x_ip) -
  U_upwind)) over the element ips, computed here independently from the
  node coordinates (x_ip by trilinear interpolation of the corner points at
  the standard SCS locations).
"""
import numpy as np
import pytest

BC = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
LR = [(0, 1), (1, 2), (2, 3), (0, 3), (4, 5), (5, 6), (6, 7), (4, 7), (0, 4), (1, 5), (2, 6), (3, 7)]
SIG = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
LOC = np.array([[0, -.25, -.25], [.25, 0, -.25], [0, .25, -.25], [-.25, 0, -.25],
                [0, -.25, .25], [.25, 0, .25], [0, .25, .25], [-.25, 0, .25],
                [-.25, -.25, 0], [.25, -.25, 0], [.25, .25, 0], [-.25, .25, 0]])


def make(oracle, meshgen, n=5, perturb=0.0, seed=0, **kw):
    m = meshgen.cavity_hex(n, perturb=perturb, seed=seed)
    return oracle.Case(m, oracle.default_params(**kw), BC), m


def test_blending_factor_linear_field(oracle, meshgen):
    c, m = make(oracle, meshgen, n=6)
    x = m["xyz"]
    U = np.zeros((c.nn, 3))
    U[:, 0] = x[:, 0]
    c.field("U")[:] = U
    g = np.zeros((c.nn, 9))
    g[:, 0] = 1.0                     # exact dU/dx
    c.field("gradU")[:] = g
    c.update_blending("U")
    b = c.field("beta")[:, 0]
    inner = (x[:, 0] > 1e-9) & (x[:, 0] < 1 - 1e-9)
    np.testing.assert_allclose(b[inner], 1.0, rtol=0, atol=1e-15)
    np.testing.assert_allclose(b[~inner], 0.0, rtol=0, atol=1e-15)


def test_blending_factor_local_extremum(oracle, meshgen):
    c, m = make(oracle, meshgen, n=6)
    x = m["xyz"]
    i0 = int(np.argmin(np.linalg.norm(x - 0.5, axis=1)))
    U = np.zeros((c.nn, 3))
    U[:, 1] = 0.1 * x[:, 2]
    U[i0, 1] = 5.0                    # strict maximum of the y-component
    c.field("U")[:] = U
    g = np.zeros((c.nn, 9))
    g[:, 3 + 2] = 0.1
    g[i0, 3:6] = [1.0, -2.0, 0.5]
    c.field("gradU")[:] = g
    c.update_blending("U")
    assert c.field("beta")[i0, 1] == 0.0


def ip_points(x, conn):
    """Physical SCS ip locations at the standard ips, [E][12][3]."""
    N = np.prod(0.5 + SIG[None, :, :] * LOC[:, None, :], axis=2)   # [12][8]
    return np.einsum("pn,enj->epj", N, x[conn])


@pytest.mark.parametrize("perturb", [0.0, 0.2])
def test_deferred_correction_exact_for_linear_velocity(oracle, meshgen, perturb):
    c, m = make(oracle, meshgen, n=5, perturb=perturb, seed=3, mu=0.0)
    rng = np.random.default_rng(4)
    x, conn = m["xyz"], m["conn"]
    Aff = rng.normal(size=(3, 3))
    U = x @ Aff.T + np.array([0.3, -0.1, 0.2])
    c.field("U")[:] = U
    c.initialize()
    c.field("gradU")[:] = np.tile(Aff.ravel(), (c.nn, 1))
    md = rng.normal(scale=0.1, size=(c.ne, 12))
    c.field("mdot")[:] = md
    c.field("gradP")[:] = 0.0
    c.field("beta")[:] = 0.0
    c.assemble_momentum()
    b0 = c.rhs(3).copy()
    c.field("beta")[:] = 1.0
    c.assemble_momentum()
    b1 = c.rhs(3).copy()
    want = np.zeros((c.nn, 3))
    xip = ip_points(x, conn)
    for e in range(c.ne):
        for ip, (l, r) in enumerate(LR):
            L, R = conn[e][l], conn[e][r]
            up = L if md[e, ip] > 0 else R
            d = md[e, ip] * ((Aff @ xip[e, ip] + np.array([0.3, -0.1, 0.2])) - U[up])
            want[L] -= d
            want[R] += d
    # wall rows are relaxed by 0.75 after assembly
    bnode = np.any((x < 1e-9) | (x > 1 - 1e-9), axis=1)
    want[bnode] *= 0.75
    got = (b1 - b0).reshape(-1, 3)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
