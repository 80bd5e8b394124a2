"""The oracle's transient BDF1 / BDF2 node terms (navierStokesAssemblerNodeTerms.cpp:
151-400, phiAssemblerNodeTerms.hpp:150-404, coefficients types.cpp:392-411),
pinned by closed forms:

* a spatially uniform velocity with no mass flow, pressure gradient or
  boundary faces leaves only the time term: b = -rho V dU/dt at every node,
  and BDF2 with a variable step reproduces dU/dt of a quadratic in time
  exactly (BDF1 that of a linear one);
* BDF1 with U_old = U assembles the steady system whose pseudo time step is
  the physical one;
* the scalar (k / omega) node terms alone give b = -rho V dphi/dt likewise.
"""
import numpy as np
import pytest

BC = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])


def no_faces(mesh):
    m = dict(mesh)
    for k in ("bf_elem", "bf_ord", "bf_bc"):
        m[k] = np.zeros(0, dtype=np.int32)
    return m


def uniform_case(oracle, meshgen, scheme, dts, values, rho=1.3, sst=False):
    """Uniform U (and k, omega) set to values[i] before time step i."""
    m = no_faces(meshgen.cavity_hex(3, perturb=0.15, seed=2))
    c = oracle.Case(m, oracle.default_params(rho=rho, p_ref_node=-1), BC)
    if sst:
        c.sst_enable(oracle.default_sst())
    c.field("rho")[:] = rho
    for dt, v in zip(dts, values):
        c.field("U")[:] = v
        if sst:
            c.field("k")[:] = v[0]
            c.field("omega")[:] = v[1]
        c.time_step(scheme, dt)
    return c, m


def uniform_state(c):
    c.field("p")[:] = 0.0
    c.initialize()
    c.field("mdot")[:] = 0.0
    c.field("div")[:] = 0.0
    c.field("gradP")[:] = 0.0


@pytest.mark.parametrize("scheme", [1, 2])
def test_momentum_time_term_closed_form(oracle, meshgen, scheme):
    rho = 1.3
    # U(t) = a + b t + q t^2 componentwise, sampled at t0 < t1 < t2 (steps 0.3, 0.2)
    a, b, q = np.array([0.2, -0.1, 0.4]), np.array([1.0, 0.5, -2.0]), np.array([3.0, -1.0, 0.5])
    t = np.array([0.0, 0.3, 0.5])
    U = [a + b * s + q * s * s for s in t]
    # time steps: first step (t0 -> t1) then (t1 -> t2); before each step U is
    # the current level, after the second step the new level is U(t2)
    c, m = uniform_case(oracle, meshgen, scheme, [0.3, 0.2], [U[0], U[1]], rho)
    c.field("U")[:] = U[2]
    uniform_state(c)
    c.assemble_momentum()
    vol = c.field("vol")
    b_ = c.rhs(3).reshape(-1, 3)
    if scheme == 2:
        dudt = b + 2 * q * t[2]                       # exact for a quadratic
    else:
        dudt = (U[2] - U[1]) / 0.2                    # backward difference
    want = -rho * vol[:, None] * dudt[None, :]
    assert np.abs(b_ - want).max() < 1e-12 * np.abs(want).max()
    # the diagonal: c0 rho V / dt (viscous part zero for the uniform field is
    # not true of the matrix; compare against the steady matrix instead)
    om = 0.2 / 0.3
    c0 = (1 + 2 * om) / (1 + om) if scheme == 2 else 1.0
    c2, m2 = uniform_case(oracle, meshgen, 1, [0.2 / c0], [U[2]], rho)   # BDF1, dt / c0
    c2.field("U")[:] = U[2]
    uniform_state(c2)
    c2.assemble_momentum()
    np.testing.assert_allclose(c2.vals(3), c.vals(3), rtol=1e-13, atol=1e-15 * np.abs(c.vals(3)).max())


def test_bdf1_with_equal_levels_is_the_steady_system(oracle, meshgen):
    m = meshgen.cavity_hex(4, perturb=0.2, seed=5)
    rng = np.random.default_rng(3)
    U = rng.normal(scale=0.3, size=(m["xyz"].shape[0], 3))
    out = []
    for transient in (False, True):
        c = oracle.Case(m, oracle.default_params(dt=0.37), BC)
        c.field("U")[:] = U
        c.initialize()
        if transient:
            c.time_step(1, 0.37)          # U_old = U
        c.assemble_momentum()
        out.append((c.vals(3).copy(), c.rhs(3).copy()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.abs(out[0][1] - out[1][1]).max() <= 1e-15 * np.abs(out[0][1]).max()


@pytest.mark.parametrize("scheme", [1, 2])
def test_scalar_time_term_closed_form(oracle, meshgen, scheme):
    rho = 0.8
    t = np.array([0.0, 0.25, 0.4])
    kq = lambda s: 1e-2 + 3e-2 * s - 4e-2 * s * s
    wq = lambda s: 5.0 - 2.0 * s + 6.0 * s * s
    vals = [np.array([kq(s), wq(s)]) for s in t]
    c, m = uniform_case(oracle, meshgen, scheme, [0.25, 0.15],
                        [np.array([0.1, 0.2, 0.3]), np.array([0.1, 0.2, 0.3])], rho, sst=True)
    # k / omega of the earlier levels, then the current level
    c.field("kold2")[:] = kq(t[0]) if scheme == 2 else 0.0
    c.field("kold")[:] = kq(t[1])
    c.field("wold2")[:] = wq(t[0]) if scheme == 2 else 0.0
    c.field("wold")[:] = wq(t[1])
    c.field("k")[:] = kq(t[2])
    c.field("omega")[:] = wq(t[2])
    uniform_state(c)
    vol = c.field("vol")
    for which, fn, dq in (("k", kq, lambda s: 3e-2 - 8e-2 * s), ("omega", wq, lambda s: -2.0 + 12.0 * s)):
        c.assemble_scalar(which, terms=1)          # node terms only
        dphidt = dq(t[2]) if scheme == 2 else (fn(t[2]) - fn(t[1])) / 0.15
        want = -rho * vol * dphidt
        assert np.abs(c.rhs(1) - want).max() < 1e-12 * np.abs(want).max(), which
