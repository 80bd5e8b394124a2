"""Pins the oracle's k-omega SST restatement (oracle/sst.c) with properties
the model and the discretisation fix (the reference ships no tests or golden
vectors for these):

* eddy viscosity: zero strain gives mu_t = rho k / omega; a strained near-wall
  node (F2 = 1) gives a1 rho k / S; mu_eff = mu + mu_t;
* F1 = 1 next to a wall and -> 0 far from it;
* production of a simple shear du/dy = g is mu_t g^2 (interior nodes), the
  supplied wall value at no-slip wall nodes;
* interior element terms are in residual form: A phi = -b for any phi (full
  stencil; the reduced one lumps absent columns into the diagonal);
* diffusion with a constant diffusivity is exact for a linear field on an
  affine mesh (b = 0 at interior nodes with no advection);
* upwind advection is conservative: row sums of A equal the mass divergence
  and the rhs telescopes to zero over the mesh;
* the k production limiter (10 D_k) and the equilibrium P_k = D_k;
* relaxation divides the diagonal by urf; omega rows at no-slip wall nodes
  keep only their diagonal and a zero rhs.

The wall treatment (omega at wall nodes, production at wall nodes, the wall
distance) is an input here, not restated: those fields are supplied.
"""
import numpy as np
import pytest
import scipy.sparse as sp

SMALL = np.finfo(np.float64).eps
BC = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])


def make_case(oracle, meshgen, n=5, perturb=0.0, seed=0, **prm):
    m = meshgen.cavity_hex(n, perturb=perturb, seed=seed)
    c = oracle.Case(m, oracle.default_params(**prm), BC)
    c.sst_enable(oracle.default_sst())
    return c, m


def state(c, rng):
    c.field("U")[:] = rng.normal(scale=0.3, size=(c.nn, 3))
    c.initialize()
    c.field("k")[:] = rng.uniform(1e-4, 1e-2, c.nn)
    c.field("omega")[:] = rng.uniform(0.5, 5.0, c.nn)
    c.field("ymin")[:] = rng.uniform(0.01, 0.3, c.nn)
    c.field("gradU")[:] = rng.normal(scale=1.0, size=(c.nn, 9))
    c.field("gradK")[:] = rng.normal(scale=1e-2, size=(c.nn, 3))
    c.field("gradW")[:] = rng.normal(scale=1.0, size=(c.nn, 3))
    c.field("mut")[:] = rng.uniform(0, 1e-2, c.nn)
    c.field("F1")[:] = rng.uniform(0, 1, c.nn)


def matrix(c):
    return sp.csr_matrix((c.vals(1).ravel().copy(), c.cols.copy(), c.row_ptr.copy()),
                         shape=(c.nn, c.nn))


def test_eddy_viscosity_closed_forms(oracle, meshgen):
    c, m = make_case(oracle, meshgen)
    state(c, np.random.default_rng(1))
    c.field("gradU")[:] = 0.0
    c.sst_update_mut()
    k, w = c.field("k"), c.field("omega")
    np.testing.assert_allclose(c.field("mut"), k / w, rtol=1e-13)
    np.testing.assert_allclose(c.field("mu"), 0.01 + k / w, rtol=1e-13)
    # near-wall node with a strong strain: F2 = 1, mu_t = a1 rho k / S
    g = np.zeros(9)
    g[1] = 50.0                      # du/dy
    c.field("gradU")[:] = g
    c.field("ymin")[:] = 1e-6
    c.sst_update_mut()
    S = 50.0                         # sqrt(2 S_ij S_ij) of a simple shear
    np.testing.assert_allclose(c.field("mut"), 0.31 * k / S, rtol=1e-12)


def test_F1_limits(oracle, meshgen):
    c, m = make_case(oracle, meshgen)
    state(c, np.random.default_rng(2))
    c.field("ymin")[:] = 1e-7
    c.sst_update_F1()
    np.testing.assert_allclose(c.field("F1"), 1.0, rtol=0, atol=1e-15)
    c.field("ymin")[:] = 1e3
    c.field("k")[:] = 1e-6
    c.field("omega")[:] = 1.0
    c.sst_update_F1()
    assert np.abs(c.field("F1")).max() < 1e-12


def test_production_simple_shear(oracle, meshgen):
    c, m = make_case(oracle, meshgen)
    rng = np.random.default_rng(3)
    state(c, rng)
    g = np.zeros(9)
    g[1] = 2.5
    c.field("gradU")[:] = g
    c.field("Pk_wall")[:] = rng.uniform(0, 1, c.nn)
    c.sst_update_Pk()
    bnode = np.zeros(c.nn, bool)
    sn = np.array([[0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [0, 4, 7, 3], [0, 3, 2, 1], [4, 5, 6, 7]])
    for e, o in zip(m["bf_elem"], m["bf_ord"]):
        bnode[m["conn"][e][sn[o]]] = True
    Pk = c.field("Pk")
    np.testing.assert_allclose(Pk[~bnode], c.field("mut")[~bnode] * 2.5 ** 2, rtol=1e-14)
    np.testing.assert_array_equal(Pk[bnode], c.field("Pk_wall")[bnode])


@pytest.mark.parametrize("which", ["k", "omega"])
def test_element_terms_residual_form(oracle, meshgen, which):
    # full stencil: the reduced one lumps absent columns into the diagonal
    c, m = make_case(oracle, meshgen, n=5, perturb=0.2, seed=4, reduced=0)
    state(c, np.random.default_rng(5))
    c.field("mdot")[:] = np.random.default_rng(6).normal(scale=0.05, size=(c.ne, 12))
    c.assemble_scalar(which, terms=4)
    A = matrix(c)
    phi = c.field(which)
    b = c.rhs(1)
    assert np.abs(b).max() > 0
    np.testing.assert_allclose(A @ phi, -b, rtol=0, atol=1e-12 * np.abs(b).max())


def test_diffusion_exact_for_linear_fields(oracle, meshgen):
    c, m = make_case(oracle, meshgen, n=6)
    state(c, np.random.default_rng(7))
    c.field("mdot")[:] = 0.0
    c.field("mut")[:] = 0.0           # constant Gamma = mu
    x = m["xyz"]
    c.field("k")[:] = 0.3 + 1.5 * x[:, 0] - 0.7 * x[:, 1] + 2.0 * x[:, 2]
    c.assemble_scalar("k", terms=4)
    interior = np.all((x > 1e-9) & (x < 1 - 1e-9), axis=1)
    b = c.rhs(1)
    A = matrix(c)
    scale = np.abs(A).max()
    assert scale > 0
    assert np.abs(b[interior]).max() < 1e-12 * scale
    # and a non-linear field is not reproduced
    c.field("k")[:] = x[:, 0] ** 2
    c.assemble_scalar("k", terms=4)
    assert np.abs(c.rhs(1)[interior]).max() > 1e-6 * scale


def test_advection_conservative(oracle, meshgen):
    c, m = make_case(oracle, meshgen, n=5, perturb=0.15, seed=8, mu=0.0)
    state(c, np.random.default_rng(9))
    c.field("mut")[:] = 0.0           # Gamma = 0: advection only
    c.field("mdot")[:] = np.random.default_rng(10).normal(size=(c.ne, 12))
    c.mass_divergence()
    c.assemble_scalar("omega", terms=4)
    A = matrix(c)
    np.testing.assert_allclose(np.asarray(A.sum(axis=1)).ravel(), c.field("div"), rtol=0,
                               atol=1e-12 * np.abs(c.field("div")).max())
    assert abs(c.rhs(1).sum()) < 1e-12 * np.abs(c.rhs(1)).max() * c.nn


def test_k_source_limiter_and_equilibrium(oracle, meshgen):
    c, m = make_case(oracle, meshgen)
    state(c, np.random.default_rng(11))
    k, w, vol = c.field("k"), c.field("omega"), c.field("vol")
    Dk = 0.09 * 1.0 * w * k
    c.field("Pk")[:] = Dk                        # equilibrium
    c.assemble_scalar("k", terms=2)
    assert np.abs(c.rhs(1)).max() < 1e-15 * np.abs(Dk * vol).max()
    diag = c.vals(1).ravel()[c.row_ptr[:-1] + c.diag_off]
    np.testing.assert_allclose(diag, 0.09 * w * vol, rtol=1e-14)
    c.field("Pk")[:] = 100.0 * Dk                # limited to 10 D_k
    c.assemble_scalar("k", terms=2)
    np.testing.assert_allclose(c.rhs(1), 9.0 * Dk * vol, rtol=1e-13)


def test_relaxation_and_omega_wall_rows(oracle, meshgen):
    c, m = make_case(oracle, meshgen, n=5, perturb=0.1, seed=12)
    state(c, np.random.default_rng(13))
    c.assemble_scalar("omega", terms=7)
    v7, b7 = c.vals(1).ravel().copy(), c.rhs(1).copy()
    c.assemble_scalar("omega", terms=15)
    v15, b15 = c.vals(1).ravel(), c.rhs(1)
    rp, doff = c.row_ptr, c.diag_off
    x = m["xyz"]
    wall = np.any((x < 1e-9) | (x > 1 - 1e-9), axis=1)
    for i in range(c.nn):
        d = rp[i] + doff[i]
        assert v15[d] == pytest.approx(v7[d] / 0.8, rel=1e-15)
        if wall[i]:
            row = v15[rp[i]:rp[i + 1]].copy()
            row[doff[i]] = 0.0
            assert np.all(row == 0.0) and b15[i] == 0.0
        else:
            assert b15[i] == b7[i]


def test_sst_iterations_stay_physical(oracle, meshgen):
    c, m = make_case(oracle, meshgen, n=6, perturb=0.1, seed=14)
    c.field("U")[:] = meshgen.initial_velocity(c.nn, 0)
    c.initialize()
    c.field("k")[:] = 1e-3
    c.field("omega")[:] = 1.0
    x = m["xyz"]
    c.field("ymin")[:] = np.maximum(np.minimum(x, 1 - x).min(axis=1), 1e-3)
    for _ in range(3):
        c.outer_iteration([20, 0.1, 200, 1e-6])
        h = c.sst_iteration([50, 1e-6])
        assert np.all(np.isfinite(h))
    assert c.field("k").min() >= SMALL and c.field("omega").min() >= SMALL
    assert np.all(np.isfinite(c.field("mut"))) and c.field("mut").min() >= 0
