"""Pins the oracle's graph, scatter, assembly and field-update restatements
with properties the mathematics fixes (the reference ships no tests or golden
vectors for these):

* CRS graph: closed-form nnz of the reduced (7-point) and full (27-point)
  stencils on n^3 grids, sorted columns, diagonal present;
* applyCoeff_: reduced-stencil lumping moves coefficients into the diagonal
  only, so block row sums equal those of the full-stencil matrix;
* momentum: residual of a linear velocity field (constant stress) vanishes at
  interior nodes; p' Laplacian rows sum to zero;
* Rhie-Chow mass flow: uniform U with a linear pressure and its exact
  gradient gives mdot = rho U . A; mass divergence of uniform flow is zero at
  interior nodes;
* Green-Gauss: exact for linear fields at interior nodes, relaxation
  (1 - (1-g)^k) after k updates;
* Krylov: CG / GMRES / BiCGStab agree with a direct sparse solve.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import rel_err


def make_case(oracle, meshgen, n=5, perturb=0.0, seed=0, **kw):
    m = meshgen.cavity_hex(n, perturb=perturb, seed=seed)
    prm = oracle.default_params(**kw)
    bc = np.array([[1.0, 0, 0], [0, 0, 0]])
    return oracle.Case(m, prm, bc), m


def interior(m, n):
    ijk = np.stack(np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1),
                               indexing="ij"), axis=-1).reshape(-1, 3)
    return np.all((ijk > 0) & (ijk < n), axis=1)


@pytest.mark.parametrize("n", [1, 2, 5])
def test_graph_closed_form(oracle, meshgen, n):
    m = meshgen.cavity_hex(n)
    L = oracle.lib()
    V = (n + 1) ** 3
    for reduced, expect in ((1, V + 6 * n * (n + 1) ** 2), (0, (3 * n + 1) ** 3)):
        rp = np.zeros(V + 1, np.int32)
        cols = oracle._ip()
        assert L.or_graph_build(V, n ** 3, oracle.iptr(m["conn"]), None, reduced, oracle.iptr(rp),
                                oracle.C.byref(cols)) == 0
        nnz = rp[-1]
        c = np.ctypeslib.as_array(cols, shape=(nnz,)).copy()
        L.or_free(cols)
        assert nnz == expect
        for i in range(V):
            row = c[rp[i]:rp[i + 1]]
            assert np.all(np.diff(row) > 0) and i in row


def test_apply_coeff_lumping_preserves_row_sums(oracle, meshgen):
    rng = np.random.default_rng(0)
    m = meshgen.cavity_hex(3)
    L = oracle.lib()
    V, E = (4 ** 3), 27
    mats = {}
    for reduced in (1, 0):
        rp = np.zeros(V + 1, np.int32)
        cp = oracle._ip()
        L.or_graph_build(V, E, oracle.iptr(m["conn"]), None, reduced, oracle.iptr(rp), oracle.C.byref(cp))
        cols = np.ctypeslib.as_array(cp, shape=(rp[-1],)).copy()
        L.or_free(cp)
        doff = np.array([np.searchsorted(cols[rp[i]:rp[i + 1]], i) for i in range(V)], np.int32)
        vals = np.zeros(rp[-1] * 9)
        b = np.zeros(V * 3)
        r2 = np.random.default_rng(1)
        for e in range(E):
            lhs = np.ascontiguousarray(r2.normal(size=24 * 24))
            rhs = np.ascontiguousarray(r2.normal(size=24))
            nodes = np.ascontiguousarray(m["conn"][e])
            L.or_apply_coeff(3, 8, oracle.iptr(nodes), oracle.dptr(lhs), oracle.dptr(rhs), V,
                             oracle.iptr(rp), oracle.iptr(cols), oracle.iptr(doff),
                             oracle.dptr(vals), oracle.dptr(b))
        blk = vals.reshape(-1, 3, 3)
        rs = np.array([blk[rp[i]:rp[i + 1]].sum(axis=0) for i in range(V)])
        mats[reduced] = (rs, b.copy(), rp, cols, blk)
    np.testing.assert_allclose(mats[1][0], mats[0][0], rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(mats[1][1], mats[0][1])
    # off-diagonal reduced entries equal the full-stencil ones at the same column
    rs, _, rp1, c1, b1 = mats[1]
    _, _, rp0, c0, b0 = mats[0]
    for i in range(V):
        for k in range(rp1[i], rp1[i + 1]):
            if c1[k] == i:
                continue
            k0 = rp0[i] + np.searchsorted(c0[rp0[i]:rp0[i + 1]], c1[k])
            np.testing.assert_array_equal(b1[k], b0[k0])


def test_momentum_residual_of_linear_field(oracle, meshgen):
    n = 5
    c, m = make_case(oracle, meshgen, n=n, perturb=0.15, seed=3)
    G = np.array([[0.3, -0.2, 0.5], [0.1, 0.4, -0.3], [-0.2, 0.25, 0.1]])
    c.field("U")[:] = m["xyz"] @ G.T
    c.field("mdot")[:] = 0.0
    c.field("gradP")[:] = 0.0
    c.field("div")[:] = 0.0
    c.assemble_momentum()
    b = c.rhs(3).reshape(-1, 3)
    inn = interior(m, n)
    assert np.abs(b[inn]).max() < 1e-14 * max(1.0, np.abs(b).max())
    assert np.abs(b[~inn]).max() > 1e-6      # walls see the stress


def test_reduced_and_full_stencil_row_sums(oracle, meshgen):
    rng = np.random.default_rng(2)
    res = {}
    for red in (1, 0):
        c, m = make_case(oracle, meshgen, n=4, perturb=0.1, seed=5, reduced=red, p_ref_node=-1,
                         urf_U=1.0)  # relaxation scales the lumped diagonal
        if red:
            res["U"] = rng.normal(size=(c.nn, 3))
            res["du"] = rng.uniform(0.5, 1.5, size=(c.nn, 3))
        c.field("U")[:] = res["U"]
        c.initialize()
        c.assemble_momentum()
        c.field("du")[:] = res["du"]   # du follows the (stencil dependent) diagonal
        c.assemble_pcorr()
        rp = c.row_ptr
        A3 = c.vals(3)
        A1 = c.vals(1)
        res[red] = (np.array([A3[rp[i]:rp[i + 1]].sum(axis=0) for i in range(c.nn)]),
                    np.array([A1[rp[i]:rp[i + 1]].sum() for i in range(c.nn)]),
                    c.rhs(3).copy(), c.rhs(1).copy())
        res["scale", red] = (np.abs(A3).max(), np.abs(A1).max())
    for k in range(4):
        scale = max(np.abs(res[0][k]).max(), res["scale", 0][k % 2] if k < 2 else 0.0)
        assert np.abs(res[1][k] - res[0][k]).max() <= 1e-12 * scale


def test_pcorr_rows_sum_to_zero(oracle, meshgen):
    rng = np.random.default_rng(4)
    c, m = make_case(oracle, meshgen, n=4, perturb=0.1, seed=1)
    c.field("U")[:] = rng.normal(size=(c.nn, 3))
    c.initialize()
    c.field("du")[:] = rng.uniform(0.5, 1.5, size=(c.nn, 3))
    c.assemble_pcorr()
    rp = c.row_ptr
    A = c.vals(1).ravel()
    rs = np.array([A[rp[i]:rp[i + 1]].sum() for i in range(c.nn)])
    ref = c.params.p_ref_node
    mask = np.ones(c.nn, bool)
    mask[ref] = False
    assert np.abs(rs[mask]).max() < 1e-13 * np.abs(A).max()


def test_rhie_chow_uniform_flow_and_divergence(oracle, meshgen):
    n = 4
    c, m = make_case(oracle, meshgen, n=n, perturb=0.15, seed=2)
    U0 = np.array([0.7, -0.3, 0.2])
    g = np.array([1.5, -0.5, 2.0])
    c.field("U")[:] = U0
    c.field("p")[:] = m["xyz"] @ g
    c.field("gradP")[:] = g
    c.field("du")[:] = np.random.default_rng(0).uniform(0.1, 1.0, size=(c.nn, 3))
    c.field("mdot")[:] = 1.0
    c.update_mdot()
    md = c.field("mdot")
    L = oracle.lib()
    for e in range(c.ne):
        co = np.ascontiguousarray(m["xyz"][m["conn"][e]])
        a = np.zeros((12, 3))
        L.or_hex_scs_areav(oracle.dptr(co), oracle.dptr(a))
        np.testing.assert_allclose(md[e], a @ U0, rtol=1e-12, atol=1e-15)
    c.mass_divergence()
    div = c.field("div")
    inn = interior(m, n)
    assert np.abs(div[inn]).max() < 1e-14


@pytest.mark.parametrize("which,shifted_key", [("U", "u_shifted"), ("p", "p_shifted")])
@pytest.mark.parametrize("shifted", [0, 1])
def test_green_gauss_linear_exact(oracle, meshgen, which, shifted_key, shifted):
    n = 5
    c, m = make_case(oracle, meshgen, n=n, **{shifted_key: shifted})
    gval = np.array([[0.3, -1.1, 0.7], [2.0, 0.5, -0.4], [-0.6, 0.2, 1.3]])
    if which == "U":
        c.field("U")[:] = m["xyz"] @ gval.T + 1.0
        grad = c.field("gradU")
        exact = gval.ravel()
    else:
        c.field("p")[:] = m["xyz"] @ gval[0] + 3.0
        grad = c.field("gradP")
        exact = gval[0]
    grad[:] = 0.0
    g = c.params.grad_urf
    inn = interior(m, n)
    for k in range(1, 4):
        c.gradient(which)
        np.testing.assert_allclose(grad[inn], np.tile(exact * (1 - (1 - g) ** k), (inn.sum(), 1)),
                                   rtol=1e-12, atol=1e-13)


def _scipy_system(c, N):
    rp, cols = c.row_ptr, c.cols
    A = c.vals(N)
    if N == 1:
        M = sp.csr_matrix((A.ravel().copy(), cols.copy(), rp.copy()), shape=(c.nn, c.nn))
    else:
        M = sp.bsr_matrix((A.copy(), cols.copy(), rp.copy()), shape=(3 * c.nn, 3 * c.nn)).tocsr()
    return M, c.rhs(N).copy()


@pytest.mark.parametrize("N,kind", [(3, 0), (3, 2), (1, 1), (1, 0), (1, 2)])
def test_krylov_against_direct(oracle, meshgen, N, kind):
    rng = np.random.default_rng(11)
    c, m = make_case(oracle, meshgen, n=4, perturb=0.1, seed=9)
    c.field("U")[:] = rng.normal(scale=0.1, size=(c.nn, 3))
    c.initialize()
    c.assemble_momentum()
    c.compute_du()
    if N == 1:
        c.field("U")[:] += rng.normal(scale=0.05, size=(c.nn, 3))
        c.assemble_pcorr()
    M, b = _scipy_system(c, N)
    xd = spla.spsolve(M.tocsc(), b)
    c.x(N)[:] = 0.0
    it, res = c.solve(N, kind, 2000, 1e-12, restart=30)
    assert res[1] <= 1e-12 * res[0] * 1.0001
    assert rel_err(c.x(N), xd) < 1e-8


def test_outer_iterations_reduce_residuals(oracle, meshgen):
    c, m = make_case(oracle, meshgen, n=6)
    c.field("U")[:] = meshgen.initial_velocity(c.nn, 0)
    c.initialize()
    hist = np.array([c.outer_iteration([20, 0.1, 200, 1e-6]) for _ in range(12)])
    assert np.all(np.isfinite(hist))
    assert hist[-1, 0] < 0.25 * hist[1, 0]
    assert hist[-1, 3] < 0.25 * hist[1, 3]
