"""Smoothed-aggregation AMG (OAGPU_BICGSTAB_SA) on the p' systems of a
wall-resolved backward-facing step (cells graded 1.1 from 4.8e-4 at the walls,
mu = 2e-5, k-omega SST, two oracle outer iterations from the inlet state):

* the prolongation equals P = (I - omega_p D_F^-1 A_F) T recomputed here from
  the device's fine matrix and aggregates (strong couplings |a_ij| >=
  0.08 sqrt(a_ii a_jj) kept, the others added to the diagonal), with
  omega_p = 4 / (3 lambda_max(D_F^-1 A_F)), and every coarse matrix equals
  the Galerkin product P^T A P of the level above (row-relative 1e-12);
* BiCGStab preconditioned by its V-cycle solves the p' system to the direct
  solution, in the FP64 and the mixed form, in fewer iterations than the
  unsmoothed V-cycle;
* outer iterations with it equal the oracle's (solved tightly).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import rel_err, row_rel_err

pytestmark = pytest.mark.gpu

MU, KIN = 2.0e-5, 1.5e-4
WIN = KIN / (10.0 * MU)
BC = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3, [0.0] * 3])
CTL, SCTL = [20, 1.0e-2, 200, 1.0e-2], [20, 1.0e-2]
STATE = ("U", "p", "mdot", "div", "gradU", "gradP", "mu", "k", "omega", "ymin", "mut", "F1", "Pk",
         "Pk_wall", "gradK", "gradW", "kprev", "msd")
THETA_FINE, THETA_COARSE = 0.08, 0.02


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def setup(mod, case, mesh, meshgen, device):
    case.set_bc_types([1, 2, 0, 3])       # inlet, outlet, no-slip walls, symmetry sides
    case.set_bc_values(p=[0.0] * 4, k=[KIN, 0, 0, 0], w=[WIN, 0, 0, 0])
    case.sst_enable(mod.default_sst(urf_k=0.9, urf_w=0.9))
    nn = mesh["xyz"].shape[0]
    y = meshgen.bfs_wall_distance(mesh["xyz"])
    w = np.full(nn, WIN)
    w[y < 1e-12] = 60.0 * MU / (0.075 * mesh["first_dy"] ** 2)
    U = np.zeros((nn, 3))
    U[:, 0] = 1.0
    for name, v in (("U", U), ("p", np.zeros(nn)), ("k", np.full(nn, KIN)), ("omega", w),
                    ("ymin", y)):
        if device:
            case.upload(name, v)
        else:
            case.field(name)[:] = v


@pytest.fixture(scope="module")
def developed(oracle, meshgen):
    m = meshgen.bfs_hex(3, 12, 96, 112, 3)
    assert abs(m["first_dy"] - 4.8e-4) < 0.01 * 4.8e-4
    print(f"[case] step {m['conn'].shape[0]} elements, first dy {m['first_dy']:.2e}", flush=True)
    c = oracle.Case(m, oracle.default_params(mu=MU, urf_U=0.7, urf_p=0.3, u_shifted=1), BC)
    setup(oracle, c, m, meshgen, False)
    c.initialize()
    for _ in range(2):
        c.outer_iteration(CTL)
        c.sst_iteration(SCTL)
    return c, m


def device_pcorr(oa, meshgen, c, m, precision):
    dm = oa.DeviceMesh(m)
    f = oa.FlowCase(dm, oa.default_params(precision, mu=MU, urf_U=0.7, urf_p=0.3, u_shifted=1), BC)
    setup(oa, f, m, meshgen, True)
    f.initialize()
    for name in STATE:
        f.upload(name, c.field(name))
    f.update_reversal()
    f.mass_divergence()
    f.assemble_momentum()
    f.assemble_pcorr()
    return dm, f


def csr(rp, cols, vals, n, nc=None):
    return sp.csr_matrix((vals, cols, rp), shape=(n, nc if nc is not None else n))


def filtered_prolongation(A, agg, nc, theta, omega):
    """P = (I - omega D_F^-1 A_F) T from the definition (dense rows)."""
    A = A.tocsr()
    n = A.shape[0]
    d = A.diagonal()
    C = A.tocoo()
    off = C.row != C.col
    strong = off & (np.abs(C.data) >= theta * np.sqrt(np.abs(d[C.row] * d[C.col])))
    weak = off & ~strong
    dF = d + np.bincount(C.row[weak], weights=C.data[weak], minlength=n)
    dF = np.where(dF > 0, dF, d)
    rows = np.concatenate([C.row[strong], np.arange(n)])
    cols = np.concatenate([C.col[strong], np.arange(n)])
    vals = np.concatenate([C.data[strong], dF])
    AF = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    T = sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, nc))
    P = T - sp.diags(omega / dF) @ (AF @ T)
    lam = abs(spla.eigs(sp.diags(1.0 / dF) @ AF, k=1, which="LM", return_eigenvectors=False,
                        tol=1e-6)[0])
    return P.tocsr(), lam


def test_sa_hierarchy_definition(oa, meshgen, developed):
    c, m = developed
    dm, f = device_pcorr(oa, meshgen, c, m, oa.OAGPU_FP64)
    vals, b = f.system(1)
    rp, cols, _ = dm.graph()
    A = csr(rp, cols, vals.ravel(), dm.nn)
    cd = f.solve(1, oa.OAGPU_BICGSTAB_SA, 200, 1e-8)
    assert cd.n_iterations > 0
    levels = f.sa_levels()
    print(f"[sa] levels {[(l[0], l[2]) for l in levels]}", flush=True)
    assert len(levels) >= 3 and levels[-1][0] <= 64
    Af = A
    for lev in range(1, len(levels)):
        agg, (prp, pc, pv), (arp, ac, av) = f.sa_level(lev)
        nf, nc = levels[lev - 1][0], levels[lev][0]
        assert agg.min() == 0 and agg.max() == nc - 1 and len(np.unique(agg)) == nc
        theta = THETA_FINE if lev == 1 else THETA_COARSE
        Pref, lam = filtered_prolongation(Af, agg, nc, theta, levels[lev - 1][3])
        # omega_p from a power iteration: within a few per cent of 4 / (3 lambda)
        assert abs(levels[lev - 1][3] * lam * 0.75 - 1.0) < 0.05, (lev, lam, levels[lev - 1][3])
        P = csr(prp, pc, pv, nf, nc)
        assert rel_err(P.toarray(), Pref.toarray()) < 1e-12, lev
        Ac = csr(arp, ac, av, nc)
        G = (P.T @ Af @ P).tocsr()
        G.sort_indices()
        Gd = G.toarray()
        err = np.abs(Ac.toarray() - Gd).max(axis=1) / np.maximum(np.abs(Gd).max(axis=1), 1e-300)
        print(f"[sa] level {lev}: {nc} rows, Galerkin row-relative {err.max():.2e}", flush=True)
        assert err.max() < 1e-12, lev
        Af = Ac
    f.close()


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_sa_solve_matches_direct(oa, meshgen, developed, precision):
    c, m = developed
    dm, f = device_pcorr(oa, meshgen, c, m, precision)
    vals, b = f.system(1)
    rp, cols, _ = dm.graph()
    A = csr(rp, cols, vals.ravel(), dm.nn).tocsc()
    x_ref = spla.spsolve(A, b.ravel())
    cd = f.solve(1, oa.OAGPU_BICGSTAB_SA, 400, 1e-10)
    x = f.solution(1).ravel()
    err = rel_err(x, x_ref)
    print(f"[sa] precision {precision}: {cd.n_iterations} iterations, error {err:.2e}", flush=True)
    assert err < 1e-7
    assert cd.solver_final_res <= 1e-10 * cd.solver_initial_res * 1.0001
    # a 1e-2 reduction: fewer iterations than the unsmoothed V-cycle
    f.set_system(1, vals, b)
    it_sa = f.solve(1, oa.OAGPU_BICGSTAB_SA, 400, 1e-2).n_iterations
    f.set_system(1, vals, b)
    it_amg = f.solve(1, oa.OAGPU_BICGSTAB_AMG, 400, 1e-2).n_iterations
    print(f"[sa] to 1e-2: smoothed {it_sa}, unsmoothed {it_amg}", flush=True)
    assert it_sa < it_amg
    f.close()


def test_sa_outer_iterations_match_oracle(oracle, oa, meshgen):
    """Flow + SST outer iterations with the p' systems solved tightly by
    BiCGStab-SA on the device and by the oracle's BiCGStab: the same state."""
    m = meshgen.bfs_hex(3, 8, 6, 6, 3)
    print(f"[case] step {m['conn'].shape[0]} elements", flush=True)
    kw = dict(mu=MU, urf_U=0.7, urf_p=0.3, u_shifted=1)
    c = oracle.Case(m, oracle.default_params(p_method=2, **kw), BC)
    setup(oracle, c, m, meshgen, False)
    c.initialize()
    dm = oa.DeviceMesh(m)
    f = oa.FlowCase(dm, oa.default_params(oa.OAGPU_FP64, p_method=oa.OAGPU_BICGSTAB_SA, **kw), BC)
    setup(oa, f, m, meshgen, True)
    f.initialize()
    ctl, sctl = [60, 1e-8, 1000, 1e-10], [60, 1e-10]
    for it in range(3):
        # the spanwise velocity's residual is rounding-sized (1e-8 of the
        # others): absolute agreement to 1e-12 there
        ho, hg = c.outer_iteration(ctl), f.outer_iteration(ctl)
        np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8, atol=1e-12)
        so, sg = c.sst_iteration(sctl), f.sst_iteration(sctl)
        np.testing.assert_allclose(sg[:2], so[:2], rtol=1e-8, atol=1e-12)
    for name in ("U", "p", "k", "omega"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name
    f.close()


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_cg_sa_cavity(oracle, oa, meshgen, precision):
    """CG with the smoothed-aggregation V-cycle on a symmetric p' system (a
    cavity): the direct solution, in fewer iterations than Jacobi-CG, and
    graph replays of its iteration blocks equal to eager launches bitwise."""
    m = meshgen.cavity_hex(16)
    print(f"[case] cavity 16^3 CG-SA precision {precision}", flush=True)
    BCc = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    out = []
    for graphs in (1, 0):
        dm = oa.DeviceMesh(m)
        f = oa.FlowCase(dm, oa.default_params(precision, use_graphs=graphs), BCc)
        f.upload("U", meshgen.initial_velocity(dm.nn, 0))
        f.initialize()
        f.outer_iteration([20, 0.1, 200, 1e-6])
        f.assemble_momentum()
        f.solve(3, oa.OAGPU_GMRES, 20, 0.1)
        f.correct(3, 1.0)
        f.update_mdot()
        f.assemble_pcorr()
        vals, b = f.system(1)
        rp, cols, _ = dm.graph()
        A = csr(rp, cols, vals.ravel(), dm.nn).tocsc()
        x_ref = spla.spsolve(A, b.ravel())
        cd = f.solve(1, oa.OAGPU_CG_SA, 400, 1e-10)
        x = f.solution(1).ravel()
        err = rel_err(x, x_ref)
        f.set_system(1, vals, b)
        it_j = f.solve(1, oa.OAGPU_CG, 2000, 1e-10).n_iterations
        print(f"[sa] graphs {graphs}: CG-SA {cd.n_iterations} iterations, error {err:.2e}; "
              f"Jacobi-CG {it_j}", flush=True)
        assert err < 1e-7 and cd.n_iterations < it_j
        out.append(x)
        f.close()
    np.testing.assert_array_equal(out[0], out[1])
