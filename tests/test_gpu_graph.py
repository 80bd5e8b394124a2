"""The whole outer iteration as one HIP graph (oagpu_graph_iterations).

* Graph replays are bitwise equal to the same device-controlled launch
  sequence run without a graph (eager=True): fields after several
  iterations, the residual history and the Krylov iteration counts -- on a
  wall-resolved step slice with k-omega SST (smoothed-aggregation
  AMG-BiCGStab p', FP64 and mixed) and on a perturbed cavity (smoothed-
  aggregation AMG-CG p', FP64 and mixed); replays run with the program's own
  synchronisations and downloads between them.
* An iteration whose solves need more slots than the graph has is rolled
  back and run uncapped: forced to one slot per solve, the results are still
  bitwise those of the eager run, and the flow reports the roll-backs.
* The device-controlled solves agree with the host-controlled ones of
  oagpu_outer_iteration / oagpu_sst_iteration (same algorithms; the Givens
  rotations and scalars are computed on the device instead of the host):
  equal Krylov iteration counts and fields within 1e-9 of each other after
  three iterations (the solves' tolerances are loose, 1e-2, so roundoff in
  the Krylov scalars only moves the fields at that level).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MU, KIN = 2.0e-5, 1.5e-4
WIN = KIN / (10.0 * MU)
BFS_BC = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3, [0.0] * 3])
CTL, SCTL = [20, 1.0e-2, 200, 1.0e-2], [20, 1.0e-2]
FIELDS = ("U", "p", "mdot", "k", "omega")


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


@pytest.fixture(scope="module")
def step_mesh(meshgen):
    m = meshgen.bfs_hex(3, 12, 24, 28, 2)
    print(f"[case] step slice {m['conn'].shape[0]} elements", flush=True)
    return m


def step_flow(oa, meshgen, m, precision):
    dm = oa.DeviceMesh(m)
    f = oa.FlowCase(dm, oa.default_params(precision, mu=MU, urf_U=0.7, urf_p=0.3, u_shifted=1,
                                          p_method=oa.OAGPU_BICGSTAB_SA), BFS_BC)
    f.set_bc_types([1, 2, 0, 3])
    f.set_bc_values(p=[0.0] * 4, k=[KIN, 0, 0, 0], w=[WIN, 0, 0, 0])
    f.sst_enable(oa.default_sst(urf_k=0.9, urf_w=0.9))
    nn = m["xyz"].shape[0]
    y = meshgen.bfs_wall_distance(m["xyz"])
    w = np.full(nn, WIN)
    w[y < 1e-12] = 60.0 * MU / (0.075 * m["first_dy"] ** 2)
    U = np.zeros((nn, 3))
    U[:, 0] = 1.0
    for name, v in (("U", U), ("p", np.zeros(nn)), ("k", np.full(nn, KIN)), ("omega", w), ("ymin", y)):
        f.upload(name, v)
    f.initialize()
    return f, dm


def cavity_flow(oa, meshgen, precision, n=14):
    m = meshgen.cavity_hex(n, perturb=0.15, seed=3)
    dm = oa.DeviceMesh(m)
    f = oa.FlowCase(dm, oa.default_params(precision, p_method=oa.OAGPU_CG_SA),
                    np.array([[1.0, 0, 0], [0, 0, 0]]))
    f.upload("U", meshgen.initial_velocity(dm.nn, 5))
    f.initialize()
    return f, dm


def fields(f, names):
    return {k: f.download(k).copy() for k in names}


def assert_bitwise(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), f"{k}: max diff {np.abs(a[k] - b[k]).max():.3e}"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_step_graph_equals_eager(oa, meshgen, step_mesh, precision):
    prec = oa.OAGPU_FP64 if precision == "fp64" else oa.OAGPU_MIXED
    print(f"[case] step slice, SST, BiCGStab-SA, {precision}", flush=True)
    fe, _ = step_flow(oa, meshgen, step_mesh, prec)
    fg, _ = step_flow(oa, meshgen, step_mesh, prec)
    he = fe.graph_iterations(CTL, SCTL, n=5, eager=True)
    hg = fg.graph_iterations(CTL, SCTL, n=2)
    fg.download("U")                        # eager work and synchronisations between replays
    hg = np.vstack([hg, fg.graph_iterations(CTL, SCTL, n=3)])
    st = fg.graph_stats()
    print(f"[graph] {st}; iterations {hg[:, [4, 5, 8, 9]].tolist()}", flush=True)
    assert st["launches"] >= 3 and st["captures"] >= 1
    assert np.array_equal(he, hg)
    assert_bitwise(fields(fe, FIELDS), fields(fg, FIELDS))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cavity_cg_sa_graph_equals_eager(oa, meshgen, precision):
    prec = oa.OAGPU_FP64 if precision == "fp64" else oa.OAGPU_MIXED
    print(f"[case] perturbed 14^3 cavity, CG-SA, {precision}", flush=True)
    fe, _ = cavity_flow(oa, meshgen, prec)
    fg, _ = cavity_flow(oa, meshgen, prec)
    he = fe.graph_iterations(CTL, None, n=6, eager=True)
    hg = np.vstack([fg.graph_iterations(CTL, None, n=3), fg.graph_iterations(CTL, None, n=3)])
    st = fg.graph_stats()
    print(f"[graph] {st}", flush=True)
    assert st["launches"] >= 4
    assert np.array_equal(he, hg)
    assert_bitwise(fields(fe, ("U", "p", "mdot")), fields(fg, ("U", "p", "mdot")))


def test_capped_iteration_rolls_back(oa, meshgen, step_mesh):
    print("[case] step slice, SST, fp64, one slot per solve", flush=True)
    fe, _ = step_flow(oa, meshgen, step_mesh, oa.OAGPU_FP64)
    fg, _ = step_flow(oa, meshgen, step_mesh, oa.OAGPU_FP64)
    he = fe.graph_iterations(CTL, SCTL, n=4, eager=True)
    h0 = fg.graph_iterations(CTL, SCTL, n=1)   # uncaptured first iteration
    fg.tune_set("graph_cap", 1)
    h1 = fg.graph_iterations(CTL, SCTL, n=3)
    st = fg.graph_stats()
    print(f"[graph] {st}", flush=True)
    assert st["redone"] >= 1
    assert np.array_equal(he, np.vstack([h0, h1]))
    assert_bitwise(fields(fe, FIELDS), fields(fg, FIELDS))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_solves_match_host_solves(oa, meshgen, step_mesh, precision):
    prec = oa.OAGPU_FP64 if precision == "fp64" else oa.OAGPU_MIXED
    print(f"[case] step slice, SST, {precision}: graph vs oagpu_outer_iteration", flush=True)
    fh, _ = step_flow(oa, meshgen, step_mesh, prec)
    fg, _ = step_flow(oa, meshgen, step_mesh, prec)
    hh = []
    for _ in range(3):
        a = fh.outer_iteration(CTL)
        b = fh.sst_iteration(SCTL)
        hh.append(np.concatenate([a, b[:2], b[2:]]))
    hh = np.array(hh)
    hg = fg.graph_iterations(CTL, SCTL, n=3)
    print(f"[iters] host {hh[:, [4, 5, 8, 9]].tolist()} graph {hg[:, [4, 5, 8, 9]].tolist()}", flush=True)
    assert np.array_equal(hh[:, [4, 5, 8, 9]], hg[:, [4, 5, 8, 9]])
    np.testing.assert_allclose(hg[:, [0, 1, 2, 3, 6, 7]], hh[:, [0, 1, 2, 3, 6, 7]], rtol=1e-9, atol=0)
    a, b = fields(fh, FIELDS), fields(fg, FIELDS)
    for k in a:
        err = np.abs(a[k] - b[k]).max() / np.abs(a[k]).max()
        assert err < 1e-9, f"{k}: {err:.3e}"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cavity_device_cg_matches_host_cg(oa, meshgen, precision):
    prec = oa.OAGPU_FP64 if precision == "fp64" else oa.OAGPU_MIXED
    print(f"[case] perturbed 14^3 cavity, CG-SA, {precision}: graph vs oagpu_outer_iteration", flush=True)
    fh, _ = cavity_flow(oa, meshgen, prec)
    fg, _ = cavity_flow(oa, meshgen, prec)
    hh = np.array([fh.outer_iteration(CTL) for _ in range(3)])
    hg = fg.graph_iterations(CTL, None, n=3)
    print(f"[iters] host {hh[:, 4:6].tolist()} graph {hg[:, 4:6].tolist()}", flush=True)
    assert np.array_equal(hh[:, 4:6], hg[:, 4:6])
    np.testing.assert_allclose(hg[:, :4], hh[:, :4], rtol=1e-9, atol=0)
    a, b = fields(fh, ("U", "p", "mdot")), fields(fg, ("U", "p", "mdot"))
    for k in a:
        err = np.abs(a[k] - b[k]).max() / np.abs(a[k]).max()
        assert err < 1e-9, f"{k}: {err:.3e}"
