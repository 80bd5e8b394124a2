"""k and omega assembled by one launch (the row's elements gathered and
their geometry computed once for both equations), and their gradients by
one launch, give the k-omega SST iterations of one launch per equation: the
omega system depends only on fields the k solve leaves unchanged (its node
terms read the k of the iteration's start).  Same scaled residuals, Krylov
iteration counts, k and omega fields and their gradients (relative 1e-12,
the FP64 parity bar: each equation's arithmetic is term for term the
single-equation one, but the compiler may contract the two instantiations'
multiply-adds differently in the last bits) over three flow + SST
iterations, on the
wall-resolved step slice (hex8, FP64 and mixed) and on the mixed-topology
duct with its tet4 / wedge6 / pyramid5 elements and inlet / outlet rows.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MU, KIN = 2.0e-5, 1.5e-4
WIN = KIN / (10.0 * MU)
BFS_BC = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3, [0.0] * 3])
CTL, SCTL = [20, 1.0e-2, 200, 1.0e-2], [20, 1.0e-2]


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def step_flow(oa, meshgen, m, precision, pair):
    dm = oa.DeviceMesh(m)
    f = oa.FlowCase(dm, oa.default_params(precision, mu=MU, urf_U=0.7, urf_p=0.3, u_shifted=1,
                                          p_method=oa.OAGPU_BICGSTAB_SA, high_res=1), BFS_BC)
    f.set_bc_types([1, 2, 0, 3])
    f.set_bc_values(p=[0.0] * 4, k=[KIN, 0, 0, 0], w=[WIN, 0, 0, 0])
    f.sst_enable(oa.default_sst(urf_k=0.9, urf_w=0.9))
    f.tune_set("scalar_pair", pair)
    nn = m["xyz"].shape[0]
    y = meshgen.bfs_wall_distance(m["xyz"])
    w = np.full(nn, WIN)
    w[y < 1e-12] = 60.0 * MU / (0.075 * m["first_dy"] ** 2)
    U = np.zeros((nn, 3))
    U[:, 0] = 1.0
    for name, v in (("U", U), ("p", np.zeros(nn)), ("k", np.full(nn, KIN)), ("omega", w), ("ymin", y)):
        f.upload(name, v)
    f.initialize()
    return f


def duct_flow(oa, meshgen, pair):
    m = meshgen.duct_mixed(6, layers=(6, 2, 3, 4, 3), length=18 * 0.2 / 6, jitter=0.1, seed=2)
    dm = oa.DeviceMesh(m)
    f = oa.FlowCase(dm, oa.default_params(oa.OAGPU_FP64, mu=1e-3, p_method=oa.OAGPU_BICGSTAB_SA),
                    np.array([[1.0, 0, 0], [0, 0, 0], [0, 0, 0]]))
    f.set_bc_types([1, 2, 0])
    f.set_bc_values(p=[0.0, 0.0, 0.0], k=[2e-3, 0, 0], w=[5.0, 0, 0])
    f.sst_enable(oa.default_sst())
    f.tune_set("scalar_pair", pair)
    nn = dm.nn
    rng = np.random.default_rng(4)
    U = np.zeros((nn, 3))
    U[:, 0] = 1.0
    f.upload("U", U + rng.normal(scale=0.05, size=(nn, 3)))
    f.upload("k", np.full(nn, 2e-3))
    f.upload("omega", np.full(nn, 5.0))
    f.upload("ymin", 0.01 + rng.random(nn) * 0.1)
    f.initialize()
    return f, m


def run(f):
    h = []
    for _ in range(3):
        a = f.outer_iteration(CTL)
        b = f.sst_iteration(SCTL)
        h.append(np.concatenate([a, b]))
    return np.array(h), {k: f.download(k).copy() for k in ("U", "p", "k", "omega", "gradK", "gradW")}


def compare(h1, s1, h2, s2):
    print(f"[iters] pair {h1[:, [4, 5, 8, 9]].tolist()} separate {h2[:, [4, 5, 8, 9]].tolist()}", flush=True)
    assert np.array_equal(h1[:, [4, 5, 8, 9]], h2[:, [4, 5, 8, 9]])
    np.testing.assert_allclose(h1[:, [0, 1, 2, 3, 6, 7]], h2[:, [0, 1, 2, 3, 6, 7]], rtol=1e-12, atol=0)
    for k in s1:
        err = np.abs(s1[k] - s2[k]).max() / np.abs(s2[k]).max()
        print(f"[diff] {k} {err:.2e}", flush=True)
        assert err < 1e-12, f"{k}: {err:.3e}"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pair_equals_separate_step(oa, meshgen, precision):
    prec = oa.OAGPU_FP64 if precision == "fp64" else oa.OAGPU_MIXED
    m = meshgen.bfs_hex(3, 12, 24, 28, 2)
    print(f"[case] step slice {m['conn'].shape[0]} hex8, SST, high resolution, {precision}", flush=True)
    h1, s1 = run(step_flow(oa, meshgen, m, prec, 1))
    h2, s2 = run(step_flow(oa, meshgen, m, prec, 0))
    compare(h1, s1, h2, s2)


def test_pair_equals_separate_mixed_topology(oa, meshgen):
    f1, m = duct_flow(oa, meshgen, 1)
    print(f"[case] mixed-topology duct {m['conn'].shape[0]} elements, SST, fp64", flush=True)
    h1, s1 = run(f1)
    f2, _ = duct_flow(oa, meshgen, 0)
    h2, s2 = run(f2)
    compare(h1, s1, h2, s2)
