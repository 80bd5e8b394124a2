"""Batched independent cases (oagpu_batch_outer_iterations): several small
lid-driven cavities of different sizes, viscosities and lid directions
advanced concurrently, each on its own stream from a pool of host threads,
give bitwise the same histories and fields as the same cases run one after
another (every reduction has a fixed order, so concurrency cannot change a
case's arithmetic); a case stops at its residual tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CTL = [20, 0.1, 200, 1e-6]
CASES = [dict(n=8, mu=0.01, th=0.0), dict(n=10, mu=0.002, th=0.7),
         dict(n=12, mu=0.05, th=2.1), dict(n=9, mu=0.001, th=-1.2), dict(n=11, mu=0.02, th=3.0)]


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def make(oa, meshgen, case, precision, p_method):
    m = meshgen.cavity_hex(case["n"])
    dm = oa.DeviceMesh(m)
    lid = [np.cos(case["th"]), 0.0, np.sin(case["th"])]     # tangential to the lid y = 1
    f = oa.FlowCase(dm, oa.default_params(precision, mu=case["mu"], p_method=p_method),
                    np.array([lid, [0.0, 0.0, 0.0]]))
    f.upload("U", meshgen.initial_velocity(dm.nn, case["n"]))
    f.initialize()
    return dm, f


@pytest.mark.parametrize("precision, method", [(0, 1), (1, 3)])
def test_batch_matches_sequential(oa, meshgen, precision, method):
    print(f"[case] batch precision={precision} method={method}", flush=True)
    iters = 4
    seq = []
    for c in CASES:
        dm, f = make(oa, meshgen, c, precision, method)
        h = np.array([f.outer_iteration(CTL) for _ in range(iters)])
        seq.append((h, f.download("U"), f.download("p"), dm, f))
    made = [make(oa, meshgen, c, precision, method) for c in CASES]
    hist, done, sec = oa.batch_outer_iterations([f for _, f in made], CTL, iters, threads=3)
    assert np.all(done == iters) and np.all(sec > 0)
    for k, (dm, f) in enumerate(made):
        np.testing.assert_array_equal(hist[k], seq[k][0])
        np.testing.assert_array_equal(f.download("U"), seq[k][1])
        np.testing.assert_array_equal(f.download("p"), seq[k][2])


def test_batch_stops_at_tolerance(oa, meshgen):
    print("[case] batch tolerance", flush=True)
    made = [make(oa, meshgen, c, 0, 1) for c in CASES]
    hist, done, sec = oa.batch_outer_iterations([f for _, f in made], CTL, 10, tol=1e30)
    assert np.all(done == 1)
    hist, done, sec = oa.batch_outer_iterations([f for _, f in made], CTL, 12, tol=0.05)
    for k in range(len(CASES)):
        worst = hist[k][:, :4].max(axis=1)
        d = done[k]
        assert 1 <= d <= 12
        assert np.all(worst[:d - 1] >= 0.05)                 # ran on while above the tolerance
        assert d == 12 or worst[d - 1] < 0.05                # stopped at the first one below
    assert len(set(done.tolist())) > 1                       # the cases stop at different times
