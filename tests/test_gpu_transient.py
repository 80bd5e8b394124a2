"""GPU parity of the transient path (oagpu_time_step: BDF1 / BDF2 time terms
in the momentum and k / omega node terms) against the oracle on identical
inputs: systems within 1e-12 (row-relative matrices, rhs per entry against
the flux scale), and time-stepped outer iterations (three steps, variable
dt, three outer iterations each) within 1e-8 on control residuals and
fields; the time levels the device moves back equal the oracle's."""
import numpy as np
import pytest

from conftest import entry_rel_err, rel_err, row_rel_err

pytestmark = pytest.mark.gpu

BC = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def pair(oracle, oa, meshgen, sst=False, **kw):
    m = meshgen.cavity_hex(6, perturb=0.15, seed=4)
    c = oracle.Case(m, oracle.default_params(**kw), BC)
    dm = oa.DeviceMesh(m)
    f = oa.FlowCase(dm, oa.default_params(oa.OAGPU_FP64, **kw), BC)
    if sst:
        c.sst_enable(oracle.default_sst())
        f.sst_enable(oa.default_sst())
    return c, dm, f, m


@pytest.mark.parametrize("scheme", [1, 2])
def test_transient_assembly(oracle, oa, meshgen, scheme):
    print(f"[case] transient assembly scheme={scheme}", flush=True)
    c, dm, f, m = pair(oracle, oa, meshgen, sst=True)
    rng = np.random.default_rng(31 + scheme)
    nn = c.nn
    levels = [rng.normal(scale=0.3, size=(nn, 3)) for _ in range(3)]
    klev = [rng.uniform(1e-3, 1e-2, nn) for _ in range(3)]
    wlev = [rng.uniform(0.5, 5.0, nn) for _ in range(3)]
    sst_vals = {"ymin": rng.uniform(0.01, 0.3, nn), "mut": rng.uniform(0, 1e-2, nn),
                "F1": rng.uniform(0, 1, nn), "Pk_wall": rng.uniform(0, 1e-3, nn),
                "gradK": rng.normal(scale=1e-2, size=(nn, 3)), "gradW": rng.normal(size=(nn, 3)),
                "kprev": rng.uniform(1e-4, 1e-2, nn), "Pk": rng.uniform(0, 1e-3, nn)}
    for step, dt in enumerate((0.3, 0.2)):
        for case, up in ((c, lambda n, v: c.field(n).__setitem__(slice(None), v)),
                         (f, lambda n, v: f.upload(n, v))):
            up("U", levels[step])
            up("k", klev[step])
            up("omega", wlev[step])
            case.time_step(scheme, dt)
    for name, v in (("U", levels[2]), ("k", klev[2]), ("omega", wlev[2]), ("p", rng.normal(size=nn))):
        c.field(name)[:] = v
        f.upload(name, v)
    for name in ("Uold", "Uold2", "kold", "wold"):
        assert np.array_equal(f.download(name), c.field(name)), name
    c.initialize()
    f.initialize()
    for name, v in sst_vals.items():
        c.field(name)[:] = v
        f.upload(name, v)
    c.assemble_momentum()
    c.compute_du()
    f.assemble_momentum()
    vals, b = f.system(3)
    assert row_rel_err(vals, c.vals(3), c.row_ptr) < 1e-12
    assert entry_rel_err(b, c.rhs(3)) < 1e-12
    for which in ("k", "omega"):
        c.assemble_scalar(which)
        f.assemble_scalar(which)
        vals, b = f.system(1)
        assert row_rel_err(vals, c.vals(1), c.row_ptr) < 1e-12, which
        assert entry_rel_err(b, c.rhs(1)) < 1e-12, which


@pytest.mark.parametrize("scheme", [1, 2])
def test_time_steps(oracle, oa, meshgen, scheme):
    print(f"[case] transient steps scheme={scheme}", flush=True)
    c, dm, f, m = pair(oracle, oa, meshgen, p_method=1)
    U0 = meshgen.initial_velocity(c.nn, 3)
    c.field("U")[:] = U0
    f.upload("U", U0)
    c.initialize()
    f.initialize()
    ctl = [20, 1e-3, 200, 1e-8]
    for dt in (0.05, 0.05, 0.08):
        c.time_step(scheme, dt)
        f.time_step(scheme, dt)
        for _ in range(3):
            ho = c.outer_iteration(ctl)
            hg = f.outer_iteration(ctl)
            np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8)
    for name in ("U", "p", "mdot", "Uold", "Uold2"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name
