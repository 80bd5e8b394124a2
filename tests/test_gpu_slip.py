"""GPU parity of slip walls (OAGPU_BC_SLIP) against the oracle: a channel
(unit cube, perturbed nodes) with a velocity inlet at x = 0, a
static-pressure outlet at x = 1 and slip walls on the four other faces.
FP64 momentum and p' systems within 1e-12 (row-relative, rhs per entry);
five outer iterations within 1e-8 on control residuals and fields.
"""
import numpy as np
import pytest

from conftest import entry_rel_err, rel_err, row_rel_err

pytestmark = pytest.mark.gpu

INLET, OUTLET, SLIP = 1, 2, 4
BC = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3])


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def pair(oracle, oa, meshgen, n=6, seed=6):
    print(f"[case] slip-wall channel n={n} seed={seed}", flush=True)
    m = meshgen.cavity_hex(n, perturb=0.15, seed=seed)
    bc = np.full(m["bf_bc"].shape, 2, dtype=np.int32)
    bc[m["bf_ord"] == 3] = 0
    bc[m["bf_ord"] == 1] = 1
    m["bf_bc"] = bc
    c = oracle.Case(m, oracle.default_params(p_method=1), BC)
    dm = oa.DeviceMesh(m)
    f = oa.FlowCase(dm, oa.default_params(oa.OAGPU_FP64), BC)
    for x in (c, f):
        x.set_bc_types([INLET, OUTLET, SLIP])
        x.set_bc_values(p=[0.0, 0.05, 0.0])
    return c, dm, f, m


def both(c, f, name, v):
    c.field(name)[:] = v
    f.upload(name, v)


def test_slip_systems_fp64(oracle, oa, meshgen):
    c, dm, f, m = pair(oracle, oa, meshgen)
    rng = np.random.default_rng(8)
    nn = c.nn
    U = rng.normal(scale=0.3, size=(nn, 3))
    U[:, 0] += 1.0
    both(c, f, "U", U)
    both(c, f, "p", rng.normal(scale=0.1, size=nn))
    c.initialize()
    f.initialize()
    for name in ("gradU", "gradP", "msd", "div"):
        assert rel_err(f.download(name), c.field(name)) < 1e-12, name
    c.assemble_momentum()
    c.compute_du()
    c.normal_relaxation()
    f.assemble_momentum()
    v, b = f.system(3)
    assert row_rel_err(v, c.vals(3), c.row_ptr) < 1e-12
    assert entry_rel_err(b, c.rhs(3)) < 1e-12
    c.assemble_pcorr()
    f.assemble_pcorr()
    v, b = f.system(1)
    assert row_rel_err(v, c.vals(1), c.row_ptr) < 1e-12
    assert entry_rel_err(b, c.rhs(1)) < 1e-12


def test_slip_outer_iterations(oracle, oa, meshgen):
    c, dm, f, m = pair(oracle, oa, meshgen, n=5, seed=2)
    U = np.zeros((c.nn, 3))
    U[:, 0] = 1.0
    both(c, f, "U", U)
    c.initialize()
    f.initialize()
    ctl = [20, 1e-3, 200, 1e-8]
    for _ in range(5):
        ho, hg = c.outer_iteration(ctl), f.outer_iteration(ctl)
        np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8, atol=1e-14)
    for name in ("U", "p", "mdot", "msd"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name
