"""GPU parity of the k-omega SST path (liboagpu.so through its C ABI)
against the oracle on identical seeded inputs.

Bars: nodal model fields (mu_t, mu_eff, F1, P_k) and the k / omega systems
within 1e-12 relative in the FP64 form; momentum with a variable effective
viscosity within 1e-12; three coupled flow + SST outer iterations within 1e-8
on fields and control residuals (Krylov counts within one iteration); the
mixed form's k / omega coefficients within 2e-5 of the FP64 form.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

BC = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
NODE_FIELDS = ("k", "omega", "ymin", "mut", "F1", "Pk_wall")


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def pair(oracle, oa, meshgen, n=6, perturb=0.0, seed=0, reduced=1, precision=0):
    print(f"[case] sst n={n} perturb={perturb} seed={seed} reduced={reduced} precision={precision}",
          flush=True)
    m = meshgen.cavity_hex(n, perturb=perturb, seed=seed)
    c = oracle.Case(m, oracle.default_params(reduced=reduced), BC)
    c.sst_enable(oracle.default_sst())
    dm = oa.DeviceMesh(m, oa.OAGPU_STENCIL_REDUCED if reduced else oa.OAGPU_STENCIL_FULL)
    f = oa.FlowCase(dm, oa.default_params(precision), BC)
    f.sst_enable(oa.default_sst())
    return c, dm, f, m


def sync_state(c, f, rng):
    nn = c.nn
    U = rng.normal(scale=0.3, size=(nn, 3))
    c.field("U")[:] = U
    f.upload("U", U)
    c.initialize()
    f.initialize()
    vals = {"k": rng.uniform(1e-4, 1e-2, nn), "omega": rng.uniform(0.5, 5.0, nn),
            "ymin": rng.uniform(0.01, 0.3, nn), "mut": rng.uniform(0, 1e-2, nn),
            "F1": rng.uniform(0, 1, nn), "Pk_wall": rng.uniform(0, 1e-3, nn),
            "gradU": rng.normal(size=(nn, 9)), "gradK": rng.normal(scale=1e-2, size=(nn, 3)),
            "gradW": rng.normal(size=(nn, 3)), "kprev": rng.uniform(1e-4, 1e-2, nn),
            "Pk": rng.uniform(0, 1e-3, nn)}
    for name, v in vals.items():
        c.field(name)[:] = v
        f.upload(name, v)
    md = rng.normal(scale=0.05, size=(c.ne, 12))
    c.field("mdot")[:] = md
    f.upload("mdot", md)
    c.mass_divergence()
    f.mass_divergence()


def test_sst_nodal_update(oracle, oa, meshgen):
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, perturb=0.1, seed=1)
    sync_state(c, f, np.random.default_rng(1))
    c.sst_update_mut()
    c.sst_update_F1()
    c.sst_update_Pk()
    f.sst_update()
    for name in ("mut", "F1", "Pk", "mu"):
        assert rel_err(f.download(name), c.field(name)) < 1e-12, name


@pytest.mark.parametrize("case", [dict(n=5), dict(n=6, perturb=0.2, seed=3),
                                  dict(n=4, perturb=0.1, seed=2, reduced=0)])
@pytest.mark.parametrize("which", ["k", "omega"])
def test_scalar_assembly_fp64(oracle, oa, meshgen, case, which):
    c, dm, f, m = pair(oracle, oa, meshgen, **case)
    sync_state(c, f, np.random.default_rng(2))
    c.assemble_scalar(which)
    f.assemble_scalar(which)
    v, b = f.system(1)
    assert rel_err(v, c.vals(1)) < 1e-12
    assert rel_err(b, c.rhs(1)) < 1e-12


def test_momentum_variable_viscosity(oracle, oa, meshgen):
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, perturb=0.15, seed=4)
    sync_state(c, f, np.random.default_rng(3))
    rng = np.random.default_rng(4)
    mu = 0.01 + rng.uniform(0, 0.05, c.nn)
    c.field("mu")[:] = mu
    f.upload("mu", mu)
    gP = rng.normal(scale=0.2, size=(c.nn, 3))
    c.field("gradP")[:] = gP
    f.upload("gradP", gP)
    c.assemble_momentum()
    f.assemble_momentum()
    v, b = f.system(3)
    assert rel_err(v, c.vals(3)) < 1e-12
    assert rel_err(b, c.rhs(3)) < 1e-12


def test_flow_and_sst_iterations(oracle, oa, meshgen):
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, perturb=0.1, seed=5)
    U = meshgen.initial_velocity(c.nn, 0)
    c.field("U")[:] = U
    f.upload("U", U)
    c.initialize()
    f.initialize()
    x = m["xyz"]
    init = {"k": np.full(c.nn, 1e-3), "omega": np.full(c.nn, 1.0),
            "ymin": np.maximum(np.minimum(x, 1 - x).min(axis=1), 1e-3)}
    for name, v in init.items():
        c.field(name)[:] = v
        f.upload(name, v)
    ctl, sctl = [20, 0.1, 200, 1e-6], [30, 1e-6]
    for it in range(3):
        ho, hg = c.outer_iteration(ctl), f.outer_iteration(ctl)
        np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8)
        so, sg = c.sst_iteration(sctl), f.sst_iteration(sctl)
        np.testing.assert_allclose(sg[:2], so[:2], rtol=1e-8)
        assert np.all(np.abs(sg[2:] - so[2:]) <= 1)
    for name in ("U", "p", "k", "omega", "mut", "mu", "gradK", "gradW"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name


def test_scalar_assembly_mixed(oracle, oa, meshgen):
    """FP32 local arithmetic summed into FP64: stated tolerance 2e-5."""
    out = {}
    for prec in (0, 1):
        c, dm, f, m = pair(oracle, oa, meshgen, n=8, perturb=0.2, seed=6, precision=prec)
        sync_state(c, f, np.random.default_rng(7))
        f.assemble_scalar("omega")
        out[prec] = f.system(1)
    assert rel_err(out[1][0], out[0][0]) < 2e-5
    assert rel_err(out[1][1], out[0][1]) < 2e-5
