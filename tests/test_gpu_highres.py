"""GPU parity of high-resolution advection (limited deferred correction)
through the C ABI against the oracle on identical seeded inputs.

Bars: blending factors of U, k and omega within 1e-12 absolute (they are
in [0, 1]); momentum and k / omega systems with random blending factors
within 1e-12 relative in the FP64 form; three high-resolution outer
iterations within 1e-8 on fields and control residuals; the mixed form's
momentum coefficients and rhs within 2e-5 of the FP64 form.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

BC = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def pair(oracle, oa, meshgen, n=6, perturb=0.0, seed=0, reduced=1, precision=0, sst=False):
    print(f"[case] high-res n={n} perturb={perturb} seed={seed} reduced={reduced} "
          f"precision={precision} sst={sst}", flush=True)
    m = meshgen.cavity_hex(n, perturb=perturb, seed=seed)
    c = oracle.Case(m, oracle.default_params(reduced=reduced, high_res=1), BC)
    dm = oa.DeviceMesh(m, oa.OAGPU_STENCIL_REDUCED if reduced else oa.OAGPU_STENCIL_FULL)
    f = oa.FlowCase(dm, oa.default_params(precision, high_res=1), BC)
    if sst:
        c.sst_enable(oracle.default_sst())
        f.sst_enable(oa.default_sst())
    return c, dm, f, m


def sync_state(c, f, rng, sst=False):
    nn = c.nn
    U = rng.normal(scale=0.3, size=(nn, 3))
    c.field("U")[:] = U
    f.upload("U", U)
    c.initialize()
    f.initialize()
    vals = {"gradU": rng.normal(size=(nn, 9)), "gradP": rng.normal(scale=0.2, size=(nn, 3)),
            "du": rng.uniform(1e-3, 2e-3, size=(nn, 3))}
    if sst:
        vals.update({"k": rng.uniform(1e-4, 1e-2, nn), "omega": rng.uniform(0.5, 5.0, nn),
                     "ymin": rng.uniform(0.01, 0.3, nn), "mut": rng.uniform(0, 1e-2, nn),
                     "F1": rng.uniform(0, 1, nn), "Pk_wall": rng.uniform(0, 1e-3, nn),
                     "gradK": rng.normal(scale=1e-2, size=(nn, 3)),
                     "gradW": rng.normal(size=(nn, 3)), "kprev": rng.uniform(1e-4, 1e-2, nn),
                     "Pk": rng.uniform(0, 1e-3, nn)})
    for name, v in vals.items():
        c.field(name)[:] = v
        f.upload(name, v)
    md = rng.normal(scale=0.05, size=(c.ne, 12))
    c.field("mdot")[:] = md
    f.upload("mdot", md)
    c.mass_divergence()
    f.mass_divergence()


@pytest.mark.parametrize("case", [dict(n=5), dict(n=6, perturb=0.2, seed=3)])
def test_blending_velocity(oracle, oa, meshgen, case):
    c, dm, f, m = pair(oracle, oa, meshgen, **case)
    sync_state(c, f, np.random.default_rng(1))
    c.update_blending("U")
    f.update_blending("U")
    b = c.field("beta")
    assert b.min() >= 0.0 and b.max() <= 1.0 and 0.0 < b.mean() < 1.0
    assert np.abs(f.download("beta") - b).max() < 1e-12


@pytest.mark.parametrize("which", ["k", "omega"])
def test_blending_scalars(oracle, oa, meshgen, which):
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, perturb=0.15, seed=2, sst=True)
    sync_state(c, f, np.random.default_rng(2), sst=True)
    c.update_blending(which)
    f.update_blending(which)
    name = "beta_k" if which == "k" else "beta_w"
    assert np.abs(f.download(name) - c.field(name)).max() < 1e-12


@pytest.mark.parametrize("case", [dict(n=5), dict(n=6, perturb=0.2, seed=3),
                                  dict(n=4, perturb=0.1, seed=1, reduced=0)])
def test_momentum_deferred_correction(oracle, oa, meshgen, case):
    c, dm, f, m = pair(oracle, oa, meshgen, **case)
    rng = np.random.default_rng(3)
    sync_state(c, f, rng)
    beta = rng.uniform(0, 1, size=(c.nn, 3))
    c.field("beta")[:] = beta
    f.upload("beta", beta)
    c.assemble_momentum()
    f.assemble_momentum()
    v, b = f.system(3)
    assert rel_err(v, c.vals(3)) < 1e-12
    assert rel_err(b, c.rhs(3)) < 1e-12


@pytest.mark.parametrize("which", ["k", "omega"])
def test_scalar_deferred_correction(oracle, oa, meshgen, which):
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, perturb=0.2, seed=4, sst=True)
    rng = np.random.default_rng(4)
    sync_state(c, f, rng, sst=True)
    for name in ("beta_k", "beta_w"):
        beta = rng.uniform(0, 1, size=c.nn)
        c.update_blending("k" if name == "beta_k" else "omega")   # allocates
        c.field(name)[:] = beta
        f.upload(name, beta)
    c.assemble_scalar(which)
    f.assemble_scalar(which)
    v, b = f.system(1)
    assert rel_err(v, c.vals(1)) < 1e-12
    assert rel_err(b, c.rhs(1)) < 1e-12


def test_highres_outer_iterations(oracle, oa, meshgen):
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, perturb=0.1, seed=5)
    U = meshgen.initial_velocity(c.nn, 0)
    c.field("U")[:] = U
    f.upload("U", U)
    c.initialize()
    f.initialize()
    assert np.abs(f.download("beta") - c.field("beta")).max() < 1e-12
    ctl = [20, 0.1, 200, 1e-6]
    for it in range(3):
        ho, hg = c.outer_iteration(ctl), f.outer_iteration(ctl)
        np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8)
    for name in ("U", "p", "beta", "gradU"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name


def test_momentum_deferred_correction_mixed(oracle, oa, meshgen):
    """FP32 local arithmetic summed into FP64: stated tolerance 2e-5."""
    out = {}
    for prec in (0, 1):
        c, dm, f, m = pair(oracle, oa, meshgen, n=8, perturb=0.2, seed=6, precision=prec)
        rng = np.random.default_rng(7)
        sync_state(c, f, rng)
        f.upload("beta", rng.uniform(0, 1, size=(c.nn, 3)))
        f.assemble_momentum()
        out[prec] = f.system(3)
    assert rel_err(out[1][0], out[0][0]) < 2e-5
    assert rel_err(out[1][1], out[0][1]) < 2e-5
