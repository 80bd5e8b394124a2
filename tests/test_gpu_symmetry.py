"""GPU parity of symmetry planes through the C ABI against the oracle on
identical seeded inputs: the lid-driven cavity with its z = 0 and z = 1
faces as symmetry planes (boundary id 2).

Bars: assembled symmetry areas and wall-node flags within 1e-15 absolute /
exact; FP64 momentum systems (projection and normal relaxation included)
and du within 1e-12 relative; velocity and pressure gradients 1e-12;
three flow outer iterations and flow + SST iterations 1e-8 on fields and
control residuals; the mixed form's momentum coefficients within 2e-5 of
the FP64 form.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

BC3 = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
TYPES = [0, 0, 3]


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def sym_mesh(meshgen, n, perturb=0.0, seed=0):
    m = meshgen.cavity_hex(n, perturb=perturb, seed=seed)
    bc = m["bf_bc"].copy()
    bc[(m["bf_ord"] == 4) | (m["bf_ord"] == 5)] = 2
    m["bf_bc"] = np.ascontiguousarray(bc)
    return m


def pair(oracle, oa, meshgen, n=6, perturb=0.0, seed=0, reduced=1, precision=0, sst=False):
    print(f"[case] symmetry n={n} perturb={perturb} seed={seed} reduced={reduced} "
          f"precision={precision} sst={sst}", flush=True)
    m = sym_mesh(meshgen, n, perturb, seed)
    c = oracle.Case(m, oracle.default_params(reduced=reduced), BC3)
    c.set_bc_types(TYPES)
    dm = oa.DeviceMesh(m, oa.OAGPU_STENCIL_REDUCED if reduced else oa.OAGPU_STENCIL_FULL)
    f = oa.FlowCase(dm, oa.default_params(precision), BC3)
    f.set_bc_types(TYPES)
    if sst:
        c.sst_enable(oracle.default_sst())
        f.sst_enable(oa.default_sst())
    return c, dm, f, m


def sync_state(c, f, rng):
    nn = c.nn
    U = rng.normal(scale=0.3, size=(nn, 3))
    p = rng.normal(scale=0.1, size=nn)
    for name, v in (("U", U), ("p", p)):
        c.field(name)[:] = v
        f.upload(name, v)
    c.initialize()
    f.initialize()
    for name, v in (("gradP", rng.normal(scale=0.2, size=(nn, 3))),
                    ("gradU", rng.normal(size=(nn, 9)))):
        c.field(name)[:] = v
        f.upload(name, v)
    md = rng.normal(scale=0.05, size=(c.ne, 12))
    c.field("mdot")[:] = md
    f.upload("mdot", md)
    c.mass_divergence()
    f.mass_divergence()


def test_symmetry_area(oracle, oa, meshgen):
    c, dm, f, m = pair(oracle, oa, meshgen, n=5, perturb=0.15, seed=1)
    assert np.abs(f.download("symA") - c.field("symA")).max() < 1e-15


@pytest.mark.parametrize("case", [dict(n=5), dict(n=6, perturb=0.2, seed=3),
                                  dict(n=4, perturb=0.1, seed=2, reduced=0)])
def test_momentum_symmetry_fp64(oracle, oa, meshgen, case):
    c, dm, f, m = pair(oracle, oa, meshgen, **case)
    sync_state(c, f, np.random.default_rng(3))
    c.assemble_momentum()
    c.compute_du()
    c.normal_relaxation()
    f.assemble_momentum()
    v, b = f.system(3)
    assert rel_err(v, c.vals(3)) < 1e-12
    assert rel_err(b, c.rhs(3)) < 1e-12
    assert rel_err(f.download("du"), c.field("du")) < 1e-12


@pytest.mark.parametrize("which", ["U", "p"])
def test_gradients_symmetry(oracle, oa, meshgen, which):
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, perturb=0.2, seed=4)
    sync_state(c, f, np.random.default_rng(4))
    name = "gradU" if which == "U" else "gradP"
    for _ in range(2):
        c.gradient(which)
        f.gradient(which)
    assert rel_err(f.download(name), c.field(name)) < 1e-12


def test_symmetry_outer_iterations(oracle, oa, meshgen):
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, perturb=0.1, seed=5)
    U = meshgen.initial_velocity(c.nn, 0)
    c.field("U")[:] = U
    f.upload("U", U)
    c.initialize()
    f.initialize()
    ctl = [20, 0.1, 200, 1e-6]
    for it in range(3):
        ho, hg = c.outer_iteration(ctl), f.outer_iteration(ctl)
        np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8)
    for name in ("U", "p", "gradU", "gradP", "mdot"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name
    # the flow stays in the symmetry planes
    A = c.field("symA")
    on = np.linalg.norm(A, axis=1) > 0
    assert np.abs(f.download("U")[on][:, 2]).max() < 0.05 * np.abs(f.download("U")).max()


def test_symmetry_sst_iterations(oracle, oa, meshgen):
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, perturb=0.1, seed=6, sst=True)
    U = meshgen.initial_velocity(c.nn, 0)
    c.field("U")[:] = U
    f.upload("U", U)
    c.initialize()
    f.initialize()
    x = m["xyz"]
    init = {"k": np.full(c.nn, 1e-3), "omega": np.full(c.nn, 1.0),
            "ymin": np.maximum(np.minimum(x[:, :2], 1 - x[:, :2]).min(axis=1), 1e-3)}
    for name, v in init.items():
        c.field(name)[:] = v
        f.upload(name, v)
    ctl, sctl = [20, 0.1, 200, 1e-6], [30, 1e-6]
    for it in range(2):
        ho, hg = c.outer_iteration(ctl), f.outer_iteration(ctl)
        np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8)
        so, sg = c.sst_iteration(sctl), f.sst_iteration(sctl)
        np.testing.assert_allclose(sg[:2], so[:2], rtol=1e-8)
    for name in ("U", "p", "k", "omega", "mut"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name


def test_momentum_symmetry_mixed(oracle, oa, meshgen):
    """FP32 local arithmetic summed into FP64: stated tolerance 2e-5."""
    out = {}
    for prec in (0, 1):
        c, dm, f, m = pair(oracle, oa, meshgen, n=8, perturb=0.2, seed=7, precision=prec)
        sync_state(c, f, np.random.default_rng(7))
        f.assemble_momentum()
        out[prec] = f.system(3)
    assert rel_err(out[1][0], out[0][0]) < 2e-5
    assert rel_err(out[1][1], out[0][1]) < 2e-5
