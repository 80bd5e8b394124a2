"""GPU parity on a mixed-topology mesh (hex8, tet4, wedge6 and pyramid5
elements in one mesh) through the C ABI against the oracle on identical
seeded inputs: the duct of meshgen.duct_mixed with a velocity inlet, a
static-pressure outlet and no-slip walls, nodes jittered.  Its walls are
sides of every element that touches them: quad sides of hex8 / wedge6 /
pyramid5 and tri sides of tet4; a second duct ends in pyramids (inlet: quad
bases) and wedges (outlet: tri sides).

Bars: CRS graph and diagonal offsets bit-exact (reduced and full stencil);
dual volumes, mass flows, mass divergence, gradients, blending factors and
FP64 momentum / p' / k / omega systems within 1e-12 relative; flow and SST
outer iterations 1e-8 on fields and control residuals; the mixed form's
momentum and p' coefficients within 2e-5 of the FP64 form.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

LAYERS = (4, 1, 2, 2, 2)
LAYERS_ENDS = (0, 2, 2, 3, 0)   # pyramid cubes at the inlet, wedge cubes at the outlet
M = 6
INLET, OUTLET, WALL = 1, 2, 0
BC = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3])
TYPES = [INLET, OUTLET, WALL]


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def duct(meshgen, jitter=0.1, seed=1, layers=LAYERS):
    L = sum(layers) * 0.2 / M
    return meshgen.duct_mixed(M, layers=layers, length=L, jitter=jitter, seed=seed)


def pair(oracle, oa, meshgen, reduced=1, precision=0, sst=False, high_res=0, seed=1, io=True,
         layers=LAYERS):
    print(f"[case] mixed duct layers={layers} reduced={reduced} precision={precision} sst={sst} "
          f"high_res={high_res} seed={seed} io={io}", flush=True)
    m = duct(meshgen, seed=seed, layers=layers)
    c = oracle.Case(m, oracle.default_params(reduced=reduced, high_res=high_res), BC)
    dm = oa.DeviceMesh(m, oa.OAGPU_STENCIL_REDUCED if reduced else oa.OAGPU_STENCIL_FULL)
    f = oa.FlowCase(dm, oa.default_params(precision, high_res=high_res), BC)
    if io:
        for x in (c, f):
            x.set_bc_types(TYPES)
            x.set_bc_values(p=[0.0, 0.0, 0.0], k=[2e-3, 0, 0], w=[5.0, 0, 0])
    if sst:
        c.sst_enable(oracle.default_sst())
        f.sst_enable(oa.default_sst())
    return c, dm, f, m


def both(c, f, name, v):
    c.field(name)[:] = v
    f.upload(name, v)


def sync_state(c, f, rng):
    nn = c.nn
    U = rng.normal(scale=0.3, size=(nn, 3))
    U[:, 0] += 1.0
    both(c, f, "U", U)
    both(c, f, "p", rng.normal(scale=0.1, size=nn))
    c.initialize()
    f.initialize()
    both(c, f, "gradP", rng.normal(scale=0.2, size=(nn, 3)))
    both(c, f, "gradU", rng.normal(size=(nn, 9)))
    both(c, f, "mdot", rng.normal(scale=0.01, size=(c.ne, 12)))
    c.mass_divergence()
    f.mass_divergence()


@pytest.mark.parametrize("reduced", [1, 0])
def test_graph_and_volumes(oracle, oa, meshgen, reduced):
    c, dm, f, m = pair(oracle, oa, meshgen, reduced=reduced, io=False)
    rp, cols, doff = dm.graph()
    np.testing.assert_array_equal(rp, c.row_ptr)
    np.testing.assert_array_equal(cols, c.cols)
    np.testing.assert_array_equal(doff, c.diag_off)
    assert rel_err(dm.dual_volume(), c.field("vol")) < 1e-13


def test_boundary_side_kinds(meshgen):
    """Both ducts carry tri sides on their boundary (tet walls; wedge outlet)."""
    for layers, want in ((LAYERS, {(1, 3)}), (LAYERS_ENDS, {(1, 3), (2, 4), (3, 4)})):
        m = duct(meshgen, layers=layers)
        kinds = set(zip(m["etype"][m["bf_elem"]].tolist(), m["bf_ord"].tolist()))
        assert want <= kinds, (layers, kinds)


@pytest.mark.parametrize("layers", [LAYERS, LAYERS_ENDS])
def test_mass_flow_divergence_gradients(oracle, oa, meshgen, layers):
    c, dm, f, m = pair(oracle, oa, meshgen, layers=layers)
    rng = np.random.default_rng(2)
    U = rng.normal(scale=0.3, size=(c.nn, 3))
    U[:, 0] += 1.0
    both(c, f, "U", U)
    both(c, f, "p", rng.normal(scale=0.1, size=c.nn))
    c.initialize()
    f.initialize()
    for name in ("mdot", "msd", "div", "gradU", "gradP"):
        assert rel_err(f.download(name), c.field(name)) < 1e-12, name
    both(c, f, "du", rng.uniform(0.01, 0.02, size=(c.nn, 3)))
    c.update_mdot()
    f.update_mdot()
    c.mass_divergence()
    f.mass_divergence()
    for name in ("mdot", "msd", "div"):
        assert rel_err(f.download(name), c.field(name)) < 1e-12, name
    for which, name in (("U", "gradU"), ("p", "gradP")):
        c.gradient(which)
        f.gradient(which)
        assert rel_err(f.download(name), c.field(name)) < 1e-12, name


@pytest.mark.parametrize("case", [dict(reduced=1), dict(reduced=0), dict(reduced=1, high_res=1),
                                  dict(reduced=1, layers=LAYERS_ENDS),
                                  dict(reduced=1, high_res=1, layers=LAYERS_ENDS)])
def test_systems_fp64(oracle, oa, meshgen, case):
    c, dm, f, m = pair(oracle, oa, meshgen, **case)
    sync_state(c, f, np.random.default_rng(3))
    if case.get("high_res"):
        c.update_blending("U")
        f.update_blending("U")
        assert rel_err(f.download("beta"), c.field("beta")) < 1e-12
    c.assemble_momentum()
    c.compute_du()
    f.assemble_momentum()
    v, b = f.system(3)
    assert rel_err(v, c.vals(3)) < 1e-12
    assert rel_err(b, c.rhs(3)) < 1e-12
    assert rel_err(f.download("du"), c.field("du")) < 1e-12
    c.assemble_pcorr()
    f.assemble_pcorr()
    v, b = f.system(1)
    assert rel_err(v, c.vals(1)) < 1e-12
    assert rel_err(b, c.rhs(1)) < 1e-12


@pytest.mark.parametrize("which", ["k", "omega"])
@pytest.mark.parametrize("layers", [LAYERS, LAYERS_ENDS])
def test_scalar_systems_fp64(oracle, oa, meshgen, which, layers):
    c, dm, f, m = pair(oracle, oa, meshgen, sst=True, layers=layers)
    rng = np.random.default_rng(5)
    sync_state(c, f, rng)
    nn = c.nn
    vals = {"k": rng.uniform(1e-4, 1e-2, nn), "omega": rng.uniform(0.5, 5.0, nn),
            "ymin": rng.uniform(0.01, 0.3, nn), "mut": rng.uniform(0, 1e-2, nn),
            "F1": rng.uniform(0, 1, nn), "Pk_wall": rng.uniform(0, 1e-3, nn),
            "gradK": rng.normal(scale=1e-2, size=(nn, 3)), "gradW": rng.normal(size=(nn, 3)),
            "kprev": rng.uniform(1e-4, 1e-2, nn), "Pk": rng.uniform(0, 1e-3, nn)}
    for name, v in vals.items():
        both(c, f, name, v)
    c.assemble_scalar(which)
    f.assemble_scalar(which)
    v, b = f.system(1)
    assert rel_err(v, c.vals(1)) < 1e-12
    assert rel_err(b, c.rhs(1)) < 1e-12


@pytest.mark.parametrize("layers", [LAYERS, LAYERS_ENDS])
def test_outer_iterations(oracle, oa, meshgen, layers):
    c, dm, f, m = pair(oracle, oa, meshgen, sst=True, layers=layers)
    U = np.zeros((c.nn, 3))
    U[:, 0] = 1.0
    both(c, f, "U", U)
    both(c, f, "p", np.zeros(c.nn))
    x = m["xyz"]
    init = {"k": np.full(c.nn, 2e-3), "omega": np.full(c.nn, 5.0),
            "ymin": np.maximum(np.minimum(x[:, 1:], 0.2 - x[:, 1:]).min(axis=1), 1e-3)}
    for name, v in init.items():
        both(c, f, name, v)
    c.initialize()
    f.initialize()
    ctl, sctl = [20, 0.1, 200, 1e-6], [30, 1e-6]
    for it in range(3):
        ho, hg = c.outer_iteration(ctl), f.outer_iteration(ctl)
        np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8, atol=1e-14)
        so, sg = c.sst_iteration(sctl), f.sst_iteration(sctl)
        np.testing.assert_allclose(sg[:2], so[:2], rtol=1e-8, atol=1e-14)
    for name in ("U", "p", "mdot", "msd", "k", "omega", "mut"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name


def test_systems_mixed_form(oracle, oa, meshgen):
    """FP32 local arithmetic summed into FP64: stated tolerance 2e-5."""
    out = {}
    for prec in (0, 1):
        c, dm, f, m = pair(oracle, oa, meshgen, precision=prec, seed=4)
        sync_state(c, f, np.random.default_rng(9))
        f.assemble_momentum()
        out[prec, 3] = f.system(3)
        f.assemble_pcorr()
        out[prec, 1] = f.system(1)
    for N in (3, 1):
        assert rel_err(out[1, N][0], out[0, N][0]) < 2e-5
        assert rel_err(out[1, N][1], out[0, N][1]) < 2e-5
