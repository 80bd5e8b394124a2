"""GPU parity of boundary sides of every side kind against the oracle: one
distorted tet4, wedge6 or pyramid5 element whose sides are all boundary
sides, typed so that tri and quad sides of each topology are velocity
inlets, static-pressure outlets, no-slip walls and symmetry planes (tri
sides of pyramids and wedges appear on no duct).  On random states, with
the standard and with the shifted face ips: side mass flows (initial and
Rhie-Chow updated), reversal flags, mass divergence, gradients, blending
factors and the FP64 momentum, p' and k / omega systems within 1e-12 of the
oracle (matrices row-relative, vectors per entry).

At a pyramid's apex the shifted face gradient is singular; both sides give
the gradient operator's fallback (a zero operator) there.
"""
import numpy as np
import pytest

from conftest import entry_rel_err, rel_err, row_rel_err

pytestmark = pytest.mark.gpu

INLET, OUTLET, WALL, SYM = 1, 2, 0, 3
TYPES = [INLET, OUTLET, WALL, SYM]
BC = np.array([[0.7, 0.2, -0.1], [0.0] * 3, [0.0] * 3, [0.0] * 3])
PHYS = {1: [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
        2: [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)],
        3: [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0.5, 0.5, 1)]}
# boundary id per side ordinal: every side kind of the topology takes an
# inlet or an outlet somewhere across the cases
SIDE_BC = {1: [[INLET, OUTLET, WALL, SYM], [OUTLET, WALL, SYM, INLET]],
           2: [[WALL, SYM, WALL, INLET, OUTLET], [INLET, OUTLET, SYM, WALL, WALL]],
           3: [[INLET, OUTLET, WALL, SYM, WALL], [WALL, SYM, OUTLET, WALL, INLET]]}


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def element_mesh(topo, variant, seed):
    rng = np.random.default_rng(seed)
    xyz = np.array(PHYS[topo], dtype=np.float64) + rng.uniform(-0.06, 0.06, size=(len(PHYS[topo]), 3))
    npe = xyz.shape[0]
    conn = np.full((1, 8), -1, dtype=np.int32)
    conn[0, :npe] = np.arange(npe)
    nsides = len(SIDE_BC[topo][variant])
    return {"xyz": xyz, "conn": conn, "etype": np.array([topo], dtype=np.int32),
            "bf_elem": np.zeros(nsides, dtype=np.int32),
            "bf_ord": np.arange(nsides, dtype=np.int32),
            "bf_bc": np.array(SIDE_BC[topo][variant], dtype=np.int32)}


def pair(oracle, oa, topo, variant, shifted, sst=False, high_res=0):
    print(f"[case] single element topo={topo} variant={variant} shifted={shifted} sst={sst} "
          f"high_res={high_res}", flush=True)
    m = element_mesh(topo, variant, 10 * topo + variant)
    kw = dict(u_shifted=shifted, ugrad_shifted=shifted, p_shifted=shifted, pgrad_shifted=shifted,
              high_res=high_res, p_ref_node=-1)
    c = oracle.Case(m, oracle.default_params(**kw), BC)
    dm = oa.DeviceMesh(m)
    f = oa.FlowCase(dm, oa.default_params(oa.OAGPU_FP64, **kw), BC)
    for x in (c, f):
        x.set_bc_types(TYPES)
        x.set_bc_values(p=[0.0, 0.05, 0.0, 0.0], k=[2e-3, 0, 0, 0], w=[5.0, 0, 0, 0])
    if sst:
        c.sst_enable(oracle.default_sst())
        f.sst_enable(oa.default_sst())
    return c, dm, f, m


def both(c, f, name, v):
    c.field(name)[:] = v
    f.upload(name, v)


def state(c, f, rng):
    nn = c.nn
    U = rng.normal(scale=0.3, size=(nn, 3))
    U[:, 0] += 0.5
    both(c, f, "U", U)
    both(c, f, "p", rng.normal(scale=0.1, size=nn))
    c.initialize()
    f.initialize()
    for name in ("msd", "div", "gradU", "gradP", "mdot"):
        assert rel_err(f.download(name), c.field(name)) < 1e-12, name
    both(c, f, "gradP", rng.normal(scale=0.2, size=(nn, 3)))
    both(c, f, "gradU", rng.normal(size=(nn, 9)))
    both(c, f, "du", rng.uniform(0.01, 0.05, size=(nn, 3)))
    both(c, f, "mdot", rng.normal(scale=0.05, size=(c.ne, 12)))
    msd = c.field("msd").copy()
    msd += rng.normal(scale=0.05, size=msd.shape) * (msd != 0)
    both(c, f, "msd", msd)
    c.update_reversal()
    f.update_reversal()
    np.testing.assert_array_equal(f.reversal(), c.reversal())
    c.mass_divergence()
    f.mass_divergence()
    for name in ("msd", "div"):
        assert rel_err(f.download(name), c.field(name)) < 1e-12, name


@pytest.mark.parametrize("topo", [1, 2, 3])
@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("shifted", [0, 1])
def test_side_fields_and_systems(oracle, oa, topo, variant, shifted):
    c, dm, f, m = pair(oracle, oa, topo, variant, shifted)
    rng = np.random.default_rng(100 + 10 * topo + variant)
    state(c, f, rng)
    c.update_mdot()
    f.update_mdot()
    assert rel_err(f.download("msd"), c.field("msd")) < 1e-12
    for which, name in (("U", "gradU"), ("p", "gradP")):
        c.gradient(which)
        f.gradient(which)
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


@pytest.mark.parametrize("topo", [1, 2, 3])
def test_side_blending(oracle, oa, topo):
    c, dm, f, m = pair(oracle, oa, topo, 0, 0, high_res=1)
    state(c, f, np.random.default_rng(7 + topo))
    c.update_blending("U")
    f.update_blending("U")
    assert rel_err(f.download("beta"), c.field("beta")) < 1e-12


@pytest.mark.parametrize("topo", [1, 2, 3])
@pytest.mark.parametrize("which", ["k", "omega"])
def test_side_scalar_systems(oracle, oa, topo, which):
    c, dm, f, m = pair(oracle, oa, topo, 1, 0, sst=True)
    rng = np.random.default_rng(40 + topo)
    state(c, f, rng)
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
    assert row_rel_err(v, c.vals(1), c.row_ptr) < 1e-12
    assert entry_rel_err(b, c.rhs(1)) < 1e-12
