"""GPU parity of static-pressure openings (OAGPU_BC_OPENING) through the C
ABI against the oracle on identical seeded inputs: a channel, the unit cube
with a velocity inlet at x = 0 (id 0), an opening at x = 1 (id 1), no-slip
walls at y = 0 / 1 (id 2) and symmetry planes at z = 0 / 1 (id 3); the
flow leaves through part of the opening and is entrained through the rest
(along a cartesian direction, or, for a zero bc_U row, the inward normal).
A mixed-topology duct (tet4 / wedge6 / pyramid5 / hex8) ending in an
opening covers tri sides.

Bars: side mass flows (initial, not clipped, and Rhie-Chow updated with the
nodal pressure) and mass divergence within 1e-12 relative; no reversal
flags; FP64 momentum, du, p' and k / omega systems 1e-12; outer
iterations 1e-8 on fields and control residuals; the mixed form's momentum
and p' systems within 2e-5 of the FP64 form.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

U0, P_OPEN, K_IN, W_IN = 1.0, 0.1, 2.0e-3, 5.0
INLET, OPENING, WALL, SYM = 1, 5, 0, 3
DIRS = {"cartesian": [-1.0, 0.3, 0.1], "normal": [0.0, 0.0, 0.0]}


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def channel(meshgen, n, perturb, seed):
    m = meshgen.cavity_hex(n, perturb=perturb, seed=seed)
    bc = np.full(m["bf_bc"].shape, 3, dtype=np.int32)
    bc[m["bf_ord"] == 3] = 0
    bc[m["bf_ord"] == 1] = 1
    bc[(m["bf_ord"] == 0) | (m["bf_ord"] == 2)] = 2
    m["bf_bc"] = np.ascontiguousarray(bc)
    return m


def setup(oracle, oa, m, BC, types, bcv, precision=0, sst=False, reduced=1):
    c = oracle.Case(m, oracle.default_params(reduced=reduced), BC)
    dm = oa.DeviceMesh(m, oa.OAGPU_STENCIL_REDUCED if reduced else oa.OAGPU_STENCIL_FULL)
    f = oa.FlowCase(dm, oa.default_params(precision), BC)
    for x in (c, f):
        x.set_bc_types(types)
        x.set_bc_values(**bcv)
    if sst:
        c.sst_enable(oracle.default_sst())
        f.sst_enable(oa.default_sst())
    return c, dm, f


def pair(oracle, oa, meshgen, n=6, perturb=0.15, seed=0, direction="cartesian", precision=0,
         sst=False, reduced=1):
    print(f"[case] opening channel n={n} perturb={perturb} seed={seed} dir={direction} "
          f"precision={precision} sst={sst} reduced={reduced}", flush=True)
    m = channel(meshgen, n, perturb, seed)
    BC = np.array([[U0, 0.0, 0.0], DIRS[direction], [0.0] * 3, [0.0] * 3])
    bcv = dict(p=[0.0, P_OPEN, 0.0, 0.0], k=[K_IN, 0, 0, 0], w=[W_IN, 0, 0, 0])
    c, dm, f = setup(oracle, oa, m, BC, [INLET, OPENING, WALL, SYM], bcv, precision, sst,
                     reduced)
    return c, dm, f, m


def both(c, f, name, v):
    c.field(name)[:] = v
    f.upload(name, v)


def two_way_u(nn, xyz, rng):
    """x-velocity leaving for y < 0.5 and entering above, plus noise"""
    U = rng.normal(scale=0.1, size=(nn, 3))
    U[:, 0] += U0 * (1.0 - 2.0 * xyz[:, 1])
    return U


def sync_state(c, f, m, rng):
    nn = c.nn
    both(c, f, "U", two_way_u(nn, m["xyz"], rng))
    both(c, f, "p", P_OPEN + rng.normal(scale=0.1, size=nn))
    c.initialize()
    f.initialize()
    both(c, f, "gradP", rng.normal(scale=0.2, size=(nn, 3)))
    both(c, f, "gradU", rng.normal(size=(nn, 9)))
    both(c, f, "du", rng.uniform(0.01, 0.05, size=(nn, 3)))
    both(c, f, "mdot", rng.normal(scale=0.05, size=(c.ne, 12)))
    msd = c.field("msd").copy()
    msd += rng.normal(scale=0.01, size=msd.shape) * (msd != 0)
    both(c, f, "msd", msd)
    c.mass_divergence()
    f.mass_divergence()


def test_side_mass_flow(oracle, oa, meshgen):
    c, dm, f, m = pair(oracle, oa, meshgen, n=5, seed=1)
    rng = np.random.default_rng(1)
    both(c, f, "U", two_way_u(c.nn, m["xyz"], rng))
    both(c, f, "p", P_OPEN + rng.normal(scale=0.1, size=c.nn))
    c.initialize()
    f.initialize()
    ref = c.field("msd")
    op = ref[m["bf_ord"] == 1]
    assert op.min() < 0.0 < op.max()
    assert rel_err(f.download("msd"), ref) < 1e-12
    assert rel_err(f.download("div"), c.field("div")) < 1e-12
    sync_state(c, f, m, rng)
    c.update_reversal()
    f.update_reversal()
    assert not f.reversal().any() and not c.reversal().any()
    c.update_mdot()
    f.update_mdot()
    assert rel_err(f.download("msd"), c.field("msd")) < 1e-12
    assert rel_err(f.download("mdot"), c.field("mdot")) < 1e-12


@pytest.mark.parametrize("case", [dict(direction="cartesian"), dict(direction="normal", seed=3),
                                  dict(n=4, seed=2, reduced=0)])
def test_systems_fp64(oracle, oa, meshgen, case):
    c, dm, f, m = pair(oracle, oa, meshgen, **case)
    sync_state(c, f, m, np.random.default_rng(3))
    c.assemble_momentum()
    c.compute_du()
    c.normal_relaxation()
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
def test_scalar_systems_fp64(oracle, oa, meshgen, which):
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, seed=5, sst=True)
    rng = np.random.default_rng(5)
    sync_state(c, f, m, rng)
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


@pytest.mark.parametrize("direction", ["cartesian", "normal"])
def test_outer_iterations(oracle, oa, meshgen, direction):
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, perturb=0.1, seed=6, direction=direction)
    x = m["xyz"]
    U = np.zeros((c.nn, 3))
    U[:, 0] = U0 * (1.0 - 1.6 * x[:, 0] * x[:, 1])
    both(c, f, "U", U)
    both(c, f, "p", np.full(c.nn, P_OPEN))
    c.initialize()
    f.initialize()
    assert c.field("msd")[m["bf_ord"] == 1].min() < 0.0
    ctl = [20, 0.1, 200, 1e-6]
    for it in range(3):
        ho, hg = c.outer_iteration(ctl), f.outer_iteration(ctl)
        np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8)
    for name in ("U", "p", "gradU", "gradP", "mdot", "msd"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name
    assert not f.reversal().any()


def test_systems_mixed(oracle, oa, meshgen):
    """FP32 local arithmetic summed into FP64: stated tolerance 2e-5."""
    out = {}
    for prec in (0, 1):
        c, dm, f, m = pair(oracle, oa, meshgen, n=8, perturb=0.2, seed=8, precision=prec)
        sync_state(c, f, m, np.random.default_rng(9))
        f.assemble_momentum()
        out[prec, 3] = f.system(3)
        f.assemble_pcorr()
        out[prec, 1] = f.system(1)
    for N in (3, 1):
        assert rel_err(out[1, N][0], out[0, N][0]) < 2e-5
        assert rel_err(out[1, N][1], out[0, N][1]) < 2e-5


@pytest.mark.parametrize("precision", [0, 1])
def test_mixed_topology_duct(oracle, oa, meshgen, precision):
    """duct_mixed ending in wedge cubes (tri sides on the opening)"""
    M = 6
    layers = (0, 2, 2, 3, 0)
    m = meshgen.duct_mixed(M, layers=layers, length=sum(layers) * 0.2 / M, jitter=0.1, seed=2)
    print(f"[case] opening mixed duct layers={layers} precision={precision}", flush=True)
    BC = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3])
    c, dm, f = setup(oracle, oa, m, BC, [INLET, OPENING, WALL],
                     dict(p=[0.0, 0.0, 0.0], k=[2e-3, 0, 0], w=[5.0, 0, 0]), precision)
    rng = np.random.default_rng(4)
    U = rng.normal(scale=0.6, size=(c.nn, 3))
    U[:, 0] += 0.3
    both(c, f, "U", U)
    both(c, f, "p", rng.normal(scale=0.1, size=c.nn))
    c.initialize()
    f.initialize()
    assert rel_err(f.download("msd"), c.field("msd")) < 1e-12
    c.assemble_momentum()
    c.compute_du()
    c.normal_relaxation()
    f.assemble_momentum()
    tol = 1e-12 if precision == 0 else 2e-5
    v, b = f.system(3)
    assert rel_err(v, c.vals(3)) < tol
    assert rel_err(b, c.rhs(3)) < tol
    c.assemble_pcorr()
    f.assemble_pcorr()
    v, b = f.system(1)
    assert rel_err(v, c.vals(1)) < tol
    assert rel_err(b, c.rhs(1)) < tol
