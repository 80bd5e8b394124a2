"""GPU parity of velocity inlets and static-pressure outlets through the C ABI
against the oracle on identical seeded inputs: a channel, the unit cube with
a velocity inlet at x = 0 (id 0), a static-pressure outlet at x = 1 (id 1)
and symmetry planes on the four other faces (id 2); one variant puts
no-slip walls on y = 0 and y = 1 instead (id 3); and a coarse
backward-facing step with the k-omega SST model.

Bars: side mass flows (initial and Rhie-Chow updated), reversal flags and
mass divergence within 1e-12 relative / exact for the flags; FP64 momentum,
p' and k / omega systems 1e-12; flow (and SST) outer iterations 1e-8 on
fields and control residuals; the mixed form's momentum and p' systems
within 2e-5 of the FP64 form.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

U0, P_OUT, K_IN, W_IN = 1.0, 0.1, 2.0e-3, 5.0
INLET, OUTLET, SYM, WALL = 1, 2, 3, 0


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def channel(meshgen, n, perturb=0.0, seed=0, walls=False):
    m = meshgen.cavity_hex(n, perturb=perturb, seed=seed)
    bc = np.full(m["bf_bc"].shape, 2, dtype=np.int32)
    bc[m["bf_ord"] == 3] = 0
    bc[m["bf_ord"] == 1] = 1
    if walls:
        bc[(m["bf_ord"] == 0) | (m["bf_ord"] == 2)] = 3
    m["bf_bc"] = np.ascontiguousarray(bc)
    return m


def pair(oracle, oa, meshgen, n=6, perturb=0.0, seed=0, reduced=1, precision=0, sst=False,
         walls=False):
    print(f"[case] inlet/outlet n={n} perturb={perturb} seed={seed} reduced={reduced} "
          f"precision={precision} sst={sst} walls={walls}", flush=True)
    m = channel(meshgen, n, perturb, seed, walls)
    BC = np.array([[U0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    types = [INLET, OUTLET, SYM, WALL]
    bcv = dict(p=[0.0, P_OUT, 0.0, 0.0], k=[K_IN, 0.0, 0.0, 0.0], w=[W_IN, 0.0, 0.0, 0.0])
    c = oracle.Case(m, oracle.default_params(reduced=reduced), BC)
    c.set_bc_types(types)
    c.set_bc_values(**bcv)
    dm = oa.DeviceMesh(m, oa.OAGPU_STENCIL_REDUCED if reduced else oa.OAGPU_STENCIL_FULL)
    f = oa.FlowCase(dm, oa.default_params(precision), BC)
    f.set_bc_types(types)
    f.set_bc_values(**bcv)
    if sst:
        c.sst_enable(oracle.default_sst())
        f.sst_enable(oa.default_sst())
    return c, dm, f, m


def both(c, f, name, v):
    c.field(name)[:] = v
    f.upload(name, v)


def sync_state(c, f, m, rng, reversed_faces=True):
    """Random flow state on both sides; with reversed_faces, some outlet
    faces are given a net inflow and flagged by the reversal update."""
    nn = c.nn
    U = rng.normal(scale=0.3, size=(nn, 3))
    U[:, 0] += U0
    both(c, f, "U", U)
    both(c, f, "p", P_OUT + rng.normal(scale=0.1, size=nn))
    c.initialize()
    f.initialize()
    both(c, f, "gradP", rng.normal(scale=0.2, size=(nn, 3)))
    both(c, f, "gradU", rng.normal(size=(nn, 9)))
    both(c, f, "du", rng.uniform(0.01, 0.05, size=(nn, 3)))
    both(c, f, "mdot", rng.normal(scale=0.05, size=(c.ne, 12)))
    msd = c.field("msd").copy()
    msd += rng.normal(scale=0.01, size=msd.shape) * (msd != 0)
    if reversed_faces:
        out = np.where(m["bf_ord"] == 1)[0]
        msd[out[::3]] = -np.abs(msd[out[::3]]) - 0.01
    both(c, f, "msd", msd)
    if reversed_faces:
        c.update_reversal()
        f.update_reversal()
    c.mass_divergence()
    f.mass_divergence()


def test_side_mass_flow_init(oracle, oa, meshgen):
    c, dm, f, m = pair(oracle, oa, meshgen, n=5, perturb=0.15, seed=1)
    rng = np.random.default_rng(1)
    U = rng.normal(scale=0.3, size=(c.nn, 3))
    U[:, 0] += U0
    both(c, f, "U", U)
    c.initialize()
    f.initialize()
    ref = c.field("msd")
    assert np.abs(ref).max() > 0
    assert rel_err(f.download("msd"), ref) < 1e-12
    assert rel_err(f.download("div"), c.field("div")) < 1e-12


@pytest.mark.parametrize("case", [dict(n=5), dict(n=6, perturb=0.2, seed=3)])
def test_reversal_and_side_update(oracle, oa, meshgen, case):
    c, dm, f, m = pair(oracle, oa, meshgen, **case)
    sync_state(c, f, m, np.random.default_rng(2))
    rf = c.reversal()
    assert rf.sum() > 0 and not rf.all()
    np.testing.assert_array_equal(f.reversal(), rf)
    assert rel_err(f.download("msd"), c.field("msd")) < 1e-12
    assert rel_err(f.download("div"), c.field("div")) < 1e-12
    c.update_mdot()
    f.update_mdot()
    assert rel_err(f.download("msd"), c.field("msd")) < 1e-12
    assert rel_err(f.download("mdot"), c.field("mdot")) < 1e-12


@pytest.mark.parametrize("case", [dict(n=5), dict(n=6, perturb=0.2, seed=3),
                                  dict(n=4, perturb=0.1, seed=2, reduced=0),
                                  dict(n=5, perturb=0.1, seed=4, walls=True)])
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
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, perturb=0.15, seed=5, sst=True, walls=True)
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


def developing_flow(c, f, m):
    """Uniform inflow speed in the whole channel, p = p_out: the flow
    develops from the symmetry / wall boundaries and the outlet pressure."""
    U = np.zeros((c.nn, 3))
    U[:, 0] = U0
    x = m["xyz"]
    U[:, 1] = 0.05 * np.sin(np.pi * x[:, 0]) * np.sin(np.pi * x[:, 2])
    both(c, f, "U", U)
    both(c, f, "p", np.full(c.nn, P_OUT))
    c.initialize()
    f.initialize()


@pytest.mark.parametrize("walls", [False, True])
def test_outer_iterations(oracle, oa, meshgen, walls):
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, perturb=0.1, seed=6, walls=walls)
    developing_flow(c, f, m)
    ctl = [20, 0.1, 200, 1e-6]
    for it in range(3):
        ho, hg = c.outer_iteration(ctl), f.outer_iteration(ctl)
        np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8)
    for name in ("U", "p", "gradU", "gradP", "mdot", "msd"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name
    np.testing.assert_array_equal(f.reversal(), c.reversal())
    # mass balance: the outlet takes what the inlet brings, to the p' tolerance
    md = f.download("msd")
    ords = m["bf_ord"]
    q_in, q_out = -md[ords == 3].sum(), md[ords == 1].sum()
    assert q_in > 0 and abs(q_out - q_in) < 0.2 * q_in


def test_sst_iterations(oracle, oa, meshgen):
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, perturb=0.1, seed=7, sst=True, walls=True)
    developing_flow(c, f, m)
    x = m["xyz"]
    init = {"k": np.full(c.nn, K_IN), "omega": np.full(c.nn, W_IN),
            "ymin": np.maximum(np.minimum(x[:, 1], 1 - x[:, 1]), 1e-3)}
    for name, v in init.items():
        both(c, f, name, v)
    ctl, sctl = [20, 0.1, 200, 1e-6], [30, 1e-6]
    for it in range(2):
        ho, hg = c.outer_iteration(ctl), f.outer_iteration(ctl)
        np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8)
        so, sg = c.sst_iteration(sctl), f.sst_iteration(sctl)
        np.testing.assert_allclose(sg[:2], so[:2], rtol=1e-8)
    for name in ("U", "p", "k", "omega", "mut"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name


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


def bfs_case(oracle, oa, meshgen, precision=0):
    """A coarse backward-facing step (inlet, outlet, no-slip floor / step /
    roof, symmetry sides) with the pitzDaily controls of the reference's
    examples (velocity / pressure / turbulence relaxation 0.7 / 0.3 / 0.9,
    linear_linear velocity interpolation)."""
    m = meshgen.bfs_hex(3, 8, 4, 4, 3, ratio=1.2)
    print(f"[case] backward-facing step {m['conn'].shape[0]} elements precision={precision}",
          flush=True)
    BC = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3, [0.0] * 3])
    types = [INLET, OUTLET, WALL, SYM]
    kin, win = 1.5e-4, 0.75
    bcv = dict(p=[0.0] * 4, k=[kin, 0, 0, 0], w=[win, 0, 0, 0])
    # p' by BiCGStab: the outlet makes the p' matrix nonsymmetric
    kw = dict(mu=1e-3, urf_U=0.7, urf_p=0.3, u_shifted=1, p_method=2)
    c = oracle.Case(m, oracle.default_params(**kw), BC)
    c.set_bc_types(types)
    c.set_bc_values(**bcv)
    c.sst_enable(oracle.default_sst(urf_k=0.9, urf_w=0.9))
    dm = oa.DeviceMesh(m)
    f = oa.FlowCase(dm, oa.default_params(precision, **kw), BC)
    f.set_bc_types(types)
    f.set_bc_values(**bcv)
    f.sst_enable(oa.default_sst(urf_k=0.9, urf_w=0.9))
    U = np.zeros((c.nn, 3))
    U[:, 0] = 1.0
    y = meshgen.bfs_wall_distance(m["xyz"])
    init = {"U": U, "p": np.zeros(c.nn), "k": np.full(c.nn, kin), "omega": np.full(c.nn, win),
            "ymin": y}
    for name, v in init.items():
        both(c, f, name, v)
    c.initialize()
    f.initialize()
    return c, dm, f, m


def test_bfs_sst_iterations(oracle, oa, meshgen):
    """Tight linear solves, p' by BiCGStab: the step is mirror-symmetric in
    z, and an unconverged Krylov solve (Jacobi-CG stalls on this
    nonsymmetric p' system) leaves a z-antisymmetric part driven by rounding
    whose sign differs between the two implementations (their states come
    out as near mirror images with equal residual norms); solved tightly,
    both give the same solution."""
    c, dm, f, m = bfs_case(oracle, oa, meshgen)
    ctl, sctl = [60, 1e-8, 1000, 1e-10], [60, 1e-10]
    for it in range(3):
        # v and w start at zero: their first residuals are rounding (1e-20)
        ho, hg = c.outer_iteration(ctl), f.outer_iteration(ctl)
        np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8, atol=1e-14)
        so, sg = c.sst_iteration(sctl), f.sst_iteration(sctl)
        np.testing.assert_allclose(sg[:2], so[:2], rtol=1e-8, atol=1e-14)
    for name in ("U", "p", "k", "omega", "mut", "msd"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name
    np.testing.assert_array_equal(f.reversal(), c.reversal())
