"""GPU parity on hex8 meshes whose elements are re-oriented: every element's
local node order is changed by a random proper cube rotation
(meshgen.relabel_hex), as in meshes from unstructured meshers, so a node is
the same local node of several elements and the row kernels walk their
extra-incidence list as well as the local-node-major one.

Bars (FP64 form, against the oracle on the same relabelled mesh):
* graph, columns, diagonal offsets and scatter maps bit-exact;
* every matrix row within 1e-12 of that row's largest coefficient; rhs,
  du, mass flow rates, divergence and gradients within 1e-12 per entry
  (entries below 1e-3 of the array's scale: 1e-15 of that scale);
* five outer iterations: control residuals within 1e-8, fields within 1e-8;
* the mixed form's coefficients within 2e-5 (row-relative) of the FP64 form.
"""
import numpy as np
import pytest

from conftest import entry_rel_err, rel_err, row_rel_err

pytestmark = pytest.mark.gpu

BC = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
CASES = [dict(n=6, perturb=0.2, seed=3, rot=11), dict(n=5, perturb=0.1, seed=1, rot=12, reduced=0),
         dict(n=7, perturb=0.15, seed=2, rot=13, u_shifted=1)]


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def pair(oracle, oa, meshgen, n, perturb, seed, rot, reduced=1, precision=0, high_res=0,
         sst=False, **kw):
    print(f"[case] rotated n={n} perturb={perturb} seed={seed} rot={rot} reduced={reduced} "
          f"precision={precision} high_res={high_res} sst={sst} {kw}", flush=True)
    m = meshgen.relabel_hex(meshgen.cavity_hex(n, perturb=perturb, seed=seed), rot)
    assert meshgen.repeated_local_nodes(m["conn"], m["xyz"].shape[0]) > 0
    c = oracle.Case(m, oracle.default_params(reduced=reduced, high_res=high_res, **kw), BC)
    dm = oa.DeviceMesh(m, oa.OAGPU_STENCIL_REDUCED if reduced else oa.OAGPU_STENCIL_FULL)
    f = oa.FlowCase(dm, oa.default_params(precision, high_res=high_res, **kw), BC)
    if sst:
        c.sst_enable(oracle.default_sst())
        f.sst_enable(oa.default_sst())
    return c, dm, f, m


def sync_state(c, f, rng, sst=False):
    nn = c.nn
    U = rng.normal(scale=0.3, size=(nn, 3))
    p = rng.normal(scale=0.1, size=nn)
    for name, v in (("U", U), ("p", p)):
        c.field(name)[:] = v
        f.upload(name, v)
    c.initialize()
    f.initialize()
    vals = {"gradP": rng.normal(scale=0.2, size=(nn, 3)), "du": rng.uniform(1e-3, 2e-3, size=(nn, 3)),
            "gradU": rng.normal(size=(nn, 9))}
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


@pytest.mark.parametrize("case", CASES)
def test_graph_and_scatter_map_bit_exact(oracle, oa, meshgen, case):
    c, dm, f, m = pair(oracle, oa, meshgen, **case)
    rp, cols, doff = dm.graph()
    assert np.array_equal(rp, c.row_ptr)
    assert np.array_equal(cols, c.cols)
    assert np.array_equal(doff, c.diag_off)
    ptr, inc, slots = dm.scatter_map()
    conn = m["conn"]
    L = oracle.lib()
    crp, ccols = np.ascontiguousarray(c.row_ptr), np.ascontiguousarray(c.cols)
    for a in range(c.nn):
        for k in range(ptr[a], ptr[a + 1]):
            e, la = inc[k] >> 3, inc[k] & 7
            assert conn[e, la] == a
            for cn in range(8):
                s = int((int(slots[k]) >> (8 * cn)) & 0xFF)
                want = L.or_graph_find(oracle.iptr(crp), oracle.iptr(ccols), a, int(conn[e, cn]))
                assert s == (0xFF if want < 0 else want)


@pytest.mark.parametrize("case", CASES)
def test_assembly_fp64(oracle, oa, meshgen, case):
    c, dm, f, m = pair(oracle, oa, meshgen, **case)
    assert rel_err(dm.dual_volume(), c.field("vol")) < 1e-13
    sync_state(c, f, np.random.default_rng(21))
    c.assemble_momentum()
    c.compute_du()
    f.assemble_momentum()
    vals, b = f.system(3)
    assert row_rel_err(vals, c.vals(3), c.row_ptr) < 1e-12
    assert entry_rel_err(b, c.rhs(3)) < 1e-12
    assert entry_rel_err(f.download("du"), c.field("du")) < 1e-12
    c.assemble_pcorr()
    f.assemble_pcorr()
    vals, b = f.system(1)
    assert row_rel_err(vals, c.vals(1), c.row_ptr) < 1e-12
    assert entry_rel_err(b, c.rhs(1)) < 1e-12


@pytest.mark.parametrize("case", CASES)
def test_field_updates(oracle, oa, meshgen, case):
    c, dm, f, m = pair(oracle, oa, meshgen, **case)
    sync_state(c, f, np.random.default_rng(22))
    c.update_mdot()
    f.update_mdot()
    assert entry_rel_err(f.download("mdot"), c.field("mdot")) < 1e-12
    c.mass_divergence()
    f.mass_divergence()
    assert rel_err(f.download("div"), c.field("div")) < 1e-12
    for w, name in (("U", "gradU"), ("p", "gradP")):
        c.gradient(w)
        f.gradient(w)
        assert entry_rel_err(f.download(name), c.field(name)) < 1e-12, name
    for step in (1, 2):
        oracle.lib().or_velocity_correction(c.h, step)
        f.velocity_correction(step)
    assert entry_rel_err(f.download("U"), c.field("U")) < 1e-12


def test_high_resolution_and_sst_rows(oracle, oa, meshgen):
    c, dm, f, m = pair(oracle, oa, meshgen, n=6, perturb=0.15, seed=4, rot=14, high_res=1,
                       sst=True)
    sync_state(c, f, np.random.default_rng(23), sst=True)
    md = np.random.default_rng(24).normal(scale=0.05, size=(c.ne, 12))
    c.field("mdot")[:] = md
    f.upload("mdot", md)
    c.mass_divergence()
    f.mass_divergence()
    for w, name in (("U", "beta"), ("k", "beta_k"), ("omega", "beta_w")):
        c.update_blending(w)
        f.update_blending(w)
        assert np.abs(f.download(name) - c.field(name)).max() < 1e-12, name
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


@pytest.mark.parametrize("case", CASES[:2])
def test_outer_iterations(oracle, oa, meshgen, case):
    c, dm, f, m = pair(oracle, oa, meshgen, **case)
    U = meshgen.initial_velocity(c.nn, 0)
    c.field("U")[:] = U
    f.upload("U", U)
    c.initialize()
    f.initialize()
    ctl = [20, 0.1, 200, 1e-6]
    for _ in range(5):
        ho = c.outer_iteration(ctl)
        hg = f.outer_iteration(ctl)
        np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8)
        assert np.abs(hg[4:] - ho[4:]).max() <= 1
    for name in ("U", "p", "mdot", "gradU", "gradP"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name


def test_mixed_coefficients(oracle, oa, meshgen):
    """FP32 local arithmetic summed in FP64 on re-oriented elements: 2e-5."""
    out = {}
    for prec in (0, 1):
        c, dm, f, m = pair(oracle, oa, meshgen, n=8, perturb=0.2, seed=21, rot=15, precision=prec)
        sync_state(c, f, np.random.default_rng(25))
        f.assemble_momentum()
        f.assemble_pcorr()
        out[prec] = (*f.system(3), *f.system(1))
    rp = c.row_ptr
    assert row_rel_err(out[1][0], out[0][0], rp) < 2e-5
    assert row_rel_err(out[1][2], out[0][2], rp) < 2e-5
    assert rel_err(out[1][1], out[0][1]) < 2e-5
    assert rel_err(out[1][3], out[0][3]) < 2e-5
