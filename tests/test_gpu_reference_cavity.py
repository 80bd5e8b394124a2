"""The reference's own cavity case on the device path: examples/cavity/mesh.e
(tests/golden/cavity_mesh.e, 20 x 20 x 1 hex8 whose elements come in four
orientations) read by openaccel_amd.exodus, with the controls of
examples/cavity/input.i: rho = 1, mu = 0.01, physical timescale 1, velocity
relaxation 0.9, pressure relaxation 0.1, upwind advection, linear_linear
velocity interpolation, the pressure level at the node nearest (0, 0, 0);
TOP a no-slip wall moving at (1, 0, 0), SIDES no-slip walls, FRONTANDBACK
symmetry planes.  Compared with the oracle on the same inputs (p' by
Jacobi-CG in both, so the Krylov iterations match):

* graph and scatter maps bit-exact;
* FP64 momentum and p' systems row-relative 1e-12, rhs per entry 1e-12;
* 10 outer iterations from the reference's initial state (U = 0, p = 0):
  control residuals within 1e-8, fields within 1e-8;
* the mixed form's control residuals within 1e-3 of the FP64 form's.
"""
import os

import numpy as np
import pytest

from conftest import entry_rel_err, rel_err, row_rel_err

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MESH = os.path.join(HERE, "golden", "cavity_mesh.e")
BC_U = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])   # TOP, SIDES, FRONTANDBACK
BC_TYPES = [0, 0, 3]
CTL = [20, 0.1, 200, 1e-6]   # linear_solver_settings: default max 20 rtol 0.1; p' max 200 rtol 1e-6


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


@pytest.fixture(scope="module")
def mesh():
    from openaccel_amd import exodus
    m = exodus.read_exodus(MESH)
    m["p_ref"] = exodus.nearest_node(m["xyz"], (0.0, 0.0, 0.0))
    return m


def params(mod, m, precision=None):
    kw = dict(rho=1.0, mu=0.01, dt=1.0, urf_U=0.9, urf_p=0.1, u_shifted=1, p_ref_node=m["p_ref"])
    if precision is None:
        return mod.default_params(**kw)
    return mod.default_params(precision, p_method=mod.OAGPU_CG, **kw)


def pair(oracle, oa, mesh, precision=0):
    print(f"[case] reference cavity mesh.e precision={precision}", flush=True)
    c = oracle.Case(mesh, params(oracle, mesh), BC_U)
    c.set_bc_types(BC_TYPES)
    dm = oa.DeviceMesh(mesh)
    f = oa.FlowCase(dm, params(oa, mesh, precision), BC_U)
    f.set_bc_types(BC_TYPES)
    return c, dm, f


def test_graph_bit_exact(oracle, oa, mesh):
    c, dm, f = pair(oracle, oa, mesh)
    rp, cols, doff = dm.graph()
    assert np.array_equal(rp, c.row_ptr) and np.array_equal(cols, c.cols)
    assert np.array_equal(doff, c.diag_off)
    assert rel_err(dm.dual_volume(), c.field("vol")) < 1e-13


def test_assembly_fp64(oracle, oa, mesh):
    c, dm, f = pair(oracle, oa, mesh)
    rng = np.random.default_rng(5)
    nn = c.nn
    U = rng.normal(scale=0.3, size=(nn, 3))
    p = rng.normal(scale=0.1, size=nn)
    for name, v in (("U", U), ("p", p)):
        c.field(name)[:] = v
        f.upload(name, v)
    c.initialize()
    f.initialize()
    for name in ("gradU", "gradP", "mdot", "div"):
        assert entry_rel_err(f.download(name), c.field(name)) < 1e-12, name
    c.assemble_momentum()
    c.compute_du()
    c.normal_relaxation()
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


def test_outer_iterations_from_rest(oracle, oa, mesh):
    c, dm, f = pair(oracle, oa, mesh)
    c.initialize()
    f.initialize()
    hist = []
    for it in range(10):
        ho = c.outer_iteration(CTL)
        hg = f.outer_iteration(CTL)
        np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8)
        assert np.abs(hg[4:] - ho[4:]).max() <= 1, it
        hist.append(hg)
    assert hist[-1][0] < hist[0][0]          # the lid drives the flow and residuals fall
    for name in ("U", "p", "mdot", "gradU", "gradP"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name


def test_mixed_form_tracks_fp64(oa, mesh):
    hist = {}
    for prec in (0, 1):
        dm = oa.DeviceMesh(mesh)
        f = oa.FlowCase(dm, params(oa, mesh, prec), BC_U)
        f.set_bc_types(BC_TYPES)
        f.initialize()
        hist[prec] = np.array([f.outer_iteration([20, 0.1, 400, 1e-8])[:4] for _ in range(8)])
    np.testing.assert_allclose(hist[1], hist[0], rtol=1e-3)


def test_history_and_system_dump(oracle, oa, mesh, tmp_path):
    """The device path's residual histories (momentum and p', the reference's
    writeHistory_ format) and a write_system dump of the last momentum system,
    read back: equal to the outer iterations' control residuals (6 digits) and
    to the oracle's system assembled from the same state (1e-12)."""
    from openaccel_amd import sysio
    c, dm, f = pair(oracle, oa, mesh)
    c.initialize()
    f.initialize()
    rdir = str(tmp_path / "0" / "residuals")
    hu = sysio.ResidualHistory(rdir, "momentum equations", ["U-momentum", "V-momentum", "W-momentum"])
    hp = sysio.ResidualHistory(rdir, "pressure correction", ["P-correction"])
    hist = []
    for it in range(1, 4):
        h = f.outer_iteration(CTL)
        c.outer_iteration(CTL)
        hu.write(it, it, int(h[4]), 0.0, h[:3])
        hp.write(it, it, int(h[5]), 0.0, h[3:4])
        hist.append(h)
    hist = np.array(hist)
    for hf, cols in ((hu, slice(0, 3)), (hp, slice(3, 4))):
        names, rows = sysio.read_history(hf.path)
        np.testing.assert_allclose(rows[:, 5:], hist[:, cols], rtol=1e-6)
        np.testing.assert_array_equal(rows[:, 0], [1, 2, 3])
    for name in ("U", "p", "mdot", "gradU", "gradP", "div"):
        f.upload(name, c.field(name))
    c.assemble_momentum()
    c.compute_du()
    c.normal_relaxation()
    f.assemble_momentum()
    a_path, b_path = sysio.write_flow_system(f, 3, str(tmp_path / "momentum_equations"), 4)
    rp, cols, vals, b = sysio.read_system(a_path, b_path)
    assert np.array_equal(rp, c.row_ptr) and np.array_equal(cols, c.cols)
    assert row_rel_err(vals, c.vals(3), c.row_ptr) < 1e-12
    assert entry_rel_err(b, c.rhs(3)) < 1e-12
