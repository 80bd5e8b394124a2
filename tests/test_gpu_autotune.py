"""The autotuner (oagpu_autotune / oagpu_tune_set): every tunable variant
gives the default variant's results -- FP64 momentum and p' systems within
1e-12 (row-relative) of the oracle for each momentum accumulation and p'
occupancy variant, Krylov solutions with every SpMV block size equal to the
default's to rounding -- and the tuner's report names variants from the
candidate sets with positive times."""
import numpy as np
import pytest

from conftest import entry_rel_err, rel_err, row_rel_err

pytestmark = pytest.mark.gpu

BC = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def pair(oracle, oa, meshgen, precision=0):
    m = meshgen.cavity_hex(10, perturb=0.15, seed=8)
    c = oracle.Case(m, oracle.default_params(), BC)
    dm = oa.DeviceMesh(m)
    f = oa.FlowCase(dm, oa.default_params(precision), BC)
    rng = np.random.default_rng(2)
    U = rng.normal(scale=0.3, size=(c.nn, 3))
    for case in (c, f):
        if case is c:
            c.field("U")[:] = U
        else:
            f.upload("U", U)
        case.initialize()
    return c, dm, f


@pytest.mark.parametrize("mom,pc", [(0, 0), (1, 1), (0, 1)])
def test_assembly_variants_match_oracle(oracle, oa, meshgen, mom, pc):
    print(f"[case] tunable variants momentum={mom} pcorr={pc}", flush=True)
    c, dm, f = pair(oracle, oa, meshgen)
    f.tune_set("momentum", mom)
    f.tune_set("pcorr", pc)
    c.assemble_momentum()
    c.compute_du()
    f.assemble_momentum()
    vals, b = f.system(3)
    assert row_rel_err(vals, c.vals(3), c.row_ptr) < 1e-12
    assert entry_rel_err(b, c.rhs(3)) < 1e-12
    c.assemble_pcorr()
    f.assemble_pcorr()
    vals, b = f.system(1)
    assert row_rel_err(vals, c.vals(1), c.row_ptr) < 1e-12
    assert entry_rel_err(b, c.rhs(1)) < 1e-12


def test_spmv_block_sizes_and_report(oracle, oa, meshgen):
    print("[case] SpMV block sizes and the tuner's report", flush=True)
    sols = []
    for blk in (64, 128, 256):
        c, dm, f = pair(oracle, oa, meshgen, precision=1)
        for k in ("spmv1", "spmv3", "spmv1_f32", "spmv3_f32"):
            f.tune_set(k, blk)
        f.assemble_momentum()
        cd3 = f.solve(3, oa.OAGPU_GMRES, 30, 1e-6)
        x3 = f.solution(3)
        f.assemble_pcorr()
        cd1 = f.solve(1, oa.OAGPU_BICGSTAB_AMG, 100, 1e-8)
        sols.append((x3, f.solution(1), cd3.n_iterations, cd1.n_iterations))
    for s in sols[1:]:
        assert rel_err(s[0], sols[0][0]) < 1e-12 and rel_err(s[1], sols[0][1]) < 1e-12
        assert s[2:] == sols[0][2:]
    rep = f.autotune()
    assert rep["momentum"][0] in (0, 1) and rep["pcorr"][0] in (0, 1)
    for k in ("spmv1", "spmv3", "spmv1_f32", "spmv3_f32"):
        assert rep[k][0] in (64, 128, 256) and rep[k][1] > 0 and rep[k][2] >= rep[k][1]
    with pytest.raises(oa.OagpuError):
        f.tune_set("spmv1", 100)


def test_occupancy_caps_bitwise_and_state_kept(oracle, oa, meshgen):
    """Capping the resident waves of the mass-flow update and the gradients
    changes no bit of their results, and the tuner, which times the
    mass-flow update on the flow's state, leaves the stored rates as they
    were."""
    print("[case] occupancy caps of the mass-flow update and gradients", flush=True)
    c, dm, f = pair(oracle, oa, meshgen)
    f.assemble_momentum()
    md0 = f.download("mdot")
    out = []
    for cap in (0, 4, 8, 16):
        f.tune_set("mdot", cap)
        f.tune_set("grad", cap)
        f.upload("mdot", md0)
        f.update_mdot()
        f.gradient("p")
        out.append((f.download("mdot"), f.download("gradP")))
    for o in out[1:]:
        np.testing.assert_array_equal(o[0], out[0][0])
        np.testing.assert_array_equal(o[1], out[0][1])
    f.upload("mdot", md0)
    rep = f.autotune()
    np.testing.assert_array_equal(f.download("mdot"), md0)
    for k in ("mdot", "grad"):
        assert rep[k][0] in (0, 4, 8, 16) and rep[k][1] > 0
    with pytest.raises(oa.OagpuError):
        f.tune_set("mdot", 12)
