"""The mixed form where FP32 cancellation is hardest: a spanwise slice of the
8M-element backward-facing step with that configuration's wall-normal grid
(96 + 112 cells graded 1.1 from each wall, first cell 4.8e-4), mu = 2e-5
and the k-omega SST model.  The flow state is two oracle outer iterations
(flow + SST) from the inlet state; on it, the device's FP64 momentum, p' and
k / omega systems equal the oracle's (row-relative 1e-12, rhs against the
flux scale 1e-12) and the mixed form's are within 2e-5 of the FP64 form's
(row-relative; rhs against the flux scale, the momentum components of a
node sharing their largest, p' against the node's SCS mass flows).
"""
import numpy as np
import pytest

from conftest import mass_rel_err, row_rel_err, rhs_rel_err

pytestmark = pytest.mark.gpu

MU, KIN = 2.0e-5, 1.5e-4
WIN = KIN / (10.0 * MU)
BC = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3, [0.0] * 3])
CTL, SCTL = [20, 1.0e-2, 200, 1.0e-2], [20, 1.0e-2]
STATE = ("U", "p", "mdot", "div", "gradU", "gradP", "mu", "k", "omega", "ymin", "mut", "F1", "Pk",
         "Pk_wall", "gradK", "gradW", "kprev", "msd")


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def setup(mod, case, mesh, meshgen, device):
    case.set_bc_types([1, 2, 0, 3])       # inlet, outlet, no-slip walls, symmetry sides
    case.set_bc_values(p=[0.0] * 4, k=[KIN, 0, 0, 0], w=[WIN, 0, 0, 0])
    case.sst_enable(mod.default_sst(urf_k=0.9, urf_w=0.9))
    nn = mesh["xyz"].shape[0]
    y = meshgen.bfs_wall_distance(mesh["xyz"])
    w = np.full(nn, WIN)
    w[y < 1e-12] = 60.0 * MU / (0.075 * mesh["first_dy"] ** 2)
    U = np.zeros((nn, 3))
    U[:, 0] = 1.0
    for name, v in (("U", U), ("p", np.zeros(nn)), ("k", np.full(nn, KIN)), ("omega", w),
                    ("ymin", y)):
        if device:
            case.upload(name, v)
        else:
            case.field(name)[:] = v


@pytest.fixture(scope="module")
def developed(oracle, meshgen):
    m = meshgen.bfs_hex(3, 12, 96, 112, 2)
    assert abs(m["first_dy"] - 4.8e-4) < 0.01 * 4.8e-4
    print(f"[case] step slice {m['conn'].shape[0]} elements, first dy {m['first_dy']:.2e}",
          flush=True)
    c = oracle.Case(m, oracle.default_params(mu=MU, urf_U=0.7, urf_p=0.3, u_shifted=1), BC)
    setup(oracle, c, m, meshgen, False)
    c.initialize()
    for _ in range(2):
        c.outer_iteration(CTL)
        c.sst_iteration(SCTL)
    return c, m


def device_systems(oa, meshgen, c, m, precision):
    dm = oa.DeviceMesh(m)
    f = oa.FlowCase(dm, oa.default_params(precision, mu=MU, urf_U=0.7, urf_p=0.3, u_shifted=1), BC)
    setup(oa, f, m, meshgen, True)
    f.initialize()                  # derived geometry; the state is replaced below
    for name in STATE:
        f.upload(name, c.field(name))
    f.update_reversal()
    f.mass_divergence()
    out = {}
    f.assemble_momentum()
    out["momentum"] = f.system(3)
    f.assemble_pcorr()
    out["pcorr"] = f.system(1)
    for which in ("k", "omega"):
        f.assemble_scalar(which)
        out[which] = f.system(1)
    f.close()
    return out


def test_step_slice_fp64_and_mixed(oracle, oa, meshgen, developed):
    c, m = developed
    for name in STATE:
        assert np.all(np.isfinite(c.field(name))), name
    # the oracle's systems about the same state (outlet reversal flags and
    # the mass divergence taken from the side mass flows, as on the device)
    c.update_reversal()
    c.mass_divergence()
    ref = {}
    c.assemble_momentum()
    c.compute_du()
    c.normal_relaxation()
    ref["momentum"] = (c.vals(3).copy(), c.rhs(3).copy(), c.field("U").copy())
    c.assemble_pcorr()
    ref["pcorr"] = (c.vals(1).copy(), c.rhs(1).copy(), np.zeros(c.nn))
    for which in ("k", "omega"):
        c.assemble_scalar(which)
        ref[which] = (c.vals(1).copy(), c.rhs(1).copy(), c.field(which).copy())
    rp, cols = c.row_ptr, c.cols
    fp64 = device_systems(oa, meshgen, c, m, oa.OAGPU_FP64)
    mixed = device_systems(oa, meshgen, c, m, oa.OAGPU_MIXED)
    errs = {}
    for key, (A, b, x) in ref.items():
        A64, b64 = fp64[key]
        Amx, bmx = mixed[key]
        if key == "pcorr":      # a node's mass imbalance against its SCS mass flows
            eb = (mass_rel_err(b64, b, c.field("mdot"), m["conn"]),
                  mass_rel_err(bmx, b64, c.field("mdot"), m["conn"]))
        else:
            blk = key == "momentum"
            eb = (rhs_rel_err(b64, b, A, rp, cols, x, blk), rhs_rel_err(bmx, b64, A64, rp, cols, x, blk))
        errs[key] = (row_rel_err(A64, A, rp), eb[0], row_rel_err(Amx, A64, rp), eb[1])
        print(f"[errors] {key}: fp64 {errs[key][0]:.2e} {errs[key][1]:.2e} "
              f"mixed {errs[key][2]:.2e} {errs[key][3]:.2e}", flush=True)
    for key, (e64A, e64b, emxA, emxb) in errs.items():
        assert e64A < 1e-12 and e64b < 1e-12, key
        if key in ("momentum", "pcorr"):
            assert emxA < 2e-5 and emxb < 2e-5, key
