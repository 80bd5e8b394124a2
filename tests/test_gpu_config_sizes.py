"""GPU parity at the workloads' own sizes and set-ups, against the oracle:

* the 32^3 lid-driven cavity at full specification (32 768 hex8, rho = 1,
  mu = 0.01, lid (1, 0, 0), U0 ~ N(0, 1e-3) seeded, p0 = 0, 20 outer
  iterations, upwind, Jacobi-GMRES momentum and Jacobi-CG p'): control
  residual history within 1e-8 and fields within 1e-8 after 20 iterations;
  the momentum and p' systems assembled from the oracle's final state
  row-relative within 1e-12, the momentum rhs within 1e-12 of each row's
  flux scale (sum_j |A_ij||U_j|), the p' rhs within 1e-12 of each node's
  mass-flow scale (sum of |mdot| through its control-volume faces);
* one assembly pass at 128^3 (2 097 152 hex8): momentum, p', mass flow
  rates (per element) and the gradients (per node) within 1e-12;
* batched independent cavities (sizes 16^3 and 24^3, Re and lid direction
  drawn as for the 96-case workload) advanced together by
  oagpu_batch_outer_iterations: each case's residual history and fields
  equal the oracle's one-case run within 1e-8 (FP64, Jacobi-CG p'); the
  mixed form with AMG-CG p' within 1e-3 of the oracle's residuals.
"""
import os
import sys

import numpy as np
import pytest

from conftest import (block_rel_err, entry_rel_err, mass_rel_err, rel_err, rhs_rel_err,
                      row_rel_err)

pytestmark = pytest.mark.gpu

BC = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def oa():
    import openaccel_amd as oa
    oa.init(0)
    return oa


def test_cavity_32_full_specification(oracle, oa, meshgen):
    print("[case] cavity 32^3, 20 outer iterations, FP64", flush=True)
    m = meshgen.cavity_hex(32)
    c = oracle.Case(m, oracle.default_params(), BC)
    dm = oa.DeviceMesh(m)
    f = oa.FlowCase(dm, oa.default_params(oa.OAGPU_FP64, p_method=oa.OAGPU_CG), BC)
    U0 = meshgen.initial_velocity(c.nn, 0)
    c.field("U")[:] = U0
    f.upload("U", U0)
    c.initialize()
    f.initialize()
    ctl = [20, 0.1, 200, 1e-6]
    for it in range(20):
        ho = c.outer_iteration(ctl)
        hg = f.outer_iteration(ctl)
        np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-8, err_msg=f"iteration {it}")
        assert np.abs(hg[4:] - ho[4:]).max() <= 1, it
    for name in ("U", "p", "mdot", "gradU", "gradP", "div"):
        assert rel_err(f.download(name), c.field(name)) < 1e-8, name
    # the systems of the 21st iteration from the oracle's state
    for name in ("U", "p", "mdot", "gradU", "gradP", "div"):
        f.upload(name, c.field(name))
    c.assemble_momentum()
    c.compute_du()
    f.assemble_momentum()
    vals, b = f.system(3)
    assert row_rel_err(vals, c.vals(3), c.row_ptr) < 1e-12
    assert rhs_rel_err(b, c.rhs(3), c.vals(3), c.row_ptr, c.cols, c.field("U")) < 1e-12
    c.assemble_pcorr()
    f.assemble_pcorr()
    vals, b = f.system(1)
    assert row_rel_err(vals, c.vals(1), c.row_ptr) < 1e-12
    assert mass_rel_err(b, c.rhs(1), c.field("mdot"), m["conn"]) < 1e-12


@pytest.mark.timeout(900)
def test_cavity_128_assembly_fp64(oracle, oa, meshgen):
    print("[case] cavity 128^3 assembly, FP64", flush=True)
    m = meshgen.cavity_hex(128)
    c = oracle.Case(m, oracle.default_params(), BC)
    dm = oa.DeviceMesh(m)
    f = oa.FlowCase(dm, oa.default_params(oa.OAGPU_FP64), BC)
    x = m["xyz"]
    # a smooth rotating state plus noise, p a smooth field
    U = np.stack([np.sin(np.pi * x[:, 1]) * np.sin(np.pi * x[:, 2]),
                  -np.sin(np.pi * x[:, 0]) * np.cos(np.pi * x[:, 2]),
                  0.3 * np.cos(np.pi * x[:, 0] * x[:, 1])], axis=1)
    U += np.random.default_rng(9).normal(scale=1e-2, size=U.shape)
    p = 0.1 * np.cos(2 * np.pi * x[:, 0]) * x[:, 1]
    for name, v in (("U", U), ("p", p)):
        c.field(name)[:] = v
        f.upload(name, v)
    c.initialize()
    f.initialize()
    for name, k in (("gradU", 9), ("gradP", 3), ("mdot", 12)):
        assert block_rel_err(f.download(name), c.field(name), k) < 1e-12, name
    print("[case] 128^3 momentum", flush=True)
    c.assemble_momentum()
    c.compute_du()
    f.assemble_momentum()
    vals, b = f.system(3)
    assert row_rel_err(vals, c.vals(3), c.row_ptr) < 1e-12
    assert rhs_rel_err(b, c.rhs(3), c.vals(3), c.row_ptr, c.cols, U) < 1e-12
    assert entry_rel_err(f.download("du"), c.field("du")) < 1e-12
    del vals, b
    print("[case] 128^3 p'", flush=True)
    c.assemble_pcorr()
    f.assemble_pcorr()
    vals, b = f.system(1)
    assert row_rel_err(vals, c.vals(1), c.row_ptr) < 1e-12
    assert entry_rel_err(b, c.rhs(1)) < 1e-12
    print("[case] 128^3 mass flow and gradient", flush=True)
    c.update_mdot()
    f.update_mdot()
    assert block_rel_err(f.download("mdot"), c.field("mdot"), 12) < 1e-12
    c.gradient("U")
    f.gradient("U")
    assert block_rel_err(f.download("gradU"), c.field("gradU"), 9) < 1e-12


def batch_cases():
    sys.path.insert(0, ROOT)
    import bench
    cases = bench.batch_cases(40, seed=7)
    pick = [c for c in cases if c["n"] == 16][:2] + [c for c in cases if c["n"] == 24][:2]
    assert len(pick) == 4
    return pick


@pytest.mark.parametrize("form", ["fp64_cg", "mixed_amg"])
def test_batched_cases_against_oracle(oracle, oa, meshgen, form):
    cases = batch_cases()
    iters = 5
    if form == "fp64_cg":
        prec, pm, ctl, tol = oa.OAGPU_FP64, oa.OAGPU_CG, [20, 0.1, 200, 1e-6], 1e-8
    else:
        prec, pm, ctl, tol = oa.OAGPU_MIXED, oa.OAGPU_CG_AMG, [20, 0.1, 400, 1e-10], 1e-3
    meshes, flows, oracles = {}, [], []
    for k, cs in enumerate(cases):
        print(f"[case] batched {form} case {k}: n={cs['n']} mu={cs['mu']:.4g} lid={cs['lid']}",
              flush=True)
        m = meshes.setdefault(cs["n"], meshgen.cavity_hex(cs["n"]))
        bc = np.array([cs["lid"], [0.0, 0.0, 0.0]])
        dm = oa.DeviceMesh(m)
        f = oa.FlowCase(dm, oa.default_params(prec, mu=cs["mu"], p_method=pm), bc)
        c = oracle.Case(m, oracle.default_params(mu=cs["mu"]), bc)
        U0 = meshgen.initial_velocity(c.nn, cs["n"])
        f.upload("U", U0)
        c.field("U")[:] = U0
        f.initialize()
        c.initialize()
        flows.append(f)
        oracles.append(c)
    hist, done, sec = oa.batch_outer_iterations(flows, ctl, iters, threads=4)
    assert np.all(done == iters)
    for k, c in enumerate(oracles):
        ho = np.array([c.outer_iteration(ctl) for _ in range(iters)])
        np.testing.assert_allclose(hist[k][:, :4], ho[:, :4], rtol=tol, err_msg=f"case {k}")
        if form == "fp64_cg":
            assert np.abs(hist[k][:, 4:] - ho[:, 4:]).max() <= 1
            for name in ("U", "p", "mdot"):
                assert rel_err(flows[k].download(name), c.field(name)) < 1e-8, (k, name)
