"""Benchmark: steady segregated outer iterations of the lid-driven cavity on
one MI355X, mixed form vs FP64 form, with assembly kernel roofline figures
and the oracle's CPU time on the same workload.

A "step" is one outer iteration (momentum assembly + GMRES solve, mass flow
update, p' assembly + AMG-preconditioned CG solve, corrections, gradients)
over the whole mesh, inputs resident in HBM.  One JSON line is printed.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--n 128] [--perturb 0]
    python bench.py --workload bfs [--bfs-scale 1.0]   # 8M-element step, k-omega SST
    python bench.py --workload duct [--duct-scale 1.0] # 3.9M-element mixed-topology duct
    python bench.py --workload batch                   # 96 independent cavities, batched
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "outer iters/s + assembly Melem/s (%HBM peak), mixed/batched vs FP64 on 1 GPU"
HBM_PEAK_GBS = 8000.0
BC = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
CTL = [20, 0.1, 200, 1.0e-6]   # momentum GMRES max_it/rtol, p' CG max_it/rtol


def algorithmic_bytes(nn, ne, nnz, amg_value_bytes=4):
    """Compulsory HBM bytes of one launch of each row-owned kernel (each array
    read once, each output written once; FP64 values, int32 indices; the AMG
    hierarchy in FP32 for the mixed form, FP64 for the FP64 form)."""
    conn, mdot, slots = 4 * 8, 8 * 12, 4 * 32
    mom = (nn * (24 + 24 + 8) + ne * (conn + mdot + slots) + nnz * 72 + nn * 24
           + nn * (8 + 8 + 24) + nn * 24)          # fused node terms + du write
    pc = nn * (24 + 24 + 8 + 24 + 24 + 8) + ne * (conn + mdot + slots) + nnz * 8 + nn * 8
    md = nn * 112 + ne * (conn + 16 * 12)
    gu = nn * (24 + 24 + 8 + 72) + ne * conn
    # fine-level AMG kernels of the V-cycle: sliced-ELL values + column
    # indices of the true nonzeros, row vectors
    v = amg_value_bytes
    amg_pre = nnz * (v + 4) + nn * v * 4           # r, dinv in; x, residual out
    amg_post = nnz * (v + 4) + nn * (v * 5 + 4)    # x, agg, r, dinv, coarse e in; z out
    return {"momentum": mom, "pcorr": pc, "mdot": md, "gradU": gu, "amg_pre": amg_pre,
            "amg_post": amg_post}


def measured_traffic(kernel, n, precision):
    """HBM bytes per launch of `kernel` measured by rocprofv3 --pmc FETCH_SIZE
    and WRITE_SIZE on this workload (profiles/traffic_<kernel>_<n>.json, with
    the copy-kernel calibration factors applied), or None if not measured."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{n}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    rec = t.get("fp64" if precision == 0 else "mixed")
    if not rec:
        return None
    return int(rec["fetch_kib"] * 1024 * t["fetch_scale"] + rec["write_kib"] * 1024 * t["write_scale"])


def run_gpu(oa, mesh, precision, steps, warmup, p_method):
    dm = oa.DeviceMesh(mesh)
    f = oa.FlowCase(dm, oa.default_params(precision, p_method=p_method), BC)
    from openaccel_amd import meshgen
    f.upload("U", meshgen.initial_velocity(dm.nn, 0))
    f.initialize()
    for _ in range(warmup):
        f.outer_iteration(CTL)
    oa.load_library().oagpu_synchronize()
    ms = {k: 0.0 for k in ("momentum", "pcorr", "mdot", "gradU", "amg_pre", "amg_post")}
    hist = []
    v0 = f.kernel_ms("amg_vcycles")
    t0 = time.perf_counter()
    for _ in range(steps):
        hist.append(f.outer_iteration(CTL))
        for k in ms:
            ms[k] += f.kernel_ms(k)
    oa.load_library().oagpu_synchronize()
    t = time.perf_counter() - t0
    for k in ms:
        ms[k] /= steps
    # launches per step: mdot is updated twice per outer iteration, the
    # fine-level AMG kernels once per V-cycle
    vps = (f.kernel_ms("amg_vcycles") - v0) / steps
    per_step = {"momentum": 1, "pcorr": 1, "mdot": 2, "gradU": 1, "amg_pre": vps, "amg_post": vps}
    return {"seconds": t, "kernel_ms": ms, "launches": per_step, "hist": np.array(hist),
            "nn": dm.nn, "ne": dm.ne, "nnz": dm.nnz}


# backward-facing step with k-omega SST: inlet U = 1 with 1 % intensity and
# viscosity ratio 10, rho = 1, mu = 2e-5, pitzDaily controls of the
# reference's examples (relaxation 0.7 / 0.3 / 0.9, linear_linear velocity
# interpolation, linear solves to rtol 1e-2)
BFS_DIMS = (30, 210, 96, 112, 170)   # nx_up, nx_dn, ny_low, ny_high, nz: 7 996 800 elements
BFS_MU, BFS_KIN = 2.0e-5, 1.5e-4
BFS_WIN = BFS_KIN / (10.0 * BFS_MU)
BFS_CTL, BFS_SCTL = [20, 1.0e-2, 200, 1.0e-2], [20, 1.0e-2]


def bfs_setup(oa_or_oracle, mesh, case, device=True):
    """Boundary types and values, SST and the initial state of the
    backward-facing step on a FlowCase (device) or an oracle Case."""
    from openaccel_amd import meshgen
    case.set_bc_types([1, 2, 0, 3])   # inlet, outlet, no-slip walls, symmetry sides
    case.set_bc_values(p=[0.0] * 4, k=[BFS_KIN, 0, 0, 0], w=[BFS_WIN, 0, 0, 0])
    sst = oa_or_oracle.default_sst(urf_k=0.9, urf_w=0.9)
    case.sst_enable(sst)
    nn = mesh["xyz"].shape[0]
    U = np.zeros((nn, 3))
    U[:, 0] = 1.0
    y = meshgen.bfs_wall_distance(mesh["xyz"])
    w = np.full(nn, BFS_WIN)
    # omega at no-slip wall nodes: 10 * 6 nu / (beta1 y1^2), y1 the wall cells' thickness
    w[y < 1e-12] = 60.0 * BFS_MU / (0.075 * mesh["first_dy"] ** 2)
    vals = {"U": U, "p": np.zeros(nn), "k": np.full(nn, BFS_KIN), "omega": w, "ymin": y}
    for name, v in vals.items():
        if device:
            case.upload(name, v)
        else:
            case.field(name)[:] = v
    case.initialize()


def bfs_params(mod, precision=None):
    kw = dict(mu=BFS_MU, urf_U=0.7, urf_p=0.3, u_shifted=1)
    if precision is None:
        return mod.default_params(**kw)
    # p' by AMG-preconditioned BiCGStab: the outlet makes it nonsymmetric
    return mod.default_params(precision, p_method=mod.OAGPU_BICGSTAB_AMG, **kw)


def run_bfs(oa, dm, mesh, precision, steps, warmup):
    bc = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3, [0.0] * 3])
    f = oa.FlowCase(dm, bfs_params(oa, precision), bc)
    bfs_setup(oa, mesh, f)
    hist = []
    for _ in range(warmup):
        hist.append(np.concatenate([f.outer_iteration(BFS_CTL), f.sst_iteration(BFS_SCTL)]))
    oa.load_library().oagpu_synchronize()
    ms = {k: 0.0 for k in ("momentum", "pcorr", "mdot", "gradU")}
    t0 = time.perf_counter()
    for _ in range(steps):
        hist.append(np.concatenate([f.outer_iteration(BFS_CTL), f.sst_iteration(BFS_SCTL)]))
        for k in ms:
            ms[k] += f.kernel_ms(k) / steps
    oa.load_library().oagpu_synchronize()
    t = time.perf_counter() - t0
    f.close()
    return {"seconds": t, "kernel_ms": ms, "hist": np.array(hist)}


def bfs_cpu_baseline(dims):
    """Oracle on a bounded sample of the backward-facing step: the same
    geometry and controls on a coarser grid, one flow + SST outer iteration."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from openaccel_amd import meshgen
    mesh = meshgen.bfs_hex(*dims)
    bc = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3, [0.0] * 3])
    c = pyoracle.Case(mesh, bfs_params(pyoracle), bc)
    bfs_setup(pyoracle, mesh, c, device=False)
    t0 = time.perf_counter()
    c.outer_iteration(BFS_CTL)
    c.sst_iteration(BFS_SCTL)
    return {"outer_s": time.perf_counter() - t0, "ne": c.ne}


def bench_bfs(args):
    import openaccel_amd as oa
    from openaccel_amd import meshgen
    oa.init(0)
    s = args.bfs_scale
    dims = tuple(max(2, int(round(d * s))) for d in BFS_DIMS)
    mesh = meshgen.bfs_hex(*dims)
    ne, nn = mesh["conn"].shape[0], mesh["xyz"].shape[0]
    print(f"[bench] backward-facing step {dims}: {ne} elements, first dy {mesh['first_dy']:.2e}",
          file=sys.stderr, flush=True)
    dm = oa.DeviceMesh(mesh)
    r64 = run_bfs(oa, dm, mesh, oa.OAGPU_FP64, args.steps, args.warmup)
    print(f"[bench] fp64 {args.steps / r64['seconds']:.3f} it/s", file=sys.stderr, flush=True)
    rmx = run_bfs(oa, dm, mesh, oa.OAGPU_MIXED, args.steps, args.warmup)
    print(f"[bench] mixed {args.steps / rmx['seconds']:.3f} it/s", file=sys.stderr, flush=True)
    it64, itmx = args.steps / r64["seconds"], args.steps / rmx["seconds"]
    cpu = None
    if not args.no_cpu:
        cd = tuple(max(2, d // 3) for d in BFS_DIMS)
        cb = bfs_cpu_baseline(cd)
        scale = cb["ne"] / ne
        cpu = {"value": round(scale / cb["outer_s"], 6), "unit": "outer iterations/s", "cores": 1,
               "kind": "port",
               "sample": f"oracle C restatement, single thread, 1 flow + SST outer iteration of the "
                         f"same step at grid {cd} ({cb['ne']} elements, {cb['outer_s']:.2f} s) scaled "
                         f"by the element ratio {scale:.5f}"}

    def hist(r):
        return {"flow_res_U_p": [[float(x) for x in h[:4]] for h in r["hist"]],
                "sst_res_k_w": [[float(x) for x in h[6:8]] for h in r["hist"]],
                "krylov_iters": [[int(h[4]), int(h[5]), int(h[8]), int(h[9])] for h in r["hist"]]}

    line = {
        "metric": METRIC,
        "value": round(itmx, 4),
        "unit": "outer iterations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * rmx["seconds"] / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 local / f64 global", "data": "synthetic",
        "config": {"workload": f"backward-facing step k-omega SST, channel 30 x 3 x 4, step 1, "
                               f"{ne} hex8 ({nn} nodes), wall cells {mesh['first_dy']:.1e} graded "
                               f"1.1; inlet U=1 (k {BFS_KIN}, omega {BFS_WIN}), outlet p=0, "
                               f"symmetry sides, rho=1, mu={BFS_MU}; one step = flow outer "
                               f"iteration + SST iteration; mixed form",
                   "dims": dims, "elements": ne, "nodes": nn},
        "fp64": {"value": round(it64, 4), "ms_per_step": round(1e3 * r64["seconds"] / args.steps, 3),
                 "kernel_ms": {k: round(v, 4) for k, v in r64["kernel_ms"].items()}, **hist(r64)},
        "mixed_over_fp64": round(itmx / it64, 3),
        "kernel_ms": {k: round(v, 4) for k, v in rmx["kernel_ms"].items()},
        **hist(rmx),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


# mixed-topology duct 1 x 0.2 x 0.2: plug inflow U = 1, static-pressure
# outlet p = 0, no-slip walls, rho = 1, mu = 1e-3, high-resolution advection
DUCT_M, DUCT_LAYERS = 64, (187, 8, 25, 75, 25)   # 3.87M elements: 60 % tet4, 21 % hex8, 15 % wedge6, 5 % pyramid5
DUCT_MU = 1.0e-3
DUCT_CTL = [20, 1.0e-2, 200, 1.0e-2]


def duct_setup(case, mesh, device=True):
    case.set_bc_types([1, 2, 0])      # inlet, outlet, no-slip walls
    case.set_bc_values(p=[0.0] * 3)
    nn = mesh["xyz"].shape[0]
    U = np.zeros((nn, 3))
    U[:, 0] = 1.0
    for name, v in (("U", U), ("p", np.zeros(nn))):
        if device:
            case.upload(name, v)
        else:
            case.field(name)[:] = v
    case.initialize()


def run_duct(oa, dm, mesh, precision, steps, warmup):
    bc = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3])
    f = oa.FlowCase(dm, oa.default_params(precision, p_method=oa.OAGPU_BICGSTAB_AMG, mu=DUCT_MU,
                                          high_res=1), bc)
    duct_setup(f, mesh)
    hist = [f.outer_iteration(DUCT_CTL) for _ in range(warmup)]
    oa.load_library().oagpu_synchronize()
    ms = {k: 0.0 for k in ("momentum", "pcorr", "mdot", "gradU")}
    t0 = time.perf_counter()
    for _ in range(steps):
        hist.append(f.outer_iteration(DUCT_CTL))
        for k in ms:
            ms[k] += f.kernel_ms(k) / steps
    oa.load_library().oagpu_synchronize()
    t = time.perf_counter() - t0
    f.close()
    return {"seconds": t, "kernel_ms": ms, "hist": np.array(hist)}


def duct_cpu_baseline(m):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from openaccel_amd import meshgen
    layers = tuple(max(1, round(x * m / DUCT_M)) for x in DUCT_LAYERS)
    mesh = meshgen.duct_mixed(m, layers=layers, length=sum(layers) * 0.2 / m, seed=1)
    bc = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3])
    c = pyoracle.Case(mesh, pyoracle.default_params(mu=DUCT_MU, high_res=1), bc)
    duct_setup(c, mesh, device=False)
    t0 = time.perf_counter()
    c.update_blending("U")
    c.outer_iteration(DUCT_CTL)
    return {"outer_s": time.perf_counter() - t0, "ne": c.ne, "layers": layers}


def bench_duct(args):
    import openaccel_amd as oa
    from openaccel_amd import meshgen
    oa.init(0)
    m = max(8, int(round(DUCT_M * args.duct_scale)))
    layers = tuple(max(1, round(x * m / DUCT_M)) for x in DUCT_LAYERS)
    mesh = meshgen.duct_mixed(m, layers=layers, length=sum(layers) * 0.2 / m, seed=1)
    ne, nn = mesh["conn"].shape[0], mesh["xyz"].shape[0]
    counts = np.bincount(mesh["etype"], minlength=4)
    print(f"[bench] mixed duct m={m} layers={layers}: {ne} elements "
          f"(hex {counts[0]}, tet {counts[1]}, wedge {counts[2]}, pyramid {counts[3]})",
          file=sys.stderr, flush=True)
    dm = oa.DeviceMesh(mesh)
    r64 = run_duct(oa, dm, mesh, oa.OAGPU_FP64, args.steps, args.warmup)
    print(f"[bench] fp64 {args.steps / r64['seconds']:.3f} it/s", file=sys.stderr, flush=True)
    rmx = run_duct(oa, dm, mesh, oa.OAGPU_MIXED, args.steps, args.warmup)
    print(f"[bench] mixed {args.steps / rmx['seconds']:.3f} it/s", file=sys.stderr, flush=True)
    it64, itmx = args.steps / r64["seconds"], args.steps / rmx["seconds"]
    cpu = None
    if not args.no_cpu:
        cb = duct_cpu_baseline(32)
        scale = cb["ne"] / ne
        cpu = {"value": round(scale / cb["outer_s"], 6), "unit": "outer iterations/s", "cores": 1,
               "kind": "port",
               "sample": f"oracle C restatement, single thread, 1 outer iteration of the same duct "
                         f"at m=32 layers {cb['layers']} ({cb['ne']} elements, {cb['outer_s']:.2f} s) "
                         f"scaled by the element ratio {scale:.5f}"}
    line = {
        "metric": METRIC,
        "value": round(itmx, 4),
        "unit": "outer iterations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * rmx["seconds"] / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 local / f64 global", "data": "synthetic",
        "config": {"workload": f"mixed-topology duct 1 x 0.2 x 0.2, {ne} elements ({counts[1]} tet4, "
                               f"{counts[0]} hex8, {counts[2]} wedge6, {counts[3]} pyramid5), node "
                               f"jitter N(0, 0.1 h); plug inflow U=1, outlet p=0, rho=1, "
                               f"mu={DUCT_MU}, high-resolution advection; mixed form",
                   "m": m, "layers": layers, "elements": ne, "nodes": nn},
        "fp64": {"value": round(it64, 4), "ms_per_step": round(1e3 * r64["seconds"] / args.steps, 3),
                 "kernel_ms": {k: round(v, 4) for k, v in r64["kernel_ms"].items()},
                 "res_U_p": [[float(x) for x in h[:4]] for h in r64["hist"]],
                 "krylov_iters": [[int(h[4]), int(h[5])] for h in r64["hist"]]},
        "mixed_over_fp64": round(itmx / it64, 3),
        "kernel_ms": {k: round(v, 4) for k, v in rmx["kernel_ms"].items()},
        "kernel_Melem_s": {k: round(ne / (v * 1e-3) / 1e6, 1) for k, v in rmx["kernel_ms"].items() if v > 0},
        "res_U_p": [[float(x) for x in h[:4]] for h in rmx["hist"]],
        "krylov_iters": [[int(h[4]), int(h[5])] for h in rmx["hist"]],
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


# batched independent cavities: sizes from {16, 24, 32, 48, 64}^3, Re
# log-uniform in [10, 1000] (mu = 1 / Re, unit lid speed and cube), the lid
# velocity a random unit vector in the lid plane, all seeded
BATCH_N, BATCH_SIZES, BATCH_ITERS = 96, (16, 24, 32, 48, 64), 20


def batch_cases(n_cases, seed=7):
    rng = np.random.default_rng(seed)
    sizes = rng.choice(BATCH_SIZES, size=n_cases)
    re = 10.0 ** rng.uniform(1.0, 3.0, size=n_cases)
    th = rng.uniform(0.0, 2.0 * np.pi, size=n_cases)
    return [dict(n=int(n), mu=float(1.0 / r), lid=[float(np.cos(t)), 0.0, float(np.sin(t))])
            for n, r, t in zip(sizes, re, th)]


def bench_batch(args):
    import openaccel_amd as oa
    from openaccel_amd import meshgen
    oa.init(0)
    cases = batch_cases(args.batch_cases)
    meshes = {n: oa.DeviceMesh(meshgen.cavity_hex(n)) for n in sorted({c["n"] for c in cases})}
    elems = sum(c["n"] ** 3 for c in cases)
    print(f"[bench] batch of {len(cases)} cavities, {elems} elements in all", file=sys.stderr, flush=True)

    def flow(c, precision):
        dm = meshes[c["n"]]
        f = oa.FlowCase(dm, oa.default_params(precision, mu=c["mu"], p_method=oa.OAGPU_CG_AMG),
                        np.array([c["lid"], [0.0, 0.0, 0.0]]))
        f.upload("U", meshgen.initial_velocity(dm.nn, c["n"]))
        f.initialize()
        return f

    iters = args.batch_iters
    # FP64, one case at a time
    seq = np.zeros(len(cases))
    oa.load_library().oagpu_synchronize()
    t0 = time.perf_counter()
    for k, c in enumerate(cases):
        f = flow(c, oa.OAGPU_FP64)
        t1 = time.perf_counter()
        for _ in range(iters):
            f.outer_iteration(CTL)
        seq[k] = time.perf_counter() - t1
        f.close()
    t_seq = seq.sum()
    print(f"[bench] fp64 one at a time: {t_seq:.2f} s of iterations ({time.perf_counter() - t0:.1f} s "
          f"with set-up)", file=sys.stderr, flush=True)
    # mixed, batched
    flows = [flow(c, oa.OAGPU_MIXED) for c in cases]
    oa.load_library().oagpu_synchronize()
    t0 = time.perf_counter()
    hist, done, sec = oa.batch_outer_iterations(flows, CTL, iters, threads=16)
    t_batch = time.perf_counter() - t0
    print(f"[bench] mixed batched: {t_batch:.2f} s", file=sys.stderr, flush=True)
    res = np.array([h[-1][:4].max() for h in hist])
    line = {
        "metric": METRIC,
        "value": round(len(cases) / t_batch, 4),
        "unit": "cases/s",
        "n_gpus": 1, "steps": iters, "warmup": 0,
        "ms_per_step": round(1e3 * t_batch / iters, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 local / f64 global", "data": "synthetic",
        "config": {"workload": f"{len(cases)} independent lid-driven cavities, sizes from "
                               f"{BATCH_SIZES}^3 hex8, Re log-uniform in [10, 1000], lid velocity a "
                               f"random unit vector in the lid plane (seeded); {iters} outer "
                               f"iterations each; mixed form batched on 16 host threads (one "
                               f"stream per case) vs FP64 one case at a time",
                   "cases": len(cases), "elements": elems, "iterations_per_case": iters},
        "fp64_one_at_a_time": {"value": round(len(cases) / t_seq, 4), "unit": "cases/s",
                               "per_case_s_p50": round(float(np.percentile(seq, 50)), 4),
                               "per_case_s_p99": round(float(np.percentile(seq, 99)), 4)},
        "batched_over_one_at_a_time": round(t_seq / t_batch, 3),
        "per_case_s_p50": round(float(np.percentile(sec, 50)), 4),
        "per_case_s_p99": round(float(np.percentile(sec, 99)), 4),
        "residual_after_iters_p50": float(np.percentile(res, 50)),
        "residual_after_iters_max": float(res.max()),
        "cpu_baseline": None,
    }
    print(json.dumps(line))


def cpu_baseline(n_sample, perturb):
    """Oracle (single-thread C restatement) on a bounded sample of the same
    workload: the same cavity at n_sample^3, one outer iteration timed after
    set-up, and its momentum assembly alone."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from openaccel_amd import meshgen
    mesh = meshgen.cavity_hex(n_sample, perturb=perturb, seed=1)
    c = pyoracle.Case(mesh, pyoracle.default_params(), BC)
    c.field("U")[:] = meshgen.initial_velocity(c.nn, 0)
    c.initialize()
    t0 = time.perf_counter()
    c.assemble_momentum()
    t_mom = time.perf_counter() - t0
    t0 = time.perf_counter()
    c.outer_iteration(CTL)
    t_it = time.perf_counter() - t0
    return {"outer_s": t_it, "mom_s": t_mom, "ne": c.ne}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--perturb", type=float, default=0.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=64, help="oracle sample grid (elements scale)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device once (calibrates HBM byte counters)")
    ap.add_argument("--p-method", choices=("cg_amg", "bicgstab_amg"), default="cg_amg",
                    help="cavity p' solver: AMG-CG (default), or AMG-BiCGStab for the p' systems "
                         "of distorted meshes, which are not symmetric")
    ap.add_argument("--batch-cases", type=int, default=BATCH_N)
    ap.add_argument("--batch-iters", type=int, default=BATCH_ITERS)
    ap.add_argument("--workload", choices=("cavity", "bfs", "duct", "batch"), default="cavity",
                    help="cavity: the lid-driven cavity (default); bfs: the 8M-element "
                         "backward-facing step with k-omega SST; duct: the 4M-element "
                         "mixed-topology duct")
    ap.add_argument("--duct-scale", type=float, default=1.0,
                    help="scale of the duct's cube counts per direction (1: 3.9M elements)")
    ap.add_argument("--bfs-scale", type=float, default=1.0,
                    help="scale of the step's cell counts per direction (1: 8M elements)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X only: the multi-GPU path (element partition + "
              "halo exchange) is designed but not implemented; run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        nb = args.copy_probe << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": nb}))
        return

    if args.workload == "bfs":
        bench_bfs(args)
        return
    if args.workload == "duct":
        bench_duct(args)
        return
    if args.workload == "batch":
        bench_batch(args)
        return

    import openaccel_amd as oa
    from openaccel_amd import meshgen
    oa.init(0)
    mesh = meshgen.cavity_hex(args.n, perturb=args.perturb, seed=1)
    print(f"[bench] mesh {args.n}^3 perturb={args.perturb}: {mesh['conn'].shape[0]} elements",
          file=sys.stderr, flush=True)
    pm = oa.OAGPU_CG_AMG if args.p_method == "cg_amg" else oa.OAGPU_BICGSTAB_AMG
    r64 = run_gpu(oa, mesh, oa.OAGPU_FP64, args.steps, args.warmup, pm)
    print(f"[bench] fp64 {args.steps / r64['seconds']:.3f} it/s {r64['kernel_ms']}", file=sys.stderr,
          flush=True)
    rmx = run_gpu(oa, mesh, oa.OAGPU_MIXED, args.steps, args.warmup, pm)
    print(f"[bench] mixed {args.steps / rmx['seconds']:.3f} it/s {rmx['kernel_ms']}", file=sys.stderr,
          flush=True)
    nn, ne, nnz = rmx["nn"], rmx["ne"], rmx["nnz"]

    def kern(r):
        ab = algorithmic_bytes(nn, ne, nnz, 4 if r is rmx else 8)
        out = {}
        for k, ms in r["kernel_ms"].items():
            gbs = ab[k] / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
            out[k] = {"ms": round(ms, 4), "GB_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                      "bytes": int(ab[k]), "launches_per_step": round(r["launches"][k], 2),
                      "ms_per_step": round(ms * r["launches"][k], 3)}
            if not k.startswith("amg"):
                out[k]["Melem_s"] = round(ne / (ms * 1e-3) / 1e6, 1) if ms > 0 else 0
        return out

    k64, kmx = kern(r64), kern(rmx)
    # the dominant kernel: largest device time per outer iteration
    dom = max(kmx, key=lambda k: kmx[k]["ms_per_step"])
    it_s_mx = args.steps / rmx["seconds"]
    it_s_64 = args.steps / r64["seconds"]
    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle, 1 thread) ...", file=sys.stderr, flush=True)
        cn = min(args.cpu_n, args.n)
        cb = cpu_baseline(cn, args.perturb)
        scale = (cn / args.n) ** 3
        cpu = {"value": round(scale / cb["outer_s"], 6), "unit": "outer iterations/s", "cores": 1,
               "kind": "port",
               "sample": f"oracle C restatement, single thread, 1 outer iteration of the same cavity "
                         f"at {cn}^3 ({cb['ne']} elements, {cb['outer_s']:.2f} s) scaled by the "
                         f"element ratio {scale:.4f} to {args.n}^3; its momentum assembly alone "
                         f"runs at {cb['ne'] / cb['mom_s'] / 1e6:.3f} Melem/s"}
    line = {
        "metric": METRIC,
        "value": round(it_s_mx, 4),
        "unit": "outer iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * rmx["seconds"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 local / f64 global",
        "data": "synthetic",
        "config": {"workload": f"lid-driven cavity, uniform{' perturbed' if args.perturb else ''} "
                               f"{args.n}^3 hex8 ({ne} elements, {nn} nodes), Re=100, mixed form; "
                               f"Jacobi-GMRES(20 it, rtol 0.1) momentum, aggregation-AMG-"
                               f"{'CG' if args.p_method == 'cg_amg' else 'BiCGStab'} "
                               f"(200 it, rtol 1e-6) p'",
                   "n": args.n, "elements": ne, "nodes": nn, "nnz": nnz},
        "fp64": {"value": round(it_s_64, 4), "ms_per_step": round(1e3 * r64["seconds"] / args.steps, 3),
                 "krylov_iters_mean": [round(float(v), 1) for v in r64["hist"][:, 4:6].mean(axis=0)],
                 "kernels": k64},
        "krylov_iters_mean": [round(float(v), 1) for v in rmx["hist"][:, 4:6].mean(axis=0)],
        "mixed_over_fp64": round(it_s_mx / it_s_64, 3),
        "kernels": kmx,
        "roofline": {"bound": "hbm", "kernel": dom, "launch_ms": kmx[dom]["ms"],
                     "achieved": kmx[dom]["GB_s"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": kmx[dom]["hbm_frac"],
                     "traffic": measured_traffic(dom, args.n, 1) if not args.perturb else None},
        "cpu_baseline": cpu,
        "residuals_last": [float(x) for x in rmx["hist"][-1][:4]],
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
