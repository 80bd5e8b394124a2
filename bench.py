"""Benchmark of the CVFEM assemble-and-solve path on one MI355X: outer
iterations/s of the mixed form against the FP64 form, per-kernel device time
and HBM roofline, the mixed form's coefficient error against the FP64 form,
and the oracle's CPU time on a bounded sample of the same workload.

A "step" is one outer iteration over the whole mesh with every input resident
in HBM: for the backward-facing step (the default workload, 8M hex8 cells,
k-omega SST) the segregated flow iteration (momentum assembly + Jacobi-GMRES,
mass flow update, p' assembly + AMG-BiCGStab, corrections, gradients) and
the SST iteration (model update, k and omega assembly + GMRES, corrections,
gradients).  One JSON line is printed.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2]          # backward-facing step
    python bench.py --workload cavity [--n 128] [--perturb 0.2 --p-method bicgstab_amg]
    python bench.py --workload cavity --to-residual 1e-8          # time to a scaled residual
    python bench.py --workload duct                               # 3.9M-cell mixed topologies
    python bench.py --workload batch                              # 96 independent cavities

Kernel device times come from the library's HIP-event clocks on the flow's
stream (oagpu_profile / oagpu_kernel_stats): every main kernel is clocked
during the timed steps (events recorded without a host wait, read after the
steps) and the one with the largest device time per step is the roofline's
kernel (its average launch time gives roofline.achieved = algorithmic bytes
per launch / launch time).
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "outer iters/s + assembly Melem/s (%HBM peak), mixed/batched vs FP64 on 1 GPU"
HBM_PEAK_GBS = 8000.0
BC = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
CTL = [20, 0.1, 200, 1.0e-6]   # momentum GMRES max_it/rtol, p' max_it/rtol


# ------------------------------------------------------------------ bytes
def algorithmic_bytes(nn, ne, nnz, value_bytes, vector_bytes, high_res=False, sa=None, scalar_eqs=1):
    """Compulsory HBM bytes of one launch of each clocked kernel on a hex8
    mesh: each array read once, each output written once, FP64 fields, int32
    indices.  value_bytes / vector_bytes: the AMG fine level's matrix values
    and vectors (FP64 form 8 / 8; the mixed form's FP32 inner solves 4 / 4).
    sa = (prolongation entries of the fine level, rows of level 1): the
    smoothed-aggregation cycle, whose fine-level post-smoothing reads P.
    scalar_eqs: equations of one k / omega launch (2: k and omega assembled
    together, the mesh and shared fields read once, phi and the system per
    equation)."""
    conn, mdot, slots = 4 * 8, 8 * 12, 4 * 32
    elem = ne * (conn + mdot + slots)
    hr = nn * (72 + 24) if high_res else 0
    mom = (nn * (24 + 24 + 8) + elem + nnz * 72 + nn * 24    # x, U, mu_eff; A, b
           + nn * (8 + 8 + 24) + nn * 24 + hr)                # V, div, grad p node terms; du
    pc = nn * (24 + 24 + 8 + 24 + 24 + 8) + elem + nnz * 8 + nn * 8
    sc = (nn * (24 + 8 + 8 + 8 + 8) + elem                          # x, V, div, mu_t, F1
          + scalar_eqs * (nn * 8 + nnz * 8 + nn * 8))               # phi; A, b per equation
    md = nn * 112 + ne * (conn + 16 * 12)
    gu = nn * (24 + 24 + 8 + 72 + 72) + ne * conn           # x, U, V, grad U in and out
    gp = nn * (24 + 8 + 8 + 24 + 24) + ne * conn
    gkw = nn * (24 + 8 + 2 * (8 + 24 + 24)) + ne * conn      # x, V; k, omega, their gradients in and out
    v, w = value_bytes, vector_bytes
    amg_pre = nnz * (v + 4) + nn * w * 4                     # A, cols; r, dinv in; x, res out
    amg_post = nnz * (v + 4) + nn * (w * 5 + 4)              # A, cols; x, r, dinv, e, z; agg
    if sa:
        npf, n1 = sa   # t = x + P e (P values, columns; e; x in, t out), then smoothing of t
        amg_post = npf * (v + 4) + n1 * w + nn * 2 * w + nnz * (v + 4) + nn * 4 * w
    spmv1 = nnz * (8 + 4) + nn * 16
    spmv3 = nnz * (72 + 4) + nn * 48
    spmv1f = nnz * (4 + 4) + nn * 8
    spmv3f = nnz * (36 + 4) + nn * 24
    return {"momentum": mom, "pcorr": pc, "scalar": sc, "mdot": md, "gradU": gu, "gradP": gp, "gradKW": gkw,
            "amg_pre": amg_pre, "amg_post": amg_post, "spmv1": spmv1, "spmv3": spmv3,
            "spmv1_f32": spmv1f, "spmv3_f32": spmv3f}


def measured_traffic(kernel, workload, precision):
    """HBM bytes per launch measured by rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE (profiles/traffic_<kernel>_<workload>.json, copy-kernel
    calibration applied), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    rec = t.get("fp64" if precision == 0 else "mixed")
    if not rec:
        return None
    return int(rec["fetch_kib"] * 1024 * t["fetch_scale"] + rec["write_kib"] * 1024 * t["write_scale"])


# ---------------------------------------------------------------- profiles
# rocprofv3 output of the default workload -> the summaries under profiles/
# (python bench.py --summarize-profiles DIR, DIR holding the runs that
# bench.py --profile-commands prints)
PROFILE_REGEX = ("k_scalar_rows|k_mdot_update|k_momentum_rows|k_pcorr_rows|k_gradient|k_spmv|"
                 "k_amg_pre_residual|k_sa_prolong|k_amg_smooth|copyBuffer")
PROFILE_COMMANDS = [
    "cd /tmp && export TMPDIR=/tmp && timeout -k 10 600 rocprofv3 --kernel-trace --stats --output-format csv "
    "-d {dir}/trace -o trace -- python3 /root/repo/bench.py --no-cpu --steps 3 --warmup 2 --graph-steps 0",
    "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 300 rocprofv3 --pmc FETCH_SIZE --output-format csv "
    f"--kernel-include-regex \"{PROFILE_REGEX}\" -d {{dir}}/fetch -o fetch -- python3 /root/repo/bench.py "
    "--steps 1 --warmup 1 --no-cpu --graph-steps 0",
    "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 300 rocprofv3 --pmc WRITE_SIZE --output-format csv "
    f"--kernel-include-regex \"{PROFILE_REGEX}\" -d {{dir}}/write -o write -- python3 /root/repo/bench.py "
    "--steps 1 --warmup 1 --no-cpu --graph-steps 0",
    "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 60 rocprofv3 --pmc FETCH_SIZE --output-format csv "
    "-d {dir}/fetchcopy -o fetchcopy -- python3 /root/repo/bench.py --copy-probe 1024",
    "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 60 rocprofv3 --pmc WRITE_SIZE --output-format csv "
    "-d {dir}/writecopy -o writecopy -- python3 /root/repo/bench.py --copy-probe 1024",
]
# clock -> (mixed form's kernel symbol, FP64 form's); None: the same symbol in
# both forms, told apart by dispatch time (the mixed form runs first)
PROFILE_SYMBOLS = {
    "scalar": (r"k_scalar_rows<float", r"k_scalar_rows<double"),
    "momentum": (r"k_momentum_rows<float", r"k_momentum_rows<double"),
    "pcorr": (r"k_pcorr_rows(_occ)?<float", r"k_pcorr_rows(_occ)?<double"),
    "mdot": (r"k_mdot_update<", None),
    "gradU": (r"k_gradient<3,", None),
    "gradP": (r"k_gradient<1,", None),
    "gradKW": (r"k_gradient<2,", None),
    "spmv1_f32": (r"k_spmv<1, float>", None),
    "spmv1": (None, r"k_spmv<1, double>"),
    "spmv3_f32": (r"k_spmv<3, float>", None),
    "spmv3": (r"k_spmv<3, double>", r"k_spmv<3, double>"),
    "amg_pre": (r"k_amg_pre_residual<float, true", r"k_amg_pre_residual<double, true"),
}


def _csv_rows(path):
    import csv
    with open(path) as fh:
        return list(csv.DictReader(fh))


def _find_csv(d, suffix):
    for root, _, files in os.walk(d):
        for fn in files:
            if fn.endswith(suffix):
                return os.path.join(root, fn)
    raise FileNotFoundError(f"no *{suffix} under {d}")


def summarize_traffic(d, workload_desc, timed=None):
    """profiles/traffic_<clock>_bfs.json from the FETCH_SIZE / WRITE_SIZE runs:
    per form, the mean counter value (KiB per dispatch) over the dispatches
    of the fine-level grid, and the copy-kernel calibration.  timed: per
    form, the kernel symbols of the timed iterations (the tuner's other
    candidates are left out)."""
    import re
    cal = {}
    for cnt in ("fetch", "write"):
        rows = _csv_rows(_find_csv(os.path.join(d, cnt + "copy"), "counter_collection.csv"))
        big = max(rows, key=lambda r: float(r["Counter_Value"]))
        cal[cnt] = (float(big["Counter_Value"]), big["Kernel_Name"])
    # 1 GiB read and written by the copy: bytes = scale * KiB * 1024
    fscale, wscale = (1 << 20) / cal["fetch"][0], (1 << 20) / cal["write"][0]
    data = {}
    for cnt in ("fetch", "write"):
        rows = _csv_rows(_find_csv(os.path.join(d, cnt), "counter_collection.csv"))
        t64 = min((int(r["Start_Timestamp"]) for r in rows if "k_momentum_rows<double" in r["Kernel_Name"]),
                  default=1 << 62)
        data[cnt] = (rows, t64)
    written = []
    for clock, pats in PROFILE_SYMBOLS.items():
        rec = {"kernel": clock, "workload": workload_desc,
               "counters": "FETCH_SIZE and WRITE_SIZE in KiB per dispatch, mean over the dispatches "
                           "of the fine-level grid in one run each",
               "fetch_scale": round(fscale, 5), "write_scale": round(wscale, 5),
               "calibration": f"{cal['fetch'][1].split('(')[0]} of 1 GiB (bench.py --copy-probe 1024): "
                              f"FETCH_SIZE {cal['fetch'][0]:.1f} KiB, WRITE_SIZE {cal['write'][0]:.1f} KiB",
               "commands": [c.format(dir=os.path.abspath(d))
                            for c in PROFILE_COMMANDS[1:]]}
        for form, k in (("mixed", 0), ("fp64", 1)):
            pat = pats[k] if pats[k] is not None else pats[1 - k]
            if pat is None:
                continue
            vals = {}
            for cnt in ("fetch", "write"):
                rows, t64 = data[cnt]
                sel = [r for r in rows if re.search(pat, r["Kernel_Name"])]
                if timed:
                    sel = [r for r in sel
                           if r["Kernel_Name"].split("(")[0].replace("void ", "") in timed[form]]
                if pats[1 - k] is None or pats[k] is None or pats[0] == pats[1]:
                    sel = [r for r in sel if (int(r["Start_Timestamp"]) >= t64) == (form == "fp64")]
                if not sel:
                    continue
                g = max(int(r["Grid_Size"]) for r in sel)
                sel = [r for r in sel if int(r["Grid_Size"]) == g]
                vals[cnt] = (sum(float(r["Counter_Value"]) for r in sel) / len(sel), len(sel), g,
                             sel[0]["Kernel_Name"].split("(")[0].replace("void ", ""))
            if len(vals) == 2:
                rec[form] = {"symbol": vals["fetch"][3], "grids": [vals["fetch"][2]],
                             "dispatches": [vals["fetch"][1], vals["write"][1]],
                             "fetch_kib": round(vals["fetch"][0], 1), "write_kib": round(vals["write"][0], 1)}
        if "mixed" in rec or "fp64" in rec:
            path = os.path.join(ROOT, "profiles", f"traffic_{clock}_bfs.json")
            with open(path, "w") as fh:
                json.dump(rec, fh, indent=1)
            written.append(path)
    return written


def summarize_trace(d, steps=3, timed=None):
    """Per-iteration kernel table of the timed outer iterations of each form
    (windows from one momentum assembly launch to the next); the kernel
    symbols met there go to timed[form] when a dict is given."""
    import re
    rows = sorted(_csv_rows(_find_csv(os.path.join(d, "trace"), "kernel_trace.csv")),
                  key=lambda r: int(r["Start_Timestamp"]))
    short = lambda n: re.sub(r"\(.*", "", n).replace("void ", "")
    marks = [i for i, r in enumerate(rows) if "k_momentum_rows" in r["Kernel_Name"]]
    wins = {"mixed": [], "fp64": []}
    for a, b in zip(marks, marks[1:]):
        form = "mixed" if "k_momentum_rows<float" in rows[a]["Kernel_Name"] else "fp64"
        busy = sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in rows[a:b]) / 1e6
        wall = (int(rows[b]["Start_Timestamp"]) - int(rows[a]["Start_Timestamp"])) / 1e6
        if b - a > 100 and wall < 2 * busy:
            wins[form].append((a, b, busy, wall))
    out = []
    for form, ws in wins.items():
        ws = ws[-steps:]
        W = len(ws)
        agg = {}
        for a, b, _, _ in ws:
            for r in rows[a:b]:
                k = short(r["Kernel_Name"])
                e = agg.setdefault(k, [0, 0.0, r["VGPR_Count"], r["Accum_VGPR_Count"], r["LDS_Block_Size"], 0])
                e[0] += 1
                e[1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
                e[5] = max(e[5], int(r["Grid_Size_X"]))
        if timed is not None:
            timed[form] = {"oagpu::" + k if not k.startswith(("oagpu::", "__")) else k for k in agg}
        busy = sum(w[2] for w in ws) / W
        wall = sum(w[3] for w in ws) / W
        nl = sum(v[0] for v in agg.values()) / W
        out.append(f"## {form}: {nl:.0f} launches / iteration, device busy {busy:.2f} ms of {wall:.2f} ms wall")
        out.append("| kernel | launches / it | avg us | ms / it | % busy | VGPR | AGPR | LDS B | largest grid |")
        out.append("|---|---|---|---|---|---|---|---|---|")
        items = sorted(agg.items(), key=lambda x: -x[1][1])
        rest = items[28:]
        for k, v in items[:28]:
            out.append(f"| `{k}` | {v[0] / W:.1f} | {v[1] / v[0]:.1f} | {v[1] / W / 1e3:.3f} | "
                       f"{100 * v[1] / W / 1e3 / busy:.1f} | {v[2]} | {v[3]} | {v[4]} | {v[5]} |")
        if rest:
            out.append(f"| ({len(rest)} more kernels) | {sum(v[0] for _, v in rest) / W:.1f} | | "
                       f"{sum(v[1] for _, v in rest) / W / 1e3:.3f} | | | | | |")
        out.append("")
    return "\n".join(out)


# ----------------------------------------------------------------- driver
def fp64_counterpart(clock):
    """The FP64 form's kernel doing the work of the mixed form's `clock`
    (its FP32 Krylov SpMV runs as the FP64 one there)."""
    return clock[:-4] if clock.endswith("_f32") else clock


def run_form(oa, f, step, steps, warmup, dominant=None, tune=True, clock_step=None, clock_steps=2):
    """Autotune the flow's launches (set-up, untimed), warm up (the first
    step builds the AMG hierarchy), then `steps` timed steps with every main
    kernel clocked (events recorded on the flow's stream, read after the
    steps).  With clock_step (graph-replayed steps, whose launches are not
    clocked) the kernels are clocked over `clock_steps` further steps of
    clock_step -- the same launch sequence issued without a graph -- right
    after the timed ones.  The dominant kernel is `dominant` or the one with
    the largest device time per clocked step.  Returns wall seconds, step
    outputs, per-kernel ms/step over the clocked steps, the dominant
    kernel's (launches, total ms) and the tuner's choices."""
    tuning = {k: [float(x) for x in v] for k, v in f.autotune().items()} if tune else None
    hist = [step() for _ in range(warmup)]
    if clock_step is None:
        f.profile("all")
    oa.load_library().oagpu_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        hist.append(step())
    oa.load_library().oagpu_synchronize()
    t = time.perf_counter() - t0
    nclk = steps
    if clock_step is not None:
        f.profile("all")
        for _ in range(clock_steps):
            clock_step()
        oa.load_library().oagpu_synchronize()
        nclk = clock_steps
    timed = {}
    for k in f.CLOCKS:
        ms, n = f.kernel_stats(k)
        if n:
            timed[k] = {"ms_per_step": ms / nclk, "launches_per_step": n / nclk, "ms": ms, "n": n}
    f.profile("none")
    if dominant is None or dominant not in timed:
        dominant = max(timed, key=lambda k: timed[k]["ms_per_step"])
    return {"seconds": t, "hist": np.array(hist), "warm": timed, "dominant": dominant,
            "dom_ms": timed[dominant]["ms"], "dom_launches": timed[dominant]["n"], "tuning": tuning}


def kernel_table(r, ab, ne):
    out = {}
    for k, v in r["warm"].items():
        e = {"ms_per_step": round(v["ms_per_step"], 3),
             "launches_per_step": round(v["launches_per_step"], 2)}
        if k in ab and v["launches_per_step"] > 0:
            per = v["ms_per_step"] / v["launches_per_step"]
            gbs = ab[k] / (per * 1e-3) / 1e9
            e.update({"ms_per_launch": round(per, 4), "bytes": int(ab[k]), "GB_s": round(gbs, 1),
                      "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)})
            if k in ("momentum", "pcorr", "scalar", "mdot", "gradU"):
                e["Melem_s"] = round(ne / (per * 1e-3) / 1e6, 1)
        out[k] = e
    return out


def roofline(r, ab, workload, precision):
    k = r["dominant"]
    per = r["dom_ms"] / max(r["dom_launches"], 1)
    gbs = ab[k] / (per * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": k, "launch_ms": round(per, 4), "launches": r["dom_launches"],
            "bytes_per_launch": int(ab[k]), "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(k, workload, precision)}


def coefficient_error(oa, dm, make_flow, setup):
    """Mixed-form momentum and p' coefficients and rhs against the FP64 form
    assembled from the same state: largest error of a matrix row relative to
    that row's largest coefficient, and of the rhs relative to its scale."""
    out = {}
    systems = {}
    for prec in (oa.OAGPU_FP64, oa.OAGPU_MIXED):
        f = make_flow(prec)
        setup(f)
        f.assemble_momentum()
        a3, b3 = f.system(3)
        f.assemble_pcorr()
        a1, b1 = f.system(1)
        systems[prec] = (a3, b3, a1, b1)
        f.close()
    rp, _, _ = dm.graph()
    starts = rp[:-1]
    for name, k in (("momentum", 0), ("pcorr", 2)):
        a, ref = systems[1][k], systems[0][k]
        d = np.abs(a - ref).reshape(a.shape[0], -1).max(axis=1)
        s = np.abs(ref).reshape(a.shape[0], -1).max(axis=1)
        err = np.maximum.reduceat(d, starts) / np.maximum(np.maximum.reduceat(s, starts), 1e-300)
        b, bref = systems[1][k + 1], systems[0][k + 1]
        out[name] = {"matrix_row_rel_max": float(err.max()),
                     "rhs_rel_max": float(np.abs(b - bref).max() / max(np.abs(bref).max(), 1e-300))}
    return out


def host_info():
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    model = platform.processor() or ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return cores, model


def oracle_module(threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.set_threads(threads)
    return pyoracle


# ---------------------------------------------------- backward-facing step
# backward-facing step with k-omega SST: inlet U = 1 with 1 % intensity and
# viscosity ratio 10, rho = 1, mu = 2e-5, pitzDaily controls of the
# reference's examples (relaxation 0.7 / 0.3 / 0.9, linear_linear velocity
# interpolation, linear solves to rtol 1e-2)
BFS_DIMS = (30, 210, 96, 112, 170)   # nx_up, nx_dn, ny_low, ny_high, nz: 7 996 800 elements
BFS_MU, BFS_KIN = 2.0e-5, 1.5e-4
BFS_WIN = BFS_KIN / (10.0 * BFS_MU)
BFS_CTL, BFS_SCTL = [20, 1.0e-2, 200, 1.0e-2], [20, 1.0e-2]
P_METHOD_NAMES = {"cg_amg": "AMG-CG (K-cycle)", "bicgstab_amg": "AMG-BiCGStab",
                  "bicgstab_sa": "smoothed-aggregation AMG-BiCGStab",
                  "cg_sa": "smoothed-aggregation AMG-CG"}
BFS_BC = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3, [0.0] * 3])


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


def bfs_params(mod, precision=None, p_method="cg_amg"):
    kw = dict(mu=BFS_MU, urf_U=0.7, urf_p=0.3, u_shifted=1)
    if precision is None:
        return mod.default_params(**kw)
    # p' by AMG-preconditioned flexible CG with K-cycle coarse levels, or by
    # AMG-preconditioned BiCGStab (the outlet and the graded cells make the
    # p' matrix nonsymmetric, by a few 1e-5 of its norm on this mesh)
    pm = {"cg_amg": mod.OAGPU_CG_AMG, "bicgstab_amg": mod.OAGPU_BICGSTAB_AMG,
          "bicgstab_sa": mod.OAGPU_BICGSTAB_SA}[p_method]
    return mod.default_params(precision, p_method=pm, **kw)


def bfs_cpu_baseline(dims, threads):
    """Oracle on a bounded sample of the backward-facing step: the same
    geometry and controls on a coarser grid, one flow + SST outer iteration
    after one untimed iteration."""
    pyoracle = oracle_module(threads)
    from openaccel_amd import meshgen
    mesh = meshgen.bfs_hex(*dims)
    c = pyoracle.Case(mesh, bfs_params(pyoracle), BFS_BC)
    bfs_setup(pyoracle, mesh, c, device=False)
    c.outer_iteration(BFS_CTL)
    c.sst_iteration(BFS_SCTL)
    t0 = time.perf_counter()
    c.outer_iteration(BFS_CTL)
    c.sst_iteration(BFS_SCTL)
    return {"outer_s": time.perf_counter() - t0, "ne": c.ne}


def graph_mode_steps(oa, f, steps, warmup=2):
    """Continue the step's flow with `steps` timed outer iterations (flow +
    SST) each replayed from one captured HIP graph, after `warmup` (the
    first runs uncaptured and sizes the iteration caps, the next captures)."""
    for _ in range(warmup):
        f.graph_iterations(BFS_CTL, BFS_SCTL, n=1)
    oa.load_library().oagpu_synchronize()
    t0 = time.perf_counter()
    h = [f.graph_iterations(BFS_CTL, BFS_SCTL, n=1)[0] for _ in range(steps)]
    oa.load_library().oagpu_synchronize()
    t = time.perf_counter() - t0
    return {"value": round(steps / t, 4), "ms_per_step": round(1e3 * t / steps, 3), "steps": steps,
            "krylov_iters": [[int(x[4]), int(x[5]), int(x[8]), int(x[9])] for x in h],
            **f.graph_stats()}


def bench_bfs_to_residual(args, oa, mesh, make_flow):
    """Wall time until every scaled residual of an outer iteration (U, V, W,
    p', k, omega) falls below the target, mixed and FP64, each outer
    iteration one graph launch (the device tests the target); a form stops
    at the iteration cap or after --time-cap seconds."""
    target, cap, tcap = args.to_residual, args.max_iters, args.time_cap
    out = {}
    forms = [{"mixed": oa.OAGPU_MIXED, "fp64": oa.OAGPU_FP64}[x] for x in args.forms.split(",")]
    for prec in forms:
        f = make_flow(prec)
        bfs_setup(oa, mesh, f)
        f.autotune()
        oa.load_library().oagpu_synchronize()
        t0 = time.perf_counter()
        hist = []
        while len(hist) < cap and time.perf_counter() - t0 < tcap:
            nb = min(args.report_every, cap - len(hist))
            if args.iteration == "graph":
                h = f.graph_iterations(BFS_CTL, BFS_SCTL, n=nb, tol=target)
            else:
                h = []
                for _ in range(nb):
                    h.append(np.concatenate([f.outer_iteration(BFS_CTL), f.sst_iteration(BFS_SCTL)]))
                    if max(h[-1][:4].max(), h[-1][6:8].max()) < target:
                        break
                h = np.array(h)
            hist.extend(h)
            if len(h) and max(h[-1][:4].max(), h[-1][6:8].max()) < target:
                break
            print(f"[bench] {'mixed' if prec else 'fp64'} {len(hist)} iterations, "
                  f"{time.perf_counter() - t0:.1f} s, residuals {np.round(np.log10(h[-1][[0, 1, 2, 3, 6, 7]]), 2).tolist()}",
                  file=sys.stderr, flush=True)
        oa.load_library().oagpu_synchronize()
        t = time.perf_counter() - t0
        hist = np.array(hist)
        res = lambda h: [float(x) for x in np.concatenate([h[:4], h[6:8]])]
        out[prec] = {"seconds": round(t, 3), "iterations": len(hist),
                     "iterations_per_s": round(len(hist) / t, 4),
                     "reached": bool(max(hist[-1, :4].max(), hist[-1, 6:8].max()) < target),
                     "final_residuals_U_V_W_p_k_omega": res(hist[-1]),
                     "history_every_50": [res(h) for h in hist[::50]],
                     **(f.graph_stats() if args.iteration == "graph" else {})}
        f.close()
    mx, f64 = out.get(oa.OAGPU_MIXED), out.get(oa.OAGPU_FP64)
    line = {"metric": f"time to scaled residual {target:g}", "value": (mx or f64)["seconds"], "unit": "s",
            "n_gpus": 1, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
            "dtype": "f32 local / f64 global", "data": "synthetic",
            "config": {"workload": f"backward-facing step k-omega SST, {mesh['conn'].shape[0]} hex8, "
                                   f"outer iterations (flow + SST"
                                   + (", each one HIP graph launch" if args.iteration == "graph" else "")
                                   + ") until "
                                   f"every scaled residual < {target:g} (cap {cap} iterations or "
                                   f"{tcap:g} s per form)"},
            "mixed": mx, "fp64": f64,
            "mixed_over_fp64_iterations_per_s": (round(mx["iterations_per_s"] / f64["iterations_per_s"], 3)
                                                 if mx and f64 else None)}
    print(json.dumps(line))


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
    dm = oa.DeviceMesh(mesh, caller_order=args.caller_order)
    nnz = dm.nnz

    def make_flow(prec):
        return oa.FlowCase(dm, bfs_params(oa, prec, args.bfs_p_method), BFS_BC)

    if args.to_residual:
        return bench_bfs_to_residual(args, oa, mesh, make_flow)
    res = {}
    for prec in (oa.OAGPU_MIXED, oa.OAGPU_FP64):
        f = make_flow(prec)
        bfs_setup(oa, mesh, f)

        clock_step = None
        if args.iteration == "graph":
            def step():  # one flow + SST outer iteration, one graph launch
                return f.graph_iterations(BFS_CTL, BFS_SCTL, n=1)[0]

            def clock_step():  # the same launches without the graph, clocked
                return f.graph_iterations(BFS_CTL, BFS_SCTL, n=1, eager=True)[0]
        else:
            def step():
                return np.concatenate([f.outer_iteration(BFS_CTL), f.sst_iteration(BFS_SCTL)])

        dom = None if prec == oa.OAGPU_MIXED else fp64_counterpart(res[oa.OAGPU_MIXED]["dominant"])
        res[prec] = run_form(oa, f, step, args.steps, args.warmup, dom, clock_step=clock_step)
        res[prec]["sa_levels"] = f.sa_levels() if args.bfs_p_method == "bicgstab_sa" else []
        res[prec]["graph"] = graph_mode_steps(oa, f, args.graph_steps) if args.graph_steps else None
        f.close()
        print(f"[bench] {'mixed' if prec else 'fp64'} {args.steps / res[prec]['seconds']:.3f} it/s, "
              f"dominant {res[prec]['dominant']}", file=sys.stderr, flush=True)
    rmx, r64 = res[oa.OAGPU_MIXED], res[oa.OAGPU_FP64]
    lv = rmx["sa_levels"]
    sa = (lv[0][1], lv[1][0]) if len(lv) > 1 else None
    ab_mx = algorithmic_bytes(nn, ne, nnz, 4, 4, sa=sa, scalar_eqs=2)
    ab_64 = algorithmic_bytes(nn, ne, nnz, 8, 8, sa=sa, scalar_eqs=2)
    it64, itmx = args.steps / r64["seconds"], args.steps / rmx["seconds"]
    print("[bench] coefficient error of the mixed form", file=sys.stderr, flush=True)
    cerr = coefficient_error(oa, dm, make_flow, lambda f: bfs_setup(oa, mesh, f))
    cpu = None
    if not args.no_cpu:
        cores, model = host_info()
        threads = args.cpu_threads or cores
        cd = tuple(max(2, int(round(d * args.cpu_scale))) for d in BFS_DIMS)
        print(f"[bench] cpu baseline (oracle, {threads} threads) at {cd}", file=sys.stderr, flush=True)
        cb = bfs_cpu_baseline(cd, threads)
        scale = cb["ne"] / ne
        cpu = {"value": round(scale / cb["outer_s"], 6), "unit": "outer iterations/s", "cores": threads,
               "kind": "port",
               "sample": f"oracle C restatement (OpenMP, {threads} threads on {model}), one flow + SST "
                         f"outer iteration of the same step at grid {cd} ({cb['ne']} elements, "
                         f"{cb['outer_s']:.2f} s) scaled by the element ratio {scale:.5f}"}

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
                               f"1.1; inlet U=1 (k {BFS_KIN}, omega {BFS_WIN:g}), outlet p=0, "
                               f"symmetry sides, rho=1, mu={BFS_MU}; one step = flow outer "
                               f"iteration + SST iteration; mixed form: FP32 local matrices into "
                               f"FP64 CRS, FP32 inner Krylov solves ("
                               f"{P_METHOD_NAMES[args.bfs_p_method]} p', GMRES "
                               f"momentum / k / omega) under FP64 refinement; "
                               + ("each outer iteration one HIP graph launch" if args.iteration == "graph"
                                  else "launches issued by the host"),
                   "dims": dims, "elements": ne, "nodes": nn, "nnz": nnz},
        "roofline": roofline(rmx, ab_mx, "bfs", 1),
        "kernels": kernel_table(rmx, ab_mx, ne),
        "autotune": {"mixed": rmx["tuning"], "fp64": r64["tuning"]},
        "fp64": {"value": round(it64, 4), "ms_per_step": round(1e3 * r64["seconds"] / args.steps, 3),
                 "roofline": roofline(r64, ab_64, "bfs", 0), "kernels": kernel_table(r64, ab_64, ne),
                 **hist(r64)},
        "mixed_over_fp64": round(itmx / it64, 3),
        "iteration": args.iteration,
        "graph_mode": ({"mixed": rmx["graph"], "fp64": r64["graph"],
                        "mixed_over_fp64": round(rmx["graph"]["value"] / r64["graph"]["value"], 3),
                        "note": "the same flows continued for the given steps, each outer iteration "
                                "(flow + SST) one HIP graph launch"}
                       if args.graph_steps else None),
        "coef_err_vs_fp64": cerr,
        "p_amg_levels": [{"rows": l[0], "entries": l[2], "p_entries": l[1], "omega_p": round(l[3], 4)}
                         for l in rmx["sa_levels"]],
        **hist(rmx),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


# ------------------------------------------------------------------ cavity
def cavity_cpu_baseline(n_sample, perturb, threads, p_method):
    """Oracle on a bounded sample of the same workload: the same cavity at
    n_sample^3, one outer iteration after one untimed one."""
    pyoracle = oracle_module(threads)
    from openaccel_amd import meshgen
    mesh = meshgen.cavity_hex(n_sample, perturb=perturb, seed=1)
    c = pyoracle.Case(mesh, pyoracle.default_params(p_method=p_method), BC)
    c.field("U")[:] = meshgen.initial_velocity(c.nn, 0)
    c.initialize()
    c.outer_iteration(CTL)
    t0 = time.perf_counter()
    c.assemble_momentum()
    t_mom = time.perf_counter() - t0
    t0 = time.perf_counter()
    c.outer_iteration(CTL)
    return {"outer_s": time.perf_counter() - t0, "mom_s": t_mom, "ne": c.ne}


def bench_cavity(args):
    import openaccel_amd as oa
    from openaccel_amd import meshgen
    oa.init(0)
    mesh = meshgen.cavity_hex(args.n, perturb=args.perturb, seed=1)
    ne = mesh["conn"].shape[0]
    print(f"[bench] cavity {args.n}^3 perturb={args.perturb}: {ne} elements", file=sys.stderr,
          flush=True)
    pm = {"cg_amg": oa.OAGPU_CG_AMG, "bicgstab_amg": oa.OAGPU_BICGSTAB_AMG,
          "bicgstab_sa": oa.OAGPU_BICGSTAB_SA, "cg_sa": oa.OAGPU_CG_SA}[args.p_method]
    dm = oa.DeviceMesh(mesh, caller_order=args.caller_order)
    nn, nnz = dm.nn, dm.nnz

    def make_flow(prec):
        return oa.FlowCase(dm, oa.default_params(prec, p_method=pm), BC)

    def setup(f):
        f.upload("U", meshgen.initial_velocity(dm.nn, 0))
        f.initialize()

    if args.to_residual:
        return bench_cavity_to_residual(args, oa, dm, make_flow, setup)
    # each outer iteration one graph launch (the smoothed-aggregation p' solvers)
    graph_mode = args.iteration == "graph" and pm in (oa.OAGPU_CG_SA, oa.OAGPU_BICGSTAB_SA)
    res = {}
    for prec in (oa.OAGPU_MIXED, oa.OAGPU_FP64):
        f = make_flow(prec)
        setup(f)
        dom = None if prec == oa.OAGPU_MIXED else fp64_counterpart(res[oa.OAGPU_MIXED]["dominant"])
        if graph_mode:
            res[prec] = run_form(oa, f, lambda: f.graph_iterations(CTL, None, n=1)[0], args.steps,
                                 args.warmup, dom,
                                 clock_step=lambda: f.graph_iterations(CTL, None, n=1, eager=True)[0])
            res[prec]["graph"] = f.graph_stats()
        else:
            res[prec] = run_form(oa, f, lambda: f.outer_iteration(CTL), args.steps, args.warmup, dom)
        f.close()
        print(f"[bench] {'mixed' if prec else 'fp64'} {args.steps / res[prec]['seconds']:.3f} it/s, "
              f"dominant {res[prec]['dominant']}", file=sys.stderr, flush=True)
    rmx, r64 = res[oa.OAGPU_MIXED], res[oa.OAGPU_FP64]
    # the mixed form's inner solves are FP32 (values and vectors)
    ab_mx = algorithmic_bytes(nn, ne, nnz, 4, 4)
    ab_64 = algorithmic_bytes(nn, ne, nnz, 8, 8)
    itmx, it64 = args.steps / rmx["seconds"], args.steps / r64["seconds"]
    cerr = coefficient_error(oa, dm, make_flow, setup)
    cpu = None
    if not args.no_cpu:
        cores, model = host_info()
        threads = args.cpu_threads or cores
        cn = min(args.cpu_n, args.n)
        print(f"[bench] cpu baseline (oracle, {threads} threads) at {cn}^3", file=sys.stderr, flush=True)
        cb = cavity_cpu_baseline(cn, args.perturb, threads, 1 if pm in (oa.OAGPU_CG_AMG, oa.OAGPU_CG_SA) else 2)
        scale = (cn / args.n) ** 3
        cpu = {"value": round(scale / cb["outer_s"], 6), "unit": "outer iterations/s",
               "cores": threads, "kind": "port",
               "sample": f"oracle C restatement (OpenMP, {threads} threads on {model}), one outer "
                         f"iteration of the same cavity at {cn}^3 ({cb['ne']} elements, "
                         f"{cb['outer_s']:.2f} s) scaled by the element ratio {scale:.4f}; its "
                         f"momentum assembly alone runs at {cb['ne'] / cb['mom_s'] / 1e6:.3f} Melem/s"}
    wl = "cavity" if not args.perturb else "cavity_perturbed"
    line = {
        "metric": METRIC, "value": round(itmx, 4), "unit": "outer iterations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * rmx["seconds"] / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 local / f64 global", "data": "synthetic",
        "config": {"workload": f"lid-driven cavity, uniform{' perturbed' if args.perturb else ''} "
                               f"{args.n}^3 hex8 ({ne} elements, {nn} nodes), Re=100, mixed form; "
                               f"Jacobi-GMRES(20 it, rtol 0.1) momentum, "
                               f"{P_METHOD_NAMES[args.p_method]} (200 it, rtol 1e-6) p'; "
                               + ("each outer iteration one HIP graph launch" if graph_mode
                                  else "launches issued by the host"),
                   "n": args.n, "elements": ne, "nodes": nn, "nnz": nnz},
        "graph": {"mixed": rmx.get("graph"), "fp64": r64.get("graph")} if graph_mode else None,
        "roofline": roofline(rmx, ab_mx, f"{wl}_{args.n}", 1),
        "kernels": kernel_table(rmx, ab_mx, ne),
        "autotune": {"mixed": rmx["tuning"], "fp64": r64["tuning"]},
        "fp64": {"value": round(it64, 4), "ms_per_step": round(1e3 * r64["seconds"] / args.steps, 3),
                 "roofline": roofline(r64, ab_64, f"{wl}_{args.n}", 0),
                 "kernels": kernel_table(r64, ab_64, ne),
                 "krylov_iters_mean": [round(float(v), 1) for v in r64["hist"][:, 4:6].mean(axis=0)]},
        "krylov_iters_mean": [round(float(v), 1) for v in rmx["hist"][:, 4:6].mean(axis=0)],
        "mixed_over_fp64": round(itmx / it64, 3),
        "coef_err_vs_fp64": cerr,
        "residuals": [[float(x) for x in h[:4]] for h in rmx["hist"]],
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def bench_cavity_to_residual(args, oa, dm, make_flow, setup):
    """Wall time until every scaled control residual of an outer iteration
    falls below the target, mixed and FP64 (p' solved to 1e-6 as in the
    reference's cavity controls)."""
    target, cap = args.to_residual, args.max_iters
    out = {}
    for prec in (oa.OAGPU_MIXED, oa.OAGPU_FP64):
        f = make_flow(prec)
        setup(f)
        oa.load_library().oagpu_synchronize()
        t0 = time.perf_counter()
        hist = []
        for _ in range(cap):
            h = f.outer_iteration(CTL)
            hist.append(h)
            if h[:4].max() < target:
                break
        oa.load_library().oagpu_synchronize()
        t = time.perf_counter() - t0
        hist = np.array(hist)
        out[prec] = {"seconds": round(t, 3), "iterations": len(hist),
                     "reached": bool(hist[-1, :4].max() < target),
                     "final_residuals": [float(x) for x in hist[-1, :4]],
                     "history_every_50": [[float(x) for x in h[:4]] for h in hist[::50]]}
        f.close()
        print(f"[bench] {'mixed' if prec else 'fp64'}: {len(hist)} iterations, {t:.2f} s",
              file=sys.stderr, flush=True)
    mx, f64 = out[oa.OAGPU_MIXED], out[oa.OAGPU_FP64]
    line = {"metric": f"time to scaled residual {target:g}", "value": mx["seconds"], "unit": "s",
            "n_gpus": 1, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
            "dtype": "f32 local / f64 global", "data": "synthetic",
            "config": {"workload": f"lid-driven cavity {args.n}^3 hex8, Re=100, lid (1,0,0), "
                                   f"U0 ~ N(0, 1e-3), outer iterations until every scaled control "
                                   f"residual < {target:g} (cap {cap})", "n": args.n},
            "mixed": mx, "fp64": f64,
            "mixed_speedup": round(f64["seconds"] / mx["seconds"], 3) if mx["seconds"] > 0 else None}
    print(json.dumps(line))


# -------------------------------------------------------------------- duct
# mixed-topology duct 1 x 0.2 x 0.2: plug inflow U = 1, static-pressure
# outlet p = 0, no-slip walls, rho = 1, mu = 1e-3, high-resolution advection
DUCT_M, DUCT_LAYERS = 64, (187, 8, 25, 75, 25)   # 3.87M elements
DUCT_MU = 1.0e-3
DUCT_CTL = [20, 1.0e-2, 200, 1.0e-2]
DUCT_BC = np.array([[1.0, 0.0, 0.0], [0.0] * 3, [0.0] * 3])


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


def duct_mesh(m):
    from openaccel_amd import meshgen
    layers = tuple(max(1, round(x * m / DUCT_M)) for x in DUCT_LAYERS)
    return meshgen.duct_mixed(m, layers=layers, length=sum(layers) * 0.2 / m, seed=1), layers


def bench_duct(args):
    import openaccel_amd as oa
    oa.init(0)
    m = max(8, int(round(DUCT_M * args.duct_scale)))
    mesh, layers = duct_mesh(m)
    ne, nn = mesh["conn"].shape[0], mesh["xyz"].shape[0]
    counts = np.bincount(mesh["etype"], minlength=4)
    print(f"[bench] mixed duct m={m} layers={layers}: {ne} elements "
          f"(hex {counts[0]}, tet {counts[1]}, wedge {counts[2]}, pyramid {counts[3]})",
          file=sys.stderr, flush=True)
    dm = oa.DeviceMesh(mesh, caller_order=args.caller_order)
    res = {}
    for prec in (oa.OAGPU_MIXED, oa.OAGPU_FP64):
        pm = {"cg_amg": oa.OAGPU_CG_AMG, "bicgstab_amg": oa.OAGPU_BICGSTAB_AMG,
              "bicgstab_sa": oa.OAGPU_BICGSTAB_SA}[args.bfs_p_method]
        f = oa.FlowCase(dm, oa.default_params(prec, p_method=pm, mu=DUCT_MU,
                                              high_res=1), DUCT_BC)
        duct_setup(f, mesh)
        res[prec] = run_form(oa, f, lambda: f.outer_iteration(DUCT_CTL), args.steps, args.warmup)
        f.close()
        print(f"[bench] {'mixed' if prec else 'fp64'} {args.steps / res[prec]['seconds']:.3f} it/s",
              file=sys.stderr, flush=True)
    rmx, r64 = res[oa.OAGPU_MIXED], res[oa.OAGPU_FP64]
    itmx, it64 = args.steps / rmx["seconds"], args.steps / r64["seconds"]
    cpu = None
    if not args.no_cpu:
        cores, model = host_info()
        threads = args.cpu_threads or cores
        pyoracle = oracle_module(threads)
        cm = 24
        cmesh, clayers = duct_mesh(cm)
        c = pyoracle.Case(cmesh, pyoracle.default_params(mu=DUCT_MU, high_res=1, p_method=2), DUCT_BC)
        duct_setup(c, cmesh, device=False)
        c.update_blending("U")
        c.outer_iteration(DUCT_CTL)
        t0 = time.perf_counter()
        c.outer_iteration(DUCT_CTL)
        ts = time.perf_counter() - t0
        scale = c.ne / ne
        cpu = {"value": round(scale / ts, 6), "unit": "outer iterations/s", "cores": threads,
               "kind": "port",
               "sample": f"oracle C restatement (OpenMP, {threads} threads on {model}), one outer "
                         f"iteration of the same duct at m={cm} ({c.ne} elements, {ts:.2f} s) scaled "
                         f"by the element ratio {scale:.5f}"}
    ms_step = {k: round(v["ms_per_step"], 3) for k, v in rmx["warm"].items()}
    line = {
        "metric": METRIC, "value": round(itmx, 4), "unit": "outer iterations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * rmx["seconds"] / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 local / f64 global", "data": "synthetic",
        "config": {"workload": f"mixed-topology duct 1 x 0.2 x 0.2, {ne} elements ({counts[1]} tet4, "
                               f"{counts[0]} hex8, {counts[2]} wedge6, {counts[3]} pyramid5), node "
                               f"jitter N(0, 0.1 h); plug inflow U=1, outlet p=0, rho=1, "
                               f"mu={DUCT_MU}, high-resolution advection; mixed form; "
                               f"{P_METHOD_NAMES[args.bfs_p_method]} p'",
                   "m": m, "layers": layers, "elements": ne, "nodes": nn},
        "fp64": {"value": round(it64, 4), "ms_per_step": round(1e3 * r64["seconds"] / args.steps, 3),
                 "kernel_ms_per_step": {k: round(v["ms_per_step"], 3) for k, v in r64["warm"].items()}},
        "mixed_over_fp64": round(itmx / it64, 3),
        "kernel_ms_per_step": ms_step,
        "assembly_Melem_s": {k: round(ne / (v["ms_per_step"] / v["launches_per_step"] * 1e-3) / 1e6, 1)
                             for k, v in rmx["warm"].items() if k in ("momentum", "pcorr", "mdot")},
        "residuals": [[float(x) for x in h[:4]] for h in rmx["hist"]],
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


# ------------------------------------------------------------------- batch
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


def batch_cpu_baseline(cases, iters, tol, threads, budget_s=20.0):
    """Oracle on a bounded sample of the batch: the cases in order, each to
    the residual target or the iteration cap, until about budget_s of CPU
    time; cases/s over the whole batch scaled by the elements sampled."""
    pyoracle = oracle_module(threads)
    from openaccel_amd import meshgen
    done_el, t_all, k = 0, 0.0, 0
    for c in cases:
        if t_all > budget_s:
            break
        m = meshgen.cavity_hex(c["n"])
        case = pyoracle.Case(m, pyoracle.default_params(mu=c["mu"], p_method=1),
                             np.array([c["lid"], [0.0, 0.0, 0.0]]))
        case.field("U")[:] = meshgen.initial_velocity(case.nn, c["n"])
        case.initialize()
        t0 = time.perf_counter()
        for _ in range(iters):
            h = case.outer_iteration(CTL)
            if tol > 0 and h[:4].max() < tol:
                break
        t_all += time.perf_counter() - t0
        done_el += c["n"] ** 3
        k += 1
    elems = sum(c["n"] ** 3 for c in cases)
    scale = done_el / elems
    _, model = host_info()
    return {"value": round(len(cases) * scale / t_all, 6), "unit": "cases/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle C restatement (OpenMP, {threads} threads on {model}), Jacobi-CG p', "
                      f"the first {k} of the {len(cases)} cases ({done_el} of {elems} elements, "
                      f"{t_all:.1f} s) each run as in the batch, scaled by the element fraction"}


def bench_batch(args):
    import openaccel_amd as oa
    from openaccel_amd import meshgen
    oa.init(0)
    cases = batch_cases(args.batch_cases)
    meshes = {n: oa.DeviceMesh(meshgen.cavity_hex(n), caller_order=args.caller_order) for n in sorted({c["n"] for c in cases})}
    elems = sum(c["n"] ** 3 for c in cases)
    print(f"[bench] batch of {len(cases)} cavities, {elems} elements in all", file=sys.stderr, flush=True)

    def flow(c, precision):
        dm = meshes[c["n"]]
        pm = {"cg_amg": oa.OAGPU_CG_AMG, "cg_sa": oa.OAGPU_CG_SA}[args.p_method]
        f = oa.FlowCase(dm, oa.default_params(precision, mu=c["mu"], p_method=pm),
                        np.array([c["lid"], [0.0, 0.0, 0.0]]))
        f.upload("U", meshgen.initial_velocity(dm.nn, c["n"]))
        f.initialize()
        return f

    iters, tol = args.batch_iters, args.batch_tol
    # FP64, one case at a time, each to the residual target (or the cap)
    seq = np.zeros(len(cases))
    seq_it = np.zeros(len(cases), dtype=int)
    for k, c in enumerate(cases):
        f = flow(c, oa.OAGPU_FP64)
        oa.load_library().oagpu_synchronize()
        t1 = time.perf_counter()
        for it in range(iters):
            h = f.outer_iteration(CTL)
            if tol > 0 and h[:4].max() < tol:
                break
        oa.load_library().oagpu_synchronize()
        seq[k] = time.perf_counter() - t1
        seq_it[k] = it + 1
        f.close()
        print(f"[bench] case {k}: {c['n']}^3, Re {1.0 / c['mu']:.0f}, {it + 1} iterations, "
              f"{seq[k]:.2f} s", file=sys.stderr, flush=True)
    t_seq = seq.sum()
    print(f"[bench] fp64 one at a time: {t_seq:.2f} s", file=sys.stderr, flush=True)

    graph_mode = args.iteration == "graph" and args.p_method == "cg_sa"

    def batched(precision, chunk=200):
        """All cases on 16 host threads, one stream each, in rounds of up to
        `chunk` outer iterations (a progress line per round); a case leaves
        once its residuals are below tol; per-case seconds sum its rounds."""
        flows = [flow(c, precision) for c in cases]
        n = len(cases)
        hist = [[] for _ in range(n)]
        done, sec = np.zeros(n, dtype=int), np.zeros(n)
        live = list(range(n))
        oa.load_library().oagpu_synchronize()
        t0 = time.perf_counter()
        run = oa.batch_graph_iterations if graph_mode else oa.batch_outer_iterations
        while live:
            step = min(chunk, iters - int(done[live[0]]))
            if step <= 0:
                break
            h, d, sc = run([flows[k] for k in live], CTL, step, tol=tol, threads=16)
            nxt = []
            for j, k in enumerate(live):
                hist[k].extend(h[j][:d[j]])
                done[k] += d[j]
                sec[k] += sc[j]
                if d[j] == step and not (tol > 0 and h[j][d[j] - 1][:4].max() < tol):
                    nxt.append(k)
            live = nxt
            print(f"[bench] {'mixed' if precision else 'fp64'} batched: {n - len(live)} of {n} cases done, "
                  f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        t = time.perf_counter() - t0
        for f in flows:
            f.close()
        width = max(len(x) for x in hist)
        H = np.zeros((n, width, len(hist[0][0])))
        for k in range(n):
            H[k, :len(hist[k])] = hist[k]
        return t, H, done, sec

    # the concurrency alone (FP64 batched), then concurrency + the mixed form
    t_b64, hist64, done64, sec64 = batched(oa.OAGPU_FP64)
    print(f"[bench] fp64 batched: {t_b64:.2f} s", file=sys.stderr, flush=True)
    t_batch, hist, done, sec = batched(oa.OAGPU_MIXED)
    print(f"[bench] mixed batched: {t_batch:.2f} s", file=sys.stderr, flush=True)
    last = np.array([hist[k][done[k] - 1][:4].max() for k in range(len(cases))])
    cpu = None
    if not args.no_cpu:
        cpu = batch_cpu_baseline(cases, iters, tol, args.cpu_threads or host_info()[0])
    line = {
        "metric": METRIC, "value": round(len(cases) / t_batch, 4), "unit": "cases/s",
        "n_gpus": 1, "steps": iters, "warmup": 0,
        "ms_per_step": round(1e3 * t_batch / max(done.max(), 1), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 local / f64 global", "data": "synthetic",
        "config": {"workload": f"{len(cases)} independent lid-driven cavities, sizes from "
                               f"{BATCH_SIZES}^3 hex8, Re log-uniform in [10, 1000], lid velocity a "
                               f"random unit vector in the lid plane (seeded); each case runs until "
                               f"its scaled residuals are below {tol:g} or {iters} outer "
                               f"iterations; mixed form batched (one stream per case, 16 host "
                               f"threads"
                               + ("; each outer iteration of a case one HIP graph launch" if graph_mode else "")
                               + ") vs FP64 one case at a time (launches issued by the host)",
                   "cases": len(cases), "elements": elems, "iteration_cap": iters},
        "fp64_one_at_a_time": {"value": round(len(cases) / t_seq, 4), "unit": "cases/s",
                               "per_case_s_p50": round(float(np.percentile(seq, 50)), 4),
                               "per_case_s_p99": round(float(np.percentile(seq, 99)), 4),
                               "reached": int((seq_it < iters).sum())},
        "fp64_batched": {"value": round(len(cases) / t_b64, 4), "unit": "cases/s",
                         "per_case_s_p50": round(float(np.percentile(sec64, 50)), 4),
                         "per_case_s_p99": round(float(np.percentile(sec64, 99)), 4),
                         "iterations_p50": float(np.percentile(done64, 50))},
        "batched_over_one_at_a_time": round(t_seq / t_batch, 3),
        "concurrency_gain_fp64": round(t_seq / t_b64, 3),
        "mixed_over_fp64_batched": round(t_b64 / t_batch, 3),
        "per_case_s_p50": round(float(np.percentile(sec, 50)), 4),
        "per_case_s_p99": round(float(np.percentile(sec, 99)), 4),
        "reached": int((last < tol).sum()) if tol > 0 else None,
        "iterations_p50": float(np.percentile(done, 50)),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("bfs", "cavity", "duct", "batch"), default="bfs",
                    help="bfs: the 8M-cell backward-facing step with k-omega SST (default); "
                         "cavity: the lid-driven cavity; duct: the 3.9M-cell mixed-topology "
                         "duct; batch: 96 independent cavities")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--iteration", choices=("graph", "eager"), default="eager",
                    help="timed steps: each outer iteration one HIP graph launch, or launched by the host")
    ap.add_argument("--graph-steps", type=int, default=5,
                    help="bfs: further timed steps of each form in graph mode (0: none)")
    ap.add_argument("--caller-order", action="store_true",
                    help="keep the generator's node numbering on the device (no locality renumbering)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0: all host cores)")
    ap.add_argument("--cpu-scale", type=float, default=0.5,
                    help="step: oracle sample grid as a fraction of the cell counts per direction")
    ap.add_argument("--cpu-n", type=int, default=64, help="cavity: oracle sample grid")
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--perturb", type=float, default=0.0)
    ap.add_argument("--p-method", choices=("cg_amg", "bicgstab_amg", "bicgstab_sa", "cg_sa"),
                    default="cg_sa",
                    help="cavity and batched cavities: p' solver")
    ap.add_argument("--bfs-p-method", choices=("cg_amg", "bicgstab_amg", "bicgstab_sa"),
                    default="bicgstab_sa",
                    help="backward-facing step and duct: p' solver")
    ap.add_argument("--to-residual", type=float, default=0.0,
                    help="cavity: time the outer iterations until every scaled residual is below this")
    ap.add_argument("--max-iters", type=int, default=4000)
    ap.add_argument("--time-cap", type=float, default=240.0,
                    help="bfs --to-residual: wall seconds per form")
    ap.add_argument("--forms", default="mixed,fp64", help="bfs --to-residual: forms to run")
    ap.add_argument("--report-every", type=int, default=25,
                    help="bfs --to-residual: iterations between progress lines")
    ap.add_argument("--batch-cases", type=int, default=BATCH_N)
    ap.add_argument("--batch-iters", type=int, default=BATCH_ITERS)
    ap.add_argument("--batch-tol", type=float, default=0.0)
    ap.add_argument("--duct-scale", type=float, default=1.0)
    ap.add_argument("--bfs-scale", type=float, default=1.0,
                    help="scale of the step's cell counts per direction (1: 8M elements)")
    ap.add_argument("--summarize-profiles", metavar="DIR", default="",
                    help="write profiles/traffic_*_bfs.json and print the per-iteration kernel table "
                         "from the rocprofv3 runs under DIR (--profile-commands lists them)")
    ap.add_argument("--profile-commands", nargs="?", const=os.path.join(ROOT, "prof_out"), default=None,
                    metavar="DIR",
                    help="print the rocprofv3 commands, writing under DIR (default prof_out), whose "
                         "output --summarize-profiles DIR reads")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device once (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X only: the multi-GPU path (element partition + "
              "halo exchange) is designed but not built; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.profile_commands:
        print("\n".join(c.format(dir=args.profile_commands) for c in PROFILE_COMMANDS))
        return
    if args.summarize_profiles:
        timed = {}
        table = summarize_trace(args.summarize_profiles, timed=timed)
        for p in summarize_traffic(args.summarize_profiles,
                                   "backward-facing step k-omega SST, 7996800 hex8 (8120619 nodes), "
                                   "bench.py defaults", timed):
            print(p, file=sys.stderr)
        print(table)
        return
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
    {"bfs": bench_bfs, "cavity": bench_cavity, "duct": bench_duct, "batch": bench_batch}[args.workload](args)


if __name__ == "__main__":
    main()
