"""openaccel_amd -- MI355X CVFEM assemble-and-solve path behind OpenAccel's
assembler / linearSystem interfaces.

The product is liboagpu.so (HIP, gfx950) with the C ABI of include/oagpu.h.
This module is a thin ctypes binding of that ABI for Python callers, tests and
bench.py.  It mirrors the reference's call sequence:

  reference (C++)                                     here
  ------------------------------------------------    ------------------------------
  mesh + nodeGraph::buildGraph_ (nodeGraph.cpp:29)     DeviceMesh(...)
  navierStokesAssembler::assemble (assembler.h:46)     FlowCase.assemble_momentum()
  pressureCorrectionAssembler::assemble                FlowCase.assemble_pcorr()
  linearSystem<N>::solve (linearSystem.h:103, 416)     FlowCase.solve(N, ...)
  flowModel::updateMassFlowRate (flowModel.cpp:5272)   FlowCase.update_mdot()
  flowModel::updateMassDivergenceField (:13541)        FlowCase.mass_divergence()
  nodeField::updateGradientField (nodeField.hpp:3253)  FlowCase.gradient("U" | "p")
  equation::correctField_ (equation.h:204)             FlowCase.correct(N, urf)
  segregatedFlowEquations::solve (:76-297)             FlowCase.outer_iteration(ctl)

There is no CPU fallback: if liboagpu.so is missing or no GPU is present the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboagpu.so")

OAGPU_FP64, OAGPU_MIXED = 0, 1
OAGPU_STENCIL_FULL, OAGPU_STENCIL_REDUCED = 0, 1
OAGPU_MESH_CALLER_ORDER = 0x100
OAGPU_GMRES, OAGPU_CG, OAGPU_BICGSTAB, OAGPU_CG_AMG, OAGPU_BICGSTAB_AMG = 0, 1, 2, 3, 4
OAGPU_BICGSTAB_SA, OAGPU_CG_SA = 5, 6
OAGPU_BC_WALL, OAGPU_BC_INLET, OAGPU_BC_OUTLET, OAGPU_BC_SYMMETRY, OAGPU_BC_SLIP = 0, 1, 2, 3, 4
OAGPU_BC_OPENING = 5
OAGPU_HEX8, OAGPU_TET4, OAGPU_WED6, OAGPU_PYR5 = 0, 1, 2, 3
OAGPU_BDF1, OAGPU_BDF2 = 1, 2
TOPOLOGY_SIZES = {OAGPU_HEX8: (8, 12), OAGPU_TET4: (4, 6), OAGPU_WED6: (6, 9), OAGPU_PYR5: (5, 12)}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class FlowParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("dt", C.c_double),
                ("urf_U", C.c_double), ("urf_p", C.c_double), ("urf_mdot", C.c_double),
                ("grad_urf", C.c_double), ("bnd_relax", C.c_double),
                ("u_shifted", C.c_int), ("ugrad_shifted", C.c_int),
                ("p_shifted", C.c_int), ("pgrad_shifted", C.c_int),
                ("p_ref_node", C.c_int), ("precision", C.c_int), ("p_method", C.c_int),
                ("use_graphs", C.c_int), ("high_res", C.c_int)]


class SstParams(C.Structure):
    _fields_ = [("urf_k", C.c_double), ("urf_w", C.c_double), ("urf_mut", C.c_double),
                ("k_shifted", C.c_int), ("w_shifted", C.c_int),
                ("kgrad_shifted", C.c_int), ("wgrad_shifted", C.c_int)]


def default_sst(**kw) -> SstParams:
    """k-omega SST numerics: relaxation 0.8 for k and omega, no eddy-viscosity
    relaxation, linear_linear interpolation and gradients."""
    s = SstParams(urf_k=0.8, urf_w=0.8, urf_mut=1.0, k_shifted=1, w_shifted=1,
                  kgrad_shifted=1, wgrad_shifted=1)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


class ControlData(C.Structure):
    _fields_ = [("n_iterations", C.c_int), ("solver_initial_res", C.c_double),
                ("solver_final_res", C.c_double), ("scaled_initial_res", C.c_double * 3)]


class OagpuError(RuntimeError):
    pass


# exported entry points and their signatures (mirrors include/oagpu.h)
_SIGS = {
    "oagpu_init": [C.c_int],
    "oagpu_last_error": [],
    "oagpu_synchronize": [],
    "oagpu_mesh_create": [C.c_int, _dp, C.c_int, _ip, _ip, C.c_int, _ip, _ip, _ip, C.c_int,
                          C.POINTER(C.c_void_p)],
    "oagpu_mesh_destroy": [C.c_void_p],
    "oagpu_mesh_sizes": [C.c_void_p, _ip, _ip, _ip, _ip, _ip],
    "oagpu_graph_copy": [C.c_void_p, _ip, _ip, _ip],
    "oagpu_scatter_map_copy": [C.c_void_p, _ip, _ip, C.POINTER(C.c_ulonglong)],
    "oagpu_mesh_volume_copy": [C.c_void_p, _dp],
    "oagpu_flow_create": [C.c_void_p, C.POINTER(FlowParams), C.c_int, _dp,
                          C.POINTER(C.c_void_p)],
    "oagpu_flow_destroy": [C.c_void_p],
    "oagpu_field_upload": [C.c_void_p, C.c_char_p, _dp, C.c_size_t],
    "oagpu_field_download": [C.c_void_p, C.c_char_p, _dp, C.c_size_t],
    "oagpu_initialize": [C.c_void_p],
    "oagpu_assemble_momentum": [C.c_void_p],
    "oagpu_assemble_pcorr": [C.c_void_p],
    "oagpu_system_download": [C.c_void_p, C.c_int, _dp, _dp],
    "oagpu_system_upload": [C.c_void_p, C.c_int, _dp, _dp],
    "oagpu_solution_download": [C.c_void_p, C.c_int, _dp],
    "oagpu_update_mdot": [C.c_void_p],
    "oagpu_mass_divergence": [C.c_void_p],
    "oagpu_gradient": [C.c_void_p, C.c_int],
    "oagpu_correct": [C.c_void_p, C.c_int, C.c_double],
    "oagpu_velocity_correction": [C.c_void_p, C.c_int],
    "oagpu_update_scales": [C.c_void_p, _dp, _dp],
    "oagpu_scaled_residual": [C.c_void_p, C.c_int, _dp],
    "oagpu_solve": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                    C.POINTER(ControlData)],
    "oagpu_element_geometry": [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp],
    "oagpu_element_geometry_noalloc": [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "oagpu_update_blending": [C.c_void_p, C.c_int],
    "oagpu_flow_set_bc_types": [C.c_void_p, C.c_int, _ip],
    "oagpu_flow_set_bc_values": [C.c_void_p, C.c_int, _dp, _dp, _dp],
    "oagpu_update_reversal": [C.c_void_p],
    "oagpu_batch_outer_iterations": [C.POINTER(C.c_void_p), C.c_int, _dp, C.c_int, C.c_double,
                                     C.c_int, _dp, _ip, _dp],
    "oagpu_reversal_download": [C.c_void_p, _ip, C.c_size_t],
    "oagpu_sst_enable": [C.c_void_p, C.POINTER(SstParams)],
    "oagpu_sst_update": [C.c_void_p],
    "oagpu_assemble_scalar": [C.c_void_p, C.c_int],
    "oagpu_correct_scalar": [C.c_void_p, C.c_int],
    "oagpu_sst_iteration": [C.c_void_p, _dp, _dp],
    "oagpu_amg_info": [C.c_void_p, _ip, _ip],
    "oagpu_amg_level_copy": [C.c_void_p, C.c_int, _ip, _ip, _ip, _dp],
    "oagpu_sa_info": [C.c_void_p, _ip, _ip, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), _dp],
    "oagpu_sa_level_copy": [C.c_void_p, C.c_int, _ip, _ip, _ip, _dp, _ip, _ip, _dp],
    "oagpu_outer_iteration": [C.c_void_p, _dp, _dp],
    "oagpu_graph_iterations": [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_int, _dp, _ip],
    "oagpu_batch_graph_iterations": [C.POINTER(C.c_void_p), C.c_int, _dp, C.c_int, C.c_double,
                                     C.c_int, _dp, _ip, _dp],
    "oagpu_graph_stats": [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                          C.POINTER(C.c_longlong), _ip],
    "oagpu_time_step": [C.c_void_p, C.c_int, C.c_double],
    "oagpu_profile": [C.c_void_p, C.c_char_p],
    "oagpu_autotune": [C.c_void_p, _dp, C.c_int],
    "oagpu_tune_set": [C.c_void_p, C.c_char_p, C.c_int],
    "oagpu_kernel_stats": [C.c_void_p, C.c_char_p, _dp, C.POINTER(C.c_longlong)],
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load liboagpu.so and declare its signatures (no GPU call is made)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise OagpuError(f"liboagpu.so not built: {path} (run python -m openaccel_amd.build)")
        L = C.CDLL(path)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_char_p if name == "oagpu_last_error" else C.c_int
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        msg = load_library().oagpu_last_error()
        raise OagpuError(f"oagpu status {rc}: {msg.decode() if msg else ''}")


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def init(device: int = 0):
    _check(load_library().oagpu_init(device))


def element_geometry(topology: int, coords, shifted: int = 0):
    """Device master-element geometry of a batch of elements of one topology:
    (areav [n][nscs][3], scv [n][npe], dndx [n][nscs][npe][3], detj [n][nscs])."""
    npe, ns = TOPOLOGY_SIZES[topology]
    c = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, npe, 3)
    n = c.shape[0]
    A = np.zeros((n, ns, 3))
    V = np.zeros((n, npe))
    G = np.zeros((n, ns, npe, 3))
    D = np.zeros((n, ns))
    _check(load_library().oagpu_element_geometry(topology, shifted, n, _d(c), _d(A), _d(V), _d(G),
                                                 _d(D)))
    return A, V, G, D


def default_params(precision: int = OAGPU_FP64, **kw) -> FlowParams:
    """Defaults of controls.h:57-290 for the lid-driven cavity
    (velocity interpolation trilinear, gradients linear_linear)."""
    p = FlowParams(rho=1.0, mu=0.01, dt=1.0, urf_U=0.9, urf_p=0.1, urf_mdot=1.0,
                   grad_urf=0.5625, bnd_relax=0.75, u_shifted=0, ugrad_shifted=1,
                   p_shifted=1, pgrad_shifted=1, p_ref_node=0, precision=precision,
                   p_method=OAGPU_CG, use_graphs=1, high_res=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def batch_outer_iterations(flows, ctl, iters, tol=0.0, threads=0):
    """Advance independent FlowCases concurrently (oagpu_batch_outer_iterations):
    returns hist [n][iters][6], iterations run [n], wall seconds [n]."""
    n = len(flows)
    hs = (C.c_void_p * max(n, 1))(*[f.h.value if hasattr(f.h, "value") else f.h for f in flows])
    c = np.ascontiguousarray(ctl, dtype=np.float64)
    hist = np.zeros((max(n, 1), max(iters, 1), 6))
    done = np.zeros(max(n, 1), np.int32)
    sec = np.zeros(max(n, 1))
    _check(load_library().oagpu_batch_outer_iterations(
        C.cast(hs, C.POINTER(C.c_void_p)), n, _d(c), iters, tol, threads, _d(hist), _i(done),
        _d(sec)))
    return hist[:n, :iters], done[:n], sec[:n]


def batch_graph_iterations(flows, ctl, iters, tol=0.0, threads=0):
    """batch_outer_iterations with every outer iteration of a case one HIP
    graph launch (oagpu_batch_graph_iterations): hist [n][iters][10]."""
    n = len(flows)
    hs = (C.c_void_p * max(n, 1))(*[f.h.value if hasattr(f.h, "value") else f.h for f in flows])
    c = np.ascontiguousarray(ctl, dtype=np.float64)
    hist = np.zeros((max(n, 1), max(iters, 1), 10))
    done = np.zeros(max(n, 1), np.int32)
    sec = np.zeros(max(n, 1))
    _check(load_library().oagpu_batch_graph_iterations(
        C.cast(hs, C.POINTER(C.c_void_p)), n, _d(c), iters, tol, threads, _d(hist), _i(done),
        _d(sec)))
    return hist[:n, :iters], done[:n], sec[:n]


class DeviceMesh:
    """Device mirror of a mesh (oagpu_mesh_create); mesh["etype"] gives the
    element topologies of a mixed mesh (absent: all hex8).  The device holds
    the mesh renumbered for locality unless caller_order is set; every array
    exchanged through the class is in the caller's numbering either way."""

    def __init__(self, mesh: dict, stencil: int = OAGPU_STENCIL_REDUCED, caller_order: bool = False):
        if caller_order:
            stencil |= OAGPU_MESH_CALLER_ORDER
        L = load_library()
        xyz = np.ascontiguousarray(mesh["xyz"], dtype=np.float64)
        conn = np.ascontiguousarray(mesh["conn"], dtype=np.int32)
        bfe = np.ascontiguousarray(mesh["bf_elem"], dtype=np.int32)
        bfo = np.ascontiguousarray(mesh["bf_ord"], dtype=np.int32)
        bfb = np.ascontiguousarray(mesh["bf_bc"], dtype=np.int32)
        et = mesh.get("etype")
        self._etype = None if et is None else np.ascontiguousarray(et, dtype=np.int32)
        h = C.c_void_p()
        _check(L.oagpu_mesh_create(xyz.shape[0], _d(xyz), conn.shape[0], _i(conn),
                                   None if self._etype is None else _i(self._etype), bfe.size,
                                   _i(bfe), _i(bfo), _i(bfb), stencil, C.byref(h)))
        self.h = h
        self.nbf = bfe.size
        nn, ne, nnz, ninc, nbinc = (C.c_int() for _ in range(5))
        _check(L.oagpu_mesh_sizes(h, C.byref(nn), C.byref(ne), C.byref(nnz), C.byref(ninc),
                                  C.byref(nbinc)))
        self.nn, self.ne, self.nnz, self.ninc, self.nbinc = (v.value for v in (nn, ne, nnz, ninc, nbinc))

    def graph(self):
        rp = np.zeros(self.nn + 1, np.int32)
        cols = np.zeros(self.nnz, np.int32)
        doff = np.zeros(self.nn, np.int32)
        _check(load_library().oagpu_graph_copy(self.h, _i(rp), _i(cols), _i(doff)))
        return rp, cols, doff

    def scatter_map(self):
        ptr = np.zeros(self.nn + 1, np.int32)
        inc = np.zeros(self.ninc, np.int32)
        slots = np.zeros(self.ninc, np.uint64)
        _check(load_library().oagpu_scatter_map_copy(
            self.h, _i(ptr), _i(inc), slots.ctypes.data_as(C.POINTER(C.c_ulonglong))))
        return ptr, inc, slots

    def dual_volume(self):
        v = np.zeros(self.nn)
        _check(load_library().oagpu_mesh_volume_copy(self.h, _d(v)))
        return v

    def close(self):
        if getattr(self, "h", None):
            load_library().oagpu_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class FlowCase:
    """Device state of one segregated-flow case (oagpu_flow)."""

    def __init__(self, mesh: DeviceMesh, params: FlowParams, bc_U):
        self.mesh = mesh
        self.params = params
        bc = np.ascontiguousarray(bc_U, dtype=np.float64).reshape(-1, 3)
        h = C.c_void_p()
        _check(load_library().oagpu_flow_create(mesh.h, C.byref(params), bc.shape[0], _d(bc),
                                                C.byref(h)))
        self.h = h

    _comp = {"Uold": 3, "Uold2": 3, "kold": 1, "kold2": 1, "wold": 1, "wold2": 1,
             "U": 3, "p": 1, "gradU": 9, "gradP": 3, "du": 3, "div": 1, "mu": 1, "k": 1,
             "omega": 1, "gradK": 3, "gradW": 3, "ymin": 1, "mut": 1, "F1": 1, "Pk": 1,
             "Pk_wall": 1, "kprev": 1, "beta": 3, "beta_k": 1, "beta_w": 1, "symA": 3}

    def _size(self, name):
        if name == "mdot":
            return self.mesh.ne * 12
        if name == "msd":
            return self.mesh.nbf * 4
        return self.mesh.nn * self._comp[name]

    def upload(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        _check(load_library().oagpu_field_upload(self.h, name.encode(), _d(a), a.size))

    def download(self, name):
        n = self._size(name)
        a = np.zeros(n)
        _check(load_library().oagpu_field_download(self.h, name.encode(), _d(a), n))
        if name == "mdot":
            return a.reshape(self.mesh.ne, 12)
        if name == "msd":
            return a.reshape(self.mesh.nbf, 4)
        k = self._comp[name]
        return a if k == 1 else a.reshape(self.mesh.nn, k)

    def initialize(self):
        _check(load_library().oagpu_initialize(self.h))

    def assemble_momentum(self):
        _check(load_library().oagpu_assemble_momentum(self.h))

    def assemble_pcorr(self):
        _check(load_library().oagpu_assemble_pcorr(self.h))

    def system(self, N):
        vals = np.zeros(self.mesh.nnz * N * N)
        b = np.zeros(self.mesh.nn * N)
        _check(load_library().oagpu_system_download(self.h, N, _d(vals), _d(b)))
        return vals.reshape(self.mesh.nnz, N, N), b

    def set_system(self, N, vals, b):
        v = np.ascontiguousarray(vals, dtype=np.float64).ravel()
        bb = np.ascontiguousarray(b, dtype=np.float64).ravel()
        _check(load_library().oagpu_system_upload(self.h, N, _d(v), _d(bb)))

    def solution(self, N):
        x = np.zeros(self.mesh.nn * N)
        _check(load_library().oagpu_solution_download(self.h, N, _d(x)))
        return x

    def update_mdot(self):
        _check(load_library().oagpu_update_mdot(self.h))

    def mass_divergence(self):
        _check(load_library().oagpu_mass_divergence(self.h))

    def gradient(self, which):
        _check(load_library().oagpu_gradient(self.h, {"U": 0, "p": 1, "k": 2, "omega": 3}[which]))

    def set_bc_types(self, types):
        """Physical type per boundary id: OAGPU_BC_WALL, OAGPU_BC_INLET,
        OAGPU_BC_OUTLET, OAGPU_BC_SYMMETRY, OAGPU_BC_SLIP or OAGPU_BC_OPENING
        (whose bc_U row is the entrainment direction; zero: the inward
        normal)."""
        t = np.ascontiguousarray(types, dtype=np.int32)
        _check(load_library().oagpu_flow_set_bc_types(self.h, len(t), _i(t)))

    def set_bc_values(self, p=None, k=None, w=None):
        """Per boundary id: outlet / opening static pressure p, inlet k and omega (None:
        unchanged)."""
        arrs = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in (p, k, w)]
        n = next(a.size for a in arrs if a is not None)
        ptr = [None if a is None else _d(a) for a in arrs]
        _check(load_library().oagpu_flow_set_bc_values(self.h, n, *ptr))

    def update_reversal(self):
        _check(load_library().oagpu_update_reversal(self.h))

    def reversal(self):
        """Reversal flag per boundary face ip [n_bfaces][4]."""
        r = np.zeros(self.mesh.nbf * 4, np.int32)
        _check(load_library().oagpu_reversal_download(self.h, _i(r), r.size))
        return r.reshape(self.mesh.nbf, 4)

    def update_blending(self, which):
        _check(load_library().oagpu_update_blending(self.h, {"U": 0, "k": 2, "omega": 3}[which]))

    # k-omega SST
    def sst_enable(self, sst: SstParams):
        self._sst = sst
        _check(load_library().oagpu_sst_enable(self.h, C.byref(sst)))

    def sst_update(self):
        _check(load_library().oagpu_sst_update(self.h))

    def assemble_scalar(self, which):
        _check(load_library().oagpu_assemble_scalar(self.h, {"k": 0, "omega": 1}[which]))

    def correct_scalar(self, which):
        _check(load_library().oagpu_correct_scalar(self.h, {"k": 0, "omega": 1}[which]))

    def sst_iteration(self, ctl):
        c = np.ascontiguousarray(ctl, dtype=np.float64)
        h = np.zeros(4)
        _check(load_library().oagpu_sst_iteration(self.h, _d(c), _d(h)))
        return h

    def correct(self, N, urf):
        _check(load_library().oagpu_correct(self.h, N, urf))

    def velocity_correction(self, step):
        _check(load_library().oagpu_velocity_correction(self.h, step))

    def scales(self):
        u, p = C.c_double(), C.c_double()
        _check(load_library().oagpu_update_scales(self.h, C.byref(u), C.byref(p)))
        return u.value, p.value

    def scaled_residual(self, N):
        out = np.zeros(N)
        _check(load_library().oagpu_scaled_residual(self.h, N, _d(out)))
        return out

    def solve(self, N, method, max_it, rtol, atol=1e-30, restart=30):
        cd = ControlData()
        _check(load_library().oagpu_solve(self.h, N, method, max_it, rtol, atol, restart,
                                          C.byref(cd)))
        return cd

    def amg_levels(self):
        """Rows per level of the AMG hierarchy (empty before the first AMG solve)."""
        n = C.c_int()
        sizes = np.zeros(16, dtype=np.int32)
        _check(load_library().oagpu_amg_info(self.h, C.byref(n), _i(sizes)))
        return [int(v) for v in sizes[:n.value]]

    def amg_level(self, level):
        """(agg of the finer level's rows, row_ptr, cols, vals) of AMG level >= 1."""
        sizes = self.amg_levels()
        nf, nc = sizes[level - 1], sizes[level]
        agg = np.zeros(nf, dtype=np.int32)
        rp = np.zeros(nc + 1, dtype=np.int32)
        _check(load_library().oagpu_amg_level_copy(self.h, level, _i(agg), _i(rp), None, None))
        cols = np.zeros(int(rp[-1]), dtype=np.int32)
        vals = np.zeros(int(rp[-1]))
        _check(load_library().oagpu_amg_level_copy(self.h, level, None, None, _i(cols), _d(vals)))
        return agg, rp, cols, vals

    def sa_levels(self):
        """(rows, prolongation entries, matrix entries, omega_p) per level of the
        smoothed-aggregation hierarchy (empty before the first SA solve)."""
        n = C.c_int()
        sizes = np.zeros(16, dtype=np.int32)
        pn = np.zeros(16, dtype=np.int64)
        an = np.zeros(16, dtype=np.int64)
        om = np.zeros(16)
        _check(load_library().oagpu_sa_info(self.h, C.byref(n), _i(sizes),
                                            pn.ctypes.data_as(C.POINTER(C.c_longlong)),
                                            an.ctypes.data_as(C.POINTER(C.c_longlong)), _d(om)))
        k = n.value
        return [(int(sizes[l]), int(pn[l]), int(an[l]), float(om[l])) for l in range(k)]

    def sa_level(self, level):
        """(agg, P as (row_ptr, cols, vals) over the rows of level - 1, and the
        matrix of level >= 1 as (row_ptr, cols, vals))."""
        lv = self.sa_levels()
        nf, pnnz = lv[level - 1][0], lv[level - 1][1]
        nc, annz = lv[level][0], lv[level][2]
        agg = np.zeros(nf, dtype=np.int32)
        prp = np.zeros(nf + 1, dtype=np.int32)
        pc = np.zeros(pnnz, dtype=np.int32)
        pv = np.zeros(pnnz)
        arp = np.zeros(nc + 1, dtype=np.int32)
        ac = np.zeros(annz, dtype=np.int32)
        av = np.zeros(annz)
        _check(load_library().oagpu_sa_level_copy(self.h, level, _i(agg), _i(prp), _i(pc), _d(pv),
                                                  _i(arp), _i(ac), _d(av)))
        return agg, (prp, pc, pv), (arp, ac, av)

    def time_step(self, scheme, dt):
        """Start a physical time step: scheme OAGPU_BDF1 or OAGPU_BDF2."""
        _check(load_library().oagpu_time_step(self.h, scheme, dt))

    def outer_iteration(self, ctl):
        c = np.ascontiguousarray(ctl, dtype=np.float64)
        h = np.zeros(6)
        _check(load_library().oagpu_outer_iteration(self.h, _d(c), _d(h)))
        return h

    def graph_iterations(self, ctl, sctl=None, n=1, tol=0.0, eager=False):
        """Up to n outer iterations (flow, then SST with sctl), each one HIP
        graph launch (eager: the same device-controlled launches without a
        graph).  Returns hist[iterations run][10]: U, V, W, P residuals,
        momentum / p' iterations, k, omega residuals, k / omega iterations."""
        c = np.ascontiguousarray(ctl, dtype=np.float64)
        sc = None if sctl is None else np.ascontiguousarray(sctl, dtype=np.float64)
        h = np.zeros((max(n, 1), 10))
        done = C.c_int()
        _check(load_library().oagpu_graph_iterations(self.h, _d(c), None if sc is None else _d(sc), n,
                                                     tol, 1 if eager else 0, _d(h), C.byref(done)))
        return h[:done.value]

    def graph_stats(self):
        """{launches, redone, captures, caps} of this flow's graph iterations."""
        a, b, c = C.c_longlong(), C.c_longlong(), C.c_longlong()
        caps = np.zeros(4, dtype=np.int32)
        _check(load_library().oagpu_graph_stats(self.h, C.byref(a), C.byref(b), C.byref(c), _i(caps)))
        return {"launches": a.value, "redone": b.value, "captures": c.value, "caps": caps.tolist()}

    CLOCKS = ("momentum", "pcorr", "scalar", "mdot", "gradU", "gradP", "amg_pre", "amg_post",
              "amg_tail", "spmv1", "spmv3", "spmv1_f32", "spmv3_f32", "gradKW")

    TUNABLES = ("momentum", "pcorr", "spmv1", "spmv3", "spmv1_f32", "spmv3_f32", "scalar", "mdot", "grad")

    def autotune(self):
        """Time the tunable launches' variants on this flow and keep the
        fastest; {kernel: (variant, ms, default ms)}."""
        r = np.zeros(3 * len(self.TUNABLES))
        _check(load_library().oagpu_autotune(self.h, _d(r), r.size))
        return {k: (r[3 * i], r[3 * i + 1], r[3 * i + 2]) for i, k in enumerate(self.TUNABLES)}

    def tune_set(self, which, variant):
        _check(load_library().oagpu_tune_set(self.h, which.encode(), int(variant)))

    def profile(self, which="all"):
        """Clock "all", "none" or one named kernel (HIP events); resets totals."""
        _check(load_library().oagpu_profile(self.h, which.encode()))

    def kernel_stats(self, which):
        """(total device ms, launches) of a clocked kernel since profile()."""
        ms, n = C.c_double(), C.c_longlong()
        _check(load_library().oagpu_kernel_stats(self.h, which.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self):
        if getattr(self, "h", None):
            load_library().oagpu_flow_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
