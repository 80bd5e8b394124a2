"""ctypes binding of the oracle (liboracle.so) and of the reference's own
Fortran master elements built into oracle/_ref/libnalu_me.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libnalu_me.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Params(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("dt", C.c_double),
                ("urf_U", C.c_double), ("urf_p", C.c_double), ("urf_mdot", C.c_double),
                ("grad_urf", C.c_double), ("bnd_relax", C.c_double),
                ("u_shifted", C.c_int), ("ugrad_shifted", C.c_int),
                ("p_shifted", C.c_int), ("pgrad_shifted", C.c_int),
                ("reduced", C.c_int), ("p_ref_node", C.c_int), ("high_res", C.c_int),
                ("p_method", C.c_int)]


class SstParams(C.Structure):
    _fields_ = [("urf_k", C.c_double), ("urf_w", C.c_double), ("urf_mut", C.c_double),
                ("k_shifted", C.c_int), ("w_shifted", C.c_int),
                ("kgrad_shifted", C.c_int), ("wgrad_shifted", C.c_int)]


def default_sst(**kw) -> SstParams:
    s = SstParams(urf_k=0.8, urf_w=0.8, urf_mut=1.0, k_shifted=1, w_shifted=1,
                  kgrad_shifted=1, wgrad_shifted=1)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.or_case_create.restype = C.c_void_p
        L.or_case_create.argtypes = [C.c_int, C.c_int, _dp, _ip, _ip, C.c_int, _ip, _ip, _ip,
                                     C.c_int, _dp, C.POINTER(Params)]
        L.or_case_destroy.argtypes = [C.c_void_p]
        for f in ("or_case_nn", "or_case_nnz"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_int
        for f in ("or_case_row_ptr", "or_case_cols", "or_case_diag_off"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = _ip
        L.or_case_field.argtypes = [C.c_void_p, C.c_char_p]
        L.or_case_field.restype = _dp
        for f in ("or_case_vals", "or_case_rhs", "or_case_x"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_int]
            getattr(L, f).restype = _dp
        for f in ("or_case_initialize", "or_dual_volume", "or_assemble_momentum",
                  "or_compute_du", "or_assemble_pcorr", "or_update_mdot",
                  "or_mass_divergence"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.or_gradient.argtypes = [C.c_void_p, C.c_int]
        L.or_correct_U.argtypes = [C.c_void_p, C.c_double]
        L.or_correct_p.argtypes = [C.c_void_p, C.c_double]
        L.or_velocity_correction.argtypes = [C.c_void_p, C.c_int]
        L.or_update_scales.argtypes = [C.c_void_p, _dp, _dp]
        L.or_solve.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                               C.c_double, C.c_int, _dp]
        L.or_solve.restype = C.c_int
        L.or_scaled_residual.argtypes = [C.c_void_p, C.c_int, _dp]
        L.or_outer_iteration.argtypes = [C.c_void_p, _dp, _dp]
        L.or_graph_build.argtypes = [C.c_int, C.c_int, _ip, _ip, C.c_int, _ip, C.POINTER(_ip)]
        L.or_topo_shape.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.or_graph_build.restype = C.c_int
        L.or_free.argtypes = [C.c_void_p]
        L.or_graph_find.argtypes = [_ip, _ip, C.c_int, C.c_int]
        L.or_graph_find.restype = C.c_int
        L.or_apply_coeff.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp, C.c_int, _ip, _ip,
                                     _ip, _dp, _dp]
        L.or_hex_scs_areav.argtypes = [_dp, _dp]
        L.or_hex_scv_vol.argtypes = [_dp, _dp]
        L.or_hex_shape.argtypes = [C.c_int, _dp, _dp]
        L.or_hex_deriv.argtypes = [C.c_int, _dp, _dp]
        L.or_hex_gradop.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.or_hex_gradop.restype = C.c_int
        L.or_hex_scs_gradop.argtypes = [C.c_int, _dp, _dp, _dp]
        L.or_hex_scs_gradop.restype = C.c_int
        L.or_hex_face_gradop.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.or_hex_face_gradop.restype = C.c_int
        L.or_hex_scs_shape.argtypes = [C.c_int, _dp]
        L.or_quad_face_areav.argtypes = [_dp, _dp]
        L.or_quad_face_shape.argtypes = [C.c_int, _dp]
        L.or_hex_lrscv.restype = _ip
        L.or_topo_npe.argtypes = [C.c_int]
        L.or_topo_nscs.argtypes = [C.c_int]
        L.or_topo_lrscv.argtypes = [C.c_int]
        L.or_topo_lrscv.restype = _ip
        L.or_topo_scs_ips.argtypes = [C.c_int, C.c_int]
        L.or_topo_scs_ips.restype = _dp
        L.or_topo_scs_areav.argtypes = [C.c_int, _dp, _dp]
        L.or_topo_scv_vol.argtypes = [C.c_int, _dp, _dp]
        L.or_topo_deriv.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.or_topo_gradop.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.or_topo_gradop.restype = C.c_int
        L.or_topo_scs_gradop.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.or_topo_scs_gradop.restype = C.c_int
        L.or_sst_enable.argtypes = [C.c_void_p, C.POINTER(SstParams)]
        for f in ("or_sst_update_mut", "or_sst_update_F1", "or_sst_update_Pk"):
            getattr(L, f).argtypes = [C.c_void_p]
        for f in ("or_assemble_scalar", "or_correct_scalar", "or_set_scalar_terms"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_int]
        L.or_sst_iteration.argtypes = [C.c_void_p, _dp, _dp]
        L.or_update_blending.argtypes = [C.c_void_p, C.c_int]
        L.or_normal_relaxation.argtypes = [C.c_void_p]
        L.or_case_set_bc_types.argtypes = [C.c_void_p, _ip]
        L.or_case_set_bc_values.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.or_case_reversal.argtypes = [C.c_void_p]
        L.or_case_reversal.restype = _ip
        L.or_update_reversal.argtypes = [C.c_void_p]
        L.or_hex_side_nodes.restype = _ip
        L.or_set_threads.argtypes = [C.c_int]
        L.or_time_step.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.or_time_step.restype = C.c_int
        _lib = L
    return _lib


def set_threads(n: int):
    """Threads of the oracle's loops (OpenMP colour classes); 1 by default."""
    lib().or_set_threads(int(n))


def default_params(**kw) -> Params:
    p = Params(rho=1.0, mu=0.01, dt=1.0, urf_U=0.9, urf_p=0.1, urf_mdot=1.0,
               grad_urf=0.5625, bnd_relax=0.75, u_shifted=0, ugrad_shifted=1,
               p_shifted=1, pgrad_shifted=1, reduced=1, p_ref_node=0, high_res=0, p_method=1)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Case:
    """Oracle state of one segregated-flow case on a hex8 mesh."""

    def __init__(self, mesh: dict, params: Params, bc_U: np.ndarray):
        L = lib()
        self.mesh = mesh
        self.params = params
        self.nn = mesh["xyz"].shape[0]
        self.ne = mesh["conn"].shape[0]
        bc = np.ascontiguousarray(bc_U, dtype=np.float64)
        self._nbc = bc.shape[0]
        et = mesh.get("etype")
        et = None if et is None else np.ascontiguousarray(et, dtype=np.int32)
        self._keep = [mesh["xyz"], mesh["conn"], mesh["bf_elem"], mesh["bf_ord"],
                      mesh["bf_bc"], bc, et]
        self.h = L.or_case_create(self.nn, self.ne, dptr(mesh["xyz"]), iptr(mesh["conn"]),
                                  None if et is None else iptr(et),
                                  mesh["bf_elem"].size, iptr(mesh["bf_elem"]),
                                  iptr(mesh["bf_ord"]), iptr(mesh["bf_bc"]),
                                  bc.shape[0], dptr(bc), C.byref(params))
        if not self.h:
            raise RuntimeError("or_case_create failed")
        self.nnz = L.or_case_nnz(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_case_destroy(self.h)
            self.h = None

    def _view(self, ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,))

    @property
    def row_ptr(self):
        return self._view(lib().or_case_row_ptr(self.h), self.nn + 1)

    @property
    def cols(self):
        return self._view(lib().or_case_cols(self.h), self.nnz)

    @property
    def diag_off(self):
        return self._view(lib().or_case_diag_off(self.h), self.nn)

    _sizes = {"U": 3, "p": 1, "gradU": 9, "gradP": 3, "du": 3, "div": 1, "vol": 1,
              "rho": 1, "mu": 1, "beta": 3, "symA": 3, "xyz": 3, "k": 1, "omega": 1, "gradK": 3,
              "gradW": 3, "ymin": 1, "mut": 1, "F1": 1, "Pk": 1, "Pk_wall": 1, "kprev": 1,
              "beta_k": 1, "beta_w": 1, "Uold": 3, "Uold2": 3, "kold": 1, "kold2": 1,
              "wold": 1, "wold2": 1}

    def field(self, name):
        ptr = lib().or_case_field(self.h, name.encode())
        if name == "mdot":
            return self._view(ptr, self.ne * 12).reshape(self.ne, 12)
        if name == "msd":
            return self._view(ptr, self.mesh["bf_elem"].size * 4).reshape(-1, 4)
        if name == "bareav":
            return self._view(ptr, self.mesh["bf_elem"].size * 12).reshape(-1, 4, 3)
        k = self._sizes[name]
        v = self._view(ptr, self.nn * k)
        return v if k == 1 else v.reshape(self.nn, k)

    def vals(self, N):
        return self._view(lib().or_case_vals(self.h, N), self.nnz * N * N).reshape(self.nnz, N, N)

    def rhs(self, N):
        return self._view(lib().or_case_rhs(self.h, N), self.nn * N)

    def x(self, N):
        return self._view(lib().or_case_x(self.h, N), self.nn * N)

    def initialize(self):
        lib().or_case_initialize(self.h)

    def assemble_momentum(self):
        lib().or_assemble_momentum(self.h)

    def compute_du(self):
        lib().or_compute_du(self.h)

    def normal_relaxation(self):
        lib().or_normal_relaxation(self.h)

    BC_WALL, BC_INLET, BC_OUTLET, BC_SYMMETRY, BC_SLIP, BC_OPENING = 0, 1, 2, 3, 4, 5

    def set_bc_values(self, p=None, k=None, w=None):
        """Per boundary id: outlet static pressure, inlet k and omega."""
        nb = self._nbc
        arr = [np.ascontiguousarray(np.zeros(nb) if v is None else v, dtype=np.float64)
               for v in (p, k, w)]
        self._bcv = arr
        lib().or_case_set_bc_values(self.h, *[a.ctypes.data_as(_dp) for a in arr])

    def update_reversal(self):
        lib().or_update_reversal(self.h)

    def reversal(self):
        nbf = self.mesh["bf_elem"].size
        ptr = lib().or_case_reversal(self.h)
        return np.ctypeslib.as_array(ptr, shape=(nbf * 4,)).reshape(nbf, 4).copy()

    def set_bc_types(self, types):
        t = np.ascontiguousarray(types, dtype=np.int32)
        if lib().or_case_set_bc_types(self.h, t.ctypes.data_as(_ip)) != 0:
            raise ValueError("unsupported boundary type")

    def assemble_pcorr(self):
        lib().or_assemble_pcorr(self.h)

    def update_mdot(self):
        lib().or_update_mdot(self.h)

    def mass_divergence(self):
        lib().or_mass_divergence(self.h)

    def gradient(self, which):
        lib().or_gradient(self.h, {"U": 0, "p": 1, "k": 2, "omega": 3}[which])

    # k-omega SST
    def sst_enable(self, sst: SstParams):
        self._sst = sst
        lib().or_sst_enable(self.h, C.byref(sst))

    def update_blending(self, which):
        lib().or_update_blending(self.h, {"U": 0, "k": 2, "omega": 3}[which])

    def sst_update_mut(self):
        lib().or_sst_update_mut(self.h)

    def sst_update_F1(self):
        lib().or_sst_update_F1(self.h)

    def sst_update_Pk(self):
        lib().or_sst_update_Pk(self.h)

    def assemble_scalar(self, which, terms=15):
        lib().or_set_scalar_terms(self.h, terms)
        lib().or_assemble_scalar(self.h, {"k": 0, "omega": 1}[which])
        lib().or_set_scalar_terms(self.h, 15)

    def correct_scalar(self, which):
        lib().or_correct_scalar(self.h, {"k": 0, "omega": 1}[which])

    def sst_iteration(self, ctl):
        c = np.ascontiguousarray(ctl, dtype=np.float64)
        h = np.zeros(4)
        lib().or_sst_iteration(self.h, dptr(c), dptr(h))
        return h

    def solve(self, N, kind, max_it, rtol, atol=1e-30, restart=30):
        res = np.zeros(2)
        it = lib().or_solve(self.h, N, kind, max_it, rtol, atol, restart, dptr(res))
        return it, res

    def scaled_residual(self, N):
        out = np.zeros(N)
        lib().or_scaled_residual(self.h, N, dptr(out))
        return out

    def scales(self):
        u = np.zeros(1)
        p = np.zeros(1)
        lib().or_update_scales(self.h, dptr(u), dptr(p))
        return float(u[0]), float(p[0])

    def time_step(self, scheme, dt):
        """Start a physical time step: scheme 1 BDF1, 2 BDF2."""
        if lib().or_time_step(self.h, scheme, dt) != 0:
            raise ValueError("bad time step")

    def outer_iteration(self, ctl):
        c = np.ascontiguousarray(ctl, dtype=np.float64)
        h = np.zeros(6)
        lib().or_outer_iteration(self.h, dptr(c), dptr(h))
        return h


# ---------------------------------------------------------------- reference
_ref = None


def ref_lib():
    """The reference's Fortran master elements (oracle/_ref), or None."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        _ref = C.CDLL(REF_PATH)
    return _ref


def _i(v):
    return C.byref(C.c_int(v))


def ref_hex_scs_areav(coords: np.ndarray) -> np.ndarray:
    """hex_scs_det (MasterElementWork.F:1) on one element, [12][3]."""
    R = ref_lib()
    c = np.ascontiguousarray(coords, dtype=np.float64)
    out = np.zeros((12, 3))
    R.hex_scs_det_(_i(1), _i(8), _i(12), dptr(c), dptr(out))
    return out


def ref_hex_scv_vol(coords):
    R = ref_lib()
    c = np.ascontiguousarray(coords, dtype=np.float64)
    vol = np.zeros(8)
    err = np.zeros(1)
    nerr = C.c_int(0)
    R.hex_scv_det_(_i(1), _i(8), _i(8), dptr(c), dptr(vol), dptr(err), C.byref(nerr))
    return vol


def ref_hex_shape(par):
    R = ref_lib()
    p = np.ascontiguousarray(par, dtype=np.float64)
    n = p.shape[0]
    out = np.zeros((n, 8))
    R.hex_shape_fcn_(_i(n), dptr(p), dptr(out))
    return out


def ref_hex_gradop(par, coords):
    """hex_derivative + hex_gradient_operator at the given points."""
    R = ref_lib()
    p = np.ascontiguousarray(par, dtype=np.float64)
    c = np.ascontiguousarray(coords, dtype=np.float64)
    n = p.shape[0]
    deriv = np.zeros((n, 8, 3))
    R.hex_derivative_(_i(n), dptr(p), dptr(deriv))
    g = np.zeros((n, 8, 3))
    det = np.zeros(n)
    err = np.zeros(1)
    nerr = C.c_int(0)
    R.hex_gradient_operator_(_i(1), _i(8), _i(n), dptr(deriv), dptr(c), dptr(g), dptr(det),
                             dptr(err), C.byref(nerr))
    return g, det


TOPO = {"hex8": 0, "tet4": 1, "wed6": 2, "pyr5": 3}
_FTN = {1: "tet", 2: "wed", 3: "pyr"}


def topo_geometry(topo: int, coords, shifted=0):
    """Oracle SCS areas [nscs][3], SCV volumes [npe], SCS gradient operator
    [nscs][npe][3] and det J [nscs] of one element."""
    L = lib()
    c = np.ascontiguousarray(coords, dtype=np.float64)
    npe, ns = L.or_topo_npe(topo), L.or_topo_nscs(topo)
    a = np.zeros((ns, 3))
    v = np.zeros(npe)
    g = np.zeros((ns, npe, 3))
    d = np.zeros(ns)
    L.or_topo_scs_areav(topo, dptr(c), dptr(a))
    L.or_topo_scv_vol(topo, dptr(c), dptr(v))
    L.or_topo_scs_gradop(topo, shifted, dptr(c), dptr(g), dptr(d))
    return a, v, g, d


def topo_deriv(topo: int, shifted: int, par):
    L = lib()
    p = np.ascontiguousarray(par, dtype=np.float64)
    npe = L.or_topo_npe(topo)
    d = np.zeros((p.shape[0], npe, 3))
    L.or_topo_deriv(topo, shifted, p.shape[0], dptr(p), dptr(d))
    return d


def ref_topo_geometry(topo: int, coords, deriv):
    """The reference Fortran *_scs_det, *_scv_det and *_gradient_operator
    (MasterElementWork.F) on one element; deriv [nint][npe][3] feeds the
    gradient operator."""
    R = ref_lib()
    c = np.ascontiguousarray(coords, dtype=np.float64)
    dv = np.ascontiguousarray(deriv, dtype=np.float64)
    npe = c.shape[0]
    ns = {1: 6, 2: 9, 3: 12}[topo]
    name = _FTN[topo]
    a = np.zeros((ns, 3))
    getattr(R, f"{name}_scs_det_")(_i(1), _i(npe), _i(ns), dptr(c), dptr(a))
    v = np.zeros(npe)
    err = np.zeros(1)
    nerr = C.c_int(0)
    getattr(R, f"{name}_scv_det_")(_i(1), _i(npe), _i(npe), dptr(c), dptr(v), dptr(err),
                                   C.byref(nerr))
    nint = dv.shape[0]
    g = np.zeros((nint, npe, 3))
    det = np.zeros(nint)
    getattr(R, f"{name}_gradient_operator_")(_i(1), _i(npe), _i(nint), dptr(dv), dptr(c),
                                             dptr(g), dptr(det), dptr(err), C.byref(nerr))
    return a, v, g, det


def topo_sides(topo: int):
    """Side-node ordinals of every side of a topology (lists of 3 or 4)."""
    L = lib()
    out = []
    for ord_ in range(L.or_topo_nsides(topo)):
        nodes = (C.c_int * 4)()
        npf = L.or_topo_side_nodes(topo, ord_, nodes)
        out.append([nodes[k] for k in range(npf)])
    return out


def side_ips(topo: int, ord_: int, shifted: int):
    """Exposed-face ip locations [npf][3] of a side in the element's parametric space."""
    L = lib()
    npf = len(topo_sides(topo)[ord_])
    par = np.zeros((npf, 3))
    L.or_topo_side_ips(topo, ord_, shifted, dptr(par))
    return par


def side_gradop(topo: int, shifted: int, ord_: int, coords):
    """Face gradient operator [npf][npe][3] and det J [npf] at a side's face ips."""
    L = lib()
    c = np.ascontiguousarray(coords, dtype=np.float64)
    npf = len(topo_sides(topo)[ord_])
    npe = L.or_topo_npe(topo)
    g = np.zeros((npf, npe, 3))
    d = np.zeros(npf)
    L.or_topo_side_gradop(topo, shifted, ord_, dptr(c), dptr(g), dptr(d))
    return g, d


def tri_face_areav(coords3):
    L = lib()
    c = np.ascontiguousarray(coords3, dtype=np.float64)
    out = np.zeros((3, 3))
    L.or_tri_face_areav(dptr(c), dptr(out))
    return out


def quad_face_areav(coords4):
    L = lib()
    c = np.ascontiguousarray(coords4, dtype=np.float64)
    out = np.zeros((4, 3))
    L.or_quad_face_areav(dptr(c), dptr(out))
    return out


def ref_tri_areav(coords3):
    """The reference Fortran tri3d_scs_det (MasterElementWork.F:1589-1697)."""
    R = ref_lib()
    c = np.ascontiguousarray(coords3, dtype=np.float64)
    out = np.zeros((3, 3))
    R.tri3d_scs_det_(_i(1), _i(3), _i(3), dptr(c), dptr(out))
    return out


def ref_quad_areav(coords4):
    R = ref_lib()
    c = np.ascontiguousarray(coords4, dtype=np.float64)
    out = np.zeros((4, 3))
    R.quad3d_scs_det_(_i(1), dptr(c), dptr(out))
    return out
