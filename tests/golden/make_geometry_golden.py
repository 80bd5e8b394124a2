"""Write tests/golden/hex8_geometry.npz from the reference's own Fortran
master elements (oracle/_ref/libnalu_me.so, built by `make -C oracle ref`
from external/nalu/src/master_element/*.F where they lie in the reference).

Inputs: 24 hexahedra -- the reference element, affinely mapped ones and
randomly perturbed ones (seeded) -- plus 6 quad faces.  Outputs: SCS area
vectors (hex_scs_det), SCV volumes (hex_scv_det), shape functions at the
standard and shifted SCS ips (hex_shape_fcn), gradient operators and
Jacobian determinants at both ip sets (hex_derivative +
hex_gradient_operator), exposed-face gradient operators, and quad face area
vectors (quad3d_scs_det).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

SCS = np.array([[0.00, -0.25, -0.25], [0.25, 0.00, -0.25], [0.00, 0.25, -0.25], [-0.25, 0.00, -0.25],
                [0.00, -0.25, 0.25], [0.25, 0.00, 0.25], [0.00, 0.25, 0.25], [-0.25, 0.00, 0.25],
                [-0.25, -0.25, 0.00], [0.25, -0.25, 0.00], [0.25, 0.25, 0.00], [-0.25, 0.25, 0.00]])
BASE = np.array([[-.5, -.5, -.5], [.5, -.5, -.5], [.5, .5, -.5], [-.5, .5, -.5],
                 [-.5, -.5, .5], [.5, -.5, .5], [.5, .5, .5], [-.5, .5, .5]])
SIDES = np.array([[0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [0, 4, 7, 3], [0, 3, 2, 1], [4, 5, 6, 7]])
FACE_LOC = np.array([
    [[-0.25, -0.50, -0.25], [0.25, -0.50, -0.25], [0.25, -0.50, 0.25], [-0.25, -0.50, 0.25]],
    [[0.50, -0.25, -0.25], [0.50, 0.25, -0.25], [0.50, 0.25, 0.25], [0.50, -0.25, 0.25]],
    [[0.25, 0.50, -0.25], [-0.25, 0.50, -0.25], [-0.25, 0.50, 0.25], [0.25, 0.50, 0.25]],
    [[-0.50, -0.25, -0.25], [-0.50, -0.25, 0.25], [-0.50, 0.25, 0.25], [-0.50, 0.25, -0.25]],
    [[-0.25, -0.25, -0.50], [-0.25, 0.25, -0.50], [0.25, 0.25, -0.50], [0.25, -0.25, -0.50]],
    [[-0.25, -0.25, 0.50], [0.25, -0.25, 0.50], [0.25, 0.25, 0.50], [-0.25, 0.25, 0.50]]])


def elements(seed=20261203):
    rng = np.random.default_rng(seed)
    out = [BASE.copy(), BASE * 2.0 + 1.0]
    for _ in range(6):  # affine maps with positive determinant
        M = np.eye(3) + 0.3 * rng.normal(size=(3, 3))
        if np.linalg.det(M) < 0.2:
            M = np.eye(3)
        out.append(BASE @ M.T + rng.normal(size=3))
    for amp in (0.05, 0.1, 0.15, 0.2):  # perturbed vertices
        for _ in range(4):
            out.append(BASE + rng.uniform(-amp, amp, size=BASE.shape))
    return np.array(out)


def main():
    if po.ref_lib() is None:
        raise SystemExit("oracle/_ref/libnalu_me.so missing: run make -C oracle ref")
    E = elements()
    n = len(E)
    shifted = SCS * 2.0
    areav = np.zeros((n, 12, 3))
    scv = np.zeros((n, 8))
    g_std = np.zeros((n, 12, 8, 3))
    d_std = np.zeros((n, 12))
    g_sh = np.zeros((n, 12, 8, 3))
    d_sh = np.zeros((n, 12))
    g_face = np.zeros((n, 6, 4, 8, 3))
    g_face_sh = np.zeros((n, 6, 4, 8, 3))
    qarea = np.zeros((n, 6, 4, 3))
    for k, co in enumerate(E):
        areav[k] = po.ref_hex_scs_areav(co)
        scv[k] = po.ref_hex_scv_vol(co)
        g_std[k], d_std[k] = po.ref_hex_gradop(SCS, co)
        g_sh[k], d_sh[k] = po.ref_hex_gradop(shifted, co)
        for f in range(6):
            g_face[k, f], _ = po.ref_hex_gradop(FACE_LOC[f], co)
            g_face_sh[k, f], _ = po.ref_hex_gradop(BASE[SIDES[f]], co)
            qarea[k, f] = po.ref_quad_areav(co[SIDES[f]])
    np.savez_compressed(
        os.path.join(os.path.dirname(os.path.abspath(__file__)), "hex8_geometry.npz"),
        coords=E, scs_areav=areav, scv_vol=scv, gradop_std=g_std, detj_std=d_std,
        gradop_shift=g_sh, detj_shift=d_sh, face_gradop=g_face, face_gradop_shift=g_face_sh,
        quad_areav=qarea, shape_std=po.ref_hex_shape(SCS), shape_shift=po.ref_hex_shape(shifted))
    print("wrote hex8_geometry.npz with", n, "elements")


if __name__ == "__main__":
    main()
