// sides.hpp -- boundary sides of any element topology on the device: the
// side table, tri3 / quad4 face shape functions and sub-face area vectors,
// and the geometry of one face ip (area vector and the parent element's
// dN/dx there, face_grad_op / shifted_face_grad_op).  Hex8 sides compute it
// in place; tet4 / wedge6 / pyramid5 sides read it from a record
// precomputed in FP64 once per mesh (k_side_geo).
//   Tri3DSCS shape functions / tri3d_scs_det: MasterElement.C:610-715,
//   MasterElementWork.F:1589-1697; exposed-face ips intgExpFace_ /
//   intgExpFaceShift_: Tet4CVFEM.C:384-403, Wed6CVFEM.C:458-480,
//   Pyr5CVFEM.C:737-757
#pragma once

#include "device_util.hpp"
#include "side_table.hpp"

namespace oagpu {

static __constant__ int c_sides[4][6][4] = OAGPU_SIDE_TABLE;

// tri face shape functions at the standard ips ((7/36, 7/36) from each node:
// 11/18 on the nearest node) and at the nodes (shifted)
static __constant__ double c_tri_shape[2][3][3] = {
    {{11.0 / 18.0, 7.0 / 36.0, 7.0 / 36.0},
     {7.0 / 36.0, 11.0 / 18.0, 7.0 / 36.0},
     {7.0 / 36.0, 7.0 / 36.0, 11.0 / 18.0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};


__device__ inline double face_shape(int npf, int sh, int fk, int q) {
    return npf == 3 ? c_tri_shape[sh][fk][q] : hex8::c_quad_shape[sh][fk][q];
}

// tri3d_scs_det: sub-face k of a tri face spans node k, mid-edge k, the
// centroid and mid-edge k-1; half the cross product of its diagonals
template <typename T>
__device__ inline void tri_face_areav(const T p[4][3], int k, T A[3]) {
    T e[3][3], cc[3], pk[3], e_a[3], e_b[3];
    for (int d = 0; d < 3; ++d) {
        for (int i = 0; i < 3; ++i) e[i][d] = (p[i][d] + p[i == 2 ? 0 : i + 1][d]) * T(0.5);
        cc[d] = (p[0][d] + p[1][d] + p[2][d]) / T(3);
        pk[d] = k == 0 ? p[0][d] : k == 1 ? p[1][d] : p[2][d];
        e_a[d] = k == 0 ? e[2][d] : k == 1 ? e[0][d] : e[1][d];
        e_b[d] = k == 0 ? e[0][d] : k == 1 ? e[1][d] : e[2][d];
    }
    const T dx13 = cc[0] - pk[0], dy13 = cc[1] - pk[1], dz13 = cc[2] - pk[2];
    const T dx24 = e_a[0] - e_b[0], dy24 = e_a[1] - e_b[1], dz24 = e_a[2] - e_b[2];
    A[0] = T(0.5) * (dz24 * dy13 - dz13 * dy24);
    A[1] = T(0.5) * (dx24 * dz13 - dx13 * dz24);
    A[2] = T(0.5) * (dy24 * dx13 - dy13 * dx24);
}

template <typename T>
__device__ inline void face_areav(int npf, const T p[4][3], int k, T A[3]) {
    if (npf == 3) tri_face_areav(p, k, A);
    else hex8::quad_face_areav(p, k, A);
}

// the sides' element topologies and the precomputed records of the
// tet4 / wedge6 / pyramid5 sides (etype null: every element hex8)
struct SideGeo {
    const int* etype;
    const long long* off;
    const double* geo;
};

inline SideGeo side_geo(const DevMesh& m) {
    return {m.mixed ? m.etype.p : nullptr, m.sgeo_off.p, m.sgeo.p};
}

// geometry of one face ip of a boundary side
template <typename T>
struct FaceIp {
    int npe, npf, gen;
    int sn[4];        // side nodes as element-local ordinals (sn[3] = -1 on tri sides)
    T s[3], Ji[9];    // hex8: the ip's parametric location and inverse Jacobian
    T A[3];           // sub-face area vector (outward)
    const double* g;  // other topologies: dN/dx [npe][3] at the ip
    __device__ void grad(int ic, T out[3]) const {
        if (gen) {
            for (int d = 0; d < 3; ++d) out[d] = T(g[3 * ic + d]);
        } else {
            grad_N(ic, s[0], s[1], s[2], Ji, out);
        }
    }
    __device__ bool on_side(int ic) const {
        return ic == sn[0] || ic == sn[1] || ic == sn[2] || ic == sn[3];
    }
    __device__ double shape(int sh, int fk, int q) const { return face_shape(npf, sh, fk, q); }
};

// face ip fk of side f (element e with nodes ce, side ordinal ord) at the
// standard or shifted location; hex8 coordinates relative to xr.  MX false:
// every element is hex8.
template <bool MX, typename T>
__device__ inline void face_ip(const SideGeo& sg, const double* xyz, const int* ce, int e, int f,
                               int ord, int fk, int shifted, const double xr[3], FaceIp<T>& fi,
                               int& bad) {
    const int t = (MX && sg.etype) ? sg.etype[e] : OAGPU_HEX8;
    for (int k = 0; k < 4; ++k) fi.sn[k] = c_sides[t][ord][k];
    fi.npf = fi.sn[3] < 0 ? 3 : 4;
    if (!MX || t == OAGPU_HEX8) {
        fi.npe = 8;
        fi.gen = 0;
        fi.g = nullptr;
        face_ip_geom(xyz, ce, ord, fk, shifted, xr, fi.s, fi.Ji, fi.A, bad);
        return;
    }
    fi.npe = t == OAGPU_TET4 ? 4 : t == OAGPU_WED6 ? 6 : 5;
    fi.gen = 1;
    const double* q = sg.geo + sg.off[f] + size_t(fk) * kSideIpStride;
    for (int d = 0; d < 3; ++d) fi.A[d] = T(q[d]);
    fi.g = q + 3 + (shifted ? 18 : 0);
}

// Rhie-Chow mass flow rate at an outlet / opening face ip (flowModel.cpp:
// 7323-7690 / 8094-8380, SIMPLE: duRhs = du): rho U_ip . A - rho du_ip .
// (grad p - Gp) A with U and du interpolated by the face shape functions
// (us: shifted), grad p by the face ip's dN/dx, Gp the side nodes' mean
// nodal pressure gradient (weights 1/npf).  side_p: the specified pressure
// pbc on the side's nodes (outlets, and the p' rhs of openings), pressures
// then entering relative to pbc (dN/dx sums to zero over the nodes); else
// the nodal pressure on every node (the opening's rate update).
template <typename T>
__device__ inline double outlet_rhie_chow(const FaceIp<T>& fi, const int* ce, int fk, int us,
                                          const double* U, const double* du, const double* p,
                                          const double* gradP, double rho, double pbc,
                                          bool side_p = true) {
    const double fw = 1.0 / fi.npf;
    double rip = 0, u[3] = {0, 0, 0}, d[3] = {0, 0, 0}, Gp[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
    for (int q = 0; q < fi.npf; ++q) {
        const double r = fi.shape(us, fk, q);
        const size_t o = 3 * size_t(ce[fi.sn[q]]);
        rip += r * rho;
        for (int j = 0; j < 3; ++j) {
            u[j] += r * U[o + j];
            d[j] += r * du[o + j];
            Gp[j] += fw * gradP[o + j];
        }
    }
    const double p0 = side_p ? pbc : 0.0;
#pragma unroll 1
    for (int ic = 0; ic < fi.npe; ++ic) {
        if (side_p && fi.on_side(ic)) continue;
        T g[3];
        fi.grad(ic, g);
        const double pv = p[ce[ic]] - p0;
        for (int j = 0; j < 3; ++j) dp[j] += double(g[j]) * pv;
    }
    double m = 0;
    for (int j = 0; j < 3; ++j) m += (rip * u[j] - rip * d[j] * (dp[j] - Gp[j])) * double(fi.A[j]);
    return m;
}

}  // namespace oagpu
