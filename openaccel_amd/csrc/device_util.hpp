// device_util.hpp -- small device helpers shared by the CVFEM kernels.
#pragma once

#include <cfloat>

#include "flow.hpp"
#include "hex8.hpp"

namespace oagpu {



inline constexpr double kSmall = DBL_EPSILON;  // src/basic/types.h:179

// node location signs: N_c(s) = prod_k (1/2 + sigma_ck s_k)
__device__ inline int node_sigma(int c, int k) {
    // hex8 node order (Hex8CVFEM.C:428-432)
    const int bx = ((c + 1) >> 1) & 1;  // nodes 1,2,5,6 at +x
    const int by = (c >> 1) & 1;        // nodes 2,3,6,7 at +y
    const int bz = (c >> 2) & 1;        // nodes 4..7 at +z
    const int b = k == 0 ? bx : (k == 1 ? by : bz);
    return 2 * b - 1;
}

template <typename T>
__device__ inline void dN_ds(int c, T s1, T s2, T s3, T d[3]) {
    const T g1 = T(node_sigma(c, 0)), g2 = T(node_sigma(c, 1)), g3 = T(node_sigma(c, 2));
    const T a = T(0.5) + g1 * s1, b = T(0.5) + g2 * s2, cc = T(0.5) + g3 * s3;
    d[0] = g1 * b * cc;
    d[1] = g2 * a * cc;
    d[2] = g3 * a * b;
}

template <typename T>
__device__ inline T N_at(int c, T s1, T s2, T s3) {
    return (T(0.5) + T(node_sigma(c, 0)) * s1) * (T(0.5) + T(node_sigma(c, 1)) * s2) *
           (T(0.5) + T(node_sigma(c, 2)) * s3);
}

// Inverse Jacobian at isoparametric point s: Ji[k*3+d] = ds_k/dx_d
// (hex_gradient_operator, MasterElementWork.F:2666-2745)
template <typename T>
__device__ inline T jac_inv(T s1, T s2, T s3, const T c[8][3], T Ji[9], int* bad) {
    T dx1 = 0, dx2 = 0, dx3 = 0, dy1 = 0, dy2 = 0, dy3 = 0, dz1 = 0, dz2 = 0, dz3 = 0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        T d[3];
        dN_ds(n, s1, s2, s3, d);
        dx1 += d[0] * c[n][0];
        dx2 += d[1] * c[n][0];
        dx3 += d[2] * c[n][0];
        dy1 += d[0] * c[n][1];
        dy2 += d[1] * c[n][1];
        dy3 += d[2] * c[n][1];
        dz1 += d[0] * c[n][2];
        dz2 += d[1] * c[n][2];
        dz3 += d[2] * c[n][2];
    }
    const T det = dx1 * (dy2 * dz3 - dz2 * dy3) + dy1 * (dz2 * dx3 - dx2 * dz3) +
                  dz1 * (dx2 * dy3 - dy2 * dx3);
    T test = det;
    if (!(test > T(0))) {
        test = T(1);
        *bad = 1;
    }
    const T den = T(1) / test;
    Ji[0] = den * (dy2 * dz3 - dz2 * dy3);
    Ji[3] = den * (dz1 * dy3 - dy1 * dz3);
    Ji[6] = den * (dy1 * dz2 - dz1 * dy2);
    Ji[1] = den * (dz2 * dx3 - dx2 * dz3);
    Ji[4] = den * (dx1 * dz3 - dz1 * dx3);
    Ji[7] = den * (dz1 * dx2 - dx1 * dz2);
    Ji[2] = den * (dx2 * dy3 - dy2 * dx3);
    Ji[5] = den * (dy1 * dx3 - dx1 * dy3);
    Ji[8] = den * (dx1 * dy2 - dy1 * dx2);
    return det;
}

// dN_c/dx at point s given the inverse Jacobian there
template <typename T>
__device__ inline void grad_N(int c, T s1, T s2, T s3, const T Ji[9], T g[3]) {
    T d[3];
    dN_ds(c, s1, s2, s3, d);
    g[0] = d[0] * Ji[0] + d[1] * Ji[3] + d[2] * Ji[6];
    g[1] = d[0] * Ji[1] + d[1] * Ji[4] + d[2] * Ji[7];
    g[2] = d[0] * Ji[2] + d[1] * Ji[5] + d[2] * Ji[8];
}

__device__ inline void load_conn(const int* conn, int e, int nd[8]) {
    const int4* p = reinterpret_cast<const int4*>(conn + 8 * size_t(e));
    const int4 a = p[0], b = p[1];
    nd[0] = a.x; nd[1] = a.y; nd[2] = a.z; nd[3] = a.w;
    nd[4] = b.x; nd[5] = b.y; nd[6] = b.z; nd[7] = b.w;
}

// Sliced-ELL row handle of the owning lane: the slice's first slot and the
// lane's position in the slice (rows of one wavefront share a slice).
struct SellRow {
    size_t base;  // (slice_ptr[row / 64] * NN) * 64 + row % 64
    int width;    // slots of the slice
};

template <int NN>
__device__ __forceinline__ SellRow sell_row(const int* slice_ptr, const int* slice_w, int row) {
    const int sl = row >> 6;
    return {size_t(slice_ptr[sl]) * NN * 64 + (row & 63), slice_w[sl]};
}

// value k of the NN-block in slot s: base + (s * NN + k) * 64
template <int NN>
__device__ __forceinline__ double* sell_val(double* vals, const SellRow& r, int s, int k) {
    return vals + r.base + size_t(s * NN + k) * 64;
}

__device__ inline int slot_of(uint64_t sl, int c) { return int((sl >> (8 * c)) & 0xFFu); }

// Item (row or element) of this lane when 64-item groups are walked in the
// order `order` (ngroups entries): a block of 64 S lanes takes S consecutive
// groups, lane l of a group its item 64 g + l.  -1 beyond the last item.
__device__ inline int ordered_item(const int* order, int ngroups, int nitems) {
    const int per = int(blockDim.x >> 6);
    const int g = int(xcd_block()) * per + int(threadIdx.x >> 6);
    if (g >= ngroups) return -1;
    const int i = order[g] * 64 + int(threadIdx.x & 63u);
    return i < nitems ? i : -1;
}

// Face ip fk of side `ord` of an element (node ids ce[0..7]): the ip
// location s (the exposed-face ip, or the side node itself when shifted),
// the inverse Jacobian there and the sub-face area vector (outward,
// quad3d_scs_det); coordinates relative to xr.  Node ids at run-time local
// indices are read from ce (memory), never from a register array.
template <typename T>
__device__ inline void face_ip_geom(const double* xyz, const int* ce, int ord, int fk,
                                    int shifted, const double xr[3], T s[3], T Ji[9], T A[3],
                                    int& bad) {
    const int* sn = hex8::c_side_nodes[ord];
    for (int d = 0; d < 3; ++d)
        s[d] = shifted ? T(hex8::c_node_loc[sn[fk]][d]) : T(hex8::c_face_loc[ord][fk][d]);
    {
        T c[8][3];
#pragma unroll
        for (int n = 0; n < 8; ++n)
            for (int d = 0; d < 3; ++d) c[n][d] = T(xyz[3 * size_t(ce[n]) + d] - xr[d]);
        jac_inv(s[0], s[1], s[2], c, Ji, &bad);
    }
    T fp[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const size_t o = 3 * size_t(ce[sn[q]]);
        for (int d = 0; d < 3; ++d) fp[q][d] = T(xyz[o + d] - xr[d]);
    }
    hex8::quad_face_areav(fp, fk, A);
}

}  // namespace oagpu
