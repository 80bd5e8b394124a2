// hex8.hpp -- device CVFEM master element for 8-node hexahedra.
//
// Restates, for gfx950 and templated on the local arithmetic type T (float in
// the mixed form, double in the FP64 form), the nalu master element the
// reference evaluates per element pass:
//   SCS area vectors   MasterElementWork.F:1-231 (hex_scs_det) with
//                      CalculateQuadArea.F (4 triangles about the facet centre)
//   SCV volumes        MasterElementWork.F:233-457 (hex_scv_det) with the
//                      24-facet Grandy volume, CalculateHexVolume.F /
//                      CalculatePolyhedralVolume.F
//   shape functions    MasterElementWork.F:2344 (hex_shape_fcn)
//   derivatives        MasterElementWork.F:2404 (hex_derivative)
//   gradient operator  MasterElementWork.F:2623 (hex_gradient_operator)
//   ip tables          Hex8CVFEM.C:264-276 (lrscv), :318-344 (ip locations),
//                      :346-377 / :434-446 (exposed-face ips), :419-426 (sides)
//   quad face areas    MasterElementWork.F:1467 (quad3d_scs_det)
// Isoparametric coordinates are in [-1/2, 1/2] as in the reference.
#pragma once

#include <hip/hip_runtime.h>

namespace oagpu {
namespace hex8 {

// SCS ip isoparametric locations: standard then edge-midpoint ("shifted")
static __constant__ double c_scs_loc[2][12][3] = {
    {{0.00, -0.25, -0.25}, {0.25, 0.00, -0.25}, {0.00, 0.25, -0.25}, {-0.25, 0.00, -0.25},
     {0.00, -0.25, 0.25},  {0.25, 0.00, 0.25},  {0.00, 0.25, 0.25},  {-0.25, 0.00, 0.25},
     {-0.25, -0.25, 0.00}, {0.25, -0.25, 0.00}, {0.25, 0.25, 0.00},  {-0.25, 0.25, 0.00}},
    {{0.00, -0.50, -0.50}, {0.50, 0.00, -0.50}, {0.00, 0.50, -0.50}, {-0.50, 0.00, -0.50},
     {0.00, -0.50, 0.50},  {0.50, 0.00, 0.50},  {0.00, 0.50, 0.50},  {-0.50, 0.00, 0.50},
     {-0.50, -0.50, 0.00}, {0.50, -0.50, 0.00}, {0.50, 0.50, 0.00},  {-0.50, 0.50, 0.00}}};

static __constant__ int c_lrscv[12][2] = {{0, 1}, {1, 2}, {2, 3}, {0, 3}, {4, 5}, {5, 6},
                                         {6, 7}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

// ips touching each node: ip index, +1 when the node is the left node
static __constant__ int c_node_ip[8][3] = {{0, 3, 8},  {0, 1, 9},  {1, 2, 10}, {2, 3, 11},
                                          {4, 7, 8},  {4, 5, 9},  {5, 6, 10}, {6, 7, 11}};
static __constant__ int c_node_ip_sign[8][3] = {{1, 1, 1},   {-1, 1, 1},   {-1, 1, 1},
                                               {-1, -1, 1},  {1, 1, -1},   {-1, 1, -1},
                                               {-1, 1, -1},  {-1, -1, -1}};

// facet of each SCS in the 27-point subdivision (MasterElementWork.F:41-53)
static __constant__ int c_edge_facet[12][4] = {
    {20, 8, 12, 26},  {24, 9, 12, 26},  {10, 12, 26, 23}, {11, 25, 26, 12},
    {13, 20, 26, 17}, {17, 14, 24, 26}, {17, 15, 23, 26}, {16, 17, 26, 25},
    {19, 20, 26, 25}, {20, 18, 24, 26}, {22, 23, 26, 24}, {21, 25, 26, 23}};

// subdivision points 8..25 as node averages (MasterElementWork.F:72-184)
static __constant__ int c_sub_nodes[18][4] = {
    {0, 1, -1, -1}, {1, 2, -1, -1}, {2, 3, -1, -1}, {3, 0, -1, -1}, {0, 1, 2, 3},
    {4, 5, -1, -1}, {5, 6, -1, -1}, {6, 7, -1, -1}, {7, 4, -1, -1}, {4, 5, 6, 7},
    {1, 5, -1, -1}, {0, 4, -1, -1}, {0, 1, 4, 5},   {3, 7, -1, -1}, {2, 6, -1, -1},
    {2, 3, 6, 7},   {1, 2, 5, 6},   {0, 3, 4, 7}};

// side ordinal -> element nodes == face ip -> nearest element node
static __constant__ int c_side_nodes[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6},
                                             {0, 4, 7, 3}, {0, 3, 2, 1}, {4, 5, 6, 7}};

// exposed-face ip locations (standard); shifted ones are the side nodes'
static __constant__ double c_face_loc[6][4][3] = {
    {{-0.25, -0.50, -0.25}, {0.25, -0.50, -0.25}, {0.25, -0.50, 0.25}, {-0.25, -0.50, 0.25}},
    {{0.50, -0.25, -0.25}, {0.50, 0.25, -0.25}, {0.50, 0.25, 0.25}, {0.50, -0.25, 0.25}},
    {{0.25, 0.50, -0.25}, {-0.25, 0.50, -0.25}, {-0.25, 0.50, 0.25}, {0.25, 0.50, 0.25}},
    {{-0.50, -0.25, -0.25}, {-0.50, -0.25, 0.25}, {-0.50, 0.25, 0.25}, {-0.50, 0.25, -0.25}},
    {{-0.25, -0.25, -0.50}, {-0.25, 0.25, -0.50}, {0.25, 0.25, -0.50}, {0.25, -0.25, -0.50}},
    {{-0.25, -0.25, 0.50}, {0.25, -0.25, 0.50}, {0.25, 0.25, 0.50}, {-0.25, 0.25, 0.50}}};

static __constant__ double c_node_loc[8][3] = {
    {-0.5, -0.5, -0.5}, {0.5, -0.5, -0.5}, {0.5, 0.5, -0.5}, {-0.5, 0.5, -0.5},
    {-0.5, -0.5, 0.5},  {0.5, -0.5, 0.5},  {0.5, 0.5, 0.5},  {-0.5, 0.5, 0.5}};

// quad face shape functions at standard / shifted ips (quad3d_shape_fcn)
static __constant__ double c_quad_shape[2][4][4] = {
    {{0.5625, 0.1875, 0.0625, 0.1875},
     {0.1875, 0.5625, 0.1875, 0.0625},
     {0.0625, 0.1875, 0.5625, 0.1875},
     {0.1875, 0.0625, 0.1875, 0.5625}},
    {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}}};

template <typename T>
__device__ inline void shape(T s1, T s2, T s3, T N[8]) {
    const T half = T(0.5), q = T(0.25), e = T(0.125);
    const T s123 = s1 * s2 * s3;
    N[0] = e + q * (-s1 - s2 - s3) + half * (s2 * s3 + s3 * s1 + s1 * s2) - s123;
    N[1] = e + q * (s1 - s2 - s3) + half * (s2 * s3 - s3 * s1 - s1 * s2) + s123;
    N[2] = e + q * (s1 + s2 - s3) + half * (-s2 * s3 - s3 * s1 + s1 * s2) - s123;
    N[3] = e + q * (-s1 + s2 - s3) + half * (-s2 * s3 + s3 * s1 - s1 * s2) + s123;
    N[4] = e + q * (-s1 - s2 + s3) + half * (-s2 * s3 - s3 * s1 + s1 * s2) + s123;
    N[5] = e + q * (s1 - s2 + s3) + half * (-s2 * s3 + s3 * s1 - s1 * s2) - s123;
    N[6] = e + q * (s1 + s2 + s3) + half * (s2 * s3 + s3 * s1 + s1 * s2) + s123;
    N[7] = e + q * (-s1 + s2 + s3) + half * (s2 * s3 - s3 * s1 - s1 * s2) - s123;
}

template <typename T>
__device__ inline void deriv(T s1, T s2, T s3, T d[8][3]) {
    const T half = T(0.5), q = T(0.25);
    const T s1s2 = s1 * s2, s2s3 = s2 * s3, s1s3 = s1 * s3;
    d[0][0] = half * (s3 + s2) - s2s3 - q;
    d[1][0] = half * (-s3 - s2) + s2s3 + q;
    d[2][0] = half * (-s3 + s2) - s2s3 + q;
    d[3][0] = half * (s3 - s2) + s2s3 - q;
    d[4][0] = half * (-s3 + s2) + s2s3 - q;
    d[5][0] = half * (s3 - s2) - s2s3 + q;
    d[6][0] = half * (s3 + s2) + s2s3 + q;
    d[7][0] = half * (-s3 - s2) - s2s3 - q;
    d[0][1] = half * (s3 + s1) - s1s3 - q;
    d[1][1] = half * (s3 - s1) + s1s3 - q;
    d[2][1] = half * (-s3 + s1) - s1s3 + q;
    d[3][1] = half * (-s3 - s1) + s1s3 + q;
    d[4][1] = half * (-s3 + s1) + s1s3 - q;
    d[5][1] = half * (-s3 - s1) - s1s3 - q;
    d[6][1] = half * (s3 + s1) + s1s3 + q;
    d[7][1] = half * (s3 - s1) - s1s3 + q;
    d[0][2] = half * (s2 + s1) - s1s2 - q;
    d[1][2] = half * (s2 - s1) + s1s2 - q;
    d[2][2] = half * (-s2 - s1) - s1s2 - q;
    d[3][2] = half * (-s2 + s1) + s1s2 - q;
    d[4][2] = half * (-s2 - s1) + s1s2 + q;
    d[5][2] = half * (-s2 + s1) - s1s2 + q;
    d[6][2] = half * (s2 + s1) + s1s2 + q;
    d[7][2] = half * (s2 - s1) - s1s2 + q;
}

// hex_gradient_operator at one point: g = dN/dx, returns det J.  A
// non-positive determinant is replaced by 1 as in the reference, and *bad set.
template <typename T>
__device__ inline T gradop(const T d[8][3], const T c[8][3], T g[8][3], int* bad) {
    T dx1 = 0, dx2 = 0, dx3 = 0, dy1 = 0, dy2 = 0, dy3 = 0, dz1 = 0, dz2 = 0, dz3 = 0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        dx1 += d[n][0] * c[n][0];
        dx2 += d[n][1] * c[n][0];
        dx3 += d[n][2] * c[n][0];
        dy1 += d[n][0] * c[n][1];
        dy2 += d[n][1] * c[n][1];
        dy3 += d[n][2] * c[n][1];
        dz1 += d[n][0] * c[n][2];
        dz2 += d[n][1] * c[n][2];
        dz3 += d[n][2] * c[n][2];
    }
    const T det = dx1 * (dy2 * dz3 - dz2 * dy3) + dy1 * (dz2 * dx3 - dx2 * dz3) +
                  dz1 * (dx2 * dy3 - dy2 * dx3);
    T test = det;
    if (!(test > T(0))) {
        test = T(1);
        *bad = 1;
    }
    const T den = T(1) / test;
    const T s1x = den * (dy2 * dz3 - dz2 * dy3), s2x = den * (dz1 * dy3 - dy1 * dz3),
            s3x = den * (dy1 * dz2 - dz1 * dy2);
    const T s1y = den * (dz2 * dx3 - dx2 * dz3), s2y = den * (dx1 * dz3 - dz1 * dx3),
            s3y = den * (dz1 * dx2 - dx1 * dz2);
    const T s1z = den * (dx2 * dy3 - dy2 * dx3), s2z = den * (dy1 * dx3 - dx1 * dy3),
            s3z = den * (dx1 * dy2 - dy1 * dx2);
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        g[n][0] = d[n][0] * s1x + d[n][1] * s2x + d[n][2] * s3x;
        g[n][1] = d[n][0] * s1y + d[n][1] * s2y + d[n][2] * s3y;
        g[n][2] = d[n][0] * s1z + d[n][1] * s2z + d[n][2] * s3z;
    }
    return det;
}

// subdivision point idx (0..26) of the element; ctr is point 26
template <typename T>
__device__ inline void sub_point(int idx, const T c[8][3], const T ctr[3], T p[3]) {
    if (idx < 8) {
        p[0] = c[idx][0];
        p[1] = c[idx][1];
        p[2] = c[idx][2];
    } else if (idx == 26) {
        p[0] = ctr[0];
        p[1] = ctr[1];
        p[2] = ctr[2];
    } else {
        const int* s = c_sub_nodes[idx - 8];
        if (s[2] < 0) {
            for (int k = 0; k < 3; ++k) p[k] = T(0.5) * (c[s[0]][k] + c[s[1]][k]);
        } else {
            for (int k = 0; k < 3; ++k)
                p[k] = T(0.25) * (c[s[0]][k] + c[s[1]][k] + c[s[2]][k] + c[s[3]][k]);
        }
    }
}

template <typename T>
__device__ inline void centroid(const T c[8][3], T ctr[3]) {
    for (int k = 0; k < 3; ++k)
        ctr[k] = T(0.125) * (c[0][k] + c[1][k] + c[2][k] + c[3][k] + c[4][k] + c[5][k] +
                             c[6][k] + c[7][k]);
}

// quadAreaByTriangleFacets on 4 points
template <typename T>
__device__ inline void quad_area(const T q[4][3], T A[3]) {
    T xm[3], r1[3], r2[3];
    for (int k = 0; k < 3; ++k) {
        xm[k] = T(0.25) * (q[0][k] + q[1][k] + q[2][k] + q[3][k]);
        A[k] = 0;
        r2[k] = q[0][k] - xm[k];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int iq = (t + 1) & 3;
        for (int k = 0; k < 3; ++k) {
            r1[k] = r2[k];
            r2[k] = q[iq][k] - xm[k];
        }
        A[0] += r1[1] * r2[2] - r2[1] * r1[2];
        A[1] += r1[2] * r2[0] - r2[2] * r1[0];
        A[2] += r1[0] * r2[1] - r2[0] * r1[1];
    }
    for (int k = 0; k < 3; ++k) A[k] *= T(0.5);
}

// area vector of SCS `ip`, oriented from its left to its right node
template <typename T>
__device__ inline void scs_areav(int ip, const T c[8][3], const T ctr[3], T A[3]) {
    T q[4][3];
    for (int n = 0; n < 4; ++n) sub_point(c_edge_facet[ip][n], c, ctr, q[n]);
    quad_area(q, A);
}

// dN/dx at SCS ip (standard or shifted location)
template <typename T>
__device__ inline T scs_gradop(int ip, int shifted, const T c[8][3], T g[8][3], int* bad) {
    T d[8][3];
    deriv(T(c_scs_loc[shifted][ip][0]), T(c_scs_loc[shifted][ip][1]),
          T(c_scs_loc[shifted][ip][2]), d);
    return gradop(d, c, g, bad);
}

template <typename T>
__device__ inline void scs_shape(int ip, int shifted, T N[8]) {
    shape(T(c_scs_loc[shifted][ip][0]), T(c_scs_loc[shifted][ip][1]),
          T(c_scs_loc[shifted][ip][2]), N);
}

// dN/dx at face ip k of side `ord` of the parent element
template <typename T>
__device__ inline T face_gradop(int ord, int k, int shifted, const T c[8][3], T g[8][3],
                                int* bad) {
    T d[8][3];
    if (shifted) {
        const int nd = c_side_nodes[ord][k];
        deriv(T(c_node_loc[nd][0]), T(c_node_loc[nd][1]), T(c_node_loc[nd][2]), d);
    } else {
        deriv(T(c_face_loc[ord][k][0]), T(c_face_loc[ord][k][1]), T(c_face_loc[ord][k][2]),
              d);
    }
    return gradop(d, c, g, bad);
}

// quad3d_scs_det: area vector of sub-face k of a quad face (outward for the
// side-ordinal node order)
template <typename T>
__device__ inline void quad_face_areav(const T p[4][3], int k, T A[3]) {
    // sub-face k spans corner p_k, edge midpoints e(k-1), e(k) and the centre,
    // e(i) = mid(p_i, p_{i+1}); selects keep p in registers for a runtime k
    T e[4][3], cc[3], pk[3], e_a[3], e_b[3];
    for (int d = 0; d < 3; ++d) {
        for (int i = 0; i < 4; ++i) e[i][d] = (p[i][d] + p[(i + 1) & 3][d]) * T(0.5);
        cc[d] = (p[0][d] + p[1][d] + p[2][d] + p[3][d]) * T(0.25);
        pk[d] = k == 0 ? p[0][d] : k == 1 ? p[1][d] : k == 2 ? p[2][d] : p[3][d];
        e_a[d] = k == 0 ? e[3][d] : k == 1 ? e[0][d] : k == 2 ? e[1][d] : e[2][d];
        e_b[d] = k == 0 ? e[0][d] : k == 1 ? e[1][d] : k == 2 ? e[2][d] : e[3][d];
    }
    const T dx13 = cc[0] - pk[0], dy13 = cc[1] - pk[1], dz13 = cc[2] - pk[2];
    const T dx24 = e_a[0] - e_b[0], dy24 = e_a[1] - e_b[1], dz24 = e_a[2] - e_b[2];
    A[0] = T(0.5) * (dz24 * dy13 - dz13 * dy24);
    A[1] = T(0.5) * (dx24 * dz13 - dx13 * dz24);
    A[2] = T(0.5) * (dy24 * dx13 - dy13 * dx24);
}

// Grandy volume of a hexahedron (24 triangular facets)
__device__ inline double grandy_volume(const double v[8][3]) {
    double c[14][3];
    for (int j = 0; j < 8; ++j)
        for (int k = 0; k < 3; ++k) c[j][k] = v[j][k];
    for (int k = 0; k < 3; ++k) {
        c[8][k] = 0.25 * (v[0][k] + v[1][k] + v[2][k] + v[3][k]);
        c[9][k] = 0.25 * (v[4][k] + v[5][k] + v[6][k] + v[7][k]);
        c[10][k] = 0.25 * (v[0][k] + v[1][k] + v[5][k] + v[4][k]);
        c[11][k] = 0.25 * (v[3][k] + v[2][k] + v[6][k] + v[7][k]);
        c[12][k] = 0.25 * (v[1][k] + v[2][k] + v[6][k] + v[5][k]);
        c[13][k] = 0.25 * (v[0][k] + v[3][k] + v[7][k] + v[4][k]);
    }
    const int F[24][3] = {{0, 8, 1},  {8, 2, 1},   {3, 2, 8},   {3, 8, 0},  {6, 9, 5},
                          {7, 9, 6},  {4, 9, 7},   {4, 5, 9},   {10, 0, 1}, {5, 10, 1},
                          {4, 10, 5}, {4, 0, 10},  {7, 6, 11},  {6, 2, 11}, {2, 3, 11},
                          {3, 7, 11}, {6, 12, 2},  {5, 12, 6},  {5, 1, 12}, {1, 2, 12},
                          {0, 4, 13}, {4, 7, 13},  {7, 3, 13},  {3, 0, 13}};
    double vol = 0.0;
#pragma unroll
    for (int t = 0; t < 24; ++t) {
        const double *P = c[F[t][0]], *Q = c[F[t][1]], *R = c[F[t][2]];
        const double x0 = P[0] + Q[0] + R[0], x1 = P[1] + Q[1] + R[1], x2 = P[2] + Q[2] + R[2];
        vol += x0 * ((Q[1] - P[1]) * (R[2] - P[2]) - (R[1] - P[1]) * (Q[2] - P[2])) -
               x1 * ((Q[0] - P[0]) * (R[2] - P[2]) - (R[0] - P[0]) * (Q[2] - P[2])) +
               x2 * ((Q[0] - P[0]) * (R[1] - P[1]) - (R[0] - P[0]) * (Q[1] - P[1]));
    }
    return vol / 18.0;
}

// SCV volumes (hex_scv_det): 8 sub-hexahedra of the 27-point subdivision
__device__ inline void scv_volumes(const double c[8][3], double vol[8]) {
    const int S[8][8] = {{0, 8, 12, 11, 19, 20, 26, 25},  {8, 1, 9, 12, 20, 18, 24, 26},
                         {12, 9, 2, 10, 26, 24, 22, 23},  {11, 12, 10, 3, 25, 26, 23, 21},
                         {19, 20, 26, 25, 4, 13, 17, 16}, {20, 18, 24, 26, 13, 5, 14, 17},
                         {26, 24, 22, 23, 17, 14, 6, 15}, {25, 26, 23, 21, 16, 17, 15, 7}};
    double ctr[3];
    centroid(c, ctr);
    for (int s = 0; s < 8; ++s) {
        double v[8][3];
        for (int n = 0; n < 8; ++n) sub_point(S[s][n], c, ctr, v[n]);
        vol[s] = grandy_volume(v);
    }
}

}  // namespace hex8
}  // namespace oagpu
