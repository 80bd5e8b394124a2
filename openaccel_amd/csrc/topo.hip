// topo.hip -- CVFEM master-element geometry of tet4, wedge6 and pyramid5
// elements (and hex8 through hex8.hpp): SCS area vectors, SCV volumes and
// SCS gradient operators of a batch of elements, one lane per element.
//
// Restates (external/nalu/src/master_element/): tet_scs_det / tet_scv_det
// (MasterElementWork.F:459-753), pyr_scs_det / pyr_scv_det (:754-1123),
// wed_scs_det / wed_scv_det (:1124-1466) with quadAreaByTriangleFacets,
// hexVolumeByTriangleFacets, bhexVolumeByTriangleFacets,
// octohedronVolumeByTriangleFacets and polyhedralVolumeByFaces
// (Calculate*.F); the *_gradient_operator routines (:3061, :3345, :3498);
// tet_derivative (:2539), wedge_derivative (Wed6CVFEM.C:59-86),
// pyr_derivative / shifted_pyr_derivative (Pyr5CVFEM.C:119-185); SCS ip
// locations of TetSCS, WedSCS, PyrSCS (standard and shifted).
#include "device_util.hpp"
#include "flow.hpp"

namespace oagpu {

namespace topo {

constexpr int kHex = 0, kTet = 1, kWed = 2, kPyr = 3;

// subdivision point = average of up to six vertices (n == 0: vertex v[0])
struct SubPt {
    int n;
    int v[6];
};

__constant__ SubPt c_tet_pts[15] = {
    {0, {0}}, {0, {1}}, {0, {2}}, {0, {3}}, {2, {0, 1}}, {2, {1, 2}}, {2, {2, 0}}, {3, {0, 1, 2}},
    {2, {2, 3}}, {2, {3, 1}}, {3, {1, 2, 3}}, {2, {0, 3}}, {3, {0, 2, 3}}, {3, {0, 1, 3}},
    {4, {0, 1, 2, 3}}};
__constant__ int c_tet_scs[6][4] = {{4, 7, 14, 13}, {7, 14, 10, 5}, {6, 12, 14, 7},
                                    {11, 13, 14, 12}, {13, 9, 10, 14}, {10, 8, 12, 14}};
__constant__ int c_tet_scv[4][8] = {{0, 4, 7, 6, 11, 13, 14, 12}, {1, 5, 7, 4, 9, 10, 14, 13},
                                    {2, 6, 7, 5, 8, 12, 14, 10}, {3, 9, 13, 11, 8, 10, 14, 12}};
__constant__ double c_tet_ip[2][6][3] = {
    {{13.0 / 36, 5.0 / 36, 5.0 / 36}, {13.0 / 36, 13.0 / 36, 5.0 / 36}, {5.0 / 36, 13.0 / 36, 5.0 / 36},
     {5.0 / 36, 5.0 / 36, 13.0 / 36}, {13.0 / 36, 5.0 / 36, 13.0 / 36}, {5.0 / 36, 13.0 / 36, 13.0 / 36}},
    {{0.5, 0, 0}, {0.5, 0.5, 0}, {0, 0.5, 0}, {0, 0, 0.5}, {0.5, 0, 0.5}, {0, 0.5, 0.5}}};

__constant__ SubPt c_wed_pts[21] = {
    {0, {0}}, {0, {1}}, {0, {2}}, {0, {3}}, {0, {4}}, {0, {5}}, {2, {0, 1}}, {2, {1, 2}},
    {2, {2, 0}}, {3, {0, 1, 2}}, {2, {3, 4}}, {2, {4, 5}}, {2, {5, 3}}, {3, {3, 4, 5}},
    {2, {1, 4}}, {2, {0, 3}}, {4, {0, 1, 4, 3}}, {2, {2, 5}}, {4, {1, 4, 5, 2}},
    {4, {5, 3, 0, 2}}, {6, {0, 1, 2, 3, 4, 5}}};
__constant__ int c_wed_scs[9][4] = {{6, 9, 20, 16},  {7, 9, 20, 18},  {9, 8, 19, 20},
                                    {10, 16, 20, 13}, {13, 11, 18, 20}, {12, 13, 20, 19},
                                    {15, 16, 20, 19}, {16, 14, 18, 20}, {19, 20, 18, 17}};
__constant__ int c_wed_scv[6][8] = {{0, 15, 16, 6, 8, 19, 20, 9},    {9, 6, 1, 7, 20, 16, 14, 18},
                                    {8, 9, 7, 2, 19, 20, 18, 17},    {19, 15, 16, 20, 12, 3, 10, 13},
                                    {20, 16, 14, 18, 13, 10, 4, 11}, {19, 20, 18, 17, 12, 13, 11, 5}};
__constant__ double c_wed_ip[2][9][3] = {
    {{5.0 / 12, 1.0 / 6, -0.5}, {5.0 / 12, 5.0 / 12, -0.5}, {1.0 / 6, 5.0 / 12, -0.5},
     {5.0 / 12, 1.0 / 6, 0.5}, {5.0 / 12, 5.0 / 12, 0.5}, {1.0 / 6, 5.0 / 12, 0.5},
     {7.0 / 36, 7.0 / 36, 0.0}, {11.0 / 18, 7.0 / 36, 0.0}, {7.0 / 36, 11.0 / 18, 0.0}},
    {{0.5, 0, -1}, {0.5, 0.5, -1}, {0, 0.5, -1}, {0.5, 0, 1}, {0.5, 0.5, 1}, {0, 0.5, 1},
     {0, 0, 0}, {1, 0, 0}, {0, 1, 0}}};

__constant__ SubPt c_pyr_pts[19] = {
    {0, {0}}, {0, {1}}, {0, {2}}, {0, {3}}, {0, {4}}, {2, {0, 1}}, {2, {1, 2}}, {2, {2, 3}},
    {2, {3, 0}}, {4, {0, 1, 2, 3}}, {2, {1, 4}}, {2, {4, 0}}, {3, {0, 1, 4}}, {2, {2, 4}},
    {3, {1, 2, 4}}, {2, {3, 4}}, {3, {3, 4, 2}}, {3, {0, 4, 3}}, {5, {0, 1, 2, 3, 4}}};
__constant__ int c_pyr_scs[12][4] = {{5, 9, 18, 12},   {6, 9, 18, 14},   {7, 9, 18, 16},
                                     {8, 17, 18, 9},   {12, 12, 18, 17}, {11, 12, 12, 17},
                                     {14, 14, 18, 12}, {10, 14, 14, 12}, {16, 16, 18, 14},
                                     {13, 16, 16, 14}, {17, 17, 18, 16}, {15, 17, 17, 16}};
__constant__ int c_pyr_scv[4][8] = {{0, 5, 9, 8, 11, 12, 18, 17}, {1, 6, 9, 5, 10, 14, 18, 12},
                                    {2, 7, 9, 6, 13, 16, 18, 14}, {3, 8, 9, 7, 15, 17, 18, 16}};
__constant__ int c_pyr_apex[10] = {4, 18, 15, 17, 11, 12, 10, 14, 13, 16};
__constant__ double c_pyr_ip[2][12][3] = {
    {{0.0, -29.0 / 63, 41.0 / 315}, {29.0 / 63, 0.0, 41.0 / 315}, {0.0, 29.0 / 63, 41.0 / 315},
     {-29.0 / 63, 0.0, 41.0 / 315}, {-2.0 / 9, -2.0 / 9, 13.0 / 45}, {-7.0 / 18, -7.0 / 18, 7.0 / 18},
     {2.0 / 9, -2.0 / 9, 13.0 / 45}, {7.0 / 18, -7.0 / 18, 7.0 / 18}, {2.0 / 9, 2.0 / 9, 13.0 / 45},
     {7.0 / 18, 7.0 / 18, 7.0 / 18}, {-2.0 / 9, 2.0 / 9, 13.0 / 45}, {-7.0 / 18, 7.0 / 18, 7.0 / 18}},
    {{0, -1, 0}, {1, 0, 0}, {0, 1, 0}, {-1, 0, 0}, {-0.5, -0.5, 0.5}, {-0.5, -0.5, 0.5},
     {0.5, -0.5, 0.5}, {0.5, -0.5, 0.5}, {0.5, 0.5, 0.5}, {0.5, 0.5, 0.5}, {-0.5, 0.5, 0.5},
     {-0.5, 0.5, 0.5}}};

// the 24 triangles of the hex / bent-hex and octohedron volume routines
__constant__ int c_hex_tris[24][3] = {
    {0, 8, 1},  {8, 2, 1},  {3, 2, 8},  {3, 8, 0},  {6, 9, 5},  {7, 9, 6},  {4, 9, 7},  {4, 5, 9},
    {10, 0, 1}, {5, 10, 1}, {4, 10, 5}, {4, 0, 10}, {7, 6, 11}, {6, 2, 11}, {2, 3, 11}, {3, 7, 11},
    {6, 12, 2}, {5, 12, 6}, {5, 1, 12}, {1, 2, 12}, {0, 4, 13}, {4, 7, 13}, {7, 3, 13}, {3, 0, 13}};
__constant__ int c_oct_tris[24][3] = {
    {1, 3, 10}, {2, 10, 3}, {2, 9, 10}, {10, 9, 1}, {4, 3, 11}, {3, 1, 11}, {11, 1, 5}, {4, 11, 5},
    {1, 12, 5}, {1, 7, 12}, {12, 7, 6}, {5, 12, 6}, {9, 8, 13}, {13, 8, 7}, {13, 7, 1}, {9, 13, 1},
    {4, 5, 0},  {5, 6, 0},  {6, 7, 0},  {7, 8, 0},  {0, 8, 9},  {0, 9, 2},  {0, 2, 3},  {0, 3, 4}};

template <int TOPO>
struct Traits;
template <>
struct Traits<kTet> {
    static constexpr int npe = 4, nscs = 6, npts = 15;
};
template <>
struct Traits<kWed> {
    static constexpr int npe = 6, nscs = 9, npts = 21;
};
template <>
struct Traits<kPyr> {
    static constexpr int npe = 5, nscs = 12, npts = 19;
};

template <int TOPO>
__device__ const SubPt* pts() {
    if constexpr (TOPO == kTet) return c_tet_pts;
    else if constexpr (TOPO == kWed) return c_wed_pts;
    else return c_pyr_pts;
}

// subdivision points in the Fortran's arithmetic: w * (sum of the vertices),
// the element centroid as a running sum of w * vertex
template <int TOPO>
__device__ void subdivide(const double c[][3], double p[][3]) {
    constexpr int npe = Traits<TOPO>::npe;
    const SubPt* P = pts<TOPO>();
    for (int i = 0; i < Traits<TOPO>::npts; ++i) {
        const int n = P[i].n;
        const double w = n == 2 ? 0.5 : n == 3 ? 1.0 / 3.0 : n == 4 ? 0.25 : n == 5 ? 0.2 : 1.0 / 6.0;
        for (int k = 0; k < 3; ++k) {
            double acc = 0.0;
            if (n == 0) {
                acc = c[P[i].v[0]][k];
            } else if (n == npe) {
                for (int j = 0; j < npe; ++j) acc = acc + w * c[j][k];
            } else {
                for (int j = 0; j < n; ++j) acc = acc + c[P[i].v[j]][k];
                acc = w * acc;
            }
            p[i][k] = acc;
        }
    }
}

__device__ double poly_volume(const double c[14][3], const int (*tri)[3]) {
    double vol = 0.0;
    for (int t = 0; t < 24; ++t) {
        const double* P = c[tri[t][0]];
        const double* Q = c[tri[t][1]];
        const double* R = c[tri[t][2]];
        const double xf0 = P[0] + Q[0] + R[0], xf1 = P[1] + Q[1] + R[1], xf2 = P[2] + Q[2] + R[2];
        vol = vol + xf0 * ((Q[1] - P[1]) * (R[2] - P[2]) - (R[1] - P[1]) * (Q[2] - P[2])) -
              xf1 * ((Q[0] - P[0]) * (R[2] - P[2]) - (R[0] - P[0]) * (Q[2] - P[2])) +
              xf2 * ((Q[0] - P[0]) * (R[1] - P[1]) - (R[0] - P[0]) * (Q[1] - P[1]));
    }
    return vol / 18.0;
}

__device__ double hex_volume(const double v[8][3], bool bent) {
    double c[14][3];
    for (int j = 0; j < 8; ++j)
        for (int k = 0; k < 3; ++k) c[j][k] = v[j][k];
    for (int k = 0; k < 3; ++k) {
        c[8][k] = 0.25 * (v[0][k] + v[1][k] + v[2][k] + v[3][k]);
        c[9][k] = bent ? 0.5 * (v[5][k] + v[7][k]) : 0.25 * (v[4][k] + v[5][k] + v[6][k] + v[7][k]);
        c[10][k] = 0.25 * (v[0][k] + v[1][k] + v[5][k] + v[4][k]);
        c[11][k] = 0.25 * (v[3][k] + v[2][k] + v[6][k] + v[7][k]);
        c[12][k] = 0.25 * (v[1][k] + v[2][k] + v[6][k] + v[5][k]);
        c[13][k] = 0.25 * (v[0][k] + v[3][k] + v[7][k] + v[4][k]);
    }
    return poly_volume(c, c_hex_tris);
}

__device__ double oct_volume(const double v[10][3]) {
    double c[14][3];
    for (int j = 0; j < 10; ++j)
        for (int k = 0; k < 3; ++k) c[j][k] = v[j][k];
    for (int k = 0; k < 3; ++k) {
        c[10][k] = 0.5 * (v[3][k] + v[9][k]);
        c[11][k] = 0.5 * (v[3][k] + v[5][k]);
        c[12][k] = 0.5 * (v[5][k] + v[7][k]);
        c[13][k] = 0.5 * (v[7][k] + v[9][k]);
    }
    return poly_volume(c, c_oct_tris);
}

template <int TOPO>
__device__ void scs_areas(const double p[][3], double* A) {
    for (int ics = 0; ics < Traits<TOPO>::nscs; ++ics) {
        const int* f = TOPO == kTet ? c_tet_scs[ics] : TOPO == kWed ? c_wed_scs[ics] : c_pyr_scs[ics];
        double q[4][3];
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 3; ++k) q[j][k] = p[f[j]][k];
        double a[3];
        hex8::quad_area(q, a);
        for (int k = 0; k < 3; ++k) A[3 * ics + k] = a[k];
    }
}

template <int TOPO>
__device__ void scv_volumes(const double p[][3], double* V) {
    double v[10][3];
    if constexpr (TOPO == kPyr) {
        for (int icv = 0; icv < 4; ++icv) {
            for (int j = 0; j < 8; ++j)
                for (int k = 0; k < 3; ++k) v[j][k] = p[c_pyr_scv[icv][j]][k];
            V[icv] = hex_volume(v, true);
        }
        for (int j = 0; j < 10; ++j)
            for (int k = 0; k < 3; ++k) v[j][k] = p[c_pyr_apex[j]][k];
        V[4] = oct_volume(v);
    } else {
        for (int icv = 0; icv < Traits<TOPO>::npe; ++icv) {
            const int* t = TOPO == kTet ? c_tet_scv[icv] : c_wed_scv[icv];
            for (int j = 0; j < 8; ++j)
                for (int k = 0; k < 3; ++k) v[j][k] = p[t[j]][k];
            V[icv] = hex_volume(v, false);
        }
    }
}

// dN/ds [npe][3] at parametric point (r, s, z)
template <int TOPO>
__device__ void deriv(int shifted, double r, double s, double z, double d[][3]) {
    if constexpr (TOPO == kTet) {
        const double t4[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int n = 0; n < 4; ++n)
            for (int k = 0; k < 3; ++k) d[n][k] = t4[n][k];
    } else if constexpr (TOPO == kWed) {
        const double t = 1.0 - r - s, xi = z;
        const double w[6][3] = {{-0.5 * (1.0 - xi), -0.5 * (1.0 - xi), -0.5 * t},
                                {0.5 * (1.0 - xi), 0.0, -0.5 * r},
                                {0.0, 0.5 * (1.0 - xi), -0.5 * s},
                                {-0.5 * (1.0 + xi), -0.5 * (1.0 + xi), 0.5 * t},
                                {0.5 * (1.0 + xi), 0.0, 0.5 * r},
                                {0.0, 0.5 * (1.0 + xi), 0.5 * s}};
        for (int n = 0; n < 6; ++n)
            for (int k = 0; k < 3; ++k) d[n][k] = w[n][k];
    } else if (!shifted) {
        const double eps = DBL_EPSILON;
        const double omt = 1.0 - z;
        const double t = fabs(omt) > eps ? z : 1.0 + copysign(eps, omt);
        const double q = 0.25 / (1.0 - t);
        const double tt = 4.0 * r * s * q * q;
        const double p5[5][3] = {{-(1.0 - s - t) * q, -(1.0 - r - t) * q, (+tt - 0.25)},
                                 {+(1.0 - s - t) * q, -(1.0 + r - t) * q, (-tt - 0.25)},
                                 {+(1.0 + s - t) * q, +(1.0 + r - t) * q, (+tt - 0.25)},
                                 {-(1.0 + s - t) * q, +(1.0 - r - t) * q, (-tt - 0.25)},
                                 {0.0, 0.0, 1.0}};
        for (int n = 0; n < 5; ++n)
            for (int k = 0; k < 3; ++k) d[n][k] = p5[n][k];
    } else {
        const double t = z;
        const double p5[5][3] = {
            {-0.25 * (1.0 - s) * (1.0 - t), -0.25 * (1.0 - r) * (1.0 - t), -0.25 * (1.0 - r) * (1.0 - s)},
            {0.25 * (1.0 - s) * (1.0 - t), -0.25 * (1.0 + r) * (1.0 - t), -0.25 * (1.0 + r) * (1.0 - s)},
            {0.25 * (1.0 + s) * (1.0 - t), 0.25 * (1.0 + r) * (1.0 - t), -0.25 * (1.0 + r) * (1.0 + s)},
            {-0.25 * (1.0 + s) * (1.0 - t), 0.25 * (1.0 - r) * (1.0 - t), -0.25 * (1.0 - r) * (1.0 + s)},
            {0.0, 0.0, 1.0}};
        for (int n = 0; n < 5; ++n)
            for (int k = 0; k < 3; ++k) d[n][k] = p5[n][k];
    }
}

// the gradient-operator algorithm of the *_gradient_operator routines
template <int NPE>
__device__ double gradop(const double d[][3], const double c[][3], double g[][3], int* bad) {
    double dx1 = 0, dx2 = 0, dx3 = 0, dy1 = 0, dy2 = 0, dy3 = 0, dz1 = 0, dz2 = 0, dz3 = 0;
    for (int n = 0; n < NPE; ++n) {
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
    const double det = dx1 * (dy2 * dz3 - dz2 * dy3) + dy1 * (dz2 * dx3 - dx2 * dz3) +
                       dz1 * (dx2 * dy3 - dy2 * dx3);
    double test = det;
    if (test <= 1.e6 * 2.2250738585072014e-308) {
        test = 1.0;
        *bad = 1;
    }
    const double den = 1.0 / test;
    const double s1x = den * (dy2 * dz3 - dz2 * dy3), s2x = den * (dz1 * dy3 - dy1 * dz3),
                 s3x = den * (dy1 * dz2 - dz1 * dy2);
    const double s1y = den * (dz2 * dx3 - dx2 * dz3), s2y = den * (dx1 * dz3 - dz1 * dx3),
                 s3y = den * (dz1 * dx2 - dx1 * dz2);
    const double s1z = den * (dx2 * dy3 - dy2 * dx3), s2z = den * (dy1 * dx3 - dx1 * dy3),
                 s3z = den * (dx1 * dy2 - dy1 * dx2);
    for (int n = 0; n < NPE; ++n) {
        g[n][0] = d[n][0] * s1x + d[n][1] * s2x + d[n][2] * s3x;
        g[n][1] = d[n][0] * s1y + d[n][1] * s2y + d[n][2] * s3y;
        g[n][2] = d[n][0] * s1z + d[n][1] * s2z + d[n][2] * s3z;
    }
    return det;
}

template <int TOPO>
__device__ const double* ip_loc(int shifted, int ip) {
    if constexpr (TOPO == kTet) return c_tet_ip[shifted][ip];
    else if constexpr (TOPO == kWed) return c_wed_ip[shifted][ip];
    else return c_pyr_ip[shifted][ip];
}

}  // namespace topo

// One lane per element: A [n][nscs][3], V [n][npe], G [n][nscs][npe][3],
// detj [n][nscs] (any of them may be null)
template <int TOPO>
__global__ void __launch_bounds__(64) k_element_geometry(int n, int shifted, const double* coords,
                                                         double* A, double* V, double* G,
                                                         double* detj, int* err) {
    using Tr = topo::Traits<TOPO>;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double c[Tr::npe][3];
    for (int j = 0; j < Tr::npe; ++j)
        for (int k = 0; k < 3; ++k) c[j][k] = coords[(size_t(e) * Tr::npe + j) * 3 + k];
    double p[Tr::npts][3];
    topo::subdivide<TOPO>(c, p);
    if (A) topo::scs_areas<TOPO>(p, A + size_t(e) * Tr::nscs * 3);
    if (V) topo::scv_volumes<TOPO>(p, V + size_t(e) * Tr::npe);
    int bad = 0;
    if (G || detj) {
        for (int ip = 0; ip < Tr::nscs; ++ip) {
            const double* loc = topo::ip_loc<TOPO>(shifted, ip);
            double d[Tr::npe][3], g[Tr::npe][3];
            topo::deriv<TOPO>(shifted, loc[0], loc[1], loc[2], d);
            const double dj = topo::gradop<Tr::npe>(d, c, g, &bad);
            if (detj) detj[size_t(e) * Tr::nscs + ip] = dj;
            if (G)
                for (int j = 0; j < Tr::npe; ++j)
                    for (int k = 0; k < 3; ++k)
                        G[((size_t(e) * Tr::nscs + ip) * Tr::npe + j) * 3 + k] = g[j][k];
        }
    }
    if (bad) *err = 1;
}

// hex8 through the row kernels' device geometry
__global__ void __launch_bounds__(64) k_element_geometry_hex(int n, int shifted, const double* coords,
                                                             double* A, double* V, double* G,
                                                             double* detj, int* err) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double c[8][3];
    for (int j = 0; j < 8; ++j)
        for (int k = 0; k < 3; ++k) c[j][k] = coords[(size_t(e) * 8 + j) * 3 + k];
    if (A) {
        double ctr[3];
        hex8::centroid(c, ctr);
        for (int ip = 0; ip < 12; ++ip) hex8::scs_areav(ip, c, ctr, A + (size_t(e) * 12 + ip) * 3);
    }
    if (V) hex8::scv_volumes(c, V + size_t(e) * 8);
    int bad = 0;
    if (G || detj)
        for (int ip = 0; ip < 12; ++ip) {
            double g[8][3];
            const double dj = hex8::scs_gradop(ip, shifted, c, g, &bad);
            if (detj) detj[size_t(e) * 12 + ip] = dj;
            if (G)
                for (int j = 0; j < 8; ++j)
                    for (int k = 0; k < 3; ++k) G[((size_t(e) * 12 + ip) * 8 + j) * 3 + k] = g[j][k];
        }
    if (bad) *err = 1;
}

int topology_npe(int topo) { return topo == 1 ? 4 : topo == 2 ? 6 : topo == 3 ? 5 : 8; }
int topology_nscs(int topo) { return topo == 1 ? 6 : topo == 2 ? 9 : 12; }

void element_geometry(int topo, int shifted, int n, const double* coords, double* A, double* V,
                      double* G, double* detj, int* err, hipStream_t s) {
    const unsigned B = 64, grid = unsigned((n + B - 1) / B);
    if (n <= 0) return;
    switch (topo) {
        case 0: k_element_geometry_hex<<<grid, B, 0, s>>>(n, shifted, coords, A, V, G, detj, err); break;
        case 1: k_element_geometry<topo::kTet><<<grid, B, 0, s>>>(n, shifted, coords, A, V, G, detj, err); break;
        case 2: k_element_geometry<topo::kWed><<<grid, B, 0, s>>>(n, shifted, coords, A, V, G, detj, err); break;
        case 3: k_element_geometry<topo::kPyr><<<grid, B, 0, s>>>(n, shifted, coords, A, V, G, detj, err); break;
        default: throw Error(OAGPU_EINVAL, "unknown element topology");
    }
    OAGPU_HIP(hipGetLastError());
}

}  // namespace oagpu
