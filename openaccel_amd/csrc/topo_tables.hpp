// topo_tables.hpp -- master-element tables of tet4, wedge6 and pyramid5
// (external/nalu/src/master_element/Tet4CVFEM.C, Wed6CVFEM.C, Pyr5CVFEM.C:
// SCS subdivision points and quads of tet_scs_det / wed_scs_det /
// pyr_scs_det in MasterElementWork.F:459-1466, SCS ip locations standard and
// shifted, lrscv_) and per-ip geometry shared by the batch geometry kernel
// and the mixed-mesh element passes: shape functions (tet_shape_fcn
// Tet4CVFEM.C:300, wedge_shape_fcn Wed6CVFEM.C:1151, pyr_shape_fcn /
// shifted_pyr_shape_fcn Pyr5CVFEM.C:1365-1420), derivatives (tet_derivative
// MasterElementWork.F:2539, wedge_derivative Wed6CVFEM.C:59-86,
// pyr_derivative / shifted_pyr_derivative Pyr5CVFEM.C:119-185), the
// *_gradient_operator algorithm and one SCS area vector.
#pragma once

#include <cfloat>

#include "hex8.hpp"

namespace oagpu {
namespace topo {

constexpr int kHex = 0, kTet = 1, kWed = 2, kPyr = 3;

// subdivision point = average of up to six vertices (n == 0: vertex v[0])
struct SubPt {
    int n;
    int v[6];
};

static __constant__ SubPt c_tet_pts[15] = {
    {0, {0}}, {0, {1}}, {0, {2}}, {0, {3}}, {2, {0, 1}}, {2, {1, 2}}, {2, {2, 0}}, {3, {0, 1, 2}},
    {2, {2, 3}}, {2, {3, 1}}, {3, {1, 2, 3}}, {2, {0, 3}}, {3, {0, 2, 3}}, {3, {0, 1, 3}},
    {4, {0, 1, 2, 3}}};
static __constant__ int c_tet_scs[6][4] = {{4, 7, 14, 13}, {7, 14, 10, 5}, {6, 12, 14, 7},
                                    {11, 13, 14, 12}, {13, 9, 10, 14}, {10, 8, 12, 14}};
static __constant__ double c_tet_ip[2][6][3] = {
    {{13.0 / 36, 5.0 / 36, 5.0 / 36}, {13.0 / 36, 13.0 / 36, 5.0 / 36}, {5.0 / 36, 13.0 / 36, 5.0 / 36},
     {5.0 / 36, 5.0 / 36, 13.0 / 36}, {13.0 / 36, 5.0 / 36, 13.0 / 36}, {5.0 / 36, 13.0 / 36, 13.0 / 36}},
    {{0.5, 0, 0}, {0.5, 0.5, 0}, {0, 0.5, 0}, {0, 0, 0.5}, {0.5, 0, 0.5}, {0, 0.5, 0.5}}};

static __constant__ SubPt c_wed_pts[21] = {
    {0, {0}}, {0, {1}}, {0, {2}}, {0, {3}}, {0, {4}}, {0, {5}}, {2, {0, 1}}, {2, {1, 2}},
    {2, {2, 0}}, {3, {0, 1, 2}}, {2, {3, 4}}, {2, {4, 5}}, {2, {5, 3}}, {3, {3, 4, 5}},
    {2, {1, 4}}, {2, {0, 3}}, {4, {0, 1, 4, 3}}, {2, {2, 5}}, {4, {1, 4, 5, 2}},
    {4, {5, 3, 0, 2}}, {6, {0, 1, 2, 3, 4, 5}}};
static __constant__ int c_wed_scs[9][4] = {{6, 9, 20, 16},  {7, 9, 20, 18},  {9, 8, 19, 20},
                                    {10, 16, 20, 13}, {13, 11, 18, 20}, {12, 13, 20, 19},
                                    {15, 16, 20, 19}, {16, 14, 18, 20}, {19, 20, 18, 17}};
static __constant__ double c_wed_ip[2][9][3] = {
    {{5.0 / 12, 1.0 / 6, -0.5}, {5.0 / 12, 5.0 / 12, -0.5}, {1.0 / 6, 5.0 / 12, -0.5},
     {5.0 / 12, 1.0 / 6, 0.5}, {5.0 / 12, 5.0 / 12, 0.5}, {1.0 / 6, 5.0 / 12, 0.5},
     {7.0 / 36, 7.0 / 36, 0.0}, {11.0 / 18, 7.0 / 36, 0.0}, {7.0 / 36, 11.0 / 18, 0.0}},
    {{0.5, 0, -1}, {0.5, 0.5, -1}, {0, 0.5, -1}, {0.5, 0, 1}, {0.5, 0.5, 1}, {0, 0.5, 1},
     {0, 0, 0}, {1, 0, 0}, {0, 1, 0}}};

static __constant__ SubPt c_pyr_pts[19] = {
    {0, {0}}, {0, {1}}, {0, {2}}, {0, {3}}, {0, {4}}, {2, {0, 1}}, {2, {1, 2}}, {2, {2, 3}},
    {2, {3, 0}}, {4, {0, 1, 2, 3}}, {2, {1, 4}}, {2, {4, 0}}, {3, {0, 1, 4}}, {2, {2, 4}},
    {3, {1, 2, 4}}, {2, {3, 4}}, {3, {3, 4, 2}}, {3, {0, 4, 3}}, {5, {0, 1, 2, 3, 4}}};
static __constant__ int c_pyr_scs[12][4] = {{5, 9, 18, 12},   {6, 9, 18, 14},   {7, 9, 18, 16},
                                     {8, 17, 18, 9},   {12, 12, 18, 17}, {11, 12, 12, 17},
                                     {14, 14, 18, 12}, {10, 14, 14, 12}, {16, 16, 18, 14},
                                     {13, 16, 16, 14}, {17, 17, 18, 16}, {15, 17, 17, 16}};
static __constant__ double c_pyr_ip[2][12][3] = {
    {{0.0, -29.0 / 63, 41.0 / 315}, {29.0 / 63, 0.0, 41.0 / 315}, {0.0, 29.0 / 63, 41.0 / 315},
     {-29.0 / 63, 0.0, 41.0 / 315}, {-2.0 / 9, -2.0 / 9, 13.0 / 45}, {-7.0 / 18, -7.0 / 18, 7.0 / 18},
     {2.0 / 9, -2.0 / 9, 13.0 / 45}, {7.0 / 18, -7.0 / 18, 7.0 / 18}, {2.0 / 9, 2.0 / 9, 13.0 / 45},
     {7.0 / 18, 7.0 / 18, 7.0 / 18}, {-2.0 / 9, 2.0 / 9, 13.0 / 45}, {-7.0 / 18, 7.0 / 18, 7.0 / 18}},
    {{0, -1, 0}, {1, 0, 0}, {0, 1, 0}, {-1, 0, 0}, {-0.5, -0.5, 0.5}, {-0.5, -0.5, 0.5},
     {0.5, -0.5, 0.5}, {0.5, -0.5, 0.5}, {0.5, 0.5, 0.5}, {0.5, 0.5, 0.5}, {-0.5, 0.5, 0.5},
     {-0.5, 0.5, 0.5}}};

static __constant__ int c_tet_lr[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
static __constant__ int c_wed_lr[9][2] = {{0, 1}, {1, 2}, {0, 2}, {3, 4}, {4, 5},
                                         {3, 5}, {0, 3}, {1, 4}, {2, 5}};
static __constant__ int c_pyr_lr[12][2] = {{0, 1}, {1, 2}, {2, 3}, {0, 3}, {0, 4}, {0, 4},
                                          {1, 4}, {1, 4}, {2, 4}, {2, 4}, {3, 4}, {3, 4}};

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
__device__ inline const SubPt* pts() {
    if constexpr (TOPO == kTet) return c_tet_pts;
    else if constexpr (TOPO == kWed) return c_wed_pts;
    else return c_pyr_pts;
}

template <int TOPO>
__device__ inline const int* scs_quad(int ip) {
    if constexpr (TOPO == kTet) return c_tet_scs[ip];
    else if constexpr (TOPO == kWed) return c_wed_scs[ip];
    else return c_pyr_scs[ip];
}

template <int TOPO>
__device__ inline const int* lr(int ip) {
    if constexpr (TOPO == kTet) return c_tet_lr[ip];
    else if constexpr (TOPO == kWed) return c_wed_lr[ip];
    else return c_pyr_lr[ip];
}

template <int TOPO>
__device__ inline const double* ip_loc(int shifted, int ip) {
    if constexpr (TOPO == kTet) return c_tet_ip[shifted][ip];
    else if constexpr (TOPO == kWed) return c_wed_ip[shifted][ip];
    else return c_pyr_ip[shifted][ip];
}

// subdivision point i: w * (sum of its vertices), the element centroid as a
// running sum of w * vertex (the Fortran's arithmetic, vertices summed in
// ascending order).  The vertex set is applied as a mask over a fully
// unrolled vertex loop: no run-time index into the register array c.
template <int TOPO, typename T>
__device__ inline void sub_point(int i, const T c[][3], T p[3]) {
    constexpr int npe = Traits<TOPO>::npe;
    const SubPt& P = pts<TOPO>()[i];
    const int n = P.n;
    unsigned mask = 0;
    for (int j = 0; j < (n == 0 ? 1 : n); ++j) mask |= 1u << P.v[j];
    const T w = n == 2 ? T(0.5) : n == 3 ? T(1.0 / 3.0) : n == 4 ? T(0.25) : n == 5 ? T(0.2) : T(1.0 / 6.0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T acc = T(0);
        if (n == npe) {
#pragma unroll
            for (int j = 0; j < npe; ++j) acc = acc + w * c[j][k];
        } else {
#pragma unroll
            for (int j = 0; j < npe; ++j)
                if (mask & (1u << j)) acc = acc + c[j][k];
            if (n != 0) acc = w * acc;
        }
        p[k] = acc;
    }
}

// row r of a register array, r known only at run time: a select over the
// unrolled rows
template <int NPE, typename T>
__device__ inline void pick_row(const T c[][3], int r, T out[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        T v = T(0);
#pragma unroll
        for (int j = 0; j < NPE; ++j) v = j == r ? c[j][d] : v;
        out[d] = v;
    }
}

template <int NPE, typename T>
__device__ inline T pick(const T v[], int r) {
    T x = T(0);
#pragma unroll
    for (int j = 0; j < NPE; ++j) x = j == r ? v[j] : x;
    return x;
}

// area vector of SCS ip (oriented from its left to its right node)
template <int TOPO, typename T>
__device__ inline void scs_area(int ip, const T c[][3], T A[3]) {
    const int* f = scs_quad<TOPO>(ip);
    T q[4][3];
#pragma unroll
    for (int j = 0; j < 4; ++j) sub_point<TOPO>(f[j], c, q[j]);
    hex8::quad_area(q, A);
}

// shape functions [npe] at parametric point (r, s, z)
template <int TOPO, typename T>
__device__ inline void shape(int shifted, T r, T s, T z, T N[]) {
    if constexpr (TOPO == kTet) {
        N[0] = T(1) - r - s - z;
        N[1] = r;
        N[2] = s;
        N[3] = z;
    } else if constexpr (TOPO == kWed) {
        const T t = T(1) - r - s;
        N[0] = T(0.5) * t * (T(1) - z);
        N[1] = T(0.5) * r * (T(1) - z);
        N[2] = T(0.5) * s * (T(1) - z);
        N[3] = T(0.5) * t * (T(1) + z);
        N[4] = T(0.5) * r * (T(1) + z);
        N[5] = T(0.5) * s * (T(1) + z);
    } else if (!shifted) {
        const T eps = sizeof(T) == 4 ? T(FLT_EPSILON) : T(DBL_EPSILON);
        const T omt = T(1) - z;
        const T t = fabs(omt) > eps ? z : T(1) + copysign(eps, omt);
        const T q = T(0.25) / (T(1) - t);
        N[0] = (T(1) - r - t) * (T(1) - s - t) * q;
        N[1] = (T(1) + r - t) * (T(1) - s - t) * q;
        N[2] = (T(1) + r - t) * (T(1) + s - t) * q;
        N[3] = (T(1) - r - t) * (T(1) + s - t) * q;
        N[4] = t;
    } else {
        N[0] = T(0.25) * (T(1) - r) * (T(1) - s) * (T(1) - z);
        N[1] = T(0.25) * (T(1) + r) * (T(1) - s) * (T(1) - z);
        N[2] = T(0.25) * (T(1) + r) * (T(1) + s) * (T(1) - z);
        N[3] = T(0.25) * (T(1) - r) * (T(1) + s) * (T(1) - z);
        N[4] = z;
    }
}

// dN/ds [npe][3] at parametric point (r, s, z)
template <int TOPO, typename T>
__device__ inline void deriv(int shifted, T r, T s, T z, T d[][3]) {
    if constexpr (TOPO == kTet) {
        const T t4[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int n = 0; n < 4; ++n)
            for (int k = 0; k < 3; ++k) d[n][k] = t4[n][k];
    } else if constexpr (TOPO == kWed) {
        const T t = T(1) - r - s, xi = z, h = T(0.5);
        const T w[6][3] = {{-h * (T(1) - xi), -h * (T(1) - xi), -h * t},
                           {h * (T(1) - xi), T(0), -h * r},
                           {T(0), h * (T(1) - xi), -h * s},
                           {-h * (T(1) + xi), -h * (T(1) + xi), h * t},
                           {h * (T(1) + xi), T(0), h * r},
                           {T(0), h * (T(1) + xi), h * s}};
        for (int n = 0; n < 6; ++n)
            for (int k = 0; k < 3; ++k) d[n][k] = w[n][k];
    } else if (!shifted) {
        const T eps = sizeof(T) == 4 ? T(FLT_EPSILON) : T(DBL_EPSILON);
        const T omt = T(1) - z;
        const T t = fabs(omt) > eps ? z : T(1) + copysign(eps, omt);
        const T q = T(0.25) / (T(1) - t);
        const T tt = T(4) * r * s * q * q, f = T(0.25);
        const T p5[5][3] = {{-(T(1) - s - t) * q, -(T(1) - r - t) * q, (+tt - f)},
                            {+(T(1) - s - t) * q, -(T(1) + r - t) * q, (-tt - f)},
                            {+(T(1) + s - t) * q, +(T(1) + r - t) * q, (+tt - f)},
                            {-(T(1) + s - t) * q, +(T(1) - r - t) * q, (-tt - f)},
                            {T(0), T(0), T(1)}};
        for (int n = 0; n < 5; ++n)
            for (int k = 0; k < 3; ++k) d[n][k] = p5[n][k];
    } else {
        const T t = z, f = T(0.25);
        const T p5[5][3] = {
            {-f * (T(1) - s) * (T(1) - t), -f * (T(1) - r) * (T(1) - t), -f * (T(1) - r) * (T(1) - s)},
            {f * (T(1) - s) * (T(1) - t), -f * (T(1) + r) * (T(1) - t), -f * (T(1) + r) * (T(1) - s)},
            {f * (T(1) + s) * (T(1) - t), f * (T(1) + r) * (T(1) - t), -f * (T(1) + r) * (T(1) + s)},
            {-f * (T(1) + s) * (T(1) - t), f * (T(1) - r) * (T(1) - t), -f * (T(1) - r) * (T(1) + s)},
            {T(0), T(0), T(1)}};
        for (int n = 0; n < 5; ++n)
            for (int k = 0; k < 3; ++k) d[n][k] = p5[n][k];
    }
}

// the gradient-operator algorithm of the *_gradient_operator routines:
// g = dN/dx [npe][3], returns det J; a determinant <= 1e6 * DBL_MIN is
// replaced by 1 and *bad set
template <int NPE, typename T>
__device__ inline T gradop(const T d[][3], const T c[][3], T g[][3], int* bad) {
    T dx1 = 0, dx2 = 0, dx3 = 0, dy1 = 0, dy2 = 0, dy3 = 0, dz1 = 0, dz2 = 0, dz3 = 0;
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
    const T det = dx1 * (dy2 * dz3 - dz2 * dy3) + dy1 * (dz2 * dx3 - dx2 * dz3) +
                  dz1 * (dx2 * dy3 - dy2 * dx3);
    T test = det;
    if (test <= T(1.e6 * 2.2250738585072014e-308)) {
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
    for (int n = 0; n < NPE; ++n) {
        g[n][0] = d[n][0] * s1x + d[n][1] * s2x + d[n][2] * s3x;
        g[n][1] = d[n][0] * s1y + d[n][1] * s2y + d[n][2] * s3y;
        g[n][2] = d[n][0] * s1z + d[n][1] * s2z + d[n][2] * s3z;
    }
    return det;
}

// geometry of SCS ip of one element (coordinates c): area vector, dN/dx at
// the standard or shifted ip (gshift) and the shape functions at the
// standard (Nc) and velocity-interpolation (Nu, shifted iff us) ips
template <int TOPO, typename T>
__device__ inline void scs_ip(int ip, int gshift, int us, const T c[][3], T A[3], T g[][3], T Nc[],
                              T Nu[], int& bad) {
    constexpr int npe = Traits<TOPO>::npe;
    scs_area<TOPO>(ip, c, A);
    {
        const double* l = ip_loc<TOPO>(gshift, ip);
        T d[npe][3];
        deriv<TOPO>(gshift, T(l[0]), T(l[1]), T(l[2]), d);
        gradop<npe>(d, c, g, &bad);
    }
    const double* l0 = ip_loc<TOPO>(0, ip);
    shape<TOPO>(0, T(l0[0]), T(l0[1]), T(l0[2]), Nc);
    const double* lu = ip_loc<TOPO>(us, ip);
    shape<TOPO>(us, T(lu[0]), T(lu[1]), T(lu[2]), Nu);
}

// Static ip geometry of a tet4 / wedge6 / pyramid5 element, precomputed in
// FP64 once per mesh (k_geo_generic): per SCS ip the area vector and dN/dx
// at the standard and at the shifted ip, [nscs][3 + 2 * npe * 3] from geo.
template <int TOPO>
__device__ constexpr int geo_stride() {
    return 3 + 6 * Traits<TOPO>::npe;
}

template <int TOPO, typename T>
__device__ inline void ip_geom(const double* geo, int ip, int gshift, int us, T A[3], T g[][3],
                               T Nc[], T Nu[]) {
    constexpr int npe = Traits<TOPO>::npe;
    const double* q = geo + size_t(ip) * geo_stride<TOPO>();
    for (int d = 0; d < 3; ++d) A[d] = T(q[d]);
    const double* gq = q + 3 + (gshift ? 3 * npe : 0);
    for (int n = 0; n < npe; ++n)
        for (int d = 0; d < 3; ++d) g[n][d] = T(gq[3 * n + d]);
    const double* l0 = ip_loc<TOPO>(0, ip);
    shape<TOPO>(0, T(l0[0]), T(l0[1]), T(l0[2]), Nc);
    const double* lu = ip_loc<TOPO>(us, ip);
    shape<TOPO>(us, T(lu[0]), T(lu[1]), T(lu[2]), Nu);
}

}  // namespace topo
}  // namespace oagpu
