/*
 * master_element.c -- oracle restatement of the nalu hex8 / quad4 CVFEM
 * master elements.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Sources restated:
 *   external/nalu/src/master_element/MasterElementWork.F
 *       hex_scs_det :1-231, hex_scv_det :233-457, quad3d_scs_det :1467-1588,
 *       hex_shape_fcn :2344-2402, hex_derivative :2404-2484,
 *       hex_gradient_operator :2623-2775, quad3d_shape_fcn :3651-3697
 *   external/nalu/src/master_element/CalculateQuadArea.F (quadAreaByTriangleFacets)
 *   external/nalu/src/master_element/CalculateHexVolume.F (hexVolumeByTriangleFacets)
 *   external/nalu/src/master_element/CalculatePolyhedralVolume.F (polyhedralVolumeByFaces)
 *   external/nalu/src/master_element/Hex8CVFEM.C ip tables :264-344, :381-426
 *   external/nalu/src/master_element/Quad43DCVFEM.C ip tables :31-48
 * The Fortran is 1-based and column-major; indices here are 0-based.
 */
#include "oracle.h"
#include <math.h>
#include <string.h>

/* L/R node pair of each SCS ip (Hex8CVFEM.C:264-276) */
static const int hex_lrscv[24] = {0, 1, 1, 2, 2, 3, 0, 3, 4, 5, 5, 6,
                                  6, 7, 4, 7, 0, 4, 1, 5, 2, 6, 3, 7};

/* SCS ip isoparametric locations, standard and shifted (Hex8CVFEM.C:318-344) */
static const double hex_scs_loc[36] = {
    0.00, -0.25, -0.25, 0.25, 0.00, -0.25, 0.00, 0.25, -0.25, -0.25, 0.00, -0.25,
    0.00, -0.25, 0.25,  0.25, 0.00, 0.25,  0.00, 0.25, 0.25,  -0.25, 0.00, 0.25,
    -0.25, -0.25, 0.00, 0.25, -0.25, 0.00, 0.25, 0.25, 0.00,  -0.25, 0.25, 0.00};
static const double hex_scs_loc_shift[36] = {
    0.00, -0.50, -0.50, 0.50, 0.00, -0.50, 0.00, 0.50, -0.50, -0.50, 0.00, -0.50,
    0.00, -0.50, 0.50,  0.50, 0.00, 0.50,  0.00, 0.50, 0.50,  -0.50, 0.00, 0.50,
    -0.50, -0.50, 0.00, 0.50, -0.50, 0.00, 0.50, 0.50, 0.00,  -0.50, 0.50, 0.00};

/* side ordinal -> element node ordinals; also the face ip -> nearest node map
 * (Hex8CVFEM.C:381-393, 419-426) */
static const int hex_side_nodes[24] = {0, 1, 5, 4, 1, 2, 6, 5, 2, 3, 7, 6,
                                       0, 4, 7, 3, 0, 3, 2, 1, 4, 5, 6, 7};

/* exposed-face ip locations (Hex8CVFEM.C:346-377) */
static const double hex_face_loc[72] = {
    -0.25, -0.50, -0.25, 0.25,  -0.50, -0.25, 0.25,  -0.50, 0.25,  -0.25, -0.50, 0.25,
    0.50,  -0.25, -0.25, 0.50,  0.25,  -0.25, 0.50,  0.25,  0.25,  0.50,  -0.25, 0.25,
    0.25,  0.50,  -0.25, -0.25, 0.50,  -0.25, -0.25, 0.50,  0.25,  0.25,  0.50,  0.25,
    -0.50, -0.25, -0.25, -0.50, -0.25, 0.25,  -0.50, 0.25,  0.25,  -0.50, 0.25,  -0.25,
    -0.25, -0.25, -0.50, -0.25, 0.25,  -0.50, 0.25,  0.25,  -0.50, 0.25,  -0.25, -0.50,
    -0.25, -0.25, 0.50,  0.25,  -0.25, 0.50,  0.25,  0.25,  0.50,  -0.25, 0.25,  0.50};

/* hex node locations (Hex8CVFEM.C:428-432); the shifted face ips are the
 * face-node locations (Hex8CVFEM.C:434-446) */
static const double hex_node_loc[24] = {-0.5, -0.5, -0.5, 0.5, -0.5, -0.5, 0.5, 0.5, -0.5,
                                        -0.5, 0.5,  -0.5, -0.5, -0.5, 0.5, 0.5, -0.5, 0.5,
                                        0.5,  0.5,  0.5,  -0.5, 0.5,  0.5};

const int *or_hex_lrscv(void) { return hex_lrscv; }
const int *or_hex_side_nodes(void) { return hex_side_nodes; }

/* The 27-point subdivision of hex_scs_det / hex_scv_det
 * (MasterElementWork.F:72-184 and :297-408). */
static void hex_subdivide(const double *c, double p[27][3])
{
    for (int j = 0; j < 8; ++j)
        for (int k = 0; k < 3; ++k) p[j][k] = c[3 * j + k];
    for (int k = 0; k < 3; ++k) {
#define N(i) c[3 * (i) + k]
        p[8][k] = 0.5 * (N(0) + N(1));
        p[9][k] = 0.5 * (N(1) + N(2));
        p[10][k] = 0.5 * (N(2) + N(3));
        p[11][k] = 0.5 * (N(3) + N(0));
        p[12][k] = 0.25 * (N(0) + N(1) + N(2) + N(3));
        p[13][k] = 0.5 * (N(4) + N(5));
        p[14][k] = 0.5 * (N(5) + N(6));
        p[15][k] = 0.5 * (N(6) + N(7));
        p[16][k] = 0.5 * (N(7) + N(4));
        p[17][k] = 0.25 * (N(4) + N(5) + N(6) + N(7));
        p[18][k] = 0.5 * (N(1) + N(5));
        p[19][k] = 0.5 * (N(0) + N(4));
        p[20][k] = 0.25 * (N(0) + N(1) + N(4) + N(5));
        p[21][k] = 0.5 * (N(3) + N(7));
        p[22][k] = 0.5 * (N(2) + N(6));
        p[23][k] = 0.25 * (N(2) + N(3) + N(6) + N(7));
        p[24][k] = 0.25 * (N(1) + N(2) + N(5) + N(6));
        p[25][k] = 0.25 * (N(0) + N(3) + N(4) + N(7));
        p[26][k] = 0.125 * (N(0) + N(1) + N(2) + N(3) + N(4) + N(5) + N(6) + N(7));
#undef N
    }
}

/* quadAreaByTriangleFacets (CalculateQuadArea.F:1-67): 4 triangles about the
 * quad's vertex average, facet table (5,1,2),(5,2,3),(5,3,4),(5,4,1) */
static void quad_area_by_triangles(const double q[4][3], double *area)
{
    double xm[3], r1[3], r2[3];
    for (int k = 0; k < 3; ++k) {
        xm[k] = 0.25 * (q[0][k] + q[1][k] + q[2][k] + q[3][k]);
        area[k] = 0.0;
        r2[k] = q[0][k] - xm[k];
    }
    static const int iq[4] = {1, 2, 3, 0};
    for (int t = 0; t < 4; ++t) {
        for (int k = 0; k < 3; ++k) {
            r1[k] = r2[k];
            r2[k] = q[iq[t]][k] - xm[k];
        }
        area[0] += r1[1] * r2[2] - r2[1] * r1[2];
        area[1] += r1[2] * r2[0] - r2[2] * r1[0];
        area[2] += r1[0] * r2[1] - r2[0] * r1[1];
    }
    for (int k = 0; k < 3; ++k) area[k] *= 0.5;
}

/* HexEdgeFacetTable (MasterElementWork.F:41-53), 0-based */
static const int hex_edge_facet[12][4] = {
    {20, 8, 12, 26},  {24, 9, 12, 26},  {10, 12, 26, 23}, {11, 25, 26, 12},
    {13, 20, 26, 17}, {17, 14, 24, 26}, {17, 15, 23, 26}, {16, 17, 26, 25},
    {19, 20, 26, 25}, {20, 18, 24, 26}, {22, 23, 26, 24}, {21, 25, 26, 23}};

void or_hex_scs_areav(const double *coords, double *areav)
{
    double p[27][3], q[4][3];
    hex_subdivide(coords, p);
    for (int ics = 0; ics < 12; ++ics) {
        for (int n = 0; n < 4; ++n)
            for (int k = 0; k < 3; ++k) q[n][k] = p[hex_edge_facet[ics][n]][k];
        quad_area_by_triangles(q, &areav[3 * ics]);
    }
}

/* polyhedralVolumeByFaces (CalculatePolyhedralVolume.F) over the 24-facet
 * table of hexVolumeByTriangleFacets (CalculateHexVolume.F) */
static const int hex_vol_facets[24][3] = {
    {0, 8, 1},   {8, 2, 1},   {3, 2, 8},   {3, 8, 0},   {6, 9, 5},   {7, 9, 6},
    {4, 9, 7},   {4, 5, 9},   {10, 0, 1},  {5, 10, 1},  {4, 10, 5},  {4, 0, 10},
    {7, 6, 11},  {6, 2, 11},  {2, 3, 11},  {3, 7, 11},  {6, 12, 2},  {5, 12, 6},
    {5, 1, 12},  {1, 2, 12},  {0, 4, 13},  {4, 7, 13},  {7, 3, 13},  {3, 0, 13}};

static double hex_volume_grandy(const double v[8][3])
{
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
    double vol = 0.0;
    for (int t = 0; t < 24; ++t) {
        const double *P = c[hex_vol_facets[t][0]], *Q = c[hex_vol_facets[t][1]],
                     *R = c[hex_vol_facets[t][2]];
        double xf[3];
        for (int k = 0; k < 3; ++k) xf[k] = P[k] + Q[k] + R[k];
        vol = vol + xf[0] * ((Q[1] - P[1]) * (R[2] - P[2]) - (R[1] - P[1]) * (Q[2] - P[2])) -
              xf[1] * ((Q[0] - P[0]) * (R[2] - P[2]) - (R[0] - P[0]) * (Q[2] - P[2])) +
              xf[2] * ((Q[0] - P[0]) * (R[1] - P[1]) - (R[0] - P[0]) * (Q[1] - P[1]));
    }
    return vol / 18.0;
}

/* HexSubcontrolNodeTable (MasterElementWork.F:270-278) */
static const int hex_scv_nodes[8][8] = {
    {0, 8, 12, 11, 19, 20, 26, 25},  {8, 1, 9, 12, 20, 18, 24, 26},
    {12, 9, 2, 10, 26, 24, 22, 23},  {11, 12, 10, 3, 25, 26, 23, 21},
    {19, 20, 26, 25, 4, 13, 17, 16}, {20, 18, 24, 26, 13, 5, 14, 17},
    {26, 24, 22, 23, 17, 14, 6, 15}, {25, 26, 23, 21, 16, 17, 15, 7}};

void or_hex_scv_vol(const double *coords, double *vol)
{
    double p[27][3], v[8][3];
    hex_subdivide(coords, p);
    for (int icv = 0; icv < 8; ++icv) {
        for (int n = 0; n < 8; ++n)
            for (int k = 0; k < 3; ++k) v[n][k] = p[hex_scv_nodes[icv][n]][k];
        vol[icv] = hex_volume_grandy(v);
    }
}

void or_hex_shape(int npts, const double *par, double *shape)
{
    const double half = 0.5, one4th = 0.25, one8th = 0.125;
    for (int j = 0; j < npts; ++j) {
        const double s1 = par[3 * j], s2 = par[3 * j + 1], s3 = par[3 * j + 2];
        double *N = &shape[8 * j];
        N[0] = one8th + one4th * (-s1 - s2 - s3) + half * (s2 * s3 + s3 * s1 + s1 * s2) - s1 * s2 * s3;
        N[1] = one8th + one4th * (s1 - s2 - s3) + half * (s2 * s3 - s3 * s1 - s1 * s2) + s1 * s2 * s3;
        N[2] = one8th + one4th * (s1 + s2 - s3) + half * (-s2 * s3 - s3 * s1 + s1 * s2) - s1 * s2 * s3;
        N[3] = one8th + one4th * (-s1 + s2 - s3) + half * (-s2 * s3 + s3 * s1 - s1 * s2) + s1 * s2 * s3;
        N[4] = one8th + one4th * (-s1 - s2 + s3) + half * (-s2 * s3 - s3 * s1 + s1 * s2) + s1 * s2 * s3;
        N[5] = one8th + one4th * (s1 - s2 + s3) + half * (-s2 * s3 + s3 * s1 - s1 * s2) - s1 * s2 * s3;
        N[6] = one8th + one4th * (s1 + s2 + s3) + half * (s2 * s3 + s3 * s1 + s1 * s2) + s1 * s2 * s3;
        N[7] = one8th + one4th * (-s1 + s2 + s3) + half * (s2 * s3 - s3 * s1 - s1 * s2) - s1 * s2 * s3;
    }
}

void or_hex_deriv(int npts, const double *par, double *deriv)
{
    const double half = 0.5, one4th = 0.25;
    for (int j = 0; j < npts; ++j) {
        const double s1 = par[3 * j], s2 = par[3 * j + 1], s3 = par[3 * j + 2];
        const double s1s2 = s1 * s2, s2s3 = s2 * s3, s1s3 = s1 * s3;
        double *d = &deriv[24 * j];
#define D(n, k) d[3 * (n) + (k)]
        D(0, 0) = half * (s3 + s2) - s2s3 - one4th;
        D(1, 0) = half * (-s3 - s2) + s2s3 + one4th;
        D(2, 0) = half * (-s3 + s2) - s2s3 + one4th;
        D(3, 0) = half * (+s3 - s2) + s2s3 - one4th;
        D(4, 0) = half * (-s3 + s2) + s2s3 - one4th;
        D(5, 0) = half * (+s3 - s2) - s2s3 + one4th;
        D(6, 0) = half * (+s3 + s2) + s2s3 + one4th;
        D(7, 0) = half * (-s3 - s2) - s2s3 - one4th;

        D(0, 1) = half * (s3 + s1) - s1s3 - one4th;
        D(1, 1) = half * (s3 - s1) + s1s3 - one4th;
        D(2, 1) = half * (-s3 + s1) - s1s3 + one4th;
        D(3, 1) = half * (-s3 - s1) + s1s3 + one4th;
        D(4, 1) = half * (-s3 + s1) + s1s3 - one4th;
        D(5, 1) = half * (-s3 - s1) - s1s3 - one4th;
        D(6, 1) = half * (s3 + s1) + s1s3 + one4th;
        D(7, 1) = half * (s3 - s1) - s1s3 + one4th;

        D(0, 2) = half * (s2 + s1) - s1s2 - one4th;
        D(1, 2) = half * (s2 - s1) + s1s2 - one4th;
        D(2, 2) = half * (-s2 - s1) - s1s2 - one4th;
        D(3, 2) = half * (-s2 + s1) + s1s2 - one4th;
        D(4, 2) = half * (-s2 - s1) + s1s2 + one4th;
        D(5, 2) = half * (-s2 + s1) - s1s2 + one4th;
        D(6, 2) = half * (s2 + s1) + s1s2 + one4th;
        D(7, 2) = half * (s2 - s1) - s1s2 + one4th;
#undef D
    }
}

int or_hex_gradop(int npts, const double *deriv, const double *c, double *gradop,
                  double *detj)
{
    const double realmin = 2.2250738585072014e-308;
    int nerr = 0;
    for (int ki = 0; ki < npts; ++ki) {
        const double *d = &deriv[24 * ki];
        double dx1 = 0, dx2 = 0, dx3 = 0, dy1 = 0, dy2 = 0, dy3 = 0, dz1 = 0, dz2 = 0, dz3 = 0;
        for (int kn = 0; kn < 8; ++kn) {
            dx1 += d[3 * kn + 0] * c[3 * kn + 0];
            dx2 += d[3 * kn + 1] * c[3 * kn + 0];
            dx3 += d[3 * kn + 2] * c[3 * kn + 0];
            dy1 += d[3 * kn + 0] * c[3 * kn + 1];
            dy2 += d[3 * kn + 1] * c[3 * kn + 1];
            dy3 += d[3 * kn + 2] * c[3 * kn + 1];
            dz1 += d[3 * kn + 0] * c[3 * kn + 2];
            dz2 += d[3 * kn + 1] * c[3 * kn + 2];
            dz3 += d[3 * kn + 2] * c[3 * kn + 2];
        }
        double det = dx1 * (dy2 * dz3 - dz2 * dy3) + dy1 * (dz2 * dx3 - dx2 * dz3) +
                     dz1 * (dx2 * dy3 - dy2 * dx3);
        detj[ki] = det;
        double test = det;
        if (test <= 1.e6 * realmin) {
            test = 1.0;
            nerr++;
        }
        const double den = 1.0 / test;
        const double s1x = den * (dy2 * dz3 - dz2 * dy3);
        const double s2x = den * (dz1 * dy3 - dy1 * dz3);
        const double s3x = den * (dy1 * dz2 - dz1 * dy2);
        const double s1y = den * (dz2 * dx3 - dx2 * dz3);
        const double s2y = den * (dx1 * dz3 - dz1 * dx3);
        const double s3y = den * (dz1 * dx2 - dx1 * dz2);
        const double s1z = den * (dx2 * dy3 - dy2 * dx3);
        const double s2z = den * (dy1 * dx3 - dx1 * dy3);
        const double s3z = den * (dx1 * dy2 - dy1 * dx2);
        double *g = &gradop[24 * ki];
        for (int kn = 0; kn < 8; ++kn) {
            g[3 * kn + 0] = d[3 * kn] * s1x + d[3 * kn + 1] * s2x + d[3 * kn + 2] * s3x;
            g[3 * kn + 1] = d[3 * kn] * s1y + d[3 * kn + 1] * s2y + d[3 * kn + 2] * s3y;
            g[3 * kn + 2] = d[3 * kn] * s1z + d[3 * kn + 1] * s2z + d[3 * kn + 2] * s3z;
        }
    }
    return nerr;
}

int or_hex_scs_gradop(int shifted, const double *coords, double *gradop, double *detj)
{
    double deriv[12 * 24];
    or_hex_deriv(12, shifted ? hex_scs_loc_shift : hex_scs_loc, deriv);
    return or_hex_gradop(12, deriv, coords, gradop, detj);
}

int or_hex_face_gradop(int shifted, int ord, const double *coords, double *gradop,
                       double *detj)
{
    double loc[12];
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d)
            loc[3 * k + d] = shifted ? hex_node_loc[3 * hex_side_nodes[4 * ord + k] + d]
                                     : hex_face_loc[12 * ord + 3 * k + d];
    double deriv[4 * 24];
    or_hex_deriv(4, loc, deriv);
    return or_hex_gradop(4, deriv, coords, gradop, detj);
}

void or_hex_scs_shape(int shifted, double *shape)
{
    or_hex_shape(12, shifted ? hex_scs_loc_shift : hex_scs_loc, shape);
}

void or_quad_face_areav(const double *q, double *area)
{
    /* quad3d_scs_det (MasterElementWork.F:1467-1588) */
    double p[4][3], e[4][3], c[3];
    for (int i = 0; i < 4; ++i)
        for (int d = 0; d < 3; ++d) p[i][d] = q[3 * i + d];
    for (int d = 0; d < 3; ++d) {
        e[0][d] = (p[0][d] + p[1][d]) * 0.5;
        e[1][d] = (p[1][d] + p[2][d]) * 0.5;
        e[2][d] = (p[2][d] + p[3][d]) * 0.5;
        e[3][d] = (p[3][d] + p[0][d]) * 0.5;
        c[d] = (p[0][d] + p[1][d] + p[2][d] + p[3][d]) * 0.25;
    }
    /* sub-face i spans corner i, edge midpoints e(i-1) and e(i), centre */
    static const int ea[4] = {3, 0, 1, 2}, eb[4] = {0, 1, 2, 3};
    for (int i = 0; i < 4; ++i) {
        const double dx13 = c[0] - p[i][0], dy13 = c[1] - p[i][1], dz13 = c[2] - p[i][2];
        const double dx24 = e[ea[i]][0] - e[eb[i]][0];
        const double dy24 = e[ea[i]][1] - e[eb[i]][1];
        const double dz24 = e[ea[i]][2] - e[eb[i]][2];
        area[3 * i + 0] = 0.5 * (dz24 * dy13 - dz13 * dy24);
        area[3 * i + 1] = 0.5 * (dx24 * dz13 - dx13 * dz24);
        area[3 * i + 2] = 0.5 * (dy24 * dx13 - dy13 * dx24);
    }
}

void or_quad_face_shape(int shifted, double *shape)
{
    /* quad3d_shape_fcn at Quad3DSCS intgLoc_ / intgLocShift_ */
    static const double loc[8] = {-0.25, -0.25, 0.25, -0.25, 0.25, 0.25, -0.25, 0.25};
    static const double locs[8] = {-0.5, -0.5, 0.5, -0.5, 0.5, 0.5, -0.5, 0.5};
    const double *l = shifted ? locs : loc;
    for (int j = 0; j < 4; ++j) {
        const double s1 = l[2 * j], s2 = l[2 * j + 1];
        shape[4 * j + 0] = 0.25 + 0.5 * (-s1 - s2) + s1 * s2;
        shape[4 * j + 1] = 0.25 + 0.5 * (s1 - s2) - s1 * s2;
        shape[4 * j + 2] = 0.25 + 0.5 * (s1 + s2) + s1 * s2;
        shape[4 * j + 3] = 0.25 + 0.5 * (-s1 + s2) - s1 * s2;
    }
}
