/*
 * topo.c -- oracle restatement of the tet4 / wedge6 / pyramid5 CVFEM master
 * elements.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restated (external/nalu/src/master_element/):
 *   SCS area vectors  MasterElementWork.F tet_scs_det :608-753,
 *                     pyr_scs_det :754-930, wed_scs_det :1124-1293, each SCS
 *                     quad by quadAreaByTriangleFacets (CalculateQuadArea.F)
 *   SCV volumes       tet_scv_det :459-607, wed_scv_det :1294-1466
 *                     (hexVolumeByTriangleFacets, CalculateHexVolume.F),
 *                     pyr_scv_det :931-1123 (bhexVolumeByTriangleFacets,
 *                     CalculateBentHexVolume.F; octohedronVolumeByTriangleFacets,
 *                     CalculateOctohedronVolume.F), all through
 *                     polyhedralVolumeByFaces (CalculatePolyhedralVolume.F)
 *   gradient operator tet_gradient_operator :3061, pyr_gradient_operator
 *                     :3345, wed_gradient_operator :3498 (one algorithm)
 *   derivatives       tet_derivative (MasterElementWork.F:2539),
 *                     wedge_derivative (Wed6CVFEM.C:59-86),
 *                     pyr_derivative / shifted_pyr_derivative (Pyr5CVFEM.C:119-185)
 *   SCS ip locations  Tet4CVFEM.C:368-382, Wed6CVFEM.C:436-456 (WedSCS),
 *                     Pyr5CVFEM.C PyrSCS constructor (intgLoc_, intgLocShift_)
 *   L/R node tables   lrscv_ of TetSCS, WedSCS, PyrSCS
 *   sides             sideNodeOrdinals_ of TetSCS (Tet4CVFEM.C:416-421),
 *                     WedSCS (Wed6CVFEM.C:494-501), PyrSCS (Pyr5CVFEM.C:759-765);
 *                     exposed-face ips intgExpFace_ (Tet4CVFEM.C:384-403,
 *                     Wed6CVFEM.C:458-480, Pyr5CVFEM.C:737-757) and
 *                     intgExpFaceShift_ (the side nodes); face_grad_op /
 *                     shifted_face_grad_op (Tet4CVFEM.C:668-745,
 *                     Wed6CVFEM.C:754-860, Pyr5CVFEM.C:1035-1160)
 *   tri3 faces        tri3d_scs_det (MasterElementWork.F:1589-1697),
 *                     Tri3DSCS shape_fcn / shifted_shape_fcn (MasterElement.C:610-715)
 *
 * Subdivision points are written as vertex weights: each is the average of
 * the listed vertices (the element centroid averages all of them).
 */
#include "oracle.h"
#include <float.h>
#include <math.h>
#include <string.h>

typedef struct {
    int nv;       /* vertices averaged (0: the point is vertex `v[0]`) */
    int v[6];
} subpt;

/* -------------------------------------------------------------- tet4 */
static const subpt tet_pts[15] = {
    {0, {0}}, {0, {1}}, {0, {2}}, {0, {3}},
    {2, {0, 1}}, {2, {1, 2}}, {2, {2, 0}}, {3, {0, 1, 2}},
    {2, {2, 3}}, {2, {3, 1}}, {3, {1, 2, 3}},
    {2, {0, 3}}, {3, {0, 2, 3}}, {3, {0, 1, 3}},
    {4, {0, 1, 2, 3}}};
static const int tet_scs[6][4] = {{4, 7, 14, 13}, {7, 14, 10, 5}, {6, 12, 14, 7},
                                  {11, 13, 14, 12}, {13, 9, 10, 14}, {10, 8, 12, 14}};
static const int tet_scv[4][8] = {{0, 4, 7, 6, 11, 13, 14, 12}, {1, 5, 7, 4, 9, 10, 14, 13},
                                  {2, 6, 7, 5, 8, 12, 14, 10}, {3, 9, 13, 11, 8, 10, 14, 12}};
static const int tet_lr[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
static const double tet_ip[6][3] = {
    {13.0 / 36, 5.0 / 36, 5.0 / 36},   {13.0 / 36, 13.0 / 36, 5.0 / 36},
    {5.0 / 36, 13.0 / 36, 5.0 / 36},   {5.0 / 36, 5.0 / 36, 13.0 / 36},
    {13.0 / 36, 5.0 / 36, 13.0 / 36},  {5.0 / 36, 13.0 / 36, 13.0 / 36}};
static const double tet_ip_shift[6][3] = {{0.5, 0, 0}, {0.5, 0.5, 0}, {0, 0.5, 0},
                                          {0, 0, 0.5}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};

/* ------------------------------------------------------------ wedge6 */
static const subpt wed_pts[21] = {
    {0, {0}}, {0, {1}}, {0, {2}}, {0, {3}}, {0, {4}}, {0, {5}},
    {2, {0, 1}}, {2, {1, 2}}, {2, {2, 0}}, {3, {0, 1, 2}},
    {2, {3, 4}}, {2, {4, 5}}, {2, {5, 3}}, {3, {3, 4, 5}},
    {2, {1, 4}}, {2, {0, 3}}, {4, {0, 1, 4, 3}}, {2, {2, 5}},
    {4, {1, 4, 5, 2}}, {4, {5, 3, 0, 2}}, {6, {0, 1, 2, 3, 4, 5}}};
static const int wed_scs[9][4] = {{6, 9, 20, 16},  {7, 9, 20, 18},  {9, 8, 19, 20},
                                  {10, 16, 20, 13}, {13, 11, 18, 20}, {12, 13, 20, 19},
                                  {15, 16, 20, 19}, {16, 14, 18, 20}, {19, 20, 18, 17}};
static const int wed_scv[6][8] = {{0, 15, 16, 6, 8, 19, 20, 9},    {9, 6, 1, 7, 20, 16, 14, 18},
                                  {8, 9, 7, 2, 19, 20, 18, 17},    {19, 15, 16, 20, 12, 3, 10, 13},
                                  {20, 16, 14, 18, 13, 10, 4, 11}, {19, 20, 18, 17, 12, 13, 11, 5}};
static const int wed_lr[9][2] = {{0, 1}, {1, 2}, {0, 2}, {3, 4}, {4, 5},
                                 {3, 5}, {0, 3}, {1, 4}, {2, 5}};
static const double wed_ip[9][3] = {
    {5.0 / 12, 1.0 / 6, -0.5},        {5.0 / 12, 5.0 / 12, -0.5},  {1.0 / 6, 5.0 / 12, -0.5},
    {5.0 / 12, 1.0 / 6, 0.5},         {5.0 / 12, 5.0 / 12, 0.5},   {1.0 / 6, 5.0 / 12, 0.5},
    {7.0 / 36, 7.0 / 36, 0.0},        {11.0 / 18, 7.0 / 36, 0.0},  {7.0 / 36, 11.0 / 18, 0.0}};
static const double wed_ip_shift[9][3] = {{0.5, 0, -1}, {0.5, 0.5, -1}, {0, 0.5, -1},
                                          {0.5, 0, 1},  {0.5, 0.5, 1},  {0, 0.5, 1},
                                          {0, 0, 0},    {1, 0, 0},      {0, 1, 0}};

/* ---------------------------------------------------------- pyramid5 */
static const subpt pyr_pts[19] = {
    {0, {0}}, {0, {1}}, {0, {2}}, {0, {3}}, {0, {4}},
    {2, {0, 1}}, {2, {1, 2}}, {2, {2, 3}}, {2, {3, 0}}, {4, {0, 1, 2, 3}},
    {2, {1, 4}}, {2, {4, 0}}, {3, {0, 1, 4}}, {2, {2, 4}}, {3, {1, 2, 4}},
    {2, {3, 4}}, {3, {3, 4, 2}}, {3, {0, 4, 3}}, {5, {0, 1, 2, 3, 4}}};
static const int pyr_scs[12][4] = {{5, 9, 18, 12},   {6, 9, 18, 14},   {7, 9, 18, 16},
                                   {8, 17, 18, 9},   {12, 12, 18, 17}, {11, 12, 12, 17},
                                   {14, 14, 18, 12}, {10, 14, 14, 12}, {16, 16, 18, 14},
                                   {13, 16, 16, 14}, {17, 17, 18, 16}, {15, 17, 17, 16}};
static const int pyr_scv_hex[4][8] = {{0, 5, 9, 8, 11, 12, 18, 17}, {1, 6, 9, 5, 10, 14, 18, 12},
                                      {2, 7, 9, 6, 13, 16, 18, 14}, {3, 8, 9, 7, 15, 17, 18, 16}};
static const int pyr_scv_apex[10] = {4, 18, 15, 17, 11, 12, 10, 14, 13, 16};
static const int pyr_lr[12][2] = {{0, 1}, {1, 2}, {2, 3}, {0, 3}, {0, 4}, {0, 4},
                                  {1, 4}, {1, 4}, {2, 4}, {2, 4}, {3, 4}, {3, 4}};
#define T29_63 (29.0 / 63.0)
#define F41_315 (41.0 / 315.0)
#define T2_9 (2.0 / 9.0)
#define T13_45 (13.0 / 45.0)
#define S7_18 (7.0 / 18.0)
static const double pyr_ip[12][3] = {
    {0.0, -T29_63, F41_315}, {T29_63, 0.0, F41_315}, {0.0, T29_63, F41_315},
    {-T29_63, 0.0, F41_315}, {-T2_9, -T2_9, T13_45}, {-S7_18, -S7_18, S7_18},
    {T2_9, -T2_9, T13_45},   {S7_18, -S7_18, S7_18}, {T2_9, T2_9, T13_45},
    {S7_18, S7_18, S7_18},   {-T2_9, T2_9, T13_45},  {-S7_18, S7_18, S7_18}};
static const double pyr_ip_shift[12][3] = {
    {0, -1, 0},       {1, 0, 0},        {0, 1, 0},        {-1, 0, 0},
    {-0.5, -0.5, 0.5}, {-0.5, -0.5, 0.5}, {0.5, -0.5, 0.5}, {0.5, -0.5, 0.5},
    {0.5, 0.5, 0.5},  {0.5, 0.5, 0.5},  {-0.5, 0.5, 0.5}, {-0.5, 0.5, 0.5}};

/* ------------------------------------------------------------ common */
static void subdivide(const subpt *pts, int np, int npe, const double *c, double (*p)[3])
{
    for (int i = 0; i < np; ++i) {
        const subpt *s = &pts[i];
        for (int k = 0; k < 3; ++k) {
            if (s->nv == 0) {
                p[i][k] = c[3 * s->v[0] + k];
                continue;
            }
            /* the Fortran forms w * (sum), w = 1/2, 1/3 (one3rd), 1/4, 1/6 (one6th) or
             * 0.2; the element centroid accumulates w * c_j one vertex at a time */
            const double w = s->nv == 2 ? 0.5 : s->nv == 3 ? 1.0 / 3.0 : s->nv == 4 ? 0.25
                           : s->nv == 5 ? 0.2 : 1.0 / 6.0;
            if (s->nv == npe) {
                double acc = 0.0;
                for (int j = 0; j < npe; ++j) acc = acc + w * c[3 * j + k];
                p[i][k] = acc;
            } else {
                double acc = 0.0;
                for (int j = 0; j < s->nv; ++j) acc = acc + c[3 * s->v[j] + k];
                p[i][k] = w * acc;
            }
        }
    }
}

/* quadAreaByTriangleFacets (CalculateQuadArea.F) */
static void quad_area(const double q[4][3], double *area)
{
    double xm[3];
    for (int k = 0; k < 3; ++k) xm[k] = 0.25 * (q[0][k] + q[1][k] + q[2][k] + q[3][k]);
    area[0] = area[1] = area[2] = 0.0;
    static const int iq[4] = {1, 2, 3, 0};
    for (int t = 0; t < 4; ++t) {
        const double *p = q[t], *r = q[iq[t]];
        const double r1[3] = {p[0] - xm[0], p[1] - xm[1], p[2] - xm[2]};
        const double r2[3] = {r[0] - xm[0], r[1] - xm[1], r[2] - xm[2]};
        area[0] += r1[1] * r2[2] - r2[1] * r1[2];
        area[1] += r1[2] * r2[0] - r2[2] * r1[0];
        area[2] += r1[0] * r2[1] - r2[0] * r1[1];
    }
    for (int k = 0; k < 3; ++k) area[k] *= 0.5;
}

/* polyhedralVolumeByFaces (CalculatePolyhedralVolume.F) */
static double poly_volume(const double (*c)[3], int nt, const int (*tri)[3])
{
    double vol = 0.0;
    for (int t = 0; t < nt; ++t) {
        const double *P = c[tri[t][0]], *Q = c[tri[t][1]], *R = c[tri[t][2]];
        double xf[3];
        for (int k = 0; k < 3; ++k) xf[k] = P[k] + Q[k] + R[k];
        vol = vol + xf[0] * ((Q[1] - P[1]) * (R[2] - P[2]) - (R[1] - P[1]) * (Q[2] - P[2])) -
              xf[1] * ((Q[0] - P[0]) * (R[2] - P[2]) - (R[0] - P[0]) * (Q[2] - P[2])) +
              xf[2] * ((Q[0] - P[0]) * (R[1] - P[1]) - (R[0] - P[0]) * (Q[1] - P[1]));
    }
    return vol / 18.0;
}

/* the 24 triangles of hexVolumeByTriangleFacets and bhexVolumeByTriangleFacets */
static const int hex_tris[24][3] = {
    {0, 8, 1},   {8, 2, 1},   {3, 2, 8},   {3, 8, 0},   {6, 9, 5},   {7, 9, 6},
    {4, 9, 7},   {4, 5, 9},   {10, 0, 1},  {5, 10, 1},  {4, 10, 5},  {4, 0, 10},
    {7, 6, 11},  {6, 2, 11},  {2, 3, 11},  {3, 7, 11},  {6, 12, 2},  {5, 12, 6},
    {5, 1, 12},  {1, 2, 12},  {0, 4, 13},  {4, 7, 13},  {7, 3, 13},  {3, 0, 13}};
/* octohedronVolumeByTriangleFacets */
static const int oct_tris[24][3] = {
    {1, 3, 10},  {2, 10, 3},  {2, 9, 10},  {10, 9, 1},  {4, 3, 11},  {3, 1, 11},
    {11, 1, 5},  {4, 11, 5},  {1, 12, 5},  {1, 7, 12},  {12, 7, 6},  {5, 12, 6},
    {9, 8, 13},  {13, 8, 7},  {13, 7, 1},  {9, 13, 1},  {4, 5, 0},   {5, 6, 0},
    {6, 7, 0},   {7, 8, 0},   {0, 8, 9},   {0, 9, 2},   {0, 2, 3},   {0, 3, 4}};

/* bent: bhexVolumeByTriangleFacets (top "face centre" = midpoint of 6 and 8) */
static double hex_volume(const double v[8][3], int bent)
{
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
    return poly_volume((const double(*)[3])c, 24, hex_tris);
}

static double oct_volume(const double v[10][3])
{
    double c[14][3];
    for (int j = 0; j < 10; ++j)
        for (int k = 0; k < 3; ++k) c[j][k] = v[j][k];
    for (int k = 0; k < 3; ++k) {
        c[10][k] = 0.5 * (v[3][k] + v[9][k]);
        c[11][k] = 0.5 * (v[3][k] + v[5][k]);
        c[12][k] = 0.5 * (v[5][k] + v[7][k]);
        c[13][k] = 0.5 * (v[7][k] + v[9][k]);
    }
    return poly_volume((const double(*)[3])c, 24, oct_tris);
}

int or_topo_npe(int topo) { return topo == OR_TET4 ? 4 : topo == OR_WED6 ? 6 : topo == OR_PYR5 ? 5 : 8; }
int or_topo_nscs(int topo) { return topo == OR_TET4 ? 6 : topo == OR_WED6 ? 9 : 12; }
const int *or_topo_lrscv(int topo)
{
    return topo == OR_TET4 ? &tet_lr[0][0] : topo == OR_WED6 ? &wed_lr[0][0]
         : topo == OR_PYR5 ? &pyr_lr[0][0] : or_hex_lrscv();
}

void or_topo_scs_areav(int topo, const double *coords, double *areav)
{
    if (topo == OR_HEX8) {
        or_hex_scs_areav(coords, areav);
        return;
    }
    double p[21][3], q[4][3];
    const int(*tab)[4];
    int n;
    if (topo == OR_TET4) {
        subdivide(tet_pts, 15, 4, coords, p);
        tab = tet_scs;
        n = 6;
    } else if (topo == OR_WED6) {
        subdivide(wed_pts, 21, 6, coords, p);
        tab = wed_scs;
        n = 9;
    } else {
        subdivide(pyr_pts, 19, 5, coords, p);
        tab = pyr_scs;
        n = 12;
    }
    for (int ics = 0; ics < n; ++ics) {
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 3; ++k) q[j][k] = p[tab[ics][j]][k];
        quad_area(q, &areav[3 * ics]);
    }
}

void or_topo_scv_vol(int topo, const double *coords, double *vol)
{
    if (topo == OR_HEX8) {
        or_hex_scv_vol(coords, vol);
        return;
    }
    double p[21][3], v[10][3];
    if (topo == OR_TET4 || topo == OR_WED6) {
        const int tet = topo == OR_TET4;
        subdivide(tet ? tet_pts : wed_pts, tet ? 15 : 21, tet ? 4 : 6, coords, p);
        const int n = tet ? 4 : 6;
        for (int icv = 0; icv < n; ++icv) {
            const int *t = tet ? tet_scv[icv] : wed_scv[icv];
            for (int j = 0; j < 8; ++j)
                for (int k = 0; k < 3; ++k) v[j][k] = p[t[j]][k];
            vol[icv] = hex_volume((const double(*)[3])v, 0);
        }
        return;
    }
    subdivide(pyr_pts, 19, 5, coords, p);
    for (int icv = 0; icv < 4; ++icv) {
        for (int j = 0; j < 8; ++j)
            for (int k = 0; k < 3; ++k) v[j][k] = p[pyr_scv_hex[icv][j]][k];
        vol[icv] = hex_volume((const double(*)[3])v, 1);
    }
    for (int j = 0; j < 10; ++j)
        for (int k = 0; k < 3; ++k) v[j][k] = p[pyr_scv_apex[j]][k];
    vol[4] = oct_volume((const double(*)[3])v);
}

/* shape-function derivatives [npts][npe][3] at parametric points */
void or_topo_deriv(int topo, int shifted, int npts, const double *par, double *deriv)
{
    const double eps = DBL_EPSILON;
    for (int j = 0; j < npts; ++j) {
        const double r = par[3 * j], s = par[3 * j + 1], z = par[3 * j + 2];
        double *d = &deriv[(size_t)j * 3 * or_topo_npe(topo)];
        if (topo == OR_TET4) {
            const double t4[12] = {-1, -1, -1, 1, 0, 0, 0, 1, 0, 0, 0, 1};
            memcpy(d, t4, sizeof(t4));
        } else if (topo == OR_WED6) {
            const double t = 1.0 - r - s, xi = z;
            const double w[18] = {-0.5 * (1.0 - xi), -0.5 * (1.0 - xi), -0.5 * t,
                                  0.5 * (1.0 - xi),  0.0,               -0.5 * r,
                                  0.0,               0.5 * (1.0 - xi),  -0.5 * s,
                                  -0.5 * (1.0 + xi), -0.5 * (1.0 + xi), 0.5 * t,
                                  0.5 * (1.0 + xi),  0.0,               0.5 * r,
                                  0.0,               0.5 * (1.0 + xi),  0.5 * s};
            memcpy(d, w, sizeof(w));
        } else if (topo == OR_PYR5 && !shifted) {
            const double omt = 1.0 - z;
            const double t = fabs(omt) > eps ? z : 1.0 + copysign(eps, omt);
            const double q = 0.25 / (1.0 - t);
            const double tt = 4.0 * r * s * q * q;
            const double p5[15] = {-(1.0 - s - t) * q, -(1.0 - r - t) * q, (+tt - 0.25),
                                   +(1.0 - s - t) * q, -(1.0 + r - t) * q, (-tt - 0.25),
                                   +(1.0 + s - t) * q, +(1.0 + r - t) * q, (+tt - 0.25),
                                   -(1.0 + s - t) * q, +(1.0 - r - t) * q, (-tt - 0.25),
                                   0.0,                0.0,                1.0};
            memcpy(d, p5, sizeof(p5));
        } else { /* shifted_pyr_derivative */
            const double t = z;
            const double p5[15] = {-0.25 * (1.0 - s) * (1.0 - t), -0.25 * (1.0 - r) * (1.0 - t),
                                   -0.25 * (1.0 - r) * (1.0 - s), 0.25 * (1.0 - s) * (1.0 - t),
                                   -0.25 * (1.0 + r) * (1.0 - t), -0.25 * (1.0 + r) * (1.0 - s),
                                   0.25 * (1.0 + s) * (1.0 - t),  0.25 * (1.0 + r) * (1.0 - t),
                                   -0.25 * (1.0 + r) * (1.0 + s), -0.25 * (1.0 + s) * (1.0 - t),
                                   0.25 * (1.0 - r) * (1.0 - t),  -0.25 * (1.0 - r) * (1.0 + s),
                                   0.0,                           0.0,
                                   1.0};
            memcpy(d, p5, sizeof(p5));
        }
    }
}

/* *_gradient_operator: one algorithm for all topologies */
int or_topo_gradop(int npe, int npts, const double *deriv, const double *c, double *gradop,
                   double *detj)
{
    const double realmin = 2.2250738585072014e-308;
    int nerr = 0;
    for (int ki = 0; ki < npts; ++ki) {
        const double *d = &deriv[(size_t)3 * npe * ki];
        double dx1 = 0, dx2 = 0, dx3 = 0, dy1 = 0, dy2 = 0, dy3 = 0, dz1 = 0, dz2 = 0, dz3 = 0;
        for (int kn = 0; kn < npe; ++kn) {
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
        const double det = dx1 * (dy2 * dz3 - dz2 * dy3) + dy1 * (dz2 * dx3 - dx2 * dz3) +
                           dz1 * (dx2 * dy3 - dy2 * dx3);
        detj[ki] = det;
        double test = det;
        if (test <= 1.e6 * realmin) {
            test = 1.0;
            nerr++;
        }
        const double den = 1.0 / test;
        const double s1x = den * (dy2 * dz3 - dz2 * dy3), s2x = den * (dz1 * dy3 - dy1 * dz3),
                     s3x = den * (dy1 * dz2 - dz1 * dy2);
        const double s1y = den * (dz2 * dx3 - dx2 * dz3), s2y = den * (dx1 * dz3 - dz1 * dx3),
                     s3y = den * (dz1 * dx2 - dx1 * dz2);
        const double s1z = den * (dx2 * dy3 - dy2 * dx3), s2z = den * (dy1 * dx3 - dx1 * dy3),
                     s3z = den * (dx1 * dy2 - dy1 * dx2);
        double *g = &gradop[(size_t)3 * npe * ki];
        for (int kn = 0; kn < npe; ++kn) {
            g[3 * kn + 0] = d[3 * kn] * s1x + d[3 * kn + 1] * s2x + d[3 * kn + 2] * s3x;
            g[3 * kn + 1] = d[3 * kn] * s1y + d[3 * kn + 1] * s2y + d[3 * kn + 2] * s3y;
            g[3 * kn + 2] = d[3 * kn] * s1z + d[3 * kn + 1] * s2z + d[3 * kn + 2] * s3z;
        }
    }
    return nerr;
}

const double *or_topo_scs_ips(int topo, int shifted)
{
    if (topo == OR_TET4) return shifted ? &tet_ip_shift[0][0] : &tet_ip[0][0];
    if (topo == OR_WED6) return shifted ? &wed_ip_shift[0][0] : &wed_ip[0][0];
    return shifted ? &pyr_ip_shift[0][0] : &pyr_ip[0][0];
}

int or_topo_scs_gradop(int topo, int shifted, const double *coords, double *gradop, double *detj)
{
    if (topo == OR_HEX8) return or_hex_scs_gradop(shifted, coords, gradop, detj);
    const int npe = or_topo_npe(topo), n = or_topo_nscs(topo);
    double deriv[12 * 8 * 3];
    or_topo_deriv(topo, shifted, n, or_topo_scs_ips(topo, shifted), deriv);
    return or_topo_gradop(npe, n, deriv, coords, gradop, detj);
}

/* ------------------------------------------------------------- sides */
/* side-node ordinals (-1: a tri side's unused fourth entry) and the
 * vertices' parametric locations */
static const int tet_sides[4][4] = {{0, 1, 3, -1}, {1, 2, 3, -1}, {0, 3, 2, -1}, {0, 2, 1, -1}};
static const int wed_sides[5][4] = {{0, 1, 4, 3}, {1, 2, 5, 4}, {0, 3, 5, 2}, {0, 2, 1, -1},
                                    {3, 4, 5, -1}};
static const int pyr_sides[5][4] = {{0, 1, 4, -1}, {1, 2, 4, -1}, {2, 3, 4, -1}, {0, 4, 3, -1},
                                    {0, 3, 2, 1}};
static const double tet_vpar[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
static const double wed_vpar[6][3] = {{0, 0, -1}, {1, 0, -1}, {0, 1, -1},
                                      {0, 0, 1},  {1, 0, 1},  {0, 1, 1}};
static const double pyr_vpar[5][3] = {{-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}, {0, 0, 1}};

int or_topo_nsides(int topo)
{
    return topo == OR_HEX8 ? 6 : topo == OR_TET4 ? 4 : 5;
}

int or_topo_side_nodes(int topo, int ord, int *nodes)
{
    const int *sn;
    if (topo == OR_HEX8) {
        sn = or_hex_side_nodes() + 4 * ord;
    } else {
        sn = topo == OR_TET4 ? tet_sides[ord] : topo == OR_WED6 ? wed_sides[ord] : pyr_sides[ord];
    }
    int npf = 0;
    for (int k = 0; k < 4; ++k) {
        nodes[k] = sn[k];
        if (sn[k] >= 0) npf++;
    }
    return npf;
}

/* exposed-face ip k of a side, in the element's parametric coordinates.
 * Standard: the centre of face node k's subface -- on a tri side the face
 * point (7/36, 7/36) from node k, i.e. weights 11/18 on node k and 7/36 on
 * the other two; on a quad side the bilinear point (+-1/2, +-1/2), weights
 * 9/16 (node k), 3/16 (its two neighbours), 1/16 (the opposite node).
 * Shifted: face node k itself.  These reproduce intgExpFace_ entry by entry. */
void or_topo_side_ips(int topo, int ord, int shifted, double *par)
{
    int sn[4];
    const int npf = or_topo_side_nodes(topo, ord, sn);
    const double(*vp)[3] = topo == OR_TET4 ? tet_vpar : topo == OR_WED6 ? wed_vpar : pyr_vpar;
    for (int k = 0; k < npf; ++k)
        for (int d = 0; d < 3; ++d) {
            double x = 0.0;
            if (shifted) {
                x = vp[sn[k]][d];
            } else if (npf == 3) {
                for (int q = 0; q < 3; ++q) x += (q == k ? 11.0 / 18.0 : 7.0 / 36.0) * vp[sn[q]][d];
            } else {
                const double w[4] = {9.0 / 16.0, 3.0 / 16.0, 1.0 / 16.0, 3.0 / 16.0};
                for (int q = 0; q < 4; ++q) x += w[(q - k + 4) % 4] * vp[sn[q]][d];
            }
            par[3 * k + d] = x;
        }
}

/* face_grad_op / shifted_face_grad_op: the element's gradient operator at
 * the side's face ips, [npf][npe][3] */
int or_topo_side_gradop(int topo, int shifted, int ord, const double *coords, double *gradop,
                        double *detj)
{
    if (topo == OR_HEX8) return or_hex_face_gradop(shifted, ord, coords, gradop, detj);
    int sn[4];
    const int npf = or_topo_side_nodes(topo, ord, sn), npe = or_topo_npe(topo);
    double par[12], deriv[4 * 8 * 3];
    or_topo_side_ips(topo, ord, shifted, par);
    or_topo_deriv(topo, shifted, npf, par, deriv);
    return or_topo_gradop(npe, npf, deriv, coords, gradop, detj);
}

/* tri3d_scs_det: the subface of node k is the quad (node k, mid-edge k,
 * centroid, mid-edge k-1); its area vector is half the cross product of the
 * diagonals (centroid - node k) x (mid-edge k-1 - mid-edge k) */
void or_tri_face_areav(const double *p, double *area)
{
    double e[3][3], c[3];
    for (int d = 0; d < 3; ++d) {
        for (int k = 0; k < 3; ++k) e[k][d] = 0.5 * (p[3 * k + d] + p[3 * ((k + 1) % 3) + d]);
        c[d] = (p[d] + p[3 + d] + p[6 + d]) / 3.0;
    }
    for (int k = 0; k < 3; ++k) {
        const int km = (k + 2) % 3;
        const double a[3] = {c[0] - p[3 * k], c[1] - p[3 * k + 1], c[2] - p[3 * k + 2]};
        const double b[3] = {e[km][0] - e[k][0], e[km][1] - e[k][1], e[km][2] - e[k][2]};
        area[3 * k + 0] = 0.5 * (a[1] * b[2] - a[2] * b[1]);
        area[3 * k + 1] = 0.5 * (a[2] * b[0] - a[0] * b[2]);
        area[3 * k + 2] = 0.5 * (a[0] * b[1] - a[1] * b[0]);
    }
}

/* Tri3DSCS shape functions [3][3]: (1 - xi - eta, xi, eta) at (7/36, 7/36),
 * (11/18, 7/36), (7/36, 11/18); shifted: at the nodes (the identity) */
void or_tri_face_shape(int shifted, double *shape)
{
    const double loc[3][2] = {{7.0 / 36.0, 7.0 / 36.0}, {11.0 / 18.0, 7.0 / 36.0},
                              {7.0 / 36.0, 11.0 / 18.0}};
    for (int k = 0; k < 3; ++k) {
        const double xi = shifted ? (k == 1) : loc[k][0], eta = shifted ? (k == 2) : loc[k][1];
        shape[3 * k + 0] = 1.0 - xi - eta;
        shape[3 * k + 1] = xi;
        shape[3 * k + 2] = eta;
    }
}
