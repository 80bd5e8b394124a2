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
#include "sides.hpp"
#include "topo_tables.hpp"

#include <mutex>

namespace oagpu {

namespace topo {

// SCV subdivisions (tet_scv_det, wed_scv_det, pyr_scv_det)
__constant__ int c_tet_scv[4][8] = {{0, 4, 7, 6, 11, 13, 14, 12}, {1, 5, 7, 4, 9, 10, 14, 13},
                                    {2, 6, 7, 5, 8, 12, 14, 10}, {3, 9, 13, 11, 8, 10, 14, 12}};
__constant__ int c_wed_scv[6][8] = {{0, 15, 16, 6, 8, 19, 20, 9},    {9, 6, 1, 7, 20, 16, 14, 18},
                                    {8, 9, 7, 2, 19, 20, 18, 17},    {19, 15, 16, 20, 12, 3, 10, 13},
                                    {20, 16, 14, 18, 13, 10, 4, 11}, {19, 20, 18, 17, 12, 13, 11, 5}};
__constant__ int c_pyr_scv[4][8] = {{0, 5, 9, 8, 11, 12, 18, 17}, {1, 6, 9, 5, 10, 14, 18, 12},
                                    {2, 7, 9, 6, 13, 16, 18, 14}, {3, 8, 9, 7, 15, 17, 18, 16}};
__constant__ int c_pyr_apex[10] = {4, 18, 15, 17, 11, 12, 10, 14, 13, 16};

// the 24 triangles of the hex / bent-hex and octohedron volume routines
__constant__ int c_hex_tris[24][3] = {
    {0, 8, 1},  {8, 2, 1},  {3, 2, 8},  {3, 8, 0},  {6, 9, 5},  {7, 9, 6},  {4, 9, 7},  {4, 5, 9},
    {10, 0, 1}, {5, 10, 1}, {4, 10, 5}, {4, 0, 10}, {7, 6, 11}, {6, 2, 11}, {2, 3, 11}, {3, 7, 11},
    {6, 12, 2}, {5, 12, 6}, {5, 1, 12}, {1, 2, 12}, {0, 4, 13}, {4, 7, 13}, {7, 3, 13}, {3, 0, 13}};
__constant__ int c_oct_tris[24][3] = {
    {1, 3, 10}, {2, 10, 3}, {2, 9, 10}, {10, 9, 1}, {4, 3, 11}, {3, 1, 11}, {11, 1, 5}, {4, 11, 5},
    {1, 12, 5}, {1, 7, 12}, {12, 7, 6}, {5, 12, 6}, {9, 8, 13}, {13, 8, 7}, {13, 7, 1}, {9, 13, 1},
    {4, 5, 0},  {5, 6, 0},  {6, 7, 0},  {7, 8, 0},  {0, 8, 9},  {0, 9, 2},  {0, 2, 3},  {0, 3, 4}};

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

// SCV volumes of the tet4 / wedge6 / pyramid5 elements of a mixed mesh,
// raw into scvw [e][0..npe) (one lane per listed element of topology TOPO)
template <int TOPO>
__global__ void __launch_bounds__(256) k_scv_generic(int n, const int* glist, const int* etype,
                                                     const int* conn, const double* xyz,
                                                     double* scvw) {
    using Tr = topo::Traits<TOPO>;
    const int i = int(xcd_block() * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int e = glist[i];
    if (etype[e] != TOPO) return;
    double c[Tr::npe][3];
    for (int j = 0; j < Tr::npe; ++j)
        for (int k = 0; k < 3; ++k) c[j][k] = xyz[3 * size_t(conn[8 * size_t(e) + j]) + k];
    double p[Tr::npts][3];
    topo::subdivide<TOPO>(c, p);
    topo::scv_volumes<TOPO>(p, scvw + 8 * size_t(e));
}

// static ip geometry of the listed elements of topology TOPO (coordinates
// relative to the element's node 0): area vector, dN/dx at the standard and
// at the shifted ip; a non-positive Jacobian sets *err
template <int TOPO>
__global__ void __launch_bounds__(256) k_geo_generic(int n, const int* glist, const int* etype,
                                                     const int* conn, const double* xyz,
                                                     const long long* off, double* geo, int* err) {
    using Tr = topo::Traits<TOPO>;
    constexpr int npe = Tr::npe;
    const int i = int(xcd_block() * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int e = glist[i];
    if (etype[e] != TOPO) return;
    double c[npe][3];
    const size_t o0 = 3 * size_t(conn[8 * size_t(e)]);
    for (int j = 0; j < npe; ++j)
        for (int k = 0; k < 3; ++k) c[j][k] = xyz[3 * size_t(conn[8 * size_t(e) + j]) + k] - xyz[o0 + k];
    int bad = 0;
    double* q = geo + off[e];
#pragma unroll 1
    for (int ip = 0; ip < Tr::nscs; ++ip, q += topo::geo_stride<TOPO>()) {
        double A[3];
        topo::scs_area<TOPO>(ip, c, A);
        for (int d = 0; d < 3; ++d) q[d] = A[d];
        for (int sh = 0; sh < 2; ++sh) {
            const double* l = topo::ip_loc<TOPO>(sh, ip);
            double dd[npe][3], g[npe][3];
            int bsh = 0;
            topo::deriv<TOPO>(sh, l[0], l[1], l[2], dd);
            topo::gradop<npe>(dd, c, g, &bsh);
            // the shifted ip at a pyramid's apex is singular (det J = 0, a
            // zero operator from the fallback); only the standard ips judge
            // the element
            if (!sh) bad |= bsh;
            for (int j = 0; j < npe; ++j)
                for (int k = 0; k < 3; ++k) q[3 + 3 * npe * sh + 3 * j + k] = g[j][k];
        }
    }
    if (bad) *err = 1;
}

void launch_geo_generic(DevMesh& m, hipStream_t s) {
    if (!m.nglist) return;
    const unsigned B = 256, G = xcd_grid(size_t(m.nglist), B);
    k_geo_generic<topo::kTet><<<G, B, 0, s>>>(m.nglist, m.glist.p, m.etype.p, m.conn.p, m.xyz.p,
                                              m.geo_off.p, m.geo.p, m.err.p);
    k_geo_generic<topo::kWed><<<G, B, 0, s>>>(m.nglist, m.glist.p, m.etype.p, m.conn.p, m.xyz.p,
                                              m.geo_off.p, m.geo.p, m.err.p);
    k_geo_generic<topo::kPyr><<<G, B, 0, s>>>(m.nglist, m.glist.p, m.etype.p, m.conn.p, m.xyz.p,
                                              m.geo_off.p, m.geo.p, m.err.p);
    OAGPU_HIP(hipGetLastError());
}

// vertex parametric locations of tet4 / wedge6 / pyramid5 (index topology - 1)
static __constant__ double c_vpar[3][6][3] = {
    {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}, {0, 0, 0}},
    {{0, 0, -1}, {1, 0, -1}, {0, 1, -1}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}},
    {{-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}, {0, 0, 1}, {0, 0, 0}}};

// Face-ip geometry of the listed boundary sides whose element has topology
// TOPO (coordinates relative to the element's node 0): per face ip k the
// sub-face area vector (tri3d_scs_det / quad3d_scs_det over the side nodes
// in side order) and dN/dx of the element (face_grad_op) at the exposed-face
// ip -- on a tri side the face point with weights 11/18 on side node k and
// 7/36 on the others, on a quad side the bilinear point with weights 9/16
// (node k), 3/16 (its neighbours), 1/16 (the opposite node), which are the
// reference's intgExpFace_ entries -- and at side node k itself
// (shifted_face_grad_op, intgExpFaceShift_).  A non-positive Jacobian at a
// standard face ip sets *err.
template <int TOPO>
__global__ void __launch_bounds__(256) k_side_geo(int n, const int* slist, const int* bf_elem,
                                                  const int* bf_ord, const int* etype,
                                                  const int* conn, const double* xyz,
                                                  const long long* off, double* geo, int* err) {
    constexpr int npe = topo::Traits<TOPO>::npe;
    const int i = int(xcd_block() * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int f = slist[i], e = bf_elem[f];
    if (etype[e] != TOPO) return;
    double c[npe][3];
    const size_t o0 = 3 * size_t(conn[8 * size_t(e)]);
    for (int j = 0; j < npe; ++j)
        for (int k = 0; k < 3; ++k) c[j][k] = xyz[3 * size_t(conn[8 * size_t(e) + j]) + k] - xyz[o0 + k];
    const int* sn = c_sides[TOPO][bf_ord[f]];
    const int npf = sn[3] < 0 ? 3 : 4;
    double fp[4][3] = {};
    for (int q = 0; q < npf; ++q)
        for (int k = 0; k < 3; ++k) fp[q][k] = c[sn[q]][k];
    const double(*vp)[3] = c_vpar[TOPO - 1];
    int bad = 0;
    double* rec = geo + off[f];
#pragma unroll 1
    for (int fk = 0; fk < npf; ++fk) {
        double* q = rec + fk * kSideIpStride;
        double A[3];
        face_areav(npf, fp, fk, A);
        for (int d = 0; d < 3; ++d) q[d] = A[d];
        for (int sh = 0; sh < 2; ++sh) {
            double l[3] = {0, 0, 0};
            for (int qn = 0; qn < npf; ++qn) {
                double w;
                if (sh) w = qn == fk ? 1.0 : 0.0;
                else if (npf == 3) w = qn == fk ? 11.0 / 18.0 : 7.0 / 36.0;
                else {
                    const int dq = (qn - fk + 4) & 3;
                    w = dq == 0 ? 9.0 / 16.0 : dq == 2 ? 1.0 / 16.0 : 3.0 / 16.0;
                }
                for (int d = 0; d < 3; ++d) l[d] += w * vp[sn[qn]][d];
            }
            double dd[npe][3], g[npe][3];
            int bsh = 0;
            topo::deriv<TOPO>(sh, l[0], l[1], l[2], dd);
            topo::gradop<npe>(dd, c, g, &bsh);
            // the shifted ip at a pyramid's apex is singular (det J = 0, a
            // zero operator from the fallback); only the standard ips judge
            // the element
            if (!sh) bad |= bsh;
            for (int j = 0; j < npe; ++j)
                for (int k = 0; k < 3; ++k) q[3 + 18 * sh + 3 * j + k] = g[j][k];
        }
    }
    if (bad) *err = 1;
}

void launch_side_geo(DevMesh& m, hipStream_t s) {
    if (!m.nslist) return;
    const unsigned B = 256, G = xcd_grid(size_t(m.nslist), B);
    k_side_geo<topo::kTet><<<G, B, 0, s>>>(m.nslist, m.slist.p, m.bf_elem.p, m.bf_ord.p, m.etype.p,
                                           m.conn.p, m.xyz.p, m.sgeo_off.p, m.sgeo.p, m.err.p);
    k_side_geo<topo::kWed><<<G, B, 0, s>>>(m.nslist, m.slist.p, m.bf_elem.p, m.bf_ord.p, m.etype.p,
                                           m.conn.p, m.xyz.p, m.sgeo_off.p, m.sgeo.p, m.err.p);
    k_side_geo<topo::kPyr><<<G, B, 0, s>>>(m.nslist, m.slist.p, m.bf_elem.p, m.bf_ord.p, m.etype.p,
                                           m.conn.p, m.xyz.p, m.sgeo_off.p, m.sgeo.p, m.err.p);
    OAGPU_HIP(hipGetLastError());
}

void launch_scv_generic(DevMesh& m, hipStream_t s) {
    if (!m.nglist) return;
    const unsigned B = 256, G = xcd_grid(size_t(m.nglist), B);
    k_scv_generic<topo::kTet><<<G, B, 0, s>>>(m.nglist, m.glist.p, m.etype.p, m.conn.p, m.xyz.p, m.scvw.p);
    k_scv_generic<topo::kWed><<<G, B, 0, s>>>(m.nglist, m.glist.p, m.etype.p, m.conn.p, m.xyz.p, m.scvw.p);
    k_scv_generic<topo::kPyr><<<G, B, 0, s>>>(m.nglist, m.glist.p, m.etype.p, m.conn.p, m.xyz.p, m.scvw.p);
    OAGPU_HIP(hipGetLastError());
}

int topology_npe(int topo) { return topo == 1 ? 4 : topo == 2 ? 6 : topo == 3 ? 5 : 8; }
int topology_nscs(int topo) { return topo == 1 ? 6 : topo == 2 ? 9 : 12; }

// ---------------------------------------------------------------------------
// Scratch of the allocation-free geometry entry point: a file-scope device
// array, one copy per GPU, held until the process ends.  Word 0 is the
// error word of the geometry kernels (non-zero: a non-positive Jacobian).
// Host code never passes `scratch` itself to a kernel (in host code the name
// is a stand-in, not a device address); it passes the address that
// hipGetSymbolAddress returned for the current device.
constexpr int kScratchWords = 64;  // 256 bytes
__device__ __align__(256) int scratch[kScratchWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[kMaxDevices] = {};

__global__ void k_fill_int(int n, int v, int* p) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = v;
}
}  // namespace

int* scratch_device_ptr() {
    int dev = 0;
    OAGPU_HIP(hipGetDevice(&dev));
    OAGPU_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        OAGPU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        OAGPU_REQUIRE(p, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = static_cast<int*>(p);
    }
    return g_scratch_ptr[dev];
}

int element_geometry_noalloc(int topo, int shifted, int n, const double* coords, double* A,
                             double* V, double* G, double* detj, hipStream_t s) {
    int* err = scratch_device_ptr();
    // clear the error word with a kernel (no memset node if the caller captures)
    k_fill_int<<<1, 64, 0, s>>>(1, 0, err);
    OAGPU_HIP(hipGetLastError());
    element_geometry(topo, shifted, n, coords, A, V, G, detj, err, s);
    int e = 0;
    OAGPU_HIP(hipMemcpyAsync(&e, err, sizeof(int), hipMemcpyDeviceToHost, s));
    OAGPU_HIP(hipStreamSynchronize(s));
    return e;
}

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
