// assemble.hip -- gfx950 kernels of the CVFEM hot path: mesh geometry,
// momentum and pressure-correction assembly, mass flow, mass divergence,
// Green-Gauss gradients and the point-wise corrections.
//
// Assembly is row-owned: one lane owns one CRS row (one mesh node).  It walks
// the node's incident elements in ascending element id (the reference's
// element order), evaluates the element geometry only at the SCS ips that
// touch the node, and adds each element's local row into the row through the
// precomputed element -> CRS scatter map, lumping columns absent from the
// reduced stencil into the diagonal block exactly as applyCoeff_ does
// (src/assemble/assembler.h:380-445).  Every row is written by one lane only:
// no atomics, no colouring, bitwise-reproducible results.
//
// Local arithmetic type T: double in the FP64 form, float in the mixed form
// (coordinates shifted to the row node before rounding so that FP32 sees
// element-sized numbers).  Global coefficients are always summed in FP64.
#include "flow.hpp"
#include "hex8.hpp"

#include <cfloat>

namespace oagpu {

using namespace hex8;

constexpr double kSmall = DBL_EPSILON;  // src/basic/types.h:179

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

__device__ inline int slot_of(uint64_t sl, int c) { return int((sl >> (8 * c)) & 0xFFu); }

// ---------------------------------------------------------------- geometry
// SCV volumes of every element (hex_scv_det), written raw to scvw
__global__ void __launch_bounds__(256) k_scv_volumes(int ne, const double* xyz, const int* conn,
                                                     double* scvw) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    int nd[8];
    load_conn(conn, e, nd);
    double c[8][3];
    for (int n = 0; n < 8; ++n)
        for (int k = 0; k < 3; ++k) c[n][k] = xyz[3 * size_t(nd[n]) + k];
    double v[8];
    scv_volumes(c, v);
    for (int n = 0; n < 8; ++n) scvw[8 * size_t(e) + n] = v[n];
}

// dual nodal volume: sum of the node's SCVs in element order
__global__ void __launch_bounds__(256) k_dual_volume(int nn, const int* inc_ptr, const int* inc,
                                                     const double* scv, double* vol) {
    const int a = xcd_block() * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    double s = 0.0;
    for (int k = inc_ptr[a]; k < inc_ptr[a + 1]; ++k) s += scv[8 * size_t(inc[k] >> 3) + (inc[k] & 7)];
    vol[a] = s;
}

// SCV volumes -> element-centre averaging weights (pressureCorrectionAssemblerElemTerms.cpp:346-368)
__global__ void __launch_bounds__(256) k_scv_weights(int ne, double* scvw) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    double* w = scvw + 8 * size_t(e);
    double V = 0.0;
    for (int n = 0; n < 8; ++n) V += w[n];
    const double inv = 1.0 / V;
    for (int n = 0; n < 8; ++n) w[n] = w[n] * inv;
}

void launch_mesh_geometry(DevMesh& m, hipStream_t s) {
    const unsigned B = 256;
    k_scv_volumes<<<xcd_grid(m.ne, B), B, 0, s>>>(m.ne, m.xyz.p, m.conn.p, m.scvw.p);
    k_dual_volume<<<xcd_grid(m.nn, B), B, 0, s>>>(m.nn, m.inc_ptr.p, m.inc.p, m.scvw.p, m.vol.p);
    k_scv_weights<<<xcd_grid(m.ne, B), B, 0, s>>>(m.ne, m.scvw.p);
    OAGPU_HIP(hipGetLastError());
}

// --------------------------------------------------------------- momentum
struct MomArgs {
    int nn;
    const double* xyz;
    const int* conn;
    const int* row_ptr;
    const int* diag_off;
    const int* inc_ptr;
    const int* inc;
    const uint64_t* inc_slots;
    const int* binc_ptr;
    const int* binc;
    const uint64_t* binc_slots;
    const int* bf_elem;
    const int* bf_ord;
    const int* bf_bc;
    const double* bc_U;
    const int* bnode;
    const double* vol;
    const double* U;
    const double* gradP;
    const double* div;
    const double* mdot;
    double rho, mu, dt, urf_inv, bnd_relax;
    int u_shifted, ugrad_shifted;
    double* vals;
    double* b;
    double* du;
    int* err;
};

// navierStokesAssembler: node terms (NodeTerms.cpp:16-149), interior element
// terms (ElemTerms.cpp:15-441), laminar no-slip walls
// (ElemBoundaryConditions.cpp:2659-2884 with IP_ZERO_NORMAL_STRESS_FIXED_VEL__
// :592-654), relaxation (phiAssembler.h:966-1006), boundary rhs relaxation
// (phiAssembler.h:1008-1152) and du = V/(a_ii + SMALL) (navierStokesAssembler.cpp:188-190).
template <typename T>
__global__ void __launch_bounds__(256) k_momentum_rows(const MomArgs a) {
    const int row = xcd_block() * blockDim.x + threadIdx.x;
    if (row >= a.nn) return;
    const int rp = a.row_ptr[row], re = a.row_ptr[row + 1], dof = a.diag_off[row];
    const double xr[3] = {a.xyz[3 * size_t(row)], a.xyz[3 * size_t(row) + 1],
                          a.xyz[3 * size_t(row) + 2]};
    int bad = 0;
    double D[9], R[3];
    for (int k = 0; k < 9; ++k) D[k] = 0.0;
    // node terms, steady false transient
    {
        const double vol = a.vol[row], dv = a.div[row];
        const double lhsfac = a.rho * vol / a.dt;
        for (int k = 0; k < 3; ++k) {
            double l = 0.0, r = 0.0;
            if (dv < 0) l += -dv;
            r -= -dv * a.U[3 * size_t(row) + k];
            r -= a.gradP[3 * size_t(row) + k] * vol;
            l += lhsfac;
            D[4 * k] = l;
            R[k] = r;
        }
    }
    for (int s = rp; s < re; ++s)
        if (s - rp != dof)
            for (int k = 0; k < 9; ++k) a.vals[9 * size_t(s) + k] = 0.0;
    const T mu = T(a.mu);

    // interior element terms
    for (int k = a.inc_ptr[row]; k < a.inc_ptr[row + 1]; ++k) {
        const int pk = a.inc[k];
        const int e = pk >> 3, la = pk & 7;
        const uint64_t sl = a.inc_slots[k];
        int nd[8];
        load_conn(a.conn, e, nd);
        T c[8][3], Uc[8][3];
        for (int n = 0; n < 8; ++n)
            for (int d = 0; d < 3; ++d) {
                c[n][d] = T(a.xyz[3 * size_t(nd[n]) + d] - xr[d]);
                Uc[n][d] = T(a.U[3 * size_t(nd[n]) + d]);
            }
        T ctr[3];
        centroid(c, ctr);
        T A[3][3], Ji[3][9], m[3];
        int ipq[3], sg[3];
        for (int q = 0; q < 3; ++q) {
            const int ip = c_node_ip[la][q];
            ipq[q] = ip;
            sg[q] = c_node_ip_sign[la][q];
            scs_areav(ip, c, ctr, A[q]);
            const int sh = a.ugrad_shifted;
            jac_inv(T(c_scs_loc[sh][ip][0]), T(c_scs_loc[sh][ip][1]), T(c_scs_loc[sh][ip][2]), c,
                    Ji[q], &bad);
            m[q] = T(a.mdot[12 * size_t(e) + ip]);
        }
        T rl[3] = {0, 0, 0};
        // upwind advection flux (deferred correction vanishes for upwind)
        for (int q = 0; q < 3; ++q) {
            const int iu = m[q] > T(0) ? c_lrscv[ipq[q]][0] : c_lrscv[ipq[q]][1];
            for (int i = 0; i < 3; ++i) {
                const T af = m[q] * Uc[iu][i];
                if (sg[q] > 0) rl[i] -= af; else rl[i] += af;
            }
        }
        for (int cn = 0; cn < 8; ++cn) {
            T acc[9];
            for (int kk = 0; kk < 9; ++kk) acc[kk] = T(0);
            for (int q = 0; q < 3; ++q) {
                const int ip = ipq[q];
                const T aL = T(0.5) * (m[q] + fabs(m[q])), aR = T(0.5) * (m[q] - fabs(m[q]));
                const T sgn = T(sg[q]);
                if (cn == c_lrscv[ip][0]) { acc[0] += sgn * aL; acc[4] += sgn * aL; acc[8] += sgn * aL; }
                if (cn == c_lrscv[ip][1]) { acc[0] += sgn * aR; acc[4] += sgn * aR; acc[8] += sgn * aR; }
                const int sh = a.ugrad_shifted;
                T g[3];
                grad_N(cn, T(c_scs_loc[sh][ip][0]), T(c_scs_loc[sh][ip][1]), T(c_scs_loc[sh][ip][2]),
                       Ji[q], g);
                const T gA = g[0] * A[q][0] + g[1] * A[q][1] + g[2] * A[q][2];
                for (int i = 0; i < 3; ++i) {
                    T ri = T(0);
                    for (int j = 0; j < 3; ++j) {
                        T blk = -mu * g[i] * A[q][j];
                        if (i == j) blk += -mu * gA;
                        acc[3 * i + j] += sgn * blk;
                        ri += blk * Uc[cn][j];
                    }
                    rl[i] -= sgn * ri;
                }
            }
            const int sp = slot_of(sl, cn);
            if (sp == 0xFF || sp == dof) {
                for (int kk = 0; kk < 9; ++kk) D[kk] += double(acc[kk]);
            } else {
                double* v = a.vals + 9 * size_t(rp + sp);
                for (int kk = 0; kk < 9; ++kk) v[kk] += double(acc[kk]);
            }
        }
        for (int i = 0; i < 3; ++i) R[i] += double(rl[i]);
    }

    // no-slip walls: the face ip of each wall side whose nearest node is this row
    for (int k = a.binc_ptr[row]; k < a.binc_ptr[row + 1]; ++k) {
        const int pk = a.binc[k];
        const int f = pk >> 2, fk = pk & 3;
        const uint64_t sl = a.binc_slots[k];
        const int e = a.bf_elem[f], ord = a.bf_ord[f], bc = a.bf_bc[f];
        int nd[8];
        load_conn(a.conn, e, nd);
        T c[8][3], Uc[8][3], bcm[8];
        for (int n = 0; n < 8; ++n) {
            bcm[n] = T(1);
            for (int d = 0; d < 3; ++d) {
                c[n][d] = T(a.xyz[3 * size_t(nd[n]) + d] - xr[d]);
                Uc[n][d] = T(a.U[3 * size_t(nd[n]) + d]);
            }
        }
        T fp[4][3];
        for (int q = 0; q < 4; ++q) {
            const int ic = c_side_nodes[ord][q];
            bcm[ic] = T(0);
            for (int d = 0; d < 3; ++d) {
                Uc[ic][d] = T(a.bc_U[3 * bc + d]);
                fp[q][d] = c[ic][d];
            }
        }
        T A[3];
        quad_face_areav(fp, fk, A);
        T g[8][3];
        face_gradop(ord, fk, a.ugrad_shifted, c, g, &bad);
        const T amag = sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
        const T nx[3] = {A[0] / amag, A[1] / amag, A[2] / amag};
        T rl[3] = {0, 0, 0};
        for (int ic = 0; ic < 8; ++ic) {
            T acc[9];
            for (int kk = 0; kk < 9; ++kk) acc[kk] = T(0);
            for (int j = 0; j < 3; ++j) {
                const T axj = A[j], dnj = g[ic][j], uxj = Uc[ic][j];
                for (int i = 0; i < 3; ++i) {
                    const T dni = g[ic][i], uxi = Uc[ic][i], nxi = nx[i];
                    const T om = T(1) - nxi * nxi;
                    T lf = -mu * dnj * axj * om;
                    acc[3 * i + i] += lf * bcm[ic];
                    rl[i] -= lf * uxi;
                    lf = -mu * dni * axj * om;
                    acc[3 * i + j] += lf * bcm[ic];
                    rl[i] -= lf * uxj;
                    for (int l = 0; l < 3; ++l) {
                        if (l == i) continue;
                        const T nn_ = nxi * nx[l];
                        lf = mu * dnj * axj * nn_;
                        acc[3 * i + l] += lf * bcm[ic];
                        rl[i] -= lf * Uc[ic][l];
                        lf = mu * g[ic][l] * axj * nn_;
                        acc[3 * i + j] += lf * bcm[ic];
                        rl[i] -= lf * uxj;
                    }
                }
            }
            const int sp = slot_of(sl, ic);
            if (sp == 0xFF || sp == dof) {
                for (int kk = 0; kk < 9; ++kk) D[kk] += double(acc[kk]);
            } else {
                double* v = a.vals + 9 * size_t(rp + sp);
                for (int kk = 0; kk < 9; ++kk) v[kk] += double(acc[kk]);
            }
        }
        for (int i = 0; i < 3; ++i) R[i] += double(rl[i]);
    }

    // relaxation of the diagonal, boundary rhs relaxation, du
    for (int k = 0; k < 3; ++k) D[4 * k] *= a.urf_inv;
    if (a.bnode[row])
        for (int k = 0; k < 3; ++k) R[k] *= a.bnd_relax;
    const double vol = a.vol[row];
    for (int k = 0; k < 3; ++k) a.du[3 * size_t(row) + k] = vol / (D[4 * k] + kSmall);
    double* dv = a.vals + 9 * size_t(rp + dof);
    for (int kk = 0; kk < 9; ++kk) dv[kk] = D[kk];
    for (int k = 0; k < 3; ++k) a.b[3 * size_t(row) + k] = R[k];
    if (bad) *a.err = 1;
}

void launch_assemble_momentum(Flow& f) {
    DevMesh& m = *f.m;
    MomArgs a;
    a.nn = m.nn;
    a.xyz = m.xyz.p;
    a.conn = m.conn.p;
    a.row_ptr = m.row_ptr.p;
    a.diag_off = m.diag_off.p;
    a.inc_ptr = m.inc_ptr.p;
    a.inc = m.inc.p;
    a.inc_slots = m.inc_slots.p;
    a.binc_ptr = m.binc_ptr.p;
    a.binc = m.binc.p;
    a.binc_slots = m.binc_slots.p;
    a.bf_elem = m.bf_elem.p;
    a.bf_ord = m.bf_ord.p;
    a.bf_bc = m.bf_bc.p;
    a.bc_U = f.bc_U.p;
    a.bnode = m.bnode.p;
    a.vol = m.vol.p;
    a.U = f.U.p;
    a.gradP = f.gradP.p;
    a.div = f.div.p;
    a.mdot = f.mdot.p;
    a.rho = f.prm.rho;
    a.mu = f.prm.mu;
    a.dt = f.prm.dt;
    a.urf_inv = 1.0 / f.prm.urf_U;
    a.bnd_relax = f.prm.bnd_relax;
    a.u_shifted = f.prm.u_shifted;
    a.ugrad_shifted = f.prm.ugrad_shifted;
    a.vals = f.A3.p;
    a.b = f.b3.p;
    a.du = f.du.p;
    a.err = m.err.p;
    const unsigned B = 256;
    if (f.prm.precision == OAGPU_MIXED)
        k_momentum_rows<float><<<xcd_grid(m.nn, B), B, 0, f.stream>>>(a);
    else
        k_momentum_rows<double><<<xcd_grid(m.nn, B), B, 0, f.stream>>>(a);
    OAGPU_HIP(hipGetLastError());
}

// ------------------------------------------------------- pressure correction
struct PcArgs {
    int nn;
    const double* xyz;
    const int* conn;
    const int* row_ptr;
    const int* diag_off;
    const int* inc_ptr;
    const int* inc;
    const uint64_t* inc_slots;
    const double* scvw;
    const double* U;
    const double* du;
    const double* p;
    const double* gradP;
    const double* mdot;
    double rho;
    int u_shifted, pgrad_shifted, p_ref;
    double* vals;
    double* b;
    int* err;
};

// pressureCorrectionAssembler interior terms (ElemTerms.cpp:17-616):
// Rhie-Chow Laplacian -rho du_ip . dN/dx A and the mass imbalance rhs, with
// the pressure-reference row (pressureCorrectionAssembler.cpp:19-63).
template <typename T>
__global__ void __launch_bounds__(256) k_pcorr_rows(const PcArgs a) {
    const int row = xcd_block() * blockDim.x + threadIdx.x;
    if (row >= a.nn) return;
    const int rp = a.row_ptr[row], re = a.row_ptr[row + 1], dof = a.diag_off[row];
    const double xr[3] = {a.xyz[3 * size_t(row)], a.xyz[3 * size_t(row) + 1],
                          a.xyz[3 * size_t(row) + 2]};
    const double pr = a.p[row];
    int bad = 0;
    double D = 0.0, R = 0.0;
    for (int s = rp; s < re; ++s)
        if (s - rp != dof) a.vals[s] = 0.0;
    const T rho = T(a.rho);
    const int su = a.u_shifted, sp_ = a.pgrad_shifted;
    for (int k = a.inc_ptr[row]; k < a.inc_ptr[row + 1]; ++k) {
        const int pk = a.inc[k];
        const int e = pk >> 3, la = pk & 7;
        const uint64_t sl = a.inc_slots[k];
        int nd[8];
        load_conn(a.conn, e, nd);
        T c[8][3], Uc[8][3], duc[8][3], pc[8], Gp[3] = {0, 0, 0};
        for (int n = 0; n < 8; ++n) {
            const size_t o = 3 * size_t(nd[n]);
            const T w = T(a.scvw[8 * size_t(e) + n]);
            for (int d = 0; d < 3; ++d) {
                c[n][d] = T(a.xyz[o + d] - xr[d]);
                Uc[n][d] = T(a.U[o + d]);
                duc[n][d] = T(a.du[o + d]);
                Gp[d] += w * T(a.gradP[o + d]);
            }
            pc[n] = T(a.p[nd[n]] - pr);
        }
        T ctr[3];
        centroid(c, ctr);
        T acc[8];
        for (int n = 0; n < 8; ++n) acc[n] = T(0);
        T rl = T(0);
        for (int q = 0; q < 3; ++q) {
            const int ip = c_node_ip[la][q];
            const T sgn = T(c_node_ip_sign[la][q]);
            T A[3], Ji[9];
            scs_areav(ip, c, ctr, A);
            const T g1 = T(c_scs_loc[sp_][ip][0]), g2 = T(c_scs_loc[sp_][ip][1]),
                    g3 = T(c_scs_loc[sp_][ip][2]);
            jac_inv(g1, g2, g3, c, Ji, &bad);
            const T u1 = T(c_scs_loc[su][ip][0]), u2 = T(c_scs_loc[su][ip][1]),
                    u3 = T(c_scs_loc[su][ip][2]);
            T uip[3] = {0, 0, 0}, duip[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
            for (int n = 0; n < 8; ++n) {
                const T r = N_at(n, u1, u2, u3);
                for (int j = 0; j < 3; ++j) {
                    uip[j] += r * Uc[n][j];
                    duip[j] += r * duc[n][j];
                }
            }
            for (int n = 0; n < 8; ++n) {
                T g[3];
                grad_N(n, g1, g2, g3, Ji, g);
                T lf = T(0);
                for (int j = 0; j < 3; ++j) {
                    dp[j] += g[j] * pc[n];
                    lf += -rho * duip[j] * g[j] * A[j];
                }
                acc[n] += sgn * lf;
            }
            T md = T(0);
            for (int j = 0; j < 3; ++j) {
                md += rho * uip[j] * A[j];
                md -= rho * duip[j] * (dp[j] - Gp[j]) * A[j];
            }
            rl -= sgn * md;
        }
        for (int n = 0; n < 8; ++n) {
            const int sp = slot_of(sl, n);
            if (sp == 0xFF || sp == dof) D += double(acc[n]);
            else a.vals[rp + sp] += double(acc[n]);
        }
        R += double(rl);
    }
    if (row == a.p_ref) {
        for (int s = rp; s < re; ++s)
            if (s - rp != dof) a.vals[s] = 0.0;
        R = 0.0;
    }
    a.vals[rp + dof] = D;
    a.b[row] = R;
    if (bad) *a.err = 1;
}

void launch_assemble_pcorr(Flow& f) {
    DevMesh& m = *f.m;
    PcArgs a;
    a.nn = m.nn;
    a.xyz = m.xyz.p;
    a.conn = m.conn.p;
    a.row_ptr = m.row_ptr.p;
    a.diag_off = m.diag_off.p;
    a.inc_ptr = m.inc_ptr.p;
    a.inc = m.inc.p;
    a.inc_slots = m.inc_slots.p;
    a.scvw = m.scvw.p;
    a.U = f.U.p;
    a.du = f.du.p;
    a.p = f.p.p;
    a.gradP = f.gradP.p;
    a.mdot = f.mdot.p;
    a.rho = f.prm.rho;
    a.u_shifted = f.prm.u_shifted;
    a.pgrad_shifted = f.prm.pgrad_shifted;
    a.p_ref = f.prm.p_ref_node;
    a.vals = f.A1.p;
    a.b = f.b1.p;
    a.err = m.err.p;
    const unsigned B = 256;
    if (f.prm.precision == OAGPU_MIXED)
        k_pcorr_rows<float><<<xcd_grid(m.nn, B), B, 0, f.stream>>>(a);
    else
        k_pcorr_rows<double><<<xcd_grid(m.nn, B), B, 0, f.stream>>>(a);
    OAGPU_HIP(hipGetLastError());
}

// --------------------------------------------------------------- mass flow
// initializeMassFlowRateInterior_ (flowModel.cpp:3303-3470): rho_ip U_ip . A
__global__ void __launch_bounds__(256) k_mdot_init(int ne, const double* xyz, const int* conn,
                                                   const double* U, double rho, int su,
                                                   double* mdot) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    int nd[8];
    load_conn(conn, e, nd);
    double c[8][3], Uc[8][3];
    for (int n = 0; n < 8; ++n)
        for (int d = 0; d < 3; ++d) {
            c[n][d] = xyz[3 * size_t(nd[n]) + d];
            Uc[n][d] = U[3 * size_t(nd[n]) + d];
        }
    double ctr[3];
    centroid(c, ctr);
    for (int ip = 0; ip < 12; ++ip) {
        double A[3];
        scs_areav(ip, c, ctr, A);
        double u[3] = {0, 0, 0}, rip = 0;
        for (int n = 0; n < 8; ++n) {
            const double r = N_at(n, c_scs_loc[su][ip][0], c_scs_loc[su][ip][1], c_scs_loc[su][ip][2]);
            rip += r * rho;
            for (int j = 0; j < 3; ++j) u[j] += r * Uc[n][j];
        }
        double t = 0;
        for (int j = 0; j < 3; ++j) t += rip * u[j] * A[j];
        mdot[12 * size_t(e) + ip] = t;
    }
}

// updateMassFlowRateInterior_ (flowModel.cpp:5272-5688): Rhie-Chow with relaxation
__global__ void __launch_bounds__(256) k_mdot_update(int ne, const double* xyz, const int* conn,
                                                     const double* scvw, const double* U,
                                                     const double* du, const double* p,
                                                     const double* gradP, double rho, double urf,
                                                     int su, int sp, double* mdot, int* err) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    int nd[8];
    load_conn(conn, e, nd);
    int bad = 0;
    double c[8][3], Uc[8][3], duc[8][3], pc[8], Gp[3] = {0, 0, 0};
    for (int n = 0; n < 8; ++n) {
        const size_t o = 3 * size_t(nd[n]);
        const double w = scvw[8 * size_t(e) + n];
        for (int d = 0; d < 3; ++d) {
            c[n][d] = xyz[o + d];
            Uc[n][d] = U[o + d];
            duc[n][d] = du[o + d];
            Gp[d] += w * gradP[o + d];
        }
        pc[n] = p[nd[n]];
    }
    double ctr[3];
    centroid(c, ctr);
    for (int ip = 0; ip < 12; ++ip) {
        double A[3], Ji[9];
        scs_areav(ip, c, ctr, A);
        const double g1 = c_scs_loc[sp][ip][0], g2 = c_scs_loc[sp][ip][1], g3 = c_scs_loc[sp][ip][2];
        jac_inv(g1, g2, g3, c, Ji, &bad);
        const double u1 = c_scs_loc[su][ip][0], u2 = c_scs_loc[su][ip][1], u3 = c_scs_loc[su][ip][2];
        double uip[3] = {0, 0, 0}, duip[3] = {0, 0, 0}, dp[3] = {0, 0, 0}, rip = 0;
        for (int n = 0; n < 8; ++n) {
            const double r = N_at(n, u1, u2, u3);
            rip += r * rho;
            double g[3];
            grad_N(n, g1, g2, g3, Ji, g);
            for (int j = 0; j < 3; ++j) {
                duip[j] += r * duc[n][j];
                uip[j] += r * Uc[n][j];
                dp[j] += g[j] * pc[n];
            }
        }
        (void)rip;  // constant density: the upwind density is rho either way
        double t = 0;
        for (int j = 0; j < 3; ++j) {
            t += rho * uip[j] * A[j];
            t -= rho * duip[j] * (dp[j] - Gp[j]) * A[j];
        }
        double* md = mdot + 12 * size_t(e) + ip;
        *md = urf * t + (1.0 - urf) * (*md);
    }
    if (bad) *err = 1;
}

void launch_mdot_init(Flow& f) {
    DevMesh& m = *f.m;
    const unsigned B = 256;
    k_mdot_init<<<xcd_grid(m.ne, B), B, 0, f.stream>>>(m.ne, m.xyz.p, m.conn.p, f.U.p, f.prm.rho,
                                                       f.prm.u_shifted, f.mdot.p);
    OAGPU_HIP(hipGetLastError());
}

void launch_update_mdot(Flow& f) {
    DevMesh& m = *f.m;
    const unsigned B = 256;
    k_mdot_update<<<xcd_grid(m.ne, B), B, 0, f.stream>>>(
        m.ne, m.xyz.p, m.conn.p, m.scvw.p, f.U.p, f.du.p, f.p.p, f.gradP.p, f.prm.rho,
        f.prm.urf_mdot, f.prm.u_shifted, f.prm.pgrad_shifted, f.mdot.p, m.err.p);
    OAGPU_HIP(hipGetLastError());
}

// updateMassDivergenceField_ interior ips (flowModel.cpp:13733-13772): divL += mdot, divR -= mdot
__global__ void __launch_bounds__(256) k_mass_divergence(int nn, const int* inc_ptr, const int* inc,
                                                         const double* mdot, double* div) {
    const int a = xcd_block() * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    double s = 0.0;
    for (int k = inc_ptr[a]; k < inc_ptr[a + 1]; ++k) {
        const int e = inc[k] >> 3, la = inc[k] & 7;
        for (int q = 0; q < 3; ++q) {
            const double m = mdot[12 * size_t(e) + c_node_ip[la][q]];
            if (c_node_ip_sign[la][q] > 0) s += m; else s -= m;
        }
    }
    div[a] = s;
}

void launch_mass_divergence(Flow& f) {
    DevMesh& m = *f.m;
    const unsigned B = 256;
    k_mass_divergence<<<xcd_grid(m.nn, B), B, 0, f.stream>>>(m.nn, m.inc_ptr.p, m.inc.p, f.mdot.p,
                                                             f.div.p);
    OAGPU_HIP(hipGetLastError());
}

// -------------------------------------------------------------- gradients
// nodeField::updateGradientField (nodeField.hpp:3253-4950), incremental
// Green-Gauss with relaxation, interior ips then wall-side ips.
template <int N>
__global__ void __launch_bounds__(256) k_gradient(int nn, const double* xyz, const int* conn,
                                                  const int* inc_ptr, const int* inc,
                                                  const int* binc_ptr, const int* binc,
                                                  const int* bf_elem, const int* bf_ord,
                                                  const double* vol, const double* phi, int shifted,
                                                  double gurf, double* grad) {
    const int a = xcd_block() * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    double G[3 * N];
    for (int k = 0; k < 3 * N; ++k) G[k] = grad[3 * N * size_t(a) + k] * (1.0 - gurf);
    double pa[N];
    for (int i = 0; i < N; ++i) pa[i] = phi[N * size_t(a) + i];
    const double iv = 1.0 / vol[a];
    const double xr[3] = {xyz[3 * size_t(a)], xyz[3 * size_t(a) + 1], xyz[3 * size_t(a) + 2]};
    for (int k = inc_ptr[a]; k < inc_ptr[a + 1]; ++k) {
        const int e = inc[k] >> 3, la = inc[k] & 7;
        int nd[8];
        load_conn(conn, e, nd);
        double c[8][3];
        for (int n = 0; n < 8; ++n)
            for (int d = 0; d < 3; ++d) c[n][d] = xyz[3 * size_t(nd[n]) + d] - xr[d];
        double ctr[3];
        centroid(c, ctr);
        for (int q = 0; q < 3; ++q) {
            const int ip = c_node_ip[la][q];
            double A[3];
            scs_areav(ip, c, ctr, A);
            double pip[N];
            for (int i = 0; i < N; ++i) pip[i] = 0.0;
            for (int n = 0; n < 8; ++n) {
                const double r = N_at(n, c_scs_loc[shifted][ip][0], c_scs_loc[shifted][ip][1],
                                      c_scs_loc[shifted][ip][2]);
                for (int i = 0; i < N; ++i) pip[i] += r * phi[N * size_t(nd[n]) + i];
            }
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < 3; ++j) {
                    const double t = gurf * (pip[i] - pa[i]) * A[j] * iv;
                    if (c_node_ip_sign[la][q] > 0) G[3 * i + j] += t; else G[3 * i + j] -= t;
                }
        }
    }
    for (int k = binc_ptr[a]; k < binc_ptr[a + 1]; ++k) {
        const int f = binc[k] >> 2, fk = binc[k] & 3;
        int nd[8];
        load_conn(conn, bf_elem[f], nd);
        const int ord = bf_ord[f];
        double fp[4][3], pf[4][N];
        for (int q = 0; q < 4; ++q) {
            const int n = nd[c_side_nodes[ord][q]];
            for (int d = 0; d < 3; ++d) fp[q][d] = xyz[3 * size_t(n) + d] - xr[d];
            for (int i = 0; i < N; ++i) pf[q][i] = phi[N * size_t(n) + i];
        }
        double A[3];
        quad_face_areav(fp, fk, A);
        for (int i = 0; i < N; ++i) {
            double pip = 0.0;
            for (int q = 0; q < 4; ++q) pip += c_quad_shape[shifted][fk][q] * pf[q][i];
            for (int j = 0; j < 3; ++j) G[3 * i + j] += gurf * (pip - pa[i]) * A[j] * iv;
        }
    }
    for (int k = 0; k < 3 * N; ++k) grad[3 * N * size_t(a) + k] = G[k];
}

void launch_gradient(Flow& f, int which) {
    DevMesh& m = *f.m;
    const unsigned B = 256;
    if (which == 0)
        k_gradient<3><<<xcd_grid(m.nn, B), B, 0, f.stream>>>(
            m.nn, m.xyz.p, m.conn.p, m.inc_ptr.p, m.inc.p, m.binc_ptr.p, m.binc.p, m.bf_elem.p,
            m.bf_ord.p, m.vol.p, f.U.p, f.prm.u_shifted, f.prm.grad_urf, f.gradU.p);
    else
        k_gradient<1><<<xcd_grid(m.nn, B), B, 0, f.stream>>>(
            m.nn, m.xyz.p, m.conn.p, m.inc_ptr.p, m.inc.p, m.binc_ptr.p, m.binc.p, m.bf_elem.p,
            m.bf_ord.p, m.vol.p, f.p.p, f.prm.p_shifted, f.prm.grad_urf, f.gradP.p);
    OAGPU_HIP(hipGetLastError());
}

// ------------------------------------------------------------ corrections
__global__ void k_axpy(size_t n, double a, const double* x, double* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

// U += f du gradP (segregatedFlowEquations.cpp:189-200, :265-276)
__global__ void k_vel_corr(size_t n, double fac, const double* du, const double* gp, double* U) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) U[i] += fac * du[i] * gp[i];
}

void launch_correct(Flow& f, int N, double urf) {
    const size_t n = size_t(f.m->nn) * N;
    const unsigned B = 256;
    k_axpy<<<unsigned((n + B - 1) / B), B, 0, f.stream>>>(n, urf, N == 3 ? f.x3.p : f.x1.p,
                                                           N == 3 ? f.U.p : f.p.p);
    OAGPU_HIP(hipGetLastError());
}

void launch_velocity_correction(Flow& f, int step) {
    const size_t n = size_t(f.m->nn) * 3;
    const unsigned B = 256;
    const double fac = (step == 1 ? 1.0 : -1.0) / f.prm.urf_p;
    k_vel_corr<<<unsigned((n + B - 1) / B), B, 0, f.stream>>>(n, fac, f.du.p, f.gradP.p, f.U.p);
    OAGPU_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ scales
// per-block max/min of |U| and p (nodeField::max/min, nodeField.hpp:6143-6260)
__global__ void __launch_bounds__(kRedBlock) k_minmax(int nn, const double* U, const double* p,
                                                      double* part) {
    __shared__ double s[4][kRedBlock];
    double umax = -DBL_MAX, umin = DBL_MAX, pmax = -DBL_MAX, pmin = DBL_MAX;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
        const double ux = U[3 * size_t(i)], uy = U[3 * size_t(i) + 1], uz = U[3 * size_t(i) + 2];
        const double mag = sqrt(ux * ux + uy * uy + uz * uz);
        umax = fmax(umax, mag);
        umin = fmin(umin, mag);
        pmax = fmax(pmax, p[i]);
        pmin = fmin(pmin, p[i]);
    }
    s[0][threadIdx.x] = umax;
    s[1][threadIdx.x] = umin;
    s[2][threadIdx.x] = pmax;
    s[3][threadIdx.x] = pmin;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            s[0][threadIdx.x] = fmax(s[0][threadIdx.x], s[0][threadIdx.x + st]);
            s[1][threadIdx.x] = fmin(s[1][threadIdx.x], s[1][threadIdx.x + st]);
            s[2][threadIdx.x] = fmax(s[2][threadIdx.x], s[2][threadIdx.x + st]);
            s[3][threadIdx.x] = fmin(s[3][threadIdx.x], s[3][threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) part[threadIdx.x * gridDim.x + blockIdx.x] = s[threadIdx.x][0];
}

static double scale_of(double maxv, double minv) {
    double sc = (maxv > 0 ? maxv : 0.0) - (minv < 0 ? minv : 0.0);
    if (sc < 1.0e-9) sc = 1.0;
    if (maxv > sc) sc = maxv;
    return sc;
}

void field_scales(Flow& f, double* U_scale, double* p_scale) {
    const int G = 128;
    k_minmax<<<G, kRedBlock, 0, f.stream>>>(f.m->nn, f.U.p, f.p.p, f.part.p);
    OAGPU_HIP(hipGetLastError());
    std::vector<double> h(4 * G);
    f.part.download(h.data(), h.size(), f.stream);
    OAGPU_HIP(hipStreamSynchronize(f.stream));
    double umax = -DBL_MAX, umin = DBL_MAX, pmax = -DBL_MAX, pmin = DBL_MAX;
    for (int i = 0; i < G; ++i) {
        umax = std::max(umax, h[i]);
        umin = std::min(umin, h[G + i]);
        pmax = std::max(pmax, h[2 * G + i]);
        pmin = std::min(pmin, h[3 * G + i]);
    }
    // nodeField::updateScale (nodeField.hpp:5993-6015); flowModel::updatePressureScale
    // (flowModel.cpp:117-128) with a constant density of scale rho (>= 1e-9) or 1
    const double us = scale_of(umax, umin);
    const double rho_s = scale_of(f.prm.rho, f.prm.rho);
    const double pa = pmax - pmin, pb = 0.5 * rho_s * us * us;
    f.U_scale = us;
    f.p_scale = pa > pb ? pa : pb;
    if (U_scale) *U_scale = f.U_scale;
    if (p_scale) *p_scale = f.p_scale;
}

// RMS over rows of b/diag per component
template <int N>
__global__ void __launch_bounds__(kRedBlock) k_scaled_res(int nn, const int* rp, const int* doff,
                                                          const double* A, const double* b,
                                                          double* part) {
    __shared__ double s[N][kRedBlock];
    double acc[N];
    for (int k = 0; k < N; ++k) acc[k] = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
        const double* d = A + size_t(N * N) * (rp[i] + doff[i]);
        for (int k = 0; k < N; ++k) {
            const double r = b[N * size_t(i) + k] / d[(N + 1) * k];
            acc[k] += r * r;
        }
    }
    for (int k = 0; k < N; ++k) s[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int k = 0; k < N; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < N) part[threadIdx.x * gridDim.x + blockIdx.x] = s[threadIdx.x][0];
}

void scaled_residual(Flow& f, int N, double* out) {
    const int G = 128;
    DevMesh& m = *f.m;
    if (N == 3)
        k_scaled_res<3><<<G, kRedBlock, 0, f.stream>>>(m.nn, m.row_ptr.p, m.diag_off.p, f.A3.p,
                                                       f.b3.p, f.part.p);
    else
        k_scaled_res<1><<<G, kRedBlock, 0, f.stream>>>(m.nn, m.row_ptr.p, m.diag_off.p, f.A1.p,
                                                       f.b1.p, f.part.p);
    OAGPU_HIP(hipGetLastError());
    std::vector<double> h(size_t(N) * G);
    f.part.download(h.data(), h.size(), f.stream);
    OAGPU_HIP(hipStreamSynchronize(f.stream));
    for (int k = 0; k < N; ++k) {
        double s = 0.0;
        for (int i = 0; i < G; ++i) s += h[size_t(k) * G + i];
        out[k] = sqrt(s / m.nn);
    }
}

}  // namespace oagpu
