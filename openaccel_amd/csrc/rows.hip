// rows.hip -- row-owned gfx950 kernels of the CVFEM hot path: momentum and
// pressure-correction assembly, Green-Gauss gradients.
//
// One lane owns one CRS row (one mesh node).  It walks the node's incident
// elements in ascending element id -- the reference's element order -- and
// evaluates each element's geometry only at the three SCS ips touching the
// node.  The element's local row is added through the precomputed
// element -> CRS scatter map; columns absent from the reduced stencil are
// lumped into the diagonal block, as applyCoeff_ does
// (src/assemble/assembler.h:380-445).  Every row is written by one lane only:
// no atomics, no colouring, bitwise-reproducible results.
//
// The incidence body is instantiated per local node LA and per ip-location
// set SH so that every ip, facet and node index is a compile-time constant:
// element data stays in registers and the shape-function derivatives at the
// fixed ip locations fold to constants.
//
// Local arithmetic type T: double in the FP64 form, float in the mixed form
// (coordinates are shifted to the row node in FP64 before rounding, so FP32
// sees element-sized numbers).  Global coefficients are summed in FP64.
#include "device_util.hpp"
#include "tables.hpp"

namespace oagpu {

// ---------------------------------------------------------- static geometry
template <int SH, int IP, int C, typename T>
__device__ __forceinline__ void dNds(T d[3]) {
    constexpr double s1 = tbl::LOC[SH][IP][0], s2 = tbl::LOC[SH][IP][1], s3 = tbl::LOC[SH][IP][2];
    constexpr double g1 = tbl::SIG[C][0], g2 = tbl::SIG[C][1], g3 = tbl::SIG[C][2];
    constexpr double a = 0.5 + g1 * s1, b = 0.5 + g2 * s2, c = 0.5 + g3 * s3;
    d[0] = T(g1 * b * c);
    d[1] = T(g2 * a * c);
    d[2] = T(g3 * a * b);
}

template <int SH, int IP, int C>
__device__ constexpr double shape_at() {
    return (0.5 + tbl::SIG[C][0] * tbl::LOC[SH][IP][0]) * (0.5 + tbl::SIG[C][1] * tbl::LOC[SH][IP][1]) *
           (0.5 + tbl::SIG[C][2] * tbl::LOC[SH][IP][2]);
}

// inverse Jacobian at SCS ip IP of location set SH: Ji[k*3+d] = ds_k/dx_d
// (hex_gradient_operator, MasterElementWork.F:2666-2745)
template <typename T, int SH, int IP>
__device__ __forceinline__ void jinv_ip(const T c[8][3], T Ji[9], int& bad) {
    T dx1 = 0, dx2 = 0, dx3 = 0, dy1 = 0, dy2 = 0, dy3 = 0, dz1 = 0, dz2 = 0, dz3 = 0;
    sfor<8>([&](auto n_) {
        constexpr int n = decltype(n_)::value;
        T d[3];
        dNds<SH, IP, n>(d);
        dx1 += d[0] * c[n][0];
        dx2 += d[1] * c[n][0];
        dx3 += d[2] * c[n][0];
        dy1 += d[0] * c[n][1];
        dy2 += d[1] * c[n][1];
        dy3 += d[2] * c[n][1];
        dz1 += d[0] * c[n][2];
        dz2 += d[1] * c[n][2];
        dz3 += d[2] * c[n][2];
    });
    const T det = dx1 * (dy2 * dz3 - dz2 * dy3) + dy1 * (dz2 * dx3 - dx2 * dz3) +
                  dz1 * (dx2 * dy3 - dy2 * dx3);
    T test = det;
    if (!(test > T(0))) {
        test = T(1);
        bad = 1;
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
}

// dN_C/dx at ip IP given the inverse Jacobian there
template <typename T, int SH, int IP, int C>
__device__ __forceinline__ void gradN(const T Ji[9], T g[3]) {
    T d[3];
    dNds<SH, IP, C>(d);
    g[0] = d[0] * Ji[0] + d[1] * Ji[3] + d[2] * Ji[6];
    g[1] = d[0] * Ji[1] + d[1] * Ji[4] + d[2] * Ji[7];
    g[2] = d[0] * Ji[2] + d[1] * Ji[5] + d[2] * Ji[8];
}

// dN_c/ds and N_c at every SCS ip of both location sets, read with a
// wave-uniform column index (scalar constant loads)
struct ShapeTables {
    double dn[2][12][8][3];
    double N[2][12][8];
};
constexpr ShapeTables make_shape_tables() {
    ShapeTables t{};
    for (int sh = 0; sh < 2; ++sh)
        for (int ip = 0; ip < 12; ++ip)
            for (int c = 0; c < 8; ++c) {
                const double s1 = tbl::LOC[sh][ip][0], s2 = tbl::LOC[sh][ip][1], s3 = tbl::LOC[sh][ip][2];
                const double g1 = tbl::SIG[c][0], g2 = tbl::SIG[c][1], g3 = tbl::SIG[c][2];
                const double a = 0.5 + g1 * s1, b = 0.5 + g2 * s2, cc = 0.5 + g3 * s3;
                t.dn[sh][ip][c][0] = g1 * b * cc;
                t.dn[sh][ip][c][1] = g2 * a * cc;
                t.dn[sh][ip][c][2] = g3 * a * b;
                t.N[sh][ip][c] = a * b * cc;
            }
    return t;
}
__constant__ ShapeTables c_shape = make_shape_tables();

template <typename T>
__device__ __forceinline__ void gradN_rt(int sh, int ip, int c, const T Ji[9], T g[3]) {
    const T d0 = T(c_shape.dn[sh][ip][c][0]), d1 = T(c_shape.dn[sh][ip][c][1]),
            d2 = T(c_shape.dn[sh][ip][c][2]);
    g[0] = d0 * Ji[0] + d1 * Ji[3] + d2 * Ji[6];
    g[1] = d0 * Ji[1] + d1 * Ji[4] + d2 * Ji[7];
    g[2] = d0 * Ji[2] + d1 * Ji[5] + d2 * Ji[8];
}

template <int IDX, typename T>
__device__ __forceinline__ void sub_pt(const T c[8][3], const T ctr[3], T p[3]) {
    if constexpr (IDX < 8) {
        for (int k = 0; k < 3; ++k) p[k] = c[IDX][k];
    } else if constexpr (IDX == 26) {
        for (int k = 0; k < 3; ++k) p[k] = ctr[k];
    } else if constexpr (tbl::SUB[IDX - 8][2] < 0) {
        constexpr int a = tbl::SUB[IDX - 8][0], b = tbl::SUB[IDX - 8][1];
        for (int k = 0; k < 3; ++k) p[k] = T(0.5) * (c[a][k] + c[b][k]);
    } else {
        constexpr int a = tbl::SUB[IDX - 8][0], b = tbl::SUB[IDX - 8][1], d = tbl::SUB[IDX - 8][2],
                      e = tbl::SUB[IDX - 8][3];
        for (int k = 0; k < 3; ++k) p[k] = T(0.25) * (c[a][k] + c[b][k] + c[d][k] + c[e][k]);
    }
}

// SCS area vector of ip IP (hex_scs_det + quadAreaByTriangleFacets), L -> R
template <typename T, int IP>
__device__ __forceinline__ void scs_area_ip(const T c[8][3], const T ctr[3], T A[3]) {
    T q[4][3];
    sub_pt<tbl::FACET[IP][0]>(c, ctr, q[0]);
    sub_pt<tbl::FACET[IP][1]>(c, ctr, q[1]);
    sub_pt<tbl::FACET[IP][2]>(c, ctr, q[2]);
    sub_pt<tbl::FACET[IP][3]>(c, ctr, q[3]);
    hex8::quad_area(q, A);
}

template <typename T>
__device__ __forceinline__ void gather_rel(const double* xyz, const int nd[8], const double xr[3],
                                           T c[8][3]) {
    sfor<8>([&](auto n_) {
        constexpr int n = decltype(n_)::value;
        const size_t o = 3 * size_t(nd[n]);
        c[n][0] = T(xyz[o] - xr[0]);
        c[n][1] = T(xyz[o + 1] - xr[1]);
        c[n][2] = T(xyz[o + 2] - xr[2]);
    });
}

template <typename T>
__device__ __forceinline__ void scatter_block(double* vals, int rp, int dof, uint64_t sl, int cn,
                                              const T acc[9], double D[9]) {
    const int sp = slot_of(sl, cn);
    if (sp == 0xFF || sp == dof) {
        for (int k = 0; k < 9; ++k) D[k] += double(acc[k]);
    } else {
        double* v = vals + 9 * size_t(rp + sp);
        for (int k = 0; k < 9; ++k) v[k] += double(acc[k]);
    }
}

// --------------------------------------------------------------- momentum
struct MomArgs {
    int nn;
    const double* xyz;
    const int* conn;
    const int* row_ptr;
    const int* diag_off;
    const int* inc_la;
    const uint64_t* slots_la;
    const int* ext_ptr;
    const int* ext;
    const uint64_t* ext_slots;
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
    int ugrad_shifted;
    double* vals;
    double* b;
    double* du;
    int* err;
};

// element e's contribution to the momentum row of its local node LA
// (navierStokesAssemblerElemTerms.cpp:256-435 restricted to the ips of LA)
template <typename T, int SH, int LA>
__device__ __forceinline__ void mom_incidence(const MomArgs& a, int e, uint64_t sl, int rp, int dof,
                                              const double xr[3], double D[9], double R[3],
                                              int& bad) {
    int nd[8];
    load_conn(a.conn, e, nd);
    T A[3][3], Ji[3][9], m[3];
    {
        T c[8][3];
        gather_rel(a.xyz, nd, xr, c);
        T ctr[3];
        hex8::centroid(c, ctr);
        sfor<3>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            constexpr int ip = tbl::NODE_IP[LA][q];
            scs_area_ip<T, ip>(c, ctr, A[q]);
            jinv_ip<T, SH, ip>(c, Ji[q], bad);
            m[q] = T(a.mdot[12 * size_t(e) + ip]);
        });
    }
    const T mu = T(a.mu);
    T rl[3] = {0, 0, 0};
    // upwind advection flux (the deferred correction vanishes for upwind)
    sfor<3>([&](auto q_) {
        constexpr int q = decltype(q_)::value;
        constexpr int ip = tbl::NODE_IP[LA][q];
        constexpr int sg = tbl::NODE_SG[LA][q];
        const int nu = m[q] > T(0) ? nd[tbl::LR[ip][0]] : nd[tbl::LR[ip][1]];
        for (int i = 0; i < 3; ++i) {
            const T af = m[q] * T(a.U[3 * size_t(nu) + i]);
            if constexpr (sg > 0) rl[i] -= af;
            else rl[i] += af;
        }
    });
#pragma unroll 1
    for (int cn = 0; cn < 8; ++cn) {
        const size_t o = 3 * size_t(a.conn[8 * size_t(e) + cn]);
        const T uc[3] = {T(a.U[o]), T(a.U[o + 1]), T(a.U[o + 2])};
        T acc[9];
        for (int k = 0; k < 9; ++k) acc[k] = T(0);
        sfor<3>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            constexpr int ip = tbl::NODE_IP[LA][q];
            constexpr T sg = T(tbl::NODE_SG[LA][q]);
            if (cn == tbl::LR[ip][0]) {
                const T aL = T(0.5) * (m[q] + fabs(m[q]));
                acc[0] += sg * aL;
                acc[4] += sg * aL;
                acc[8] += sg * aL;
            }
            if (cn == tbl::LR[ip][1]) {
                const T aR = T(0.5) * (m[q] - fabs(m[q]));
                acc[0] += sg * aR;
                acc[4] += sg * aR;
                acc[8] += sg * aR;
            }
            T g[3];
            gradN_rt(SH, ip, cn, Ji[q], g);
            const T gA = g[0] * A[q][0] + g[1] * A[q][1] + g[2] * A[q][2];
            for (int i = 0; i < 3; ++i) {
                T ri = T(0);
                for (int j = 0; j < 3; ++j) {
                    T blk = -mu * g[i] * A[q][j];
                    if (i == j) blk += -mu * gA;
                    acc[3 * i + j] += sg * blk;
                    ri += blk * uc[j];
                }
                rl[i] -= sg * ri;
            }
        });
        scatter_block(a.vals, rp, dof, sl, cn, acc, D);
    }
    for (int i = 0; i < 3; ++i) R[i] += double(rl[i]);
}

// laminar no-slip wall: the face ip of side (f, fk) nearest to this row
// (navierStokesAssemblerElemBoundaryConditions.cpp:2659-2884,
// IP_ZERO_NORMAL_STRESS_FIXED_VEL__ :592-654)
template <typename T>
__device__ __forceinline__ void mom_wall(const MomArgs& a, int f, int fk, uint64_t sl, int rp,
                                         int dof, const double xr[3], double D[9], double R[3],
                                         int& bad) {
    const int e = a.bf_elem[f], ord = a.bf_ord[f], bc = a.bf_bc[f];
    const int* sn = hex8::c_side_nodes[ord];
    // face ip location: exposed-face ip, or the side node itself when shifted
    T s[3];
    for (int d = 0; d < 3; ++d)
        s[d] = a.ugrad_shifted ? T(hex8::c_node_loc[sn[fk]][d]) : T(hex8::c_face_loc[ord][fk][d]);
    T Ji[9], A[3];
    {
        int nd[8];
        load_conn(a.conn, e, nd);
        T c[8][3];
        gather_rel(a.xyz, nd, xr, c);
        jac_inv(s[0], s[1], s[2], c, Ji, &bad);
        T fp[4][3];
        for (int q = 0; q < 4; ++q) {
            const size_t o = 3 * size_t(a.conn[8 * size_t(e) + sn[q]]);
            for (int d = 0; d < 3; ++d) fp[q][d] = T(a.xyz[o + d] - xr[d]);
        }
        hex8::quad_face_areav(fp, fk, A);
    }
    const T amag = sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
    const T nx[3] = {A[0] / amag, A[1] / amag, A[2] / amag};
    const T mu = T(a.mu);
    const T ub[3] = {T(a.bc_U[3 * bc]), T(a.bc_U[3 * bc + 1]), T(a.bc_U[3 * bc + 2])};
    T rl[3] = {0, 0, 0};
#pragma unroll 1
    for (int ic = 0; ic < 8; ++ic) {
        const bool onface = ic == sn[0] || ic == sn[1] || ic == sn[2] || ic == sn[3];
        const T bcm = onface ? T(0) : T(1);
        const size_t o = 3 * size_t(a.conn[8 * size_t(e) + ic]);
        T uc[3];
        for (int d = 0; d < 3; ++d) uc[d] = onface ? ub[d] : T(a.U[o + d]);
        T g[3];
        grad_N(ic, s[0], s[1], s[2], Ji, g);
        T acc[9];
        for (int k = 0; k < 9; ++k) acc[k] = T(0);
        for (int j = 0; j < 3; ++j) {
            const T axj = A[j], dnj = g[j], uxj = uc[j];
            for (int i = 0; i < 3; ++i) {
                const T dni = g[i], uxi = uc[i], nxi = nx[i];
                const T om = T(1) - nxi * nxi;
                T lf = -mu * dnj * axj * om;
                acc[3 * i + i] += lf * bcm;
                rl[i] -= lf * uxi;
                lf = -mu * dni * axj * om;
                acc[3 * i + j] += lf * bcm;
                rl[i] -= lf * uxj;
                for (int l = 0; l < 3; ++l) {
                    if (l == i) continue;
                    const T nn_ = nxi * nx[l];
                    lf = mu * dnj * axj * nn_;
                    acc[3 * i + l] += lf * bcm;
                    rl[i] -= lf * uc[l];
                    lf = mu * g[l] * axj * nn_;
                    acc[3 * i + j] += lf * bcm;
                    rl[i] -= lf * uxj;
                }
            }
        }
        scatter_block(a.vals, rp, dof, sl, ic, acc, D);
    }
    for (int i = 0; i < 3; ++i) R[i] += double(rl[i]);
}

// navierStokesAssembler for one row: node terms (NodeTerms.cpp:16-149),
// interior element terms, no-slip walls, relaxation (phiAssembler.h:966-1006),
// boundary rhs relaxation (phiAssembler.h:1008-1152, navierStokesAssembler.cpp:238)
// and du = V/(a_ii + SMALL) (navierStokesAssembler.cpp:188-190)
template <typename T, int SH>
__global__ void __launch_bounds__(256) k_momentum_rows(const MomArgs a) {
    const int row = xcd_block() * blockDim.x + threadIdx.x;
    if (row >= a.nn) return;
    const int rp = a.row_ptr[row], re = a.row_ptr[row + 1], dof = a.diag_off[row];
    const double xr[3] = {a.xyz[3 * size_t(row)], a.xyz[3 * size_t(row) + 1],
                          a.xyz[3 * size_t(row) + 2]};
    int bad = 0;
    double D[9], R[3];
    for (int k = 0; k < 9; ++k) D[k] = 0.0;
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
    // incidences local-node major: wave-uniform control flow
    sfor<8>([&](auto la_) {
        constexpr int LA = decltype(la_)::value;
        const size_t o = size_t(LA) * a.nn + row;
        const int e = a.inc_la[o];
        if (e >= 0) mom_incidence<T, SH, LA>(a, e, a.slots_la[o], rp, dof, xr, D, R, bad);
    });
    for (int k = a.ext_ptr[row]; k < a.ext_ptr[row + 1]; ++k) {
        const int pk = a.ext[k];
        const int e = pk >> 3;
        const uint64_t sl = a.ext_slots[k];
        switch (pk & 7) {
            case 0: mom_incidence<T, SH, 0>(a, e, sl, rp, dof, xr, D, R, bad); break;
            case 1: mom_incidence<T, SH, 1>(a, e, sl, rp, dof, xr, D, R, bad); break;
            case 2: mom_incidence<T, SH, 2>(a, e, sl, rp, dof, xr, D, R, bad); break;
            case 3: mom_incidence<T, SH, 3>(a, e, sl, rp, dof, xr, D, R, bad); break;
            case 4: mom_incidence<T, SH, 4>(a, e, sl, rp, dof, xr, D, R, bad); break;
            case 5: mom_incidence<T, SH, 5>(a, e, sl, rp, dof, xr, D, R, bad); break;
            case 6: mom_incidence<T, SH, 6>(a, e, sl, rp, dof, xr, D, R, bad); break;
            default: mom_incidence<T, SH, 7>(a, e, sl, rp, dof, xr, D, R, bad); break;
        }
    }
    for (int k = a.binc_ptr[row]; k < a.binc_ptr[row + 1]; ++k) {
        const int pk = a.binc[k];
        mom_wall<T>(a, pk >> 2, pk & 3, a.binc_slots[k], rp, dof, xr, D, R, bad);
    }
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
    a.inc_la = m.inc_la.p;
    a.slots_la = m.slots_la.p;
    a.ext_ptr = m.ext_ptr.p;
    a.ext = m.ext.p;
    a.ext_slots = m.ext_slots.p;
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
    a.ugrad_shifted = f.prm.ugrad_shifted;
    a.vals = f.A3.p;
    a.b = f.b3.p;
    a.du = f.du.p;
    a.err = m.err.p;
    const unsigned B = 256, G = xcd_grid(m.nn, B);
    const bool mixed = f.prm.precision == OAGPU_MIXED;
    if (a.ugrad_shifted) {
        if (mixed) k_momentum_rows<float, 1><<<G, B, 0, f.stream>>>(a);
        else k_momentum_rows<double, 1><<<G, B, 0, f.stream>>>(a);
    } else {
        if (mixed) k_momentum_rows<float, 0><<<G, B, 0, f.stream>>>(a);
        else k_momentum_rows<double, 0><<<G, B, 0, f.stream>>>(a);
    }
    OAGPU_HIP(hipGetLastError());
}

// ------------------------------------------------------- pressure correction
struct PcArgs {
    int nn;
    const double* xyz;
    const int* conn;
    const int* row_ptr;
    const int* diag_off;
    const int* inc_la;
    const uint64_t* slots_la;
    const int* ext_ptr;
    const int* ext;
    const uint64_t* ext_slots;
    const double* scvw;
    const double* U;
    const double* du;
    const double* p;
    const double* gradP;
    double rho;
    int u_shifted, p_ref;
    double* vals;
    double* b;
    int* err;
};

// element e's contribution to the p' row of its local node LA
// (pressureCorrectionAssemblerElemTerms.cpp:390-610 restricted to LA's ips)
template <typename T, int SP, int LA>
__device__ __forceinline__ void pc_incidence(const PcArgs& a, int e, uint64_t sl, int rp, int dof,
                                             const double xr[3], double pr, double& D, double& R,
                                             int& bad) {
    int nd[8];
    load_conn(a.conn, e, nd);
    T A[3][3], Ji[3][9];
    {
        T c[8][3];
        gather_rel(a.xyz, nd, xr, c);
        T ctr[3];
        hex8::centroid(c, ctr);
        sfor<3>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            constexpr int ip = tbl::NODE_IP[LA][q];
            scs_area_ip<T, ip>(c, ctr, A[q]);
            jinv_ip<T, SP, ip>(c, Ji[q], bad);
        });
    }
    const bool us = a.u_shifted != 0;
    T uip[3][3], duip[3][3], dp[3][3], Gp[3] = {0, 0, 0};
    for (int q = 0; q < 3; ++q)
        for (int j = 0; j < 3; ++j) uip[q][j] = duip[q][j] = dp[q][j] = T(0);
#pragma unroll 1
    for (int n = 0; n < 8; ++n) {
        const int node = a.conn[8 * size_t(e) + n];
        const size_t o = 3 * size_t(node);
        const T w = T(a.scvw[8 * size_t(e) + n]);
        const T un[3] = {T(a.U[o]), T(a.U[o + 1]), T(a.U[o + 2])};
        const T dn[3] = {T(a.du[o]), T(a.du[o + 1]), T(a.du[o + 2])};
        const T pn = T(a.p[node] - pr);
        for (int j = 0; j < 3; ++j) Gp[j] += w * T(a.gradP[o + j]);
        sfor<3>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            constexpr int ip = tbl::NODE_IP[LA][q];
            const T r = T(c_shape.N[us ? 1 : 0][ip][n]);
            T g[3];
            gradN_rt(SP, ip, n, Ji[q], g);
            for (int j = 0; j < 3; ++j) {
                uip[q][j] += r * un[j];
                duip[q][j] += r * dn[j];
                dp[q][j] += g[j] * pn;
            }
        });
    }
    const T rho = T(a.rho);
    T rl = T(0);
    sfor<3>([&](auto q_) {
        constexpr int q = decltype(q_)::value;
        constexpr T sg = T(tbl::NODE_SG[LA][q]);
        T md = T(0);
        for (int j = 0; j < 3; ++j) {
            md += rho * uip[q][j] * A[q][j];
            md -= rho * duip[q][j] * (dp[q][j] - Gp[j]) * A[q][j];
        }
        rl -= sg * md;
    });
#pragma unroll 1
    for (int n = 0; n < 8; ++n) {
        T acc = T(0);
        sfor<3>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            constexpr int ip = tbl::NODE_IP[LA][q];
            constexpr T sg = T(tbl::NODE_SG[LA][q]);
            T g[3];
            gradN_rt(SP, ip, n, Ji[q], g);
            T lf = T(0);
            for (int j = 0; j < 3; ++j) lf += -rho * duip[q][j] * g[j] * A[q][j];
            acc += sg * lf;
        });
        const int sp = slot_of(sl, n);
        if (sp == 0xFF || sp == dof) D += double(acc);
        else a.vals[rp + sp] += double(acc);
    }
    R += double(rl);
}

// pressureCorrectionAssembler interior terms for one row, with the
// pressure-reference row (pressureCorrectionAssembler.cpp:19-63)
template <typename T, int SP>
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
    sfor<8>([&](auto la_) {
        constexpr int LA = decltype(la_)::value;
        const size_t o = size_t(LA) * a.nn + row;
        const int e = a.inc_la[o];
        if (e >= 0) pc_incidence<T, SP, LA>(a, e, a.slots_la[o], rp, dof, xr, pr, D, R, bad);
    });
    for (int k = a.ext_ptr[row]; k < a.ext_ptr[row + 1]; ++k) {
        const int pk = a.ext[k];
        const int e = pk >> 3;
        const uint64_t sl = a.ext_slots[k];
        switch (pk & 7) {
            case 0: pc_incidence<T, SP, 0>(a, e, sl, rp, dof, xr, pr, D, R, bad); break;
            case 1: pc_incidence<T, SP, 1>(a, e, sl, rp, dof, xr, pr, D, R, bad); break;
            case 2: pc_incidence<T, SP, 2>(a, e, sl, rp, dof, xr, pr, D, R, bad); break;
            case 3: pc_incidence<T, SP, 3>(a, e, sl, rp, dof, xr, pr, D, R, bad); break;
            case 4: pc_incidence<T, SP, 4>(a, e, sl, rp, dof, xr, pr, D, R, bad); break;
            case 5: pc_incidence<T, SP, 5>(a, e, sl, rp, dof, xr, pr, D, R, bad); break;
            case 6: pc_incidence<T, SP, 6>(a, e, sl, rp, dof, xr, pr, D, R, bad); break;
            default: pc_incidence<T, SP, 7>(a, e, sl, rp, dof, xr, pr, D, R, bad); break;
        }
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
    a.inc_la = m.inc_la.p;
    a.slots_la = m.slots_la.p;
    a.ext_ptr = m.ext_ptr.p;
    a.ext = m.ext.p;
    a.ext_slots = m.ext_slots.p;
    a.scvw = m.scvw.p;
    a.U = f.U.p;
    a.du = f.du.p;
    a.p = f.p.p;
    a.gradP = f.gradP.p;
    a.rho = f.prm.rho;
    a.u_shifted = f.prm.u_shifted;
    a.p_ref = f.prm.p_ref_node;
    a.vals = f.A1.p;
    a.b = f.b1.p;
    a.err = m.err.p;
    const unsigned B = 256, G = xcd_grid(m.nn, B);
    const bool mixed = f.prm.precision == OAGPU_MIXED;
    if (f.prm.pgrad_shifted) {
        if (mixed) k_pcorr_rows<float, 1><<<G, B, 0, f.stream>>>(a);
        else k_pcorr_rows<double, 1><<<G, B, 0, f.stream>>>(a);
    } else {
        if (mixed) k_pcorr_rows<float, 0><<<G, B, 0, f.stream>>>(a);
        else k_pcorr_rows<double, 0><<<G, B, 0, f.stream>>>(a);
    }
    OAGPU_HIP(hipGetLastError());
}

// --------------------------------------------------------------- mass flow
// initializeMassFlowRateInterior_ (flowModel.cpp:3303-3470): rho_ip U_ip . A
template <int US>
__global__ void __launch_bounds__(256) k_mdot_init(int ne, const double* xyz, const int* conn,
                                                   const double* U, double rho, double* mdot) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    int nd[8];
    load_conn(conn, e, nd);
    const double xr[3] = {xyz[3 * size_t(nd[0])], xyz[3 * size_t(nd[0]) + 1],
                          xyz[3 * size_t(nd[0]) + 2]};
    double c[8][3];
    gather_rel(xyz, nd, xr, c);
    double ctr[3];
    hex8::centroid(c, ctr);
    sfor<12>([&](auto ip_) {
        constexpr int ip = decltype(ip_)::value;
        double A[3];
        scs_area_ip<double, ip>(c, ctr, A);
        double u[3] = {0, 0, 0}, rip = 0;
#pragma unroll 1
        for (int n = 0; n < 8; ++n) {
            const double r = c_shape.N[US][ip][n];
            const size_t o = 3 * size_t(conn[8 * size_t(e) + n]);
            rip += r * rho;
            for (int j = 0; j < 3; ++j) u[j] += r * U[o + j];
        }
        double t = 0;
        for (int j = 0; j < 3; ++j) t += rip * u[j] * A[j];
        mdot[12 * size_t(e) + ip] = t;
    });
}

// updateMassFlowRateInterior_ (flowModel.cpp:5272-5688): Rhie-Chow with relaxation
template <int US, int SP>
__global__ void __launch_bounds__(256) k_mdot_update(int ne, const double* xyz, const int* conn,
                                                     const double* scvw, const double* U,
                                                     const double* du, const double* p,
                                                     const double* gradP, double rho, double urf,
                                                     double* mdot, int* err) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    int nd[8];
    load_conn(conn, e, nd);
    int bad = 0;
    const double xr[3] = {xyz[3 * size_t(nd[0])], xyz[3 * size_t(nd[0]) + 1],
                          xyz[3 * size_t(nd[0]) + 2]};
    const double pr = p[nd[0]];
    double c[8][3];
    gather_rel(xyz, nd, xr, c);
    double ctr[3];
    hex8::centroid(c, ctr);
    double Gp[3] = {0, 0, 0};
#pragma unroll 1
    for (int n = 0; n < 8; ++n) {
        const double w = scvw[8 * size_t(e) + n];
        const size_t o = 3 * size_t(conn[8 * size_t(e) + n]);
        for (int d = 0; d < 3; ++d) Gp[d] += w * gradP[o + d];
    }
    sfor<12>([&](auto ip_) {
        constexpr int ip = decltype(ip_)::value;
        double A[3], Ji[9];
        scs_area_ip<double, ip>(c, ctr, A);
        jinv_ip<double, SP, ip>(c, Ji, bad);
        double uip[3] = {0, 0, 0}, duip[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
#pragma unroll 1
        for (int n = 0; n < 8; ++n) {
            const double r = c_shape.N[US][ip][n];
            const int node = conn[8 * size_t(e) + n];
            const size_t o = 3 * size_t(node);
            double g[3];
            gradN_rt(SP, ip, n, Ji, g);
            const double pn = p[node] - pr;
            for (int j = 0; j < 3; ++j) {
                duip[j] += r * du[o + j];
                uip[j] += r * U[o + j];
                dp[j] += g[j] * pn;
            }
        }
        // constant density: the upwind density is rho whatever the flow direction
        double t = 0;
        for (int j = 0; j < 3; ++j) {
            t += rho * uip[j] * A[j];
            t -= rho * duip[j] * (dp[j] - Gp[j]) * A[j];
        }
        double* md = mdot + 12 * size_t(e) + ip;
        *md = urf * t + (1.0 - urf) * (*md);
    });
    if (bad) *err = 1;
}

void launch_mdot_init(Flow& f) {
    DevMesh& m = *f.m;
    const unsigned B = 256, G = xcd_grid(m.ne, B);
    if (f.prm.u_shifted)
        k_mdot_init<1><<<G, B, 0, f.stream>>>(m.ne, m.xyz.p, m.conn.p, f.U.p, f.prm.rho, f.mdot.p);
    else
        k_mdot_init<0><<<G, B, 0, f.stream>>>(m.ne, m.xyz.p, m.conn.p, f.U.p, f.prm.rho, f.mdot.p);
    OAGPU_HIP(hipGetLastError());
}

void launch_update_mdot(Flow& f) {
    DevMesh& m = *f.m;
    const unsigned B = 256, G = xcd_grid(m.ne, B);
    auto args = [&](auto kern) {
        kern<<<G, B, 0, f.stream>>>(m.ne, m.xyz.p, m.conn.p, m.scvw.p, f.U.p, f.du.p, f.p.p,
                                    f.gradP.p, f.prm.rho, f.prm.urf_mdot, f.mdot.p, m.err.p);
    };
    const int us = f.prm.u_shifted ? 1 : 0, sp = f.prm.pgrad_shifted ? 1 : 0;
    if (us && sp) args(k_mdot_update<1, 1>);
    else if (us) args(k_mdot_update<1, 0>);
    else if (sp) args(k_mdot_update<0, 1>);
    else args(k_mdot_update<0, 0>);
    OAGPU_HIP(hipGetLastError());
}

// -------------------------------------------------------------- gradients
// nodeField::updateGradientField (nodeField.hpp:3253-4950): incremental
// Green-Gauss with relaxation, interior ips then wall-side ips
template <int N, int LA>
__device__ __forceinline__ void grad_incidence(const double* xyz, const int* conn, const double* phi,
                                               int e, const double xr[3], const double pa[N],
                                               double iv, double gurf, int sh, double G[3 * N]) {
    int nd[8];
    load_conn(conn, e, nd);
    double c[8][3];
    gather_rel(xyz, nd, xr, c);
    double ctr[3];
    hex8::centroid(c, ctr);
    double pip[3][N];
    for (int q = 0; q < 3; ++q)
        for (int i = 0; i < N; ++i) pip[q][i] = 0.0;
    sfor<8>([&](auto n_) {
        constexpr int n = decltype(n_)::value;
        double pn[N];
        for (int i = 0; i < N; ++i) pn[i] = phi[N * size_t(nd[n]) + i];
        sfor<3>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            constexpr int ip = tbl::NODE_IP[LA][q];
            const double r = sh ? shape_at<1, ip, n>() : shape_at<0, ip, n>();
            for (int i = 0; i < N; ++i) pip[q][i] += r * pn[i];
        });
    });
    sfor<3>([&](auto q_) {
        constexpr int q = decltype(q_)::value;
        constexpr int ip = tbl::NODE_IP[LA][q];
        double A[3];
        scs_area_ip<double, ip>(c, ctr, A);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < 3; ++j) {
                const double t = gurf * (pip[q][i] - pa[i]) * A[j] * iv;
                if constexpr (tbl::NODE_SG[LA][q] > 0) G[3 * i + j] += t;
                else G[3 * i + j] -= t;
            }
    });
}

template <int N>
__global__ void __launch_bounds__(256) k_gradient(int nn, const double* xyz, const int* conn,
                                                  const int* inc_la, const int* ext_ptr,
                                                  const int* ext,
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
    sfor<8>([&](auto la_) {
        constexpr int LA = decltype(la_)::value;
        const int e = inc_la[size_t(LA) * nn + a];
        if (e >= 0) grad_incidence<N, LA>(xyz, conn, phi, e, xr, pa, iv, gurf, shifted, G);
    });
    for (int k = ext_ptr[a]; k < ext_ptr[a + 1]; ++k) {
        const int e = ext[k] >> 3;
        switch (ext[k] & 7) {
            case 0: grad_incidence<N, 0>(xyz, conn, phi, e, xr, pa, iv, gurf, shifted, G); break;
            case 1: grad_incidence<N, 1>(xyz, conn, phi, e, xr, pa, iv, gurf, shifted, G); break;
            case 2: grad_incidence<N, 2>(xyz, conn, phi, e, xr, pa, iv, gurf, shifted, G); break;
            case 3: grad_incidence<N, 3>(xyz, conn, phi, e, xr, pa, iv, gurf, shifted, G); break;
            case 4: grad_incidence<N, 4>(xyz, conn, phi, e, xr, pa, iv, gurf, shifted, G); break;
            case 5: grad_incidence<N, 5>(xyz, conn, phi, e, xr, pa, iv, gurf, shifted, G); break;
            case 6: grad_incidence<N, 6>(xyz, conn, phi, e, xr, pa, iv, gurf, shifted, G); break;
            default: grad_incidence<N, 7>(xyz, conn, phi, e, xr, pa, iv, gurf, shifted, G); break;
        }
    }
    for (int k = binc_ptr[a]; k < binc_ptr[a + 1]; ++k) {
        const int f = binc[k] >> 2, fk = binc[k] & 3;
        const int e = bf_elem[f], ord = bf_ord[f];
        double fp[4][3], pf[4][N];
        for (int q = 0; q < 4; ++q) {
            const int n = conn[8 * size_t(e) + hex8::c_side_nodes[ord][q]];
            for (int d = 0; d < 3; ++d) fp[q][d] = xyz[3 * size_t(n) + d] - xr[d];
            for (int i = 0; i < N; ++i) pf[q][i] = phi[N * size_t(n) + i];
        }
        double A[3];
        hex8::quad_face_areav(fp, fk, A);
        for (int i = 0; i < N; ++i) {
            double pip = 0.0;
            for (int q = 0; q < 4; ++q) pip += hex8::c_quad_shape[shifted][fk][q] * pf[q][i];
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
            m.nn, m.xyz.p, m.conn.p, m.inc_la.p, m.ext_ptr.p, m.ext.p, m.binc_ptr.p, m.binc.p,
            m.bf_elem.p, m.bf_ord.p, m.vol.p, f.U.p, f.prm.u_shifted, f.prm.grad_urf, f.gradU.p);
    else
        k_gradient<1><<<xcd_grid(m.nn, B), B, 0, f.stream>>>(
            m.nn, m.xyz.p, m.conn.p, m.inc_la.p, m.ext_ptr.p, m.ext.p, m.binc_ptr.p, m.binc.p,
            m.bf_elem.p, m.bf_ord.p, m.vol.p, f.p.p, f.prm.p_shifted, f.prm.grad_urf, f.gradP.p);
    OAGPU_HIP(hipGetLastError());
}

}  // namespace oagpu
