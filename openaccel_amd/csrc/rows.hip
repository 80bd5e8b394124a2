// rows.hip -- row-owned gfx950 kernels of the CVFEM hot path: momentum and
// pressure-correction assembly, Green-Gauss gradients, and the
// element-centric mass flow kernels.
//
// One lane owns one CRS row (one mesh node).  It walks the node's incident
// elements and evaluates each element's geometry only at the three SCS ips
// touching the node.  The element's local row is added through the
// precomputed element -> CRS scatter map; columns absent from the reduced
// stencil are lumped into the diagonal block, as applyCoeff_ does
// (src/assemble/assembler.h:380-445).  Every row is written by one lane only:
// no atomics, no colouring, bitwise-reproducible results.
//
// Incidences are visited local node by local node (la = 0..7): la, the ips
// touching it and the ip geometry constants are wave-uniform, read with
// scalar loads from one constant table, so a single compact loop body serves
// all eight local nodes (the instruction working set stays in the I-cache)
// and every lane follows the same path.
//
// Local arithmetic type T: double in the FP64 form, float in the mixed form
// (coordinates are shifted to the row node in FP64 before rounding, so FP32
// sees element-sized numbers).  Global coefficients are summed in FP64.
#include "device_util.hpp"
#include "sides.hpp"
#include "topo_tables.hpp"
#include "tables.hpp"

#include <type_traits>

namespace oagpu {

// hex8 master-element constants (Hex8CVFEM.C ip tables, MasterElementWork.F
// facet and subdivision tables) precomputed at compile time
struct HexTables {
    double dn[2][12][8][3];  // dN_c/ds at the SCS ips, standard [0] and shifted [1]
    double N[2][12][8];      // N_c at the SCS ips
    double fw[12][4][8];     // the 4 facet points of each SCS as node weights
    int node_ip[8][3];       // SCS ips touching each node
    int node_sg[8][3];       // +1 if the node is the ip's left node
    int lr[12][2];           // left / right node of each SCS ip
};

constexpr HexTables make_hex_tables() {
    HexTables t{};
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
    for (int ip = 0; ip < 12; ++ip)
        for (int k = 0; k < 4; ++k) {
            const int idx = tbl::FACET[ip][k];
            for (int n = 0; n < 8; ++n) t.fw[ip][k][n] = 0.0;
            if (idx < 8) {
                t.fw[ip][k][idx] = 1.0;
            } else if (idx == 26) {
                for (int n = 0; n < 8; ++n) t.fw[ip][k][n] = 0.125;
            } else {
                const int* sub = tbl::SUB[idx - 8];
                const int m = sub[2] < 0 ? 2 : 4;
                for (int q = 0; q < m; ++q) t.fw[ip][k][sub[q]] = m == 2 ? 0.5 : 0.25;
            }
        }
    for (int n = 0; n < 8; ++n)
        for (int q = 0; q < 3; ++q) {
            t.node_ip[n][q] = tbl::NODE_IP[n][q];
            t.node_sg[n][q] = tbl::NODE_SG[n][q];
        }
    for (int ip = 0; ip < 12; ++ip) {
        t.lr[ip][0] = tbl::LR[ip][0];
        t.lr[ip][1] = tbl::LR[ip][1];
    }
    return t;
}
__constant__ HexTables c_hex = make_hex_tables();
// the same tables as compile-time constants: loads with constant indices
// fold to immediates and zero weights drop out
static constexpr HexTables kHex = make_hex_tables();

// Canonical SCS.  The rotations of the reference cube map every directed
// edge la -> (neighbour across one of la's SCS) onto the edge 0 -> 1.
// Relabelling an element's nodes by such a rotation moves no physical point
// (areas, ip locations and physical shape-function gradients are those of
// the original element) and keeps the Jacobian positive, so each of a row's
// (incidence, ip) pairs is computed as SCS ip 0 of a relabelled element,
// with every master-element constant known at compile time.  For one
// incidence the three relabellings differ by the rotation about the 0-6
// diagonal that cycles the axes (x -> y -> z), kRho below, so the element
// is gathered once and its registers are permuted between the ips.  Only the
// node permutation, the original ip id and its orientation come from a
// wave-uniform table.
constexpr int kRho[8] = {0, 3, 7, 4, 1, 2, 6, 5};  // new node k = old node kRho[k]
constexpr int kCanIp[3] = {0, 3, 8};               // node 0's ips (towards 1, 3, 4), unrotated
constexpr int kFar[4] = {2, 5, 6, 7};             // nodes sharing no SCS with node 0

struct RotTables {
    int perm[24][8];  // [3 la + q]: original local node of canonical node k
    int ip[24];       // original ip of canonical ip 0
    int sg[24];       // +1 if la is that ip's left node
};

constexpr RotTables make_rot_tables() {
    RotTables t{};
    constexpr int P[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {2, 1, 0}, {1, 0, 2}};
    for (int la = 0; la < 8; ++la) {
        int perm0[8] = {};
        bool found = false;
        for (int p = 0; p < 6 && !found; ++p)
            for (int s = 0; s < 8 && !found; ++s) {
                // (Q x)_i = sgn_i x[P[p][i]]; a rotation iff det Q = +1
                const int sgn[3] = {(s & 1) ? -1 : 1, (s & 2) ? -1 : 1, (s & 4) ? -1 : 1};
                if ((p < 3 ? 1 : -1) * sgn[0] * sgn[1] * sgn[2] != 1) continue;
                bool hit = true;
                for (int i = 0; i < 3; ++i)
                    if (sgn[i] * tbl::SIG[la][P[p][i]] != tbl::SIG[0][i]) hit = false;
                if (!hit) continue;
                for (int k = 0; k < 8; ++k)
                    for (int j = 0; j < 8; ++j) {
                        bool eq = true;
                        for (int i = 0; i < 3; ++i)
                            if (sgn[i] * tbl::SIG[j][P[p][i]] != tbl::SIG[k][i]) eq = false;
                        if (eq) perm0[k] = j;
                    }
                found = true;
            }
        int cur[8] = {};
        for (int k = 0; k < 8; ++k) cur[k] = perm0[k];
        for (int q = 0; q < 3; ++q) {
            const int r = 3 * la + q;
            for (int k = 0; k < 8; ++k) t.perm[r][k] = cur[k];
            for (int ip = 0; ip < 12; ++ip) {
                if (tbl::LR[ip][0] == la && tbl::LR[ip][1] == cur[1]) t.ip[r] = ip, t.sg[r] = 1;
                if (tbl::LR[ip][1] == la && tbl::LR[ip][0] == cur[1]) t.ip[r] = ip, t.sg[r] = -1;
            }
            int nxt[8] = {};
            for (int k = 0; k < 8; ++k) nxt[k] = cur[kRho[k]];
            for (int k = 0; k < 8; ++k) cur[k] = nxt[k];
        }
    }
    return t;
}
static constexpr RotTables kRot = make_rot_tables();
__constant__ RotTables c_rot = make_rot_tables();

constexpr bool scs_partners(int a, int b) {
    for (int ip = 0; ip < 12; ++ip)
        if ((tbl::LR[ip][0] == a && tbl::LR[ip][1] == b) || (tbl::LR[ip][1] == a && tbl::LR[ip][0] == b))
            return true;
    return false;
}
constexpr bool rot_tables_valid() {
    for (int r = 0; r < 24; ++r) {
        const int la = r / 3, ip = kRot.ip[r];
        const int* pm = kRot.perm[r];
        if (pm[0] != la) return false;
        int seen = 0;
        for (int k = 0; k < 8; ++k) seen |= 1 << pm[k];
        if (seen != 0xFF) return false;
        if (tbl::LR[ip][kRot.sg[r] > 0 ? 0 : 1] != la || tbl::LR[ip][kRot.sg[r] > 0 ? 1 : 0] != pm[1])
            return false;
        if (!scs_partners(la, pm[3]) || !scs_partners(la, pm[4])) return false;
        for (int p = 0; p < 4; ++p)
            if (scs_partners(la, pm[kFar[p]])) return false;
        // a proper relabelling: the edges 0-1, 0-3, 0-4 map onto a
        // right-handed triple, and every node lands where the map sends it
        int d[3][3] = {};
        for (int i = 0; i < 3; ++i) {
            d[0][i] = tbl::SIG[pm[1]][i] - tbl::SIG[la][i];
            d[1][i] = tbl::SIG[pm[3]][i] - tbl::SIG[la][i];
            d[2][i] = tbl::SIG[pm[4]][i] - tbl::SIG[la][i];
        }
        const int det = d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) -
                        d[0][1] * (d[1][0] * d[2][2] - d[1][2] * d[2][0]) +
                        d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
        if (det != 8) return false;
        for (int k = 0; k < 8; ++k)
            for (int i = 0; i < 3; ++i) {
                // SIG[pm[k]] = SIG[la] + sum over set bits of k's position of the edges
                const int x = tbl::SIG[k][0] > 0, y = tbl::SIG[k][1] > 0, z = tbl::SIG[k][2] > 0;
                if (tbl::SIG[pm[k]][i] != tbl::SIG[la][i] + x * d[0][i] + y * d[1][i] + z * d[2][i])
                    return false;
            }
    }
    return kRot.perm[0][6] == 6 && kRot.perm[0][1] == 1;  // edge 0 -> 1: the identity
}
static_assert(rot_tables_valid(), "hex8 relabelling tables");

template <typename T>
__device__ __forceinline__ T recip(T x) {
    return T(1) / x;
}
template <>
__device__ __forceinline__ float recip<float>(float x) {
    return __builtin_amdgcn_rcpf(x);  // 1 ulp; the mixed form's local arithmetic
}

template <typename T>
__device__ __forceinline__ void jinv_finish(T dx1, T dx2, T dx3, T dy1, T dy2, T dy3, T dz1, T dz2,
                                            T dz3, T Ji[9], int& bad);

// inverse Jacobian at SCS ip `ip` of location set `sh`: Ji[k*3+d] = ds_k/dx_d
// (hex_gradient_operator, MasterElementWork.F:2666-2745)
template <typename T>
__device__ __forceinline__ void jinv_scs(int sh, int ip, const T c[8][3], T Ji[9], int& bad) {
    T dx1 = 0, dx2 = 0, dx3 = 0, dy1 = 0, dy2 = 0, dy3 = 0, dz1 = 0, dz2 = 0, dz3 = 0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const T d0 = T(c_hex.dn[sh][ip][n][0]), d1 = T(c_hex.dn[sh][ip][n][1]),
                d2 = T(c_hex.dn[sh][ip][n][2]);
        dx1 += d0 * c[n][0];
        dx2 += d1 * c[n][0];
        dx3 += d2 * c[n][0];
        dy1 += d0 * c[n][1];
        dy2 += d1 * c[n][1];
        dy3 += d2 * c[n][1];
        dz1 += d0 * c[n][2];
        dz2 += d1 * c[n][2];
        dz3 += d2 * c[n][2];
    }
    jinv_finish(dx1, dx2, dx3, dy1, dy2, dy3, dz1, dz2, dz3, Ji, bad);
}

// inverse of the Jacobian d(x,y,z)/d(s1,s2,s3) given column by column
// (dx1 = dx/ds1, ...); a non-positive determinant flags the element
template <typename T>
__device__ __forceinline__ void jinv_finish(T dx1, T dx2, T dx3, T dy1, T dy2, T dy3, T dz1, T dz2,
                                            T dz3, T Ji[9], int& bad) {
    const T det = dx1 * (dy2 * dz3 - dz2 * dy3) + dy1 * (dz2 * dx3 - dx2 * dz3) +
                  dz1 * (dx2 * dy3 - dy2 * dx3);
    T test = det;
    if (!(test > T(0))) {
        test = T(1);
        bad = 1;
    }
    const T den = recip(test);
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

// dN_c/dx at SCS ip `ip`
template <typename T>
__device__ __forceinline__ void gradN_scs(int sh, int ip, int c, const T Ji[9], T g[3]) {
    const T d0 = T(c_hex.dn[sh][ip][c][0]), d1 = T(c_hex.dn[sh][ip][c][1]),
            d2 = T(c_hex.dn[sh][ip][c][2]);
    g[0] = d0 * Ji[0] + d1 * Ji[3] + d2 * Ji[6];
    g[1] = d0 * Ji[1] + d1 * Ji[4] + d2 * Ji[7];
    g[2] = d0 * Ji[2] + d1 * Ji[5] + d2 * Ji[8];
}

// SCS area vector of ip `ip` (hex_scs_det + quadAreaByTriangleFacets), L -> R
template <typename T>
__device__ __forceinline__ void area_scs(int ip, const T c[8][3], T A[3]) {
    T q[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) {
            T v = T(0);
#pragma unroll
            for (int n = 0; n < 8; ++n) v += T(c_hex.fw[ip][k][n]) * c[n][d];
            q[k][d] = v;
        }
    hex8::quad_area(q, A);
}

// ---- compile-time ip versions (canonical incidences).  Every table index is
// a constant after unrolling, so the weights become immediates and the
// zero weights (half of them at the edge-midpoint ips) drop out.
template <typename T>
__device__ __forceinline__ void area_c(int ip, const T c[8][3], T A[3]) {
    T q[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            T v = T(0);
#pragma unroll
            for (int n = 0; n < 8; ++n)
                if (kHex.fw[ip][k][n] != 0.0) v += T(kHex.fw[ip][k][n]) * c[n][d];
            q[k][d] = v;
        }
    hex8::quad_area(q, A);
}

template <int SH, typename T>
__device__ __forceinline__ void jinv_c(int ip, const T c[8][3], T Ji[9], int& bad) {
    T s[3][3];  // s[d][k] = dx_d / ds_k
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            T v = T(0);
#pragma unroll
            for (int n = 0; n < 8; ++n)
                if (kHex.dn[SH][ip][n][k] != 0.0) v += T(kHex.dn[SH][ip][n][k]) * c[n][d];
            s[d][k] = v;
        }
    jinv_finish(s[0][0], s[0][1], s[0][2], s[1][0], s[1][1], s[1][2], s[2][0], s[2][1], s[2][2], Ji,
                bad);
}

// dN_n/dx at ip
template <int SH, typename T>
__device__ __forceinline__ void gradN_c(int ip, int n, const T Ji[9], T g[3]) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        T v = T(0);
#pragma unroll
        for (int d = 0; d < 3; ++d)
            if (kHex.dn[SH][ip][n][d] != 0.0) v += T(kHex.dn[SH][ip][n][d]) * Ji[3 * d + j];
        g[j] = v;
    }
}

// viscous 3x3 block of one column at one ip, added into acc:
// -mu (g_i A_j + (g . A) delta_ij)
template <typename T>
__device__ __forceinline__ void visc_block(T mu, const T g[3], const T A[3], T acc[9]) {
    const T gA = g[0] * A[0] + g[1] * A[1] + g[2] * A[2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) acc[3 * i + j] -= mu * (g[i] * A[j] + (i == j ? gA : T(0)));
}

template <typename T>
__device__ __forceinline__ void gather_rel(const double* xyz, const int nd[8], const double xr[3],
                                           T c[8][3]) {
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const size_t o = 3 * size_t(nd[n]);
        c[n][0] = T(xyz[o] - xr[0]);
        c[n][1] = T(xyz[o + 1] - xr[1]);
        c[n][2] = T(xyz[o + 2] - xr[2]);
    }
}

// Row slots whose off-diagonal 3x3 blocks are summed in the wavefront's LDS
// tile (a reduced-stencil hex row has at most 7 slots); further slots of
// longer rows are read-modify-written in HBM.
constexpr int kLdsSlots = 7;

// LT: type of the LDS partial sums (lt = this lane's tile column, stride 64
// per value) or void (all off-diagonal blocks read-modify-written in HBM)
template <typename T, typename LT>
__device__ __forceinline__ void scatter_block(double* __restrict__ vals, const SellRow& sr, int dof,
                                              uint64_t sl, int cn, const T acc[9], double D[9],
                                              LT* lt) {
    const int sp = slot_of(sl, cn);
    if (sp == 0xFF || sp == dof) {
        for (int k = 0; k < 9; ++k) D[k] += double(acc[k]);
        return;
    }
    if constexpr (!std::is_void_v<LT>) {
        if (sp < kLdsSlots) {
            LT* t = lt + sp * 9 * 64;
            for (int k = 0; k < 9; ++k) t[64 * k] += LT(acc[k]);
            return;
        }
    }
    double* v = sell_val<9>(vals, sr, sp, 0);
    for (int k = 0; k < 9; ++k) v[64 * k] += double(acc[k]);
}

// Scalar rows: slots < kLdsSlots accumulate in the wavefront's LDS tile
// (lt = this lane's column, stride 64), longer rows in HBM.
__device__ __forceinline__ void scalar_scatter(double* vals, const SellRow& sr, double* lt, int sp,
                                               double v) {
    if (sp < kLdsSlots) lt[sp * 64] += v;
    else *sell_val<1>(vals, sr, sp, 0) += v;
}

// zero the off-diagonal slots of a scalar row (LDS tile and HBM)
__device__ __forceinline__ void scalar_row_zero(double* vals, const SellRow& sr, double* lt, int dof) {
#pragma unroll 1
    for (int s = 0; s < sr.width; ++s) {
        if (s == dof) continue;
        if (s < kLdsSlots) lt[s * 64] = 0.0;
        else *sell_val<1>(vals, sr, s, 0) = 0.0;
    }
}

// store the LDS part of a scalar row's off-diagonal slots
__device__ __forceinline__ void scalar_row_store(double* vals, const SellRow& sr, const double* lt,
                                                 int dof) {
    const int ns = sr.width < kLdsSlots ? sr.width : kLdsSlots;
#pragma unroll 1
    for (int s = 0; s < ns; ++s)
        if (s != dof) *sell_val<1>(vals, sr, s, 0) = lt[s * 64];
}

// --------------------------------------------------------------- momentum
struct MomArgs {
    int nn, nslice;
    const int* slice_order;
    const double* xyz;
    const int* conn;
    const int* slice_ptr;
    const int* slice_w;
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
    const double* mueff;
    const double* beta;    // blending factors [nn][3] (high resolution) or null
    const double* gradU;
    const int* bc_type;    // physical type per boundary id
    const double* symA;    // assembled symmetry area [nn][3], null without symmetry
    int typed;             // some side is not a no-slip wall (bc_type decides per side)
    const double* msd;     // side mass flow per boundary face ip, null without inlets / outlets
    const int* rfl;        // outlet reversal flag per boundary face ip
    const int* gen_ptr;    // tet4 / wedge6 / pyramid5 incidences (mixed meshes)
    const int* gen;
    const uint64_t* gen_slots;
    const double* geo;          // static ip geometry of the non-hex elements
    const long long* geo_off;   // [ne] offset into geo
    SideGeo sg;                 // boundary sides of the non-hex elements
    const int* etype;
    double rho, mu, dt, urf_inv, bnd_relax;
    int ugrad_shifted, us;
    // transient: velocity of the previous two time levels (null when steady)
    // and the BDF coefficients times rho / dt (dt above is then dt / c0)
    const double* Uold;
    const double* Uold2;
    double tc[3];
    double* vals;
    double* b;
    double* du;
    int* err;
};

// SCS ip 0 of the relabelled element r = (la, q) of element e added to the
// momentum row of la (navierStokesAssemblerElemTerms.cpp:256-435 for one
// ip; the row node is its left node).  The explicit viscous flux uses the
// velocity gradient at the ip (the sum over all columns of block times U is
// -mu (grad U + grad U^T) . A), and the columns lumped into the diagonal need
// no blocks of their own: the shape-function gradients sum to zero, so the
// diagonal takes minus the gradients of the kept columns.
template <typename T, int SH, bool HR, int US, typename LT>
__device__ __forceinline__ void mom_scs(const MomArgs& a, int r, int e, const int nd[8],
                                        const T c[8][3], uint64_t sl, const SellRow& sr, int dof,
                                        double D[9], double R[3], int& bad, LT* lt) {
    constexpr int IP = 0, NB = 1;
    const T m = T(c_rot.sg[r]) * T(a.mdot[12 * size_t(e) + c_rot.ip[r]]);
    T A[3], Ji[9], rl[3];
    area_c(IP, c, A);
    jinv_c<SH>(IP, c, Ji, bad);
    {
        // upwind advection plus the limited deferred correction
        // (navierStokesAssemblerElemTerms.cpp:312-340)
        const bool fw = m > T(0);
        const size_t nu = size_t(fw ? nd[0] : nd[NB]);
        T up[3] = {T(a.U[3 * nu]), T(a.U[3 * nu + 1]), T(a.U[3 * nu + 2])};
        if constexpr (HR) {
            T dx[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                T xi = T(0);
#pragma unroll
                for (int n = 0; n < 8; ++n) xi += T(kHex.N[0][IP][n]) * c[n][d];
                dx[d] = xi - (fw ? c[0][d] : c[NB][d]);
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                T dc = T(0);
#pragma unroll
                for (int j = 0; j < 3; ++j) dc += T(a.beta[3 * nu + i]) * dx[j] * T(a.gradU[9 * nu + 3 * i + j]);
                up[i] += dc;
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) rl[i] = -m * up[i];
    }
    // effective viscosity at the ip (velocity shape functions US; their zero
    // weights and those nodes' loads drop out); a constant viscosity needs no
    // interpolation: the shape functions sum to one
    T mu = T(a.mu);
    if (a.mueff) {
        mu = T(0);
#pragma unroll
        for (int n = 0; n < 8; ++n)
            if (kHex.N[US][IP][n] != 0.0) mu += T(kHex.N[US][IP][n]) * T(a.mueff[nd[n]]);
    }
    // explicit viscous flux + mu (G + G^T) . A, G_ij = du_i/dx_j at the ip
    // from the parametric derivatives P_id = sum_n u_n,i dN_n/ds_d
    {
        T P[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i) P[i][0] = P[i][1] = P[i][2] = T(0);
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            if (kHex.dn[SH][IP][n][0] == 0.0 && kHex.dn[SH][IP][n][1] == 0.0 &&
                kHex.dn[SH][IP][n][2] == 0.0)
                continue;
            const size_t o = 3 * size_t(nd[n]);
            const T u[3] = {T(a.U[o]), T(a.U[o + 1]), T(a.U[o + 2])};
#pragma unroll
            for (int d = 0; d < 3; ++d)
                if (kHex.dn[SH][IP][n][d] != 0.0)
#pragma unroll
                    for (int i = 0; i < 3; ++i) P[i][d] += T(kHex.dn[SH][IP][n][d]) * u[i];
        }
        T G[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) G[i][j] = P[i][0] * Ji[j] + P[i][1] * Ji[3 + j] + P[i][2] * Ji[6 + j];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            rl[i] += mu * ((G[i][0] + G[0][i]) * A[0] + (G[i][1] + G[1][i]) * A[1] +
                           (G[i][2] + G[2][i]) * A[2]);
    }
    T gD[3] = {T(0), T(0), T(0)};  // gradient sum of the diagonal's columns
#pragma unroll
    for (int p = 1; p < 5; ++p) {  // the row node's SCS neighbours 1 (across ip 0), 3, 4
        if (p == 2) continue;
        T g[3], acc[9];
        gradN_c<SH>(IP, p, Ji, g);
#pragma unroll
        for (int k = 0; k < 9; ++k) acc[k] = T(0);
        visc_block(mu, g, A, acc);
        if (p == NB) {  // the node across this SCS
            const T adv = T(0.5) * (m - fabs(m));
            acc[0] += adv;
            acc[4] += adv;
            acc[8] += adv;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) gD[d] -= g[d];
        scatter_block(a.vals, sr, dof, sl, c_rot.perm[r][p], acc, D, lt);
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {  // kept only by full-stencil rows
        const int k = kFar[p];
        if (kHex.dn[SH][IP][k][0] == 0.0 && kHex.dn[SH][IP][k][1] == 0.0 && kHex.dn[SH][IP][k][2] == 0.0)
            continue;  // zero block at this ip
        const int cn = c_rot.perm[r][k];
        if (slot_of(sl, cn) == 0xFF) continue;
        T g[3], acc[9];
        gradN_c<SH>(IP, k, Ji, g);
#pragma unroll
        for (int q = 0; q < 9; ++q) acc[q] = T(0);
        visc_block(mu, g, A, acc);
#pragma unroll
        for (int d = 0; d < 3; ++d) gD[d] -= g[d];
        scatter_block(a.vals, sr, dof, sl, cn, acc, D, lt);
    }
    {  // the row node's own column and the lumped ones
        T acc[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) acc[k] = T(0);
        visc_block(mu, gD, A, acc);
        const T adv = T(0.5) * (m + fabs(m));
        acc[0] += adv;
        acc[4] += adv;
        acc[8] += adv;
#pragma unroll
        for (int k = 0; k < 9; ++k) D[k] += double(acc[k]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) R[i] += double(rl[i]);
}

// element e's contribution to the momentum row of its local node la: the
// element is gathered once, relabelled so that la is node 0, and its three
// SCS ips touching la are added as ip 0 of the three relabellings
template <typename T, int SH, bool HR, typename LT>
__device__ __forceinline__ void mom_incidence(const MomArgs& a, int la, int e, uint64_t sl,
                                              const SellRow& sr, int dof, const double xr[3],
                                              double D[9], double R[3], int& bad, LT* lt) {
    const int* ce = a.conn + 8 * size_t(e);
    int nd[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) nd[k] = ce[c_rot.perm[3 * la][k]];
    T c[8][3];
    gather_rel(a.xyz, nd, xr, c);
#pragma unroll 1
    for (int q = 0; q < 3; ++q) {
        if (a.us) mom_scs<T, SH, HR, 1>(a, 3 * la + q, e, nd, c, sl, sr, dof, D, R, bad, lt);
        else mom_scs<T, SH, HR, 0>(a, 3 * la + q, e, nd, c, sl, sr, dof, D, R, bad, lt);
        int t[8];
        T u[8][3];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            t[k] = nd[kRho[k]];
            u[k][0] = c[kRho[k]][0];
            u[k][1] = c[kRho[k]][1];
            u[k][2] = c[kRho[k]][2];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            nd[k] = t[k];
            c[k][0] = u[k][0];
            c[k][1] = u[k][1];
            c[k][2] = u[k][2];
        }
    }
}

// the SCS ips of a tet4 / wedge6 / pyramid5 element touching the momentum
// row of its local node la (navierStokesAssemblerElemTerms.cpp:256-435 per
// ip): area vector and mass flow taken from the row's side (the row node
// on the left), upwind advection with the limited deferred correction on
// the diagonal and the other node's column, the viscous blocks
// -mu (grad N . A I + grad N (x) A) of every column with their explicit
// flux on the rhs
template <typename T, int TOPO, bool HR, typename LT>
__device__ void mom_generic(const MomArgs& a, int la, int e, uint64_t sl, const SellRow& sr, int dof,
                            const double xr[3], double D[9], double R[3], int& bad, LT* lt) {
    constexpr int npe = topo::Traits<TOPO>::npe;
    const int* ce = a.conn + 8 * size_t(e);
    const double* geo = a.geo + a.geo_off[e];
    T c[npe][3];
    if constexpr (HR)
        for (int k = 0; k < npe; ++k)
            for (int d = 0; d < 3; ++d) c[k][d] = T(a.xyz[3 * size_t(ce[k]) + d] - xr[d]);
#pragma unroll 1
    for (int ip = 0; ip < topo::Traits<TOPO>::nscs; ++ip) {
        const int* lr = topo::lr<TOPO>(ip);
        if (lr[0] != la && lr[1] != la) continue;
        const bool left = lr[0] == la;
        const int o = left ? lr[1] : lr[0];
        const T sg = left ? T(1) : T(-1);
        T A[3], g[npe][3], Nc[npe], Nu[npe];
        topo::ip_geom<TOPO>(geo, ip, a.ugrad_shifted, a.us, A, g, Nc, Nu);
        for (int d = 0; d < 3; ++d) A[d] *= sg;
        const T m = sg * T(a.mdot[12 * size_t(e) + ip]);
        T mu = T(a.mu);
        if (a.mueff) {
            mu = T(0);
            for (int k = 0; k < npe; ++k) mu += Nu[k] * T(a.mueff[ce[k]]);
        }
        const int ul = m > T(0) ? la : o;
        const size_t nu = size_t(ce[ul]);
        T up[3] = {T(a.U[3 * nu]), T(a.U[3 * nu + 1]), T(a.U[3 * nu + 2])};
        if constexpr (HR) {
            T dx[3], cu[3];
            topo::pick_row<npe>(c, ul, cu);
            for (int d = 0; d < 3; ++d) {
                T xi = T(0);
                for (int k = 0; k < npe; ++k) xi += Nc[k] * c[k][d];
                dx[d] = xi - cu[d];
            }
            for (int i = 0; i < 3; ++i) {
                T dc = T(0);
                for (int j = 0; j < 3; ++j) dc += T(a.beta[3 * nu + i]) * dx[j] * T(a.gradU[9 * nu + 3 * i + j]);
                up[i] += dc;
            }
        }
        const T aS = T(0.5) * (m + fabs(m)), aO = T(0.5) * (m - fabs(m));
        for (int i = 0; i < 3; ++i) {
            R[i] -= double(m * up[i]);
            D[4 * i] += double(aS);
        }
        T rl[3] = {0, 0, 0};
#pragma unroll
        for (int ic = 0; ic < npe; ++ic) {
            T acc[9] = {};
            visc_block(mu, g[ic], A, acc);
            const size_t oc = 3 * size_t(ce[ic]);
            const T uc[3] = {T(a.U[oc]), T(a.U[oc + 1]), T(a.U[oc + 2])};
            for (int i = 0; i < 3; ++i)
                rl[i] -= acc[3 * i] * uc[0] + acc[3 * i + 1] * uc[1] + acc[3 * i + 2] * uc[2];
            if (ic == o)
                for (int i = 0; i < 3; ++i) acc[4 * i] += aO;
            scatter_block(a.vals, sr, dof, sl, ic, acc, D, lt);
        }
        for (int i = 0; i < 3; ++i) R[i] += double(rl[i]);
    }
}

// viscous face terms of the face ip of side (f, fk) nearest to this row.
// ZNS: zero normal stress (the stress projected off the face normal:
// IP_ZERO_NORMAL_STRESS_FIXED_VEL__ / IP_ZERO_NORMAL_STRESS__,
// navierStokesAssemblerElemBoundaryConditions.cpp:592-654 / :440-500),
// else the full stress (IP_FULL_STRESS_FIXED_VEL__ / IP_FULL_STRESS__
// :344-392 / :296-343; incompressible, so no divU term).  fixed: the side's
// nodes carry the boundary velocity bc_U and their columns are dropped.
//   no-slip wall      ZNS, fixed   (:2659-2884)
//   velocity inlet    full, fixed  (:791-1124)
//   pressure outlet   ZNS, free    (:1468-1775)
//   symmetry plane    full, free   (:2082-2375)
//   opening, outflow  ZNS, free    (:3505-3510)
//   opening, inflow   ZNS without the grad(U)^T terms
//                     (IP_ZERO_NORMAL_STRESS_NO_GRADU_TRANSPOSE__ :471-519),
//                     free, after the explicit entrainment flux msd Iu d_i,
//                     Iu = U_ip . n / (d . n + SMALL) (:3511-3528)
// entrain: the opening's inflow case (ZNS only; the transpose terms are
// scaled by tr = 0).  The entrainment direction d is the boundary id's bc_U
// row normalised, or for a zero row the inward unit normal of the face ip
// (velocity.cpp:4813-4970).
template <typename T, bool ZNS, bool MX, typename LT>
__device__ __forceinline__ void mom_face(const MomArgs& a, int f, int fk, bool fixed, uint64_t sl,
                                         const SellRow& sr, int dof, const double xr[3],
                                         double D[9], double R[3], int& bad, LT* lt,
                                         bool entrain = false) {
    const T tr = entrain ? T(0) : T(1);
    const int e = a.bf_elem[f], ord = a.bf_ord[f], bc = a.bf_bc[f];
    const int* ce = a.conn + 8 * size_t(e);
    FaceIp<T> fi;
    face_ip<MX>(a.sg, a.xyz, ce, e, f, ord, fk, a.ugrad_shifted, xr, fi, bad);
    T nx[3] = {0, 0, 0};
    if constexpr (ZNS) {
        const T amag = sqrt(fi.A[0] * fi.A[0] + fi.A[1] * fi.A[1] + fi.A[2] * fi.A[2]);
        for (int d = 0; d < 3; ++d) nx[d] = fi.A[d] / amag;
        if (entrain) {  // FP64 in both forms, like the other advective fluxes
            double Ad[3], d[3], uip[3] = {0, 0, 0};
            for (int j = 0; j < 3; ++j) Ad[j] = double(fi.A[j]);
            const double am = sqrt(Ad[0] * Ad[0] + Ad[1] * Ad[1] + Ad[2] * Ad[2]);
            const double* v = a.bc_U + 3 * bc;
            const double n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            if (n2 > 0.0) {
                const double nv = sqrt(n2);
                for (int j = 0; j < 3; ++j) d[j] = v[j] / nv;
            } else {
                for (int j = 0; j < 3; ++j) d[j] = -Ad[j] / am;
            }
            for (int q = 0; q < fi.npf; ++q) {
                const double r = fi.shape(a.us, fk, q);
                const size_t o = 3 * size_t(ce[fi.sn[q]]);
                for (int j = 0; j < 3; ++j) uip[j] += r * a.U[o + j];
            }
            double num = 0, den = 0;
            for (int j = 0; j < 3; ++j) {
                const double nj = Ad[j] / am;
                num += uip[j] * nj;
                den += d[j] * nj;
            }
            const double tm = a.msd[4 * size_t(f) + fk], Iu = num / (den + kSmall);
            for (int i = 0; i < 3; ++i) R[i] -= tm * Iu * d[i];
        }
    }
    T mu = T(a.mu);
    if (a.mueff) {
        mu = T(0);
        for (int q = 0; q < fi.npf; ++q) mu += T(fi.shape(a.us, fk, q)) * T(a.mueff[ce[fi.sn[q]]]);
    }
    const T ub[3] = {T(a.bc_U[3 * bc]), T(a.bc_U[3 * bc + 1]), T(a.bc_U[3 * bc + 2])};
    const T* A = fi.A;
    T rl[3] = {0, 0, 0};
#pragma unroll 1
    for (int ic = 0; ic < fi.npe; ++ic) {
        const bool mask = fixed && fi.on_side(ic);
        const T bcm = mask ? T(0) : T(1);
        const size_t o = 3 * size_t(ce[ic]);
        T uc[3];
        for (int d = 0; d < 3; ++d) uc[d] = mask ? ub[d] : T(a.U[o + d]);
        T g[3];
        fi.grad(ic, g);
        T acc[9];
        for (int k = 0; k < 9; ++k) acc[k] = T(0);
        for (int j = 0; j < 3; ++j) {
            const T axj = A[j], dnj = g[j], uxj = uc[j];
            for (int i = 0; i < 3; ++i) {
                const T dni = g[i], uxi = uc[i], nxi = nx[i];
                const T om = ZNS ? T(1) - nxi * nxi : T(1);
                T lf = -mu * dnj * axj * om;
                acc[3 * i + i] += lf * bcm;
                rl[i] -= lf * uxi;
                lf = -mu * dni * axj * om * tr;
                acc[3 * i + j] += lf * bcm;
                rl[i] -= lf * uxj;
                if constexpr (ZNS) {
                    for (int l = 0; l < 3; ++l) {
                        if (l == i) continue;
                        const T nn_ = nxi * nx[l];
                        lf = mu * dnj * axj * nn_;
                        acc[3 * i + l] += lf * bcm;
                        rl[i] -= lf * uc[l];
                        lf = mu * g[l] * axj * nn_ * tr;
                        acc[3 * i + j] += lf * bcm;
                        rl[i] -= lf * uxj;
                    }
                }
            }
        }
        scatter_block(a.vals, sr, dof, sl, ic, acc, D, lt);
    }
    for (int i = 0; i < 3; ++i) R[i] += double(rl[i]);
}

// the boundary face ip (f, fk) nearest to this row, by the physical type of
// its boundary; inlets add the explicit advective flux msd U_bc
// (IP_EXPLICIT_ADVECTIVE_FLUX__), outlets the implicit upwind one
// (IP_IMPLICIT_UPWIND_ADVECTIVE_FLUX__: msd on the diagonal, msd U_row on
// the rhs) unless the ip is reversed, when it takes no terms (a slip wall);
// openings the outlet's terms where msd > 0, else the entrainment terms.
// OP: openings may occur (a kernel variant of its own: the entrainment path
// costs registers in the others).
template <typename T, bool MX, bool OP, typename LT>
__device__ __forceinline__ void mom_boundary(const MomArgs& a, int pk, uint64_t sl,
                                             const SellRow& sr, int dof, int row,
                                             const double xr[3], double D[9], double R[3],
                                             int& bad, LT* lt) {
    const int f = pk >> 2, fk = pk & 3;
    const int t = a.typed ? a.bc_type[a.bf_bc[f]] : OAGPU_BC_WALL;
    if (t == OAGPU_BC_SLIP) return;  // slip wall: no face terms
    if (t == OAGPU_BC_SYMMETRY) {
        mom_face<T, false, MX>(a, f, fk, false, sl, sr, dof, xr, D, R, bad, lt);
    } else if (t == OAGPU_BC_INLET) {
        const double tm = a.msd[4 * size_t(f) + fk];
        const int bc = a.bf_bc[f];
        for (int i = 0; i < 3; ++i) R[i] -= tm * a.bc_U[3 * bc + i];
        mom_face<T, false, MX>(a, f, fk, true, sl, sr, dof, xr, D, R, bad, lt);
    } else if (t == OAGPU_BC_OUTLET || (OP && t == OAGPU_BC_OPENING)) {
        if (t == OAGPU_BC_OUTLET && a.rfl[4 * size_t(f) + fk] == 1) return;
        const double tm = a.msd[4 * size_t(f) + fk];
        if constexpr (OP) {
            if (t == OAGPU_BC_OPENING && !(tm > 0.0)) {
                mom_face<T, true, MX>(a, f, fk, false, sl, sr, dof, xr, D, R, bad, lt, true);
                return;
            }
        }
        for (int i = 0; i < 3; ++i) {
            R[i] -= tm * a.U[3 * size_t(row) + i];
            D[4 * i] += tm;
        }
        mom_face<T, true, MX>(a, f, fk, false, sl, sr, dof, xr, D, R, bad, lt);
    } else {
        mom_face<T, true, MX>(a, f, fk, true, sl, sr, dof, xr, D, R, bad, lt);
    }
}

// navierStokesAssembler for one row: node terms (NodeTerms.cpp:16-149),
// interior element terms, boundary face ips, relaxation (phiAssembler.h:966-1006),
// boundary rhs relaxation (phiAssembler.h:1008-1152, navierStokesAssembler.cpp:238)
// and du = V/(a_ii + SMALL) (navierStokesAssembler.cpp:188-190)
// LDS: 0 = off-diagonal blocks read-modify-written in HBM (256-lane blocks);
// 1 / 2 = summed in a per-wavefront LDS tile in float / double and written
// once (64-lane blocks, one wavefront each).
template <typename T, int SH, int LDS, bool HR, bool MX, bool OP = false>
__global__ void __launch_bounds__(LDS ? 64 : 256) k_momentum_rows(const MomArgs a) {
    using LT = std::conditional_t<LDS == 0, void, std::conditional_t<LDS == 2, double, float>>;
    using LS = std::conditional_t<LDS == 2, double, float>;
    __shared__ LS tile[LDS ? kLdsSlots * 9 * 64 : 1];
    LT* lt = nullptr;
    if constexpr (LDS != 0) lt = tile + threadIdx.x;
    const int row = ordered_item(a.slice_order, a.nslice, a.nn);
    if (row < 0) return;
    const SellRow sr = sell_row<9>(a.slice_ptr, a.slice_w, row);
    const int dof = a.diag_off[row];
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
            const double Uk = a.U[3 * size_t(row) + k];
            r -= -dv * Uk;
            r -= a.gradP[3 * size_t(row) + k] * vol;
            l += lhsfac;
            if (a.Uold)  // BDF1 / BDF2 time derivative of the previous levels
                r -= vol * (a.tc[0] * Uk + a.tc[1] * a.Uold[3 * size_t(row) + k] +
                            a.tc[2] * a.Uold2[3 * size_t(row) + k]);
            D[4 * k] = l;
            R[k] = r;
        }
    }
    // off-diagonal slots (and the slice's padding slots) start from zero
#pragma unroll 1
    for (int s = 0; s < sr.width; ++s) {
        if (s == dof) continue;
        if constexpr (LDS != 0) {
            if (s < kLdsSlots) {
                for (int k = 0; k < 9; ++k) lt[(s * 9 + k) * 64] = LS(0);
                continue;
            }
        }
        for (int k = 0; k < 9; ++k) *sell_val<9>(a.vals, sr, s, k) = 0.0;
    }
    // local node la is the (scalar, wave-uniform) loop counter; the rare extra
    // incidences of unstructured nodes follow in a second loop
#pragma unroll 1
    for (int la = 0; la < 8; ++la) {
        const size_t o = size_t(la) * a.nn + row;
        const int e = a.inc_la[o];
        if (e >= 0) mom_incidence<T, SH, HR>(a, la, e, a.slots_la[o], sr, dof, xr, D, R, bad, lt);
    }
#pragma unroll 1
    for (int k = a.ext_ptr[row]; k < a.ext_ptr[row + 1]; ++k)
        mom_incidence<T, SH, HR>(a, a.ext[k] & 7, a.ext[k] >> 3, a.ext_slots[k], sr, dof, xr, D, R,
                                 bad, lt);
    if constexpr (MX) {
#pragma unroll 1
        for (int k = a.gen_ptr[row]; k < a.gen_ptr[row + 1]; ++k) {
            const int e = a.gen[k] >> 3, la = a.gen[k] & 7, t = a.etype[e];
            if (t == OAGPU_TET4)
                mom_generic<T, topo::kTet, HR>(a, la, e, a.gen_slots[k], sr, dof, xr, D, R, bad, lt);
            else if (t == OAGPU_WED6)
                mom_generic<T, topo::kWed, HR>(a, la, e, a.gen_slots[k], sr, dof, xr, D, R, bad, lt);
            else
                mom_generic<T, topo::kPyr, HR>(a, la, e, a.gen_slots[k], sr, dof, xr, D, R, bad, lt);
        }
    }
#pragma unroll 1
    for (int k = a.binc_ptr[row]; k < a.binc_ptr[row + 1]; ++k)
        mom_boundary<T, MX, OP>(a, a.binc[k], a.binc_slots[k], sr, dof, row, xr, D, R, bad, lt);
    if constexpr (LDS != 0) {
        const int ns = sr.width < kLdsSlots ? sr.width : kLdsSlots;
#pragma unroll 1
        for (int s = 0; s < ns; ++s)
            if (s != dof)
                for (int k = 0; k < 9; ++k)  // streamed out: kept out of the caches
                    __builtin_nontemporal_store(double(lt[(s * 9 + k) * 64]),
                                                sell_val<9>(a.vals, sr, s, k));
    }
    for (int k = 0; k < 3; ++k) D[4 * k] *= a.urf_inv;
    if (a.bnode[row])
        for (int k = 0; k < 3; ++k) R[k] *= a.bnd_relax;
    double nrm[3] = {0, 0, 0};
    bool sym = false;
    if (a.symA) {  // applySymmetryConditions_ (navierStokesAssembler.cpp:387-471)
        const double* sa = a.symA + 3 * size_t(row);
        const double amag = sqrt(sa[0] * sa[0] + sa[1] * sa[1] + sa[2] * sa[2]);
        sym = amag > 0.0;
        if (sym) {
            double dot = 0.0;
            for (int j = 0; j < 3; ++j) {
                nrm[j] = sa[j] / amag;
                dot += R[j] * nrm[j];
            }
            for (int j = 0; j < 3; ++j) R[j] -= dot * nrm[j];
        }
    }
    const double vol = a.vol[row];
    for (int k = 0; k < 3; ++k) a.du[3 * size_t(row) + k] = vol / (D[4 * k] + kSmall);
    if (sym) {  // assembleNormalRelaxation (:473-696) with urf 0.75, after du
        double mod[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double v = 0.0;
                for (int k = 0; k < 3; ++k) v += D[3 * i + k] * nrm[k] * nrm[j];
                mod[3 * i + j] = v;
            }
        for (int k = 0; k < 9; ++k) D[k] += (1.0 - 0.75) * mod[k];
    }
    for (int kk = 0; kk < 9; ++kk) __builtin_nontemporal_store(D[kk], sell_val<9>(a.vals, sr, dof, kk));
    for (int k = 0; k < 3; ++k) a.b[3 * size_t(row) + k] = R[k];
    if (bad) *a.err = 1;
}

void launch_assemble_momentum(Flow& f) {
    DevMesh& m = *f.m;
    MomArgs a;
    a.nn = m.nn;
    a.nslice = m.nslice;
    a.slice_order = m.slice_order.p;
    a.xyz = m.xyz.p;
    a.conn = m.conn.p;
    a.slice_ptr = m.slice_ptr.p;
    a.slice_w = m.slice_w.p;
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
    a.sg = side_geo(m);
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
    a.Uold = a.Uold2 = nullptr;
    a.tc[0] = a.tc[1] = a.tc[2] = 0.0;
    if (f.tscheme) {
        double c[3];
        bdf_coeff(f, c);
        a.dt = f.tdt / c[0];
        for (int k = 0; k < 3; ++k) a.tc[k] = c[k] * f.prm.rho / f.tdt;
        a.Uold = f.Uold.p;
        a.Uold2 = f.Uold2.p;
    }
    a.urf_inv = 1.0 / f.prm.urf_U;
    a.bnd_relax = f.prm.bnd_relax;
    a.ugrad_shifted = f.prm.ugrad_shifted;
    a.us = f.prm.u_shifted ? 1 : 0;
    a.mueff = f.sst_on ? f.mueff.p : nullptr;  // laminar: the constant mu
    a.beta = f.prm.high_res ? f.beta.p : nullptr;
    a.gradU = f.gradU.p;
    a.bc_type = f.bc_type.p;
    a.symA = f.has_sym ? f.symA.p : nullptr;
    a.msd = f.has_io ? f.msd.p : nullptr;
    a.typed = f.has_sym || f.has_io || f.has_slip;
    a.rfl = f.has_io ? f.rfl.p : nullptr;
    a.gen_ptr = m.gen_ptr.p;
    a.gen = m.gen.p;
    a.gen_slots = m.gen_slots.p;
    a.geo = m.geo.p;
    a.geo_off = m.geo_off.p;
    a.etype = m.etype.p;
    a.vals = f.A3.p;
    a.b = f.b3.p;
    a.du = f.du.p;
    a.err = f.err.p;
    // off-diagonal blocks summed in a per-wavefront LDS tile: double in the
    // FP64 form, float in the mixed form (the element contributions are FP32
    // there; the at most four of them per block are summed in FP32 before the
    // single FP64 store)
    const bool mixed = f.prm.precision == OAGPU_MIXED;
    auto go = [&](auto kern, unsigned B) {
        clock_begin(f, kClkMomentum);
        kern<<<xcd_grid(m.nn, B), B, 0, f.stream>>>(a);
        clock_end(f, kClkMomentum);
    };
    const bool hr = a.beta != nullptr;
    // the off-diagonal blocks accumulated in HBM instead of the LDS tile (an
    // autotuning candidate; hex-only meshes, upwind advection, no openings)
    if (f.tune.mom_tile == 0 && !hr && !m.mixed && !f.has_opening) {
        if (a.ugrad_shifted) go(mixed ? k_momentum_rows<float, 1, 0, false, false>
                                      : k_momentum_rows<double, 1, 0, false, false>, 256);
        else go(mixed ? k_momentum_rows<float, 0, 0, false, false>
                      : k_momentum_rows<double, 0, 0, false, false>, 256);
        OAGPU_HIP(hipGetLastError());
        return;
    }
    auto pick = [&](auto mx, auto op) {
        constexpr bool MX = decltype(mx)::value, OP = decltype(op)::value;
        if (a.ugrad_shifted) {
            if (mixed) go(hr ? k_momentum_rows<float, 1, 1, true, MX, OP> : k_momentum_rows<float, 1, 1, false, MX, OP>, 64);
            else go(hr ? k_momentum_rows<double, 1, 2, true, MX, OP> : k_momentum_rows<double, 1, 2, false, MX, OP>, 64);
        } else {
            if (mixed) go(hr ? k_momentum_rows<float, 0, 1, true, MX, OP> : k_momentum_rows<float, 0, 1, false, MX, OP>, 64);
            else go(hr ? k_momentum_rows<double, 0, 2, true, MX, OP> : k_momentum_rows<double, 0, 2, false, MX, OP>, 64);
        }
    };
    if (f.has_opening) {
        if (m.mixed) pick(std::true_type{}, std::true_type{});
        else pick(std::false_type{}, std::true_type{});
    } else {
        if (m.mixed) pick(std::true_type{}, std::false_type{});
        else pick(std::false_type{}, std::false_type{});
    }
    OAGPU_HIP(hipGetLastError());
}

// ------------------------------------------------------- pressure correction
struct PcArgs {
    int nn, nslice;
    const int* slice_order;
    const double* xyz;
    const int* conn;
    const int* slice_ptr;
    const int* slice_w;
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
    // inlet / outlet face ips (boundary-row kernel only)
    const int* binc_ptr;
    const int* binc;
    const uint64_t* binc_slots;
    const int* bf_elem;
    const int* bf_ord;
    const int* bf_bc;
    const int* bc_type;
    const double* bc_U;
    const double* bc_p;
    const int* rfl;
    const int* gen_ptr;    // tet4 / wedge6 / pyramid5 incidences (mixed meshes)
    const int* gen;
    const uint64_t* gen_slots;
    const double* geo;          // static ip geometry of the non-hex elements
    const long long* geo_off;   // [ne] offset into geo
    SideGeo sg;                 // boundary sides of the non-hex elements
    const int* etype;
    double rho;
    int u_shifted, p_ref;
    double* vals;
    double* b;
    int* err;
};

// rotate the registers of a relabelled element to the next relabelling
template <typename V, int W>
__device__ __forceinline__ void rotate_nodes(V v[8][W]) {
    V t[8][W];
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int j = 0; j < W; ++j) t[k][j] = v[kRho[k]][j];
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int j = 0; j < W; ++j) v[k][j] = t[k][j];
}

// SCS ip 0 of the relabelled element r = (la, q) added to the p' row of la
// (pressureCorrectionAssemblerElemTerms.cpp:390-610 for one ip; the row node
// is its left node).  The columns lumped into the diagonal take minus the
// kept columns' coefficients: the shape-function gradients sum to zero.
// Nodal values: in registers, rotated with the coordinates (REG, the FP32
// form), or read per ip from L1/L2 (the FP64 form, whose registers are full).
template <typename T, int SP, bool REG, int US>
__device__ __forceinline__ void pc_scs(const PcArgs& a, int r, const int nd[8], double pr,
                                       const T c[8][3], const T un[8][3], const T dn[8][3],
                                       const T pn[8][1], const T Gp[3], uint64_t sl,
                                       const SellRow& sr, int dof, double& D, double& R, int& bad,
                                       double* lt) {
    constexpr int IP = 0;
    T A[3], Ji[9];
    area_c(IP, c, A);
    jinv_c<SP>(IP, c, Ji, bad);
    T uip[3] = {0, 0, 0}, duip[3] = {0, 0, 0}, P[3] = {0, 0, 0};
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        // velocity interpolation US (1: edge midpoint, two non-zero weights);
        // the zero weights drop out, and with them their nodes' loads
        const bool wz = kHex.N[US][IP][n] == 0.0;
        const bool gz = kHex.dn[SP][IP][n][0] == 0.0 && kHex.dn[SP][IP][n][1] == 0.0 &&
                        kHex.dn[SP][IP][n][2] == 0.0;
        const T w = T(kHex.N[US][IP][n]);
        if (!wz) {
            T u[3], du[3];
            if constexpr (REG) {
                for (int j = 0; j < 3; ++j) u[j] = un[n][j], du[j] = dn[n][j];
            } else {
                const size_t o = 3 * size_t(nd[n]);
                for (int j = 0; j < 3; ++j) u[j] = T(a.U[o + j]), du[j] = T(a.du[o + j]);
            }
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                uip[j] += w * u[j];
                duip[j] += w * du[j];
            }
        }
        if (!gz) {
            T pv;
            if constexpr (REG) pv = pn[n][0];
            else pv = T(a.p[nd[n]] - pr);
#pragma unroll
            for (int d = 0; d < 3; ++d)
                if (kHex.dn[SP][IP][n][d] != 0.0) P[d] += T(kHex.dn[SP][IP][n][d]) * pv;
        }
    }
    const T rho = T(a.rho);
    T md = T(0);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const T dp = P[0] * Ji[j] + P[1] * Ji[3 + j] + P[2] * Ji[6 + j];
        md += rho * uip[j] * A[j];
        md -= rho * duip[j] * (dp - Gp[j]) * A[j];
    }
    R -= double(md);
    T lD = T(0);
    auto column = [&](int k) {
        T g[3];
        gradN_c<SP>(IP, k, Ji, g);
        T lf = T(0);
#pragma unroll
        for (int j = 0; j < 3; ++j) lf += -rho * duip[j] * g[j] * A[j];
        lD -= lf;
        return lf;
    };
#pragma unroll
    for (int k = 1; k < 5; ++k) {  // the row node's SCS neighbours 1 (across ip 0), 3, 4
        if (k == 2) continue;
        const T lf = column(k);
        const int sp = slot_of(sl, c_rot.perm[r][k]);
        if (sp == 0xFF || sp == dof) D += double(lf);
        else scalar_scatter(a.vals, sr, lt, sp, double(lf));
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {  // kept only by full-stencil rows
        const int k = kFar[p];
        if (kHex.dn[SP][IP][k][0] == 0.0 && kHex.dn[SP][IP][k][1] == 0.0 && kHex.dn[SP][IP][k][2] == 0.0)
            continue;  // zero coefficient at this ip
        const int sp = slot_of(sl, c_rot.perm[r][k]);
        if (sp == 0xFF) continue;
        const T lf = column(k);
        if (sp == dof) D += double(lf);
        else scalar_scatter(a.vals, sr, lt, sp, double(lf));
    }
    D += double(lD);
}

// element e's contribution to the p' row of its local node la: gathered
// once, relabelled so that la is node 0, its three SCS ips touching la added
// as ip 0 of the three relabellings
// Nodal values in registers only for the FP32 form with the full
// interpolation (all 8 nodes weighted at every ip); with the shifted one an
// ip reads just its edge's two nodes, cheaper from L1 than 8 nodes' registers.
template <typename T, int SP, int US>
__device__ __forceinline__ void pc_incidence_us(const PcArgs& a, int la, int e, uint64_t sl,
                                                const SellRow& sr, int dof, const double xr[3],
                                                double pr, double& D, double& R, int& bad,
                                                double* lt) {
    const int* ce = a.conn + 8 * size_t(e);
    int nd[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) nd[k] = ce[c_rot.perm[3 * la][k]];
    constexpr bool REG = sizeof(T) == 4 && US == 0;
    T c[8][3], un[REG ? 8 : 1][3], dn[REG ? 8 : 1][3], pn[REG ? 8 : 1][1], Gp[3] = {0, 0, 0};
    gather_rel(a.xyz, nd, xr, c);
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const size_t o = 3 * size_t(nd[n]);
        const T w = T(a.scvw[8 * size_t(e) + c_rot.perm[3 * la][n]]);
#pragma unroll
        for (int j = 0; j < 3; ++j) Gp[j] += w * T(a.gradP[o + j]);
        if constexpr (REG) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                un[n][j] = T(a.U[o + j]);
                dn[n][j] = T(a.du[o + j]);
            }
            pn[n][0] = T(a.p[nd[n]] - pr);
        }
    }
#pragma unroll 1
    for (int q = 0; q < 3; ++q) {
        pc_scs<T, SP, REG, US>(a, 3 * la + q, nd, pr, c, un, dn, pn, Gp, sl, sr, dof, D, R, bad, lt);
        rotate_nodes<T, 3>(c);
        if constexpr (REG) {
            rotate_nodes<T, 3>(un);
            rotate_nodes<T, 3>(dn);
            rotate_nodes<T, 1>(pn);
        } else {
            int t[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) t[k] = nd[kRho[k]];
#pragma unroll
            for (int k = 0; k < 8; ++k) nd[k] = t[k];
        }
    }
}

template <typename T, int SP>
__device__ __forceinline__ void pc_incidence(const PcArgs& a, int la, int e, uint64_t sl,
                                             const SellRow& sr, int dof, const double xr[3],
                                             double pr, double& D, double& R, int& bad,
                                             double* lt) {
    if (a.u_shifted) pc_incidence_us<T, SP, 1>(a, la, e, sl, sr, dof, xr, pr, D, R, bad, lt);
    else pc_incidence_us<T, SP, 0>(a, la, e, sl, sr, dof, xr, pr, D, R, bad, lt);
}

// the SCS ips of a tet4 / wedge6 / pyramid5 element touching the p' row of
// its local node la (pressureCorrectionAssemblerElemTerms.cpp:390-610 per
// ip), from the row's side: -rho du_ip grad N . A per column, the
// Rhie-Chow mass flow on the rhs (pressures relative to the row's)
template <typename T, int TOPO>
__device__ void pc_generic(const PcArgs& a, int la, int e, uint64_t sl, const SellRow& sr, int dof,
                           const double xr[3], double pr, int sp, double& D, double& R, int& bad,
                           double* lt) {
    constexpr int npe = topo::Traits<TOPO>::npe;
    const int* ce = a.conn + 8 * size_t(e);
    const double* geo = a.geo + a.geo_off[e];
    T Gp[3] = {0, 0, 0};
    for (int k = 0; k < npe; ++k) {
        const size_t o = 3 * size_t(ce[k]);
        const T w = T(a.scvw[8 * size_t(e) + k]);
        for (int d = 0; d < 3; ++d) Gp[d] += w * T(a.gradP[o + d]);
    }
    const T rho = T(a.rho);
#pragma unroll 1
    for (int ip = 0; ip < topo::Traits<TOPO>::nscs; ++ip) {
        const int* lr = topo::lr<TOPO>(ip);
        if (lr[0] != la && lr[1] != la) continue;
        const T sg = lr[0] == la ? T(1) : T(-1);
        T A[3], g[npe][3], Nc[npe], Nu[npe];
        topo::ip_geom<TOPO>(geo, ip, sp, a.u_shifted, A, g, Nc, Nu);
        for (int d = 0; d < 3; ++d) A[d] *= sg;
        T uip[3] = {0, 0, 0}, duip[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
        for (int k = 0; k < npe; ++k) {
            const size_t o = 3 * size_t(ce[k]);
            const T pn = T(a.p[ce[k]] - pr);
            for (int j = 0; j < 3; ++j) {
                uip[j] += Nu[k] * T(a.U[o + j]);
                duip[j] += Nu[k] * T(a.du[o + j]);
                dp[j] += g[k][j] * pn;
            }
        }
        T md = T(0);
        for (int j = 0; j < 3; ++j) {
            md += rho * uip[j] * A[j];
            md -= rho * duip[j] * (dp[j] - Gp[j]) * A[j];
        }
        R -= double(md);
#pragma unroll
        for (int ic = 0; ic < npe; ++ic) {
            T lf = T(0);
            for (int j = 0; j < 3; ++j) lf += -rho * duip[j] * g[ic][j] * A[j];
            const int s = slot_of(sl, ic);
            if (s == 0xFF || s == dof) D += double(lf);
            else scalar_scatter(a.vals, sr, lt, s, double(lf));
        }
    }
}

// inlet / outlet / opening face ip (f, fk) nearest to this p' row, in FP64
// in both forms.  Inlet (pressureCorrectionAssemblerElemBoundaryConditions.
// cpp:923-1393): the specified mass flow rho U_bc . A leaves the rhs.
// Outlet (:1895-2530), unless reversed, and opening (:3026-3560), at every
// ip: the Rhie-Chow pressure term -rho du_ip grad N . A for the columns off
// the side (the side's pressure is specified) and the Rhie-Chow mass flow
// on the rhs.
template <int SP>
__device__ __forceinline__ void pc_boundary(const PcArgs& a, int pk, uint64_t sl,
                                            const SellRow& sr, int dof, const double xr[3],
                                            double& D, double& R, int& bad) {
    const int f = pk >> 2, fk = pk & 3;
    const int bc = a.bf_bc[f], t = a.bc_type[bc];
    if (t != OAGPU_BC_INLET && t != OAGPU_BC_OUTLET && t != OAGPU_BC_OPENING) return;
    if (t == OAGPU_BC_OUTLET && a.rfl[4 * size_t(f) + fk] != 0) return;
    const int e = a.bf_elem[f], ord = a.bf_ord[f];
    const int* ce = a.conn + 8 * size_t(e);
    FaceIp<double> fi;
    face_ip<true>(a.sg, a.xyz, ce, e, f, ord, fk, t == OAGPU_BC_INLET ? a.u_shifted : SP, xr, fi,
                  bad);
    const int us = a.u_shifted ? 1 : 0;
    double rip = 0;
    for (int q = 0; q < fi.npf; ++q) rip += fi.shape(us, fk, q) * a.rho;
    if (t == OAGPU_BC_INLET) {
        double m = 0;
        for (int j = 0; j < 3; ++j) m += (rip * a.bc_U[3 * bc + j]) * fi.A[j];
        R -= m;
        return;
    }
    double du[3] = {0, 0, 0};
    for (int q = 0; q < fi.npf; ++q) {
        const double r = fi.shape(us, fk, q);
        for (int j = 0; j < 3; ++j) du[j] += r * a.du[3 * size_t(ce[fi.sn[q]]) + j];
    }
#pragma unroll 1
    for (int ic = 0; ic < fi.npe; ++ic) {
        if (fi.on_side(ic)) continue;
        double g[3];
        fi.grad(ic, g);
        double lf = 0;
        for (int j = 0; j < 3; ++j) lf += -rip * du[j] * g[j] * fi.A[j];
        const int sp = slot_of(sl, ic);
        if (sp == 0xFF || sp == dof) D += lf;
        else *sell_val<1>(a.vals, sr, sp, 0) += lf;
    }
    R -= outlet_rhie_chow(fi, ce, fk, us, a.U, a.du, a.p, a.gradP, a.rho, a.bc_p[bc]);
}

// The inlet / outlet / opening face ips of the p' rows in io_row (each row
// one lane, its stored row updated in place), after k_pcorr_rows.  A
// pressure-reference row exists only without outlets and openings, where
// the only face term is the inlet rhs, which the reference row drops: it is
// skipped.
template <int SP>
__global__ void __launch_bounds__(256) k_pcorr_boundary(const PcArgs a, int nrow,
                                                        const int* io_row) {
    const int i = int(xcd_block() * blockDim.x + threadIdx.x);
    if (i >= nrow) return;
    const int row = io_row[i];
    if (row == a.p_ref) return;
    const SellRow sr = sell_row<1>(a.slice_ptr, a.slice_w, row);
    const int dof = a.diag_off[row];
    const double xr[3] = {a.xyz[3 * size_t(row)], a.xyz[3 * size_t(row) + 1],
                          a.xyz[3 * size_t(row) + 2]};
    int bad = 0;
    double D = 0.0, R = 0.0;
#pragma unroll 1
    for (int k = a.binc_ptr[row]; k < a.binc_ptr[row + 1]; ++k)
        pc_boundary<SP>(a, a.binc[k], a.binc_slots[k], sr, dof, xr, D, R, bad);
    *sell_val<1>(a.vals, sr, dof, 0) += D;
    a.b[row] += R;
    if (bad) *a.err = 1;
}

// pressureCorrectionAssembler interior terms for one row, with the
// pressure-reference row (pressureCorrectionAssembler.cpp:19-63; none when
// an outlet specifies the pressure level)
// At most 2 waves per SIMD: a smaller window of rows in flight keeps more of
// their gathered nodal data in the XCD's L2 (FP32 form 1.68 -> 1.59 ms on the
// 128^3 cavity against the compiler's 3 waves).
template <typename T, int SP, bool MX>
__device__ __forceinline__ void pcorr_row(const PcArgs& a, double* lt) {
    const int row = ordered_item(a.slice_order, a.nslice, a.nn);
    if (row < 0) return;
    const SellRow sr = sell_row<1>(a.slice_ptr, a.slice_w, row);
    const int dof = a.diag_off[row];
    const double xr[3] = {a.xyz[3 * size_t(row)], a.xyz[3 * size_t(row) + 1],
                          a.xyz[3 * size_t(row) + 2]};
    const double pr = a.p[row];
    int bad = 0;
    double D = 0.0, R = 0.0;
    scalar_row_zero(a.vals, sr, lt, dof);
#pragma unroll 1
    for (int la = 0; la < 8; ++la) {
        const size_t o = size_t(la) * a.nn + row;
        const int e = a.inc_la[o];
        if (e >= 0) pc_incidence<T, SP>(a, la, e, a.slots_la[o], sr, dof, xr, pr, D, R, bad, lt);
    }
#pragma unroll 1
    for (int k = a.ext_ptr[row]; k < a.ext_ptr[row + 1]; ++k)
        pc_incidence<T, SP>(a, a.ext[k] & 7, a.ext[k] >> 3, a.ext_slots[k], sr, dof, xr, pr, D, R,
                            bad, lt);
    if constexpr (MX) {
#pragma unroll 1
        for (int k = a.gen_ptr[row]; k < a.gen_ptr[row + 1]; ++k) {
            const int e = a.gen[k] >> 3, la = a.gen[k] & 7, t = a.etype[e];
            if (t == OAGPU_TET4)
                pc_generic<T, topo::kTet>(a, la, e, a.gen_slots[k], sr, dof, xr, pr, SP, D, R, bad, lt);
            else if (t == OAGPU_WED6)
                pc_generic<T, topo::kWed>(a, la, e, a.gen_slots[k], sr, dof, xr, pr, SP, D, R, bad, lt);
            else
                pc_generic<T, topo::kPyr>(a, la, e, a.gen_slots[k], sr, dof, xr, pr, SP, D, R, bad, lt);
        }
    }
    if (row == a.p_ref) {
        scalar_row_zero(a.vals, sr, lt, dof);
        R = 0.0;
    }
    scalar_row_store(a.vals, sr, lt, dof);
    *sell_val<1>(a.vals, sr, dof, 0) = D;
    a.b[row] = R;
    if (bad) *a.err = 1;
}

template <typename T, int SP, bool MX>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) k_pcorr_rows(const PcArgs a) {
    __shared__ double tile[kLdsSlots * 64];
    pcorr_row<T, SP, MX>(a, tile + threadIdx.x);
}

// the same with the compiler's occupancy (an autotuning candidate)
template <typename T, int SP, bool MX>
__global__ void __launch_bounds__(64) k_pcorr_rows_occ(const PcArgs a) {
    __shared__ double tile[kLdsSlots * 64];
    pcorr_row<T, SP, MX>(a, tile + threadIdx.x);
}

void launch_assemble_pcorr(Flow& f) {
    DevMesh& m = *f.m;
    PcArgs a;
    a.nn = m.nn;
    a.nslice = m.nslice;
    a.slice_order = m.slice_order.p;
    a.xyz = m.xyz.p;
    a.conn = m.conn.p;
    a.slice_ptr = m.slice_ptr.p;
    a.slice_w = m.slice_w.p;
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
    a.binc_ptr = m.binc_ptr.p;
    a.binc = m.binc.p;
    a.binc_slots = m.binc_slots.p;
    a.bf_elem = m.bf_elem.p;
    a.bf_ord = m.bf_ord.p;
    a.bf_bc = m.bf_bc.p;
    a.sg = side_geo(m);
    a.bc_type = f.bc_type.p;
    a.bc_U = f.bc_U.p;
    a.bc_p = f.bc_p.p;
    a.rfl = f.rfl.p;
    a.gen_ptr = m.gen_ptr.p;
    a.gen = m.gen.p;
    a.gen_slots = m.gen_slots.p;
    a.geo = m.geo.p;
    a.geo_off = m.geo_off.p;
    a.etype = m.etype.p;
    a.rho = f.prm.rho;
    a.u_shifted = f.prm.u_shifted;
    // domain::pressureLevelRequired: none with an outlet or an opening
    a.p_ref = f.has_outlet || f.has_opening ? -1 : f.prm.p_ref_node;
    a.vals = f.A1.p;
    a.b = f.b1.p;
    a.err = f.err.p;
    const unsigned B = 64, G = xcd_grid(m.nn, B);
    const bool mixed = f.prm.precision == OAGPU_MIXED;
    const bool occ = f.tune.pc_occupancy != 0;
    auto pick = [&](auto mx) {
        constexpr bool MX = decltype(mx)::value;
        if (f.prm.pgrad_shifted) {
            if (mixed) (occ ? k_pcorr_rows_occ<float, 1, MX> : k_pcorr_rows<float, 1, MX>)<<<G, B, 0, f.stream>>>(a);
            else (occ ? k_pcorr_rows_occ<double, 1, MX> : k_pcorr_rows<double, 1, MX>)<<<G, B, 0, f.stream>>>(a);
        } else {
            if (mixed) (occ ? k_pcorr_rows_occ<float, 0, MX> : k_pcorr_rows<float, 0, MX>)<<<G, B, 0, f.stream>>>(a);
            else (occ ? k_pcorr_rows_occ<double, 0, MX> : k_pcorr_rows<double, 0, MX>)<<<G, B, 0, f.stream>>>(a);
        }
    };
    clock_begin(f, kClkPcorr);
    if (m.mixed) pick(std::true_type{});
    else pick(std::false_type{});
    clock_end(f, kClkPcorr);
    OAGPU_HIP(hipGetLastError());
    if (f.has_io) {
        const unsigned GB = xcd_grid(size_t(f.nio_row), 256);
        if (f.prm.pgrad_shifted) k_pcorr_boundary<1><<<GB, 256, 0, f.stream>>>(a, f.nio_row, f.io_row.p);
        else k_pcorr_boundary<0><<<GB, 256, 0, f.stream>>>(a, f.nio_row, f.io_row.p);
        OAGPU_HIP(hipGetLastError());
    }
}

// ------------------------------------------------------ k / omega (SST)
namespace sstc {  // shearStressTransportModel.h:21-37, RANSModel.h:20-22, turbulenceModel.h:20
constexpr double kSigK1 = 0.85, kSigK2 = 1.0, kSigW1 = 0.5, kSigW2 = 0.856;
constexpr double kBeta1 = 0.075, kBeta2 = 0.0828, kGam1 = 0.5532, kGam2 = 0.4403;
constexpr double kA1 = 0.31, kBetaStar = 0.09, kProdLim = 10.0, kCmu = 0.09;
}  // namespace sstc

struct ScArgs {
    int nn, nslice;
    const int* slice_order;
    const int* wnode;
    const double* xyz;
    const int* conn;
    const int* slice_ptr;
    const int* slice_w;
    const int* diag_off;
    const int* inc_la;
    const uint64_t* slots_la;
    const int* ext_ptr;
    const int* ext;
    const uint64_t* ext_slots;
    const int* bnode;
    const double* vol;
    const double* div;
    const double* mdot;
    const double* phi;  // k or omega
    const double* k;
    const double* w;
    const double* kprev;
    const double* gradU;
    const double* gradK;
    const double* gradW;
    const double* ymin;
    const double* mut;
    const double* F1;
    const double* Pk;
    const double* beta;   // blending factor of phi (high resolution) or null
    const double* gphi;   // grad phi [nn][3]
    // inlet / outlet face ips (boundary-row kernel only)
    const int* binc_ptr;
    const int* binc;
    const uint64_t* binc_slots;
    const int* bf_elem;
    const int* bf_ord;
    const int* bf_bc;
    const int* bc_type;
    const double* bc_phi;  // inlet k or omega per boundary id
    const double* msd;
    const int* rfl;
    const int* gen_ptr;    // tet4 / wedge6 / pyramid5 incidences (mixed meshes)
    const int* gen;
    const uint64_t* gen_slots;
    const double* geo;          // static ip geometry of the non-hex elements
    const long long* geo_off;   // [ne] offset into geo
    SideGeo sg;                 // boundary sides of the non-hex elements
    const int* etype;
    double rho, mu, dt, urf_inv, sig1, sig2;
    const double* phiold;   // transient: phi of the previous two time levels
    const double* phiold2;
    double tc[3];
    int which, us;
    double* vals;
    double* b;
    int* err;
};

// node terms of the k / omega rows (turbulentKineticEnergySSTAssemblerNodeTerms.cpp:13-128,
// turbulentEddyFrequencySSTAssemblerNodeTerms.cpp:13-219), steady, FP64
__device__ void sst_node_terms(const ScArgs& a, int i, double& lhs, double& rhs) {
    using namespace sstc;
    const double rho = a.rho, vol = a.vol[i], div = a.div[i], w = a.w[i];
    const double phi = a.phi[i];
    double l = rho * vol / a.dt, r = 0.0;
    if (a.phiold) r -= vol * (a.tc[0] * phi + a.tc[1] * a.phiold[i] + a.tc[2] * a.phiold2[i]);
    if (div < 0) l += -div;
    r -= -div * phi;
    if (a.which == 0) {
        const double Dk = kBetaStar * rho * w * a.k[i];
        double Pk = a.Pk[i];
        if (Pk > kProdLim * Dk) Pk = kProdLim * Dk;
        r += (Pk - Dk) * vol;
        l += kBetaStar * rho * w * vol;
    } else {
        const double mu = a.mu, y = a.ymin[i], kPre = a.kprev[i];
        double g[9];
        for (int q = 0; q < 9; ++q) g[q] = a.gradU[9 * size_t(i) + q];
        double P = 0.0, s = 0.0;
        for (int p = 0; p < 3; ++p)
            for (int j = 0; j < 3; ++j) {
                P += g[3 * p + j] * (g[3 * p + j] + g[3 * j + p]);
                const double rs = 0.5 * (g[3 * p + j] + g[3 * j + p]);
                s += rs * rs;
            }
        P = fmax(P, 0.0);
        const double sij = sqrt(2.0 * s);
        const double arg2 = fmin(fmax(2 * sqrt(kPre) / (kCmu * w * y + kSmall),
                                      500 * mu / (y * y * rho * w + kSmall)),
                                 100.0);
        const double F2 = tanh(arg2 * arg2);
        P = fmin(P, ((kProdLim * kBetaStar) / kA1) * w * fmax(kA1 * w, F2 * sqrt(2.0) * sij));
        double cross = 0.0;
        for (int j = 0; j < 3; ++j) cross += a.gradK[3 * size_t(i) + j] * a.gradW[3 * size_t(i) + j];
        const double F1 = a.F1[i], om = 1.0 - F1;
        const double beta = F1 * kBeta1 + om * kBeta2;
        const double gamma = F1 * kGam1 + om * kGam2;
        const double sigmaD = 2.0 * om * kSigW2;
        const double Pw = gamma * rho * P;
        const double Dw = beta * rho * w * w;
        const double Sw = sigmaD * rho * cross / (w + kSmall);
        r += (Pw - Dw + Sw) * vol;
        l += (2.0 * beta * rho * w + fmax(Sw / (w + kSmall), 0.0)) * vol;
    }
    lhs = l;
    rhs = r;
}

// SCS ip 0 of the relabelled element r = (la, q) added to the k / omega row
// of la (phiAssemblerElemTerms.hpp:171-406 for one ip; the row node is its
// left node): upwind advection with the limited deferred correction and
// diffusion with Gamma = mu + sigma(F1) mu_t interpolated to the ip.  The
// explicit diffusion flux uses grad phi at the ip; the lumped columns take
// minus the kept columns' coefficients (the shape-function gradients sum to
// zero).  Nodal values: in registers, rotated with the coordinates (REG),
// or read per ip from L1/L2.  NE equations (k, or k and omega) share the
// ip's geometry and mass flow; each keeps its own arithmetic, term for term
// that of a single equation.
template <typename T, int SG, bool HR, bool REG, int NE, int US>
__device__ __forceinline__ void sc_scs(const ScArgs* as, int r, int e, const int nd[8],
                                       const T c[8][3], const T (*ph)[8][1], const T (*gm)[8][1],
                                       uint64_t sl, const SellRow& sr, int dof, double* D,
                                       double* R, int& bad, double* const* lt) {
    constexpr int IP = 0;
    const ScArgs& a = as[0];
    const T m = T(c_rot.sg[r]) * T(a.mdot[12 * size_t(e) + c_rot.ip[r]]);
    T A[3], Ji[9];
    area_c(IP, c, A);
    jinv_c<SG>(IP, c, Ji, bad);
    T Gip[NE], P[NE][3];
#pragma unroll
    for (int q = 0; q < NE; ++q) Gip[q] = P[q][0] = P[q][1] = P[q][2] = T(0);
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        // the interpolation US's zero weights and the gradient's zero rows drop
        // out, and with them their nodes' loads
        const bool wz = kHex.N[US][IP][n] == 0.0;
        const bool gz = kHex.dn[SG][IP][n][0] == 0.0 && kHex.dn[SG][IP][n][1] == 0.0 &&
                        kHex.dn[SG][IP][n][2] == 0.0;
        double F1 = 0.0, mt = 0.0;
        if constexpr (!REG) {
            if (!wz) {
                F1 = a.F1[nd[n]];
                mt = a.mut[nd[n]];
            }
        }
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            if (!wz) {
                T gv;
                if constexpr (REG) gv = gm[q][n][0];
                else gv = T(a.mu + mt * (F1 * as[q].sig1 + (1.0 - F1) * as[q].sig2));
                Gip[q] += T(kHex.N[US][IP][n]) * gv;
            }
            if (!gz) {
                T pv;
                if constexpr (REG) pv = ph[q][n][0];
                else pv = T(as[q].phi[nd[n]]);
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    if (kHex.dn[SG][IP][n][d] != 0.0) P[q][d] += T(kHex.dn[SG][IP][n][d]) * pv;
            }
        }
    }
    // upwind advection plus the limited deferred correction
    // (phiAssemblerElemTerms.hpp:291-340)
    const bool fw = m > T(0);
    const size_t nu = size_t(fw ? nd[0] : nd[1]);
    T rl[NE];
#pragma unroll
    for (int q = 0; q < NE; ++q) {
        T up = T(as[q].phi[nu]);
        if constexpr (HR) {
            T dc = T(0);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                T xi = T(0);
#pragma unroll
                for (int n = 0; n < 8; ++n) xi += T(kHex.N[0][IP][n]) * c[n][d];
                dc += T(as[q].beta[nu]) * (xi - (fw ? c[0][d] : c[1][d])) * T(as[q].gphi[3 * nu + d]);
            }
            up += dc;
        }
        rl[q] = -(m * up);
        // explicit diffusion flux: + Gamma grad phi . A
        T f = T(0);
#pragma unroll
        for (int j = 0; j < 3; ++j) f += (P[q][0] * Ji[j] + P[q][1] * Ji[3 + j] + P[q][2] * Ji[6 + j]) * A[j];
        rl[q] += Gip[q] * f;
    }
    T lD[NE];
#pragma unroll
    for (int q = 0; q < NE; ++q) lD[q] = T(0);
    T lf[NE];
    auto column = [&](int k) {
        T g[3];
        gradN_c<SG>(IP, k, Ji, g);
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            lf[q] = T(0);
#pragma unroll
            for (int j = 0; j < 3; ++j) lf[q] += -Gip[q] * g[j] * A[j];
            lD[q] -= lf[q];
        }
    };
#pragma unroll
    for (int k = 1; k < 5; ++k) {  // the row node's SCS neighbours 1 (across ip 0), 3, 4
        if (k == 2) continue;
        column(k);
        const int sp = slot_of(sl, c_rot.perm[r][k]);
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            if (k == 1) lf[q] += T(0.5) * (m - fabs(m));
            if (sp == 0xFF || sp == dof) D[q] += double(lf[q]);
            else scalar_scatter(as[q].vals, sr, lt[q], sp, double(lf[q]));
        }
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {  // kept only by full-stencil rows
        const int k = kFar[p];
        if (kHex.dn[SG][IP][k][0] == 0.0 && kHex.dn[SG][IP][k][1] == 0.0 && kHex.dn[SG][IP][k][2] == 0.0)
            continue;  // zero coefficient at this ip
        const int sp = slot_of(sl, c_rot.perm[r][k]);
        if (sp == 0xFF) continue;
        column(k);
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            if (sp == dof) D[q] += double(lf[q]);
            else scalar_scatter(as[q].vals, sr, lt[q], sp, double(lf[q]));
        }
    }
#pragma unroll
    for (int q = 0; q < NE; ++q) {
        D[q] += double(lD[q] + T(0.5) * (m + fabs(m)));
        R[q] += double(rl[q]);
    }
}

// element e's contribution to the k / omega row(s) of its local node la:
// gathered once, relabelled so that la is node 0, its three SCS ips touching
// la added as ip 0 of the three relabellings
template <typename T, int SG, bool HR, int NE>
__device__ __forceinline__ void sc_incidence(const ScArgs* as, int la, int e, uint64_t sl,
                                             const SellRow& sr, int dof, const double xr[3],
                                             double* D, double* R, int& bad, double* const* lt) {
    constexpr bool REG = sizeof(T) == 4;
    const ScArgs& a = as[0];
    const int* ce = a.conn + 8 * size_t(e);
    int nd[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) nd[k] = ce[c_rot.perm[3 * la][k]];
    T c[8][3], ph[NE][REG ? 8 : 1][1], gm[NE][REG ? 8 : 1][1];
    gather_rel(a.xyz, nd, xr, c);
    if constexpr (REG) {
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            const double F1 = a.F1[nd[n]], mt = a.mut[nd[n]];
#pragma unroll
            for (int q = 0; q < NE; ++q) {
                ph[q][n][0] = T(as[q].phi[nd[n]]);
                gm[q][n][0] = T(a.mu + mt * (F1 * as[q].sig1 + (1.0 - F1) * as[q].sig2));
            }
        }
    }
#pragma unroll 1
    for (int q3 = 0; q3 < 3; ++q3) {
        if (a.us)
            sc_scs<T, SG, HR, REG, NE, 1>(as, 3 * la + q3, e, nd, c, reinterpret_cast<const T(*)[8][1]>(ph),
                                          reinterpret_cast<const T(*)[8][1]>(gm), sl, sr, dof, D, R, bad, lt);
        else
            sc_scs<T, SG, HR, REG, NE, 0>(as, 3 * la + q3, e, nd, c, reinterpret_cast<const T(*)[8][1]>(ph),
                                          reinterpret_cast<const T(*)[8][1]>(gm), sl, sr, dof, D, R, bad, lt);
        rotate_nodes<T, 3>(c);
        if constexpr (REG) {
#pragma unroll
            for (int q = 0; q < NE; ++q) {
                rotate_nodes<T, 1>(ph[q]);
                rotate_nodes<T, 1>(gm[q]);
            }
        }
        int t[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) t[k] = nd[kRho[k]];
#pragma unroll
        for (int k = 0; k < 8; ++k) nd[k] = t[k];
    }
}

// the SCS ips of a tet4 / wedge6 / pyramid5 element touching the k / omega
// row(s) of its local node la (phiAssemblerElemTerms.hpp:171-406 per ip),
// from the row's side: upwind advection with the limited deferred
// correction, diffusion -Gamma_ip grad N . A per column with its explicit
// flux on the rhs; NE equations share the ip's geometry and mass flow
template <typename T, int TOPO, bool HR, int NE>
__device__ void sc_generic(const ScArgs* as, int la, int e, uint64_t sl, const SellRow& sr, int dof,
                           const double xr[3], int sg_shift, double* D, double* R, int& bad,
                           double* const* lt) {
    constexpr int npe = topo::Traits<TOPO>::npe;
    const ScArgs& a = as[0];
    const int* ce = a.conn + 8 * size_t(e);
    const double* geo = a.geo + a.geo_off[e];
    T c[npe][3], ph[NE][npe], gm[NE][npe];
    for (int k = 0; k < npe; ++k) {
        if constexpr (HR)
            for (int d = 0; d < 3; ++d) c[k][d] = T(a.xyz[3 * size_t(ce[k]) + d] - xr[d]);
        const double F1 = a.F1[ce[k]], mt = a.mut[ce[k]];
        for (int q = 0; q < NE; ++q) {
            ph[q][k] = T(as[q].phi[ce[k]]);
            gm[q][k] = T(a.mu + mt * (F1 * as[q].sig1 + (1.0 - F1) * as[q].sig2));
        }
    }
#pragma unroll 1
    for (int ip = 0; ip < topo::Traits<TOPO>::nscs; ++ip) {
        const int* lr = topo::lr<TOPO>(ip);
        if (lr[0] != la && lr[1] != la) continue;
        const bool left = lr[0] == la;
        const int o = left ? lr[1] : lr[0];
        const T sgn = left ? T(1) : T(-1);
        T A[3], g[npe][3], Nc[npe], Nu[npe];
        topo::ip_geom<TOPO>(geo, ip, sg_shift, a.us, A, g, Nc, Nu);
        for (int d = 0; d < 3; ++d) A[d] *= sgn;
        const T m = sgn * T(a.mdot[12 * size_t(e) + ip]);
        const int ul = m > T(0) ? la : o;
        const T aS = T(0.5) * (m + fabs(m)), aO = T(0.5) * (m - fabs(m));
        for (int q = 0; q < NE; ++q) {
            T Gip = T(0);
            for (int k = 0; k < npe; ++k) Gip += Nu[k] * gm[q][k];
            T up = topo::pick<npe>(ph[q], ul);
            if constexpr (HR) {
                const size_t nu = size_t(ce[ul]);
                T dc = T(0), cu[3];
                topo::pick_row<npe>(c, ul, cu);
                for (int d = 0; d < 3; ++d) {
                    T xi = T(0);
                    for (int k = 0; k < npe; ++k) xi += Nc[k] * c[k][d];
                    dc += T(as[q].beta[nu]) * (xi - cu[d]) * T(as[q].gphi[3 * nu + d]);
                }
                up += dc;
            }
            T rl = -(m * up), lD = aS;
#pragma unroll
            for (int ic = 0; ic < npe; ++ic) {
                T lf = T(0);
                for (int j = 0; j < 3; ++j) lf += -Gip * g[ic][j] * A[j];
                rl -= lf * ph[q][ic];
                if (ic == o) lf += aO;
                const int s = slot_of(sl, ic);
                if (s == 0xFF || s == dof) lD += lf;
                else scalar_scatter(as[q].vals, sr, lt[q], s, double(lf));
            }
            D[q] += double(lD);
            R[q] += double(rl);
        }
    }
}

// inlet / outlet / opening face ip (f, fk) nearest to this k / omega row,
// FP64 in both forms, skipped while the ip is reversed.  Inlet
// (phiAssemblerElemBoundaryConditions.hpp:894-1229, fixed value): the
// advective flux msd phi_bc and the diffusive flux with phi_bc on the
// side's nodes (their columns dropped), Gamma interpolated from the side's
// nodes.  Outlet (:1231-1437, zero gradient) and opening (:1439-1760, whose
// ips are never reversed): implicit upwind, msd on the diagonal and
// msd phi_row on the rhs.
template <int SG>
__device__ __forceinline__ void sc_boundary(const ScArgs& a, int pk, uint64_t sl,
                                            const SellRow& sr, int dof, int row,
                                            const double xr[3], bool offd, double& D, double& R,
                                            int& bad) {
    const int f = pk >> 2, fk = pk & 3;
    const int bc = a.bf_bc[f], t = a.bc_type[bc];
    if (t != OAGPU_BC_INLET && t != OAGPU_BC_OUTLET && t != OAGPU_BC_OPENING) return;
    if (a.rfl[4 * size_t(f) + fk] != 0) return;
    const double tm = a.msd[4 * size_t(f) + fk];
    if (t != OAGPU_BC_INLET) {
        R -= tm * a.phi[row];
        D += tm;
        return;
    }
    const double pbc = a.bc_phi[bc];
    R -= tm * pbc;
    const int e = a.bf_elem[f], ord = a.bf_ord[f];
    const int* ce = a.conn + 8 * size_t(e);
    FaceIp<double> fi;
    face_ip<true>(a.sg, a.xyz, ce, e, f, ord, fk, SG, xr, fi, bad);
    double Gb = 0;
    for (int q = 0; q < fi.npf; ++q) {
        const int n = ce[fi.sn[q]];
        const double F1 = a.F1[n];
        Gb += fi.shape(a.us, fk, q) * (a.mu + a.mut[n] * (F1 * a.sig1 + (1.0 - F1) * a.sig2));
    }
#pragma unroll 1
    for (int ic = 0; ic < fi.npe; ++ic) {
        double g[3];
        fi.grad(ic, g);
        double lf = 0;
        for (int j = 0; j < 3; ++j) lf += -Gb * g[j] * fi.A[j];
        if (fi.on_side(ic)) {
            R -= lf * pbc;
            continue;
        }
        R -= lf * a.phi[ce[ic]];
        const int sp = slot_of(sl, ic);
        if (sp == 0xFF || sp == dof) D += lf;
        else if (offd) *sell_val<1>(a.vals, sr, sp, 0) += lf;
    }
}

// The inlet / outlet / opening face ips of the k / omega rows in io_row (each row one
// lane, its stored row updated in place), after k_scalar_rows: the diagonal
// part relaxed by urf, the rhs part by the boundary-row factor 0.75 (the
// rows are boundary rows).  Omega rows fixed at no-slip wall nodes keep
// only the diagonal part (their off-diagonals and rhs are zero).
template <int SG>
__global__ void __launch_bounds__(256) k_scalar_boundary(const ScArgs a, int nrow,
                                                         const int* io_row) {
    const int i = int(xcd_block() * blockDim.x + threadIdx.x);
    if (i >= nrow) return;
    const int row = io_row[i];
    const bool fixed = a.which == 1 && a.wnode[row];
    const SellRow sr = sell_row<1>(a.slice_ptr, a.slice_w, row);
    const int dof = a.diag_off[row];
    const double xr[3] = {a.xyz[3 * size_t(row)], a.xyz[3 * size_t(row) + 1],
                          a.xyz[3 * size_t(row) + 2]};
    int bad = 0;
    double D = 0.0, R = 0.0;
#pragma unroll 1
    for (int k = a.binc_ptr[row]; k < a.binc_ptr[row + 1]; ++k)
        sc_boundary<SG>(a, a.binc[k], a.binc_slots[k], sr, dof, row, xr, !fixed, D, R, bad);
    *sell_val<1>(a.vals, sr, dof, 0) += D * a.urf_inv;
    if (!fixed) a.b[row] += R * 0.75;
    if (bad) *a.err = 1;
}

// the k / omega equations assembled by one launch: each its own arguments
// (system, nodal values, diffusivity, relaxation), the mesh and the shared
// fields those of eq[0]
template <int NE>
struct ScSet {
    ScArgs eq[NE];
};

// one k / omega row of each of the NE equations: node terms and sources,
// interior terms, relaxation (phiAssembler.h:966-1006), boundary rhs
// relaxation x 0.75 (turbulentKineticEnergyAssembler.h:44-48,
// phiAssembler.h:1008-1152), and for omega the no-slip wall rows reduced to
// their diagonal with b = 0.  With NE = 2 (k and omega) the row's elements
// are gathered and their geometry computed once for both: the omega system
// depends only on fields the k solve does not change (its node terms use
// the k of the iteration's start, kprev).
template <typename T, int SG, bool HR, bool MX, int NE>
__global__ void __launch_bounds__(64) k_scalar_rows(const ScSet<NE> set) {
    __shared__ double tile[NE * kLdsSlots * 64];
    const ScArgs* as = set.eq;
    const ScArgs& a = as[0];
    double* lt[NE];
#pragma unroll
    for (int q = 0; q < NE; ++q) lt[q] = tile + q * kLdsSlots * 64 + threadIdx.x;
    const int row = ordered_item(a.slice_order, a.nslice, a.nn);
    if (row < 0) return;
    const SellRow sr = sell_row<1>(a.slice_ptr, a.slice_w, row);
    const int dof = a.diag_off[row];
    const double xr[3] = {a.xyz[3 * size_t(row)], a.xyz[3 * size_t(row) + 1],
                          a.xyz[3 * size_t(row) + 2]};
    int bad = 0;
    double D[NE], R[NE];
#pragma unroll
    for (int q = 0; q < NE; ++q) {
        sst_node_terms(as[q], row, D[q], R[q]);
        scalar_row_zero(as[q].vals, sr, lt[q], dof);
    }
#pragma unroll 1
    for (int la = 0; la < 8; ++la) {
        const size_t o = size_t(la) * a.nn + row;
        const int e = a.inc_la[o];
        if (e >= 0) sc_incidence<T, SG, HR, NE>(as, la, e, a.slots_la[o], sr, dof, xr, D, R, bad, lt);
    }
#pragma unroll 1
    for (int k = a.ext_ptr[row]; k < a.ext_ptr[row + 1]; ++k)
        sc_incidence<T, SG, HR, NE>(as, a.ext[k] & 7, a.ext[k] >> 3, a.ext_slots[k], sr, dof, xr, D, R,
                                    bad, lt);
    if constexpr (MX) {
#pragma unroll 1
        for (int k = a.gen_ptr[row]; k < a.gen_ptr[row + 1]; ++k) {
            const int e = a.gen[k] >> 3, la = a.gen[k] & 7, t = a.etype[e];
            if (t == OAGPU_TET4)
                sc_generic<T, topo::kTet, HR, NE>(as, la, e, a.gen_slots[k], sr, dof, xr, SG, D, R, bad, lt);
            else if (t == OAGPU_WED6)
                sc_generic<T, topo::kWed, HR, NE>(as, la, e, a.gen_slots[k], sr, dof, xr, SG, D, R, bad, lt);
            else
                sc_generic<T, topo::kPyr, HR, NE>(as, la, e, a.gen_slots[k], sr, dof, xr, SG, D, R, bad, lt);
        }
    }
#pragma unroll
    for (int q = 0; q < NE; ++q) {
        const ScArgs& aq = as[q];
        D[q] *= aq.urf_inv;
        if (a.bnode[row]) R[q] *= 0.75;
        if (aq.which == 1 && a.wnode[row]) {  // omega rows fixed at no-slip wall nodes
            scalar_row_zero(aq.vals, sr, lt[q], dof);
            R[q] = 0.0;
        }
        scalar_row_store(aq.vals, sr, lt[q], dof);
        *sell_val<1>(aq.vals, sr, dof, 0) = D[q];
        aq.b[row] = R[q];
    }
    if (bad) *a.err = 1;
}

// the arguments of the k (which 0) or omega (1) system, written to vals / b
static ScArgs scalar_args(Flow& f, int which, double* vals, double* b) {
    DevMesh& m = *f.m;
    ScArgs a;
    a.nn = m.nn;
    a.nslice = m.nslice;
    a.slice_order = m.slice_order.p;
    a.wnode = f.wnode.p;
    a.xyz = m.xyz.p;
    a.conn = m.conn.p;
    a.slice_ptr = m.slice_ptr.p;
    a.slice_w = m.slice_w.p;
    a.diag_off = m.diag_off.p;
    a.inc_la = m.inc_la.p;
    a.slots_la = m.slots_la.p;
    a.ext_ptr = m.ext_ptr.p;
    a.ext = m.ext.p;
    a.ext_slots = m.ext_slots.p;
    a.bnode = m.bnode.p;
    a.vol = m.vol.p;
    a.div = f.div.p;
    a.mdot = f.mdot.p;
    a.phi = which == 0 ? f.k.p : f.om.p;
    a.k = f.k.p;
    a.w = f.om.p;
    a.kprev = f.kprev.p;
    a.phiold = a.phiold2 = nullptr;
    a.tc[0] = a.tc[1] = a.tc[2] = 0.0;
    a.gradU = f.gradU.p;
    a.gradK = f.gradK.p;
    a.gradW = f.gradW.p;
    a.ymin = f.ymin.p;
    a.mut = f.mut.p;
    a.F1 = f.F1.p;
    a.Pk = f.Pk.p;
    a.beta = f.prm.high_res ? (which == 0 ? f.beta_k.p : f.beta_w.p) : nullptr;
    a.gphi = which == 0 ? f.gradK.p : f.gradW.p;
    a.binc_ptr = m.binc_ptr.p;
    a.binc = m.binc.p;
    a.binc_slots = m.binc_slots.p;
    a.bf_elem = m.bf_elem.p;
    a.bf_ord = m.bf_ord.p;
    a.bf_bc = m.bf_bc.p;
    a.sg = side_geo(m);
    a.bc_type = f.bc_type.p;
    a.bc_phi = which == 0 ? f.bc_k.p : f.bc_w.p;
    a.msd = f.msd.p;
    a.rfl = f.rfl.p;
    a.gen_ptr = m.gen_ptr.p;
    a.gen = m.gen.p;
    a.gen_slots = m.gen_slots.p;
    a.geo = m.geo.p;
    a.geo_off = m.geo_off.p;
    a.etype = m.etype.p;
    a.rho = f.prm.rho;
    a.mu = f.prm.mu;
    a.dt = f.prm.dt;
    if (f.tscheme) {
        OAGPU_REQUIRE(f.kold.p, "k-omega SST enabled after the first time step");
        double c[3];
        bdf_coeff(f, c);
        a.dt = f.tdt / c[0];
        for (int k = 0; k < 3; ++k) a.tc[k] = c[k] * f.prm.rho / f.tdt;
        a.phiold = which == 0 ? f.kold.p : f.wold.p;
        a.phiold2 = which == 0 ? f.kold2.p : f.wold2.p;
    }
    a.urf_inv = 1.0 / (which == 0 ? f.sst.urf_k : f.sst.urf_w);
    a.sig1 = which == 0 ? sstc::kSigK1 : sstc::kSigW1;
    a.sig2 = which == 0 ? sstc::kSigK2 : sstc::kSigW2;
    a.which = which;
    a.us = (which == 0 ? f.sst.k_shifted : f.sst.w_shifted) ? 1 : 0;
    a.vals = vals;
    a.b = b;
    a.err = f.err.p;
    return a;
}

static bool scalar_gs(const Flow& f, int which) {
    return (which == 0 ? f.sst.kgrad_shifted : f.sst.wgrad_shifted) != 0;
}

// the boundary rows (inlets, outlets, openings) of one scalar system
static void scalar_boundary(Flow& f, const ScArgs& a, bool gs) {
    if (!f.has_io) return;
    const unsigned GB = xcd_grid(size_t(f.nio_row), 256);
    if (gs) k_scalar_boundary<1><<<GB, 256, 0, f.stream>>>(a, f.nio_row, f.io_row.p);
    else k_scalar_boundary<0><<<GB, 256, 0, f.stream>>>(a, f.nio_row, f.io_row.p);
    OAGPU_HIP(hipGetLastError());
}

template <int NE>
static void scalar_rows(Flow& f, const ScSet<NE>& set, bool gs) {
    DevMesh& m = *f.m;
    const unsigned B = 64, G = xcd_grid(m.nn, B);
    const bool mixed = f.prm.precision == OAGPU_MIXED;
    const bool hr = set.eq[0].beta != nullptr;
    const unsigned lds = occupancy_lds(B, NE * kLdsSlots * 64 * sizeof(double), f.tune.cap_scalar);
    auto go = [&](auto kern) {
        clock_begin(f, kClkScalar);
        kern<<<G, B, lds, f.stream>>>(set);
        clock_end(f, kClkScalar);
    };
    auto pick = [&](auto mx) {
        constexpr bool MX = decltype(mx)::value;
        if (gs) {
            if (mixed) go(hr ? k_scalar_rows<float, 1, true, MX, NE> : k_scalar_rows<float, 1, false, MX, NE>);
            else go(hr ? k_scalar_rows<double, 1, true, MX, NE> : k_scalar_rows<double, 1, false, MX, NE>);
        } else {
            if (mixed) go(hr ? k_scalar_rows<float, 0, true, MX, NE> : k_scalar_rows<float, 0, false, MX, NE>);
            else go(hr ? k_scalar_rows<double, 0, true, MX, NE> : k_scalar_rows<double, 0, false, MX, NE>);
        }
    };
    if (m.mixed) pick(std::true_type{});
    else pick(std::false_type{});
    OAGPU_HIP(hipGetLastError());
}

void launch_assemble_scalar(Flow& f, int which) {
    OAGPU_REQUIRE(f.sst_on, "SST not enabled on this flow");
    OAGPU_REQUIRE(which == 0 || which == 1, "scalar system: 0 k, 1 omega");
    ScSet<1> set;
    set.eq[0] = scalar_args(f, which, f.A1.p, f.b1.p);
    scalar_rows<1>(f, set, scalar_gs(f, which));
    scalar_boundary(f, set.eq[0], scalar_gs(f, which));
}

bool scalar_pair_ok(const Flow& f) {
    return f.sst_on && f.tune.scalar_pair != 0 && scalar_gs(f, 0) == scalar_gs(f, 1) &&
           (f.sst.k_shifted != 0) == (f.sst.w_shifted != 0);
}

void launch_assemble_scalar_pair(Flow& f) {
    OAGPU_REQUIRE(scalar_pair_ok(f), "k and omega cannot share one assembly pass on this flow");
    DevMesh& m = *f.m;
    if (f.A1w.n < m.nsell) f.A1w.alloc(m.nsell);
    if (f.b1w.n < size_t(m.nn)) f.b1w.alloc(size_t(m.nn));
    ScSet<2> set;
    set.eq[0] = scalar_args(f, 0, f.A1.p, f.b1.p);
    set.eq[1] = scalar_args(f, 1, f.A1w.p, f.b1w.p);
    const bool gs = scalar_gs(f, 0);
    scalar_rows<2>(f, set, gs);
    scalar_boundary(f, set.eq[0], gs);
    scalar_boundary(f, set.eq[1], gs);
}

// --------------------------------------------------------------- mass flow
// initializeMassFlowRateInterior_ (flowModel.cpp:3303-3470): rho_ip U_ip . A
template <int US>
__global__ void __launch_bounds__(256) k_mdot_init(int ne, const double* xyz, const int* conn,
                                                   const int* etype, const double* U, double rho,
                                                   double* mdot) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    if (etype && etype[e] != OAGPU_HEX8) return;  // k_mdot_generic
    int nd[8];
    load_conn(conn, e, nd);
    const double xr[3] = {xyz[3 * size_t(nd[0])], xyz[3 * size_t(nd[0]) + 1],
                          xyz[3 * size_t(nd[0]) + 2]};
    double c[8][3];
    gather_rel(xyz, nd, xr, c);
#pragma unroll 1
    for (int ip = 0; ip < 12; ++ip) {
        double A[3];
        area_scs(ip, c, A);
        double u[3] = {0, 0, 0}, rip = 0;
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            const double r = c_hex.N[US][ip][n];
            const size_t o = 3 * size_t(nd[n]);
            rip += r * rho;
            for (int j = 0; j < 3; ++j) u[j] += r * U[o + j];
        }
        double t = 0;
        for (int j = 0; j < 3; ++j) t += rip * u[j] * A[j];
        mdot[12 * size_t(e) + ip] = t;
    }
}

// updateMassFlowRateInterior_ (flowModel.cpp:5272-5688): Rhie-Chow with relaxation
// One wavefront per 64-element group: the group's 64 x 12 mass flow rates
// are one contiguous span, staged through LDS so that HBM sees coalesced
// reads and writes instead of 8-byte accesses at a 96-byte lane stride.
template <int US, int SP>
__global__ void __launch_bounds__(256) k_mdot_update(int ne, int negroup, const int* egroup_order,
                                                     const int* etype,
                                                     const double* xyz, const int* conn,
                                                     const double* scvw, const double* U,
                                                     const double* du, const double* p,
                                                     const double* gradP, double rho, double urf,
                                                     double* mdot, int* err) {
    __shared__ double sm[4 * 64 * 12];
    const int w = int(threadIdx.x >> 6), l = int(threadIdx.x & 63u);
    const int g = int(xcd_block()) * 4 + w;
    int e0 = 0, cnt = 0;
    if (g < negroup) {
        e0 = egroup_order[g] * 64;
        cnt = ne - e0 < 64 ? ne - e0 : 64;
    }
    double* st = sm + w * 64 * 12;
    for (int k = l; k < 12 * cnt; k += 64) st[k] = mdot[12 * size_t(e0) + k];
    __syncthreads();
    int bad = 0;
    if (l < cnt && !(etype && etype[e0 + l] != OAGPU_HEX8)) {  // others: k_mdot_generic
        const int e = e0 + l;
        int nd[8];
        load_conn(conn, e, nd);
        const double xr[3] = {xyz[3 * size_t(nd[0])], xyz[3 * size_t(nd[0]) + 1],
                              xyz[3 * size_t(nd[0]) + 2]};
        const double pr = p[nd[0]];
        double c[8][3];
        gather_rel(xyz, nd, xr, c);
        double Gp[3] = {0, 0, 0};
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            const double wv = scvw[8 * size_t(e) + n];
            const size_t o = 3 * size_t(nd[n]);
            for (int d = 0; d < 3; ++d) Gp[d] += wv * gradP[o + d];
        }
#pragma unroll 1
        for (int ip = 0; ip < 12; ++ip) {
            double A[3], Ji[9];
            area_scs(ip, c, A);
            jinv_scs(SP, ip, c, Ji, bad);
            double uip[3] = {0, 0, 0}, duip[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
            if constexpr (US == 1) {
                // shifted interpolation: the ip is its edge's midpoint, where the
                // only non-zero shape functions are those of the edge's two nodes
                // (c_lrscv, in node order: the terms of the full sum, same order)
#pragma unroll
                for (int s2 = 0; s2 < 2; ++s2) {
                    const int n = hex8::c_lrscv[ip][s2];
                    const double r = c_hex.N[US][ip][n];
                    const size_t o = 3 * size_t(nd[n]);
                    for (int j = 0; j < 3; ++j) {
                        duip[j] += r * du[o + j];
                        uip[j] += r * U[o + j];
                    }
                }
            }
#pragma unroll 2
            for (int n = 0; n < 8; ++n) {
                const int node = nd[n];
                double gv[3];
                gradN_scs(SP, ip, n, Ji, gv);
                const double pn = p[node] - pr;
                if constexpr (US != 1) {
                    const double r = c_hex.N[US][ip][n];
                    const size_t o = 3 * size_t(node);
                    for (int j = 0; j < 3; ++j) {
                        duip[j] += r * du[o + j];
                        uip[j] += r * U[o + j];
                    }
                }
                for (int j = 0; j < 3; ++j) dp[j] += gv[j] * pn;
            }
            // constant density: the upwind density is rho whatever the flow direction
            double t = 0;
            for (int j = 0; j < 3; ++j) {
                t += rho * uip[j] * A[j];
                t -= rho * duip[j] * (dp[j] - Gp[j]) * A[j];
            }
            double& md = st[12 * l + ip];
            md = urf * t + (1.0 - urf) * md;
        }
    }
    __syncthreads();
    for (int k = l; k < 12 * cnt; k += 64) mdot[12 * size_t(e0) + k] = st[k];
    if (bad) *err = 1;
}

// mass flow rates of the tet4 / wedge6 / pyramid5 elements of a mixed mesh,
// one lane per listed element of topology TOPO: init rho U_ip . A
// (initializeMassFlowRateInterior_), else the Rhie-Chow rate relaxed with
// urf (updateMassFlowRateInterior_), as k_mdot_init / k_mdot_update
template <int TOPO>
__global__ void __launch_bounds__(256) k_mdot_generic(int n, const int* glist, const int* etype,
                                                      const double* xyz, const int* conn,
                                                      const double* scvw, const double* U,
                                                      const double* du, const double* p,
                                                      const double* gradP, double rho, double urf,
                                                      int us, int sp, int init, double* mdot,
                                                      int* err) {
    using Tr = topo::Traits<TOPO>;
    constexpr int npe = Tr::npe;
    const int i = int(xcd_block() * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int e = glist[i];
    if (etype[e] != TOPO) return;
    int nd[npe];
    for (int k = 0; k < npe; ++k) nd[k] = conn[8 * size_t(e) + k];
    const double xr[3] = {xyz[3 * size_t(nd[0])], xyz[3 * size_t(nd[0]) + 1], xyz[3 * size_t(nd[0]) + 2]};
    double c[npe][3];
    for (int k = 0; k < npe; ++k)
        for (int d = 0; d < 3; ++d) c[k][d] = xyz[3 * size_t(nd[k]) + d] - xr[d];
    double Gp[3] = {0, 0, 0};
    const double pr = p[nd[0]];
    if (!init)
        for (int k = 0; k < npe; ++k)
            for (int d = 0; d < 3; ++d) Gp[d] += scvw[8 * size_t(e) + k] * gradP[3 * size_t(nd[k]) + d];
    int bad = 0;
#pragma unroll 1
    for (int ip = 0; ip < Tr::nscs; ++ip) {
        double A[3], g[npe][3], Nc[npe], Nu[npe];
        topo::scs_ip<TOPO>(ip, sp, us, c, A, g, Nc, Nu, bad);
        double uip[3] = {0, 0, 0}, duip[3] = {0, 0, 0}, dp[3] = {0, 0, 0}, rip = 0;
        for (int k = 0; k < npe; ++k) {
            const size_t o = 3 * size_t(nd[k]);
            rip += Nu[k] * rho;
            const double pn = p[nd[k]] - pr;
            for (int j = 0; j < 3; ++j) {
                uip[j] += Nu[k] * U[o + j];
                duip[j] += Nu[k] * du[o + j];
                dp[j] += g[k][j] * pn;
            }
        }
        double t = 0;
        if (init) {
            for (int j = 0; j < 3; ++j) t += rip * uip[j] * A[j];
            mdot[12 * size_t(e) + ip] = t;
        } else {
            for (int j = 0; j < 3; ++j) {
                t += rho * uip[j] * A[j];
                t -= rho * duip[j] * (dp[j] - Gp[j]) * A[j];
            }
            double& md = mdot[12 * size_t(e) + ip];
            md = urf * t + (1.0 - urf) * md;
        }
    }
    if (bad) *err = 1;
}

static void launch_mdot_generic(Flow& f, bool init) {
    DevMesh& m = *f.m;
    if (!m.nglist) return;
    const unsigned B = 256, G = xcd_grid(size_t(m.nglist), B);
    auto go = [&](auto kern) {
        kern<<<G, B, 0, f.stream>>>(m.nglist, m.glist.p, m.etype.p, m.xyz.p, m.conn.p, m.scvw.p, f.U.p,
                                    f.du.p, f.p.p, f.gradP.p, f.prm.rho, f.prm.urf_mdot,
                                    f.prm.u_shifted ? 1 : 0, f.prm.pgrad_shifted ? 1 : 0,
                                    init ? 1 : 0, f.mdot.p, f.err.p);
    };
    go(k_mdot_generic<topo::kTet>);
    go(k_mdot_generic<topo::kWed>);
    go(k_mdot_generic<topo::kPyr>);
    OAGPU_HIP(hipGetLastError());
}

void launch_mdot_init(Flow& f) {
    DevMesh& m = *f.m;
    const unsigned B = 256, G = xcd_grid(m.ne, B);
    const int* et = m.mixed ? m.etype.p : nullptr;
    if (f.prm.u_shifted)
        k_mdot_init<1><<<G, B, 0, f.stream>>>(m.ne, m.xyz.p, m.conn.p, et, f.U.p, f.prm.rho, f.mdot.p);
    else
        k_mdot_init<0><<<G, B, 0, f.stream>>>(m.ne, m.xyz.p, m.conn.p, et, f.U.p, f.prm.rho, f.mdot.p);
    OAGPU_HIP(hipGetLastError());
    launch_mdot_generic(f, true);
    launch_side_mdot(f, true);
}

void launch_update_mdot(Flow& f) {
    DevMesh& m = *f.m;
    const unsigned B = 256, G = xcd_grid(m.ne, B);
    const unsigned lds = occupancy_lds(B, 4 * 64 * 12 * sizeof(double), f.tune.cap_mdot);
    auto args = [&](auto kern) {
        kern<<<G, B, lds, f.stream>>>(m.ne, m.negroup, m.egroup_order.p, m.mixed ? m.etype.p : nullptr,
                                    m.xyz.p, m.conn.p, m.scvw.p,
                                    f.U.p, f.du.p, f.p.p, f.gradP.p, f.prm.rho, f.prm.urf_mdot,
                                    f.mdot.p, f.err.p);
    };
    const int us = f.prm.u_shifted ? 1 : 0, sp = f.prm.pgrad_shifted ? 1 : 0;
    clock_begin(f, kClkMdot);
    if (us && sp) args(k_mdot_update<1, 1>);
    else if (us) args(k_mdot_update<1, 0>);
    else if (sp) args(k_mdot_update<0, 1>);
    else args(k_mdot_update<0, 0>);
    clock_end(f, kClkMdot);
    OAGPU_HIP(hipGetLastError());
    launch_mdot_generic(f, false);
    launch_side_mdot(f, false);
}

// -------------------------------------------------------------- gradients
// nodeField::updateGradientField (nodeField.hpp:3253-4950): incremental
// Green-Gauss with relaxation, interior ips then wall-side ips

// component i of node n: N = 2 are two scalar fields (phi, phi2; k and
// omega, whose gradients share the element gathers and area vectors), else
// the N components of phi interleaved
template <int N>
__device__ __forceinline__ double fval(const double* phi, const double* phi2, size_t n, int i) {
    if constexpr (N == 2) return i == 0 ? phi[n] : phi2[n];
    else return phi[N * n + i];
}
// gradient entry k (3 per component) of node a
template <int N>
__device__ __forceinline__ double* gslot(double* grad, double* grad2, size_t a, int k) {
    if constexpr (N == 2) return k < 3 ? grad + 3 * a + k : grad2 + 3 * a + (k - 3);
    else return grad + 3 * N * a + k;
}

template <int N, int SH>
__device__ __forceinline__ void grad_incidence(const double* xyz, const int* conn, const double* phi,
                                               const double* phi2, int la, int e, const double xr[3],
                                               const double pa[N], double iv, double gurf,
                                               double G[3 * N]) {
    // the element relabelled so that la is node 0: its ips are 0, 3, 8, each
    // with the row node on the left
    const int* ce = conn + 8 * size_t(e);
    int nd[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) nd[k] = ce[c_rot.perm[3 * la][k]];
    double c[8][3];
    gather_rel(xyz, nd, xr, c);
    double pip[3][N];
    for (int q = 0; q < 3; ++q)
        for (int i = 0; i < N; ++i) pip[q][i] = 0.0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        double pn[N];
        for (int i = 0; i < N; ++i) pn[i] = fval<N>(phi, phi2, size_t(nd[n]), i);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double r = kHex.N[SH][kCanIp[q]][n];
            if (r != 0.0)
                for (int i = 0; i < N; ++i) pip[q][i] += r * pn[i];
        }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        double A[3];
        area_c(kCanIp[q], c, A);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < 3; ++j) G[3 * i + j] += gurf * (pip[q][i] - pa[i]) * A[j] * iv;
    }
}

// the same for a tet4 / wedge6 / pyramid5 element: its SCS ips with the row
// node (local la) on either side, area vectors oriented away from the row
template <int N, int SH, int TOPO>
__device__ void grad_generic(const double* xyz, const int* conn, const double* phi, const double* phi2,
                             int la, int e, const double xr[3], const double pa[N], double iv, double gurf,
                             double G[3 * N]) {
    constexpr int npe = topo::Traits<TOPO>::npe;
    int nd[npe];
    double c[npe][3];
    for (int k = 0; k < npe; ++k) {
        nd[k] = conn[8 * size_t(e) + k];
        for (int d = 0; d < 3; ++d) c[k][d] = xyz[3 * size_t(nd[k]) + d] - xr[d];
    }
#pragma unroll 1
    for (int ip = 0; ip < topo::Traits<TOPO>::nscs; ++ip) {
        const int* lr = topo::lr<TOPO>(ip);
        if (lr[0] != la && lr[1] != la) continue;
        const double sg = lr[0] == la ? 1.0 : -1.0;
        double A[3], Nv[npe];
        topo::scs_area<TOPO>(ip, c, A);
        const double* l = topo::ip_loc<TOPO>(SH, ip);
        topo::shape<TOPO>(SH, l[0], l[1], l[2], Nv);
        for (int i = 0; i < N; ++i) {
            double pip = 0.0;
            for (int k = 0; k < npe; ++k) pip += Nv[k] * fval<N>(phi, phi2, size_t(nd[k]), i);
            for (int j = 0; j < 3; ++j) G[3 * i + j] += gurf * (pip - pa[i]) * (sg * A[j]) * iv;
        }
    }
}

template <int N, int SH, bool MX>
__global__ void __launch_bounds__(256) k_gradient(int nn, int nslice, const int* slice_order,
                                                  const double* xyz, const int* conn,
                                                  const int* inc_la, const uint64_t* slots_la,
                                                  const int* ext_ptr, const int* ext,
                                                  const uint64_t* ext_slots, const int* binc_ptr,
                                                  const int* binc, const int* bf_elem,
                                                  const int* bf_ord, const int* bf_bc,
                                                  const int* bc_type, const double* symA,
                                                  const double* vol, const double* phi,
                                                  const int* gen_ptr, const int* gen,
                                                  const int* etype, double gurf, double* grad,
                                                  const double* phi2, double* grad2) {
    const int a = ordered_item(slice_order, nslice, nn);
    if (a < 0) return;
    double G[3 * N];
    for (int k = 0; k < 3 * N; ++k) G[k] = *gslot<N>(grad, grad2, size_t(a), k) * (1.0 - gurf);
    double pa[N];
    for (int i = 0; i < N; ++i) pa[i] = fval<N>(phi, phi2, size_t(a), i);
    const double iv = 1.0 / vol[a];
    const double xr[3] = {xyz[3 * size_t(a)], xyz[3 * size_t(a) + 1], xyz[3 * size_t(a) + 2]};
#pragma unroll 1
    for (int la = 0; la < 8; ++la) {
        const int e = inc_la[size_t(la) * nn + a];
        if (e >= 0) grad_incidence<N, SH>(xyz, conn, phi, phi2, la, e, xr, pa, iv, gurf, G);
    }
#pragma unroll 1
    for (int k = ext_ptr[a]; k < ext_ptr[a + 1]; ++k)
        grad_incidence<N, SH>(xyz, conn, phi, phi2, ext[k] & 7, ext[k] >> 3, xr, pa, iv, gurf, G);
    if constexpr (MX)
#pragma unroll 1
        for (int k = gen_ptr[a]; k < gen_ptr[a + 1]; ++k) {
            const int e = gen[k] >> 3, la = gen[k] & 7, t = etype[e];
            if (t == OAGPU_TET4) grad_generic<N, SH, topo::kTet>(xyz, conn, phi, phi2, la, e, xr, pa, iv, gurf, G);
            else if (t == OAGPU_WED6) grad_generic<N, SH, topo::kWed>(xyz, conn, phi, phi2, la, e, xr, pa, iv, gurf, G);
            else grad_generic<N, SH, topo::kPyr>(xyz, conn, phi, phi2, la, e, xr, pa, iv, gurf, G);
        }
#pragma unroll 1
    for (int k = binc_ptr[a]; k < binc_ptr[a + 1]; ++k) {
        const int f = binc[k] >> 2, fk = binc[k] & 3;
        if (bc_type && bc_type[bf_bc[f]] == OAGPU_BC_SYMMETRY) continue;  // no face term
        const int e = bf_elem[f], ord = bf_ord[f];
        const int* sn = c_sides[MX ? etype[e] : OAGPU_HEX8][ord];
        const int npf = sn[3] < 0 ? 3 : 4;
        double fp[4][3] = {}, pf[4][N] = {};
        for (int q = 0; q < npf; ++q) {
            const int n = conn[8 * size_t(e) + sn[q]];
            for (int d = 0; d < 3; ++d) fp[q][d] = xyz[3 * size_t(n) + d] - xr[d];
            for (int i = 0; i < N; ++i) pf[q][i] = fval<N>(phi, phi2, size_t(n), i);
        }
        double A[3];
        face_areav(npf, fp, fk, A);
        for (int i = 0; i < N; ++i) {
            double pip = 0.0;
            for (int q = 0; q < npf; ++q) pip += face_shape(npf, SH, fk, q) * pf[q][i];
            for (int j = 0; j < 3; ++j) G[3 * i + j] += gurf * (pip - pa[i]) * A[j] * iv;
        }
    }
    if constexpr (N != 2) {  // (k and omega: no symmetry correction)
    if (symA) {  // symmetry planes: nodeVectorField.cpp:46-140, nodeScalarField.cpp:45-110
        const double* sa = symA + 3 * size_t(a);
        const double amag = sqrt(sa[0] * sa[0] + sa[1] * sa[1] + sa[2] * sa[2]);
        if (amag > 0.0) {
            const double n[3] = {sa[0] / amag, sa[1] / amag, sa[2] / amag};
            if constexpr (N == 1) {
                const double dot = G[0] * n[0] + G[1] * n[1] + G[2] * n[2];
                for (int j = 0; j < 3; ++j) G[j] -= dot * n[j];
            } else {
                double ng[3] = {0, 0, 0}, gn[3] = {0, 0, 0}, ngn = 0.0;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        ng[i] += G[3 * i + j] * n[j];
                        gn[j] += G[3 * i + j] * n[i];
                    }
                for (int j = 0; j < 3; ++j) ngn += gn[j] * n[j];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        G[3 * i + j] -= ng[i] * n[j];
                        G[3 * i + j] -= n[i] * gn[j];
                        G[3 * i + j] += 2.0 * n[i] * n[j] * ngn;
                    }
            }
        }
    }
    }
    for (int k = 0; k < 3 * N; ++k) *gslot<N>(grad, grad2, size_t(a), k) = G[k];
}

// ---------------------------------------------------------- blending factor
// nodeField::updateBlendingFactorField (nodeField.hpp:2069-3250) with
// updateMinMaxFields (:1843-2068) for one node per lane: min / max over the
// nodes sharing an SCS with it, then the smooth Barth-Jespersen limiter over
// those SCS ips and the boundary-face ips nearest to it (blendingFactorMax 1).
// FP64, absolute coordinates, as the reference.
__device__ __forceinline__ void bj_limit(double& beta, double mx, double mn, double phi,
                                         const double g[3], const double dx[3]) {
    double rgrad = g[0] * dx[0] + g[1] * dx[1] + g[2] * dx[2];
    rgrad = rgrad > 0 ? rgrad + kSmall : rgrad - kSmall;
    const double bl = fmax((mx - phi) / rgrad, (mn - phi) / rgrad);
    const double b12 = fmax(-(mx - phi) / rgrad, -(mn - phi) / rgrad);
    double y = fmin(bl, b12);
    y = (y * y + 2.0 * y) / (y * y + y + 2.0);
    if (rgrad > 100 * kSmall || rgrad < -100 * kSmall) beta = fmin(beta, y);
}

// a tet4 / wedge6 / pyramid5 incidence of the blending pass: pass 0 folds
// the SCS neighbours' values into mn / mx, pass 1 limits with the offset of
// each SCS ip touching the node
template <int N, int TOPO>
__device__ void blend_generic(const double* xyz, const int* conn, const double* phi, int la, int e,
                              bool limit, const double xa[3], const double pa[N], double mn[N],
                              double mx[N], double b[N], double g[N][3]) {
    constexpr int npe = topo::Traits<TOPO>::npe;
#pragma unroll 1
    for (int ip = 0; ip < topo::Traits<TOPO>::nscs; ++ip) {
        const int* lr = topo::lr<TOPO>(ip);
        if (lr[0] != la && lr[1] != la) continue;
        if (!limit) {
            const int other = conn[8 * size_t(e) + (lr[0] == la ? lr[1] : lr[0])];
            for (int i = 0; i < N; ++i) {
                const double v = phi[N * size_t(other) + i];
                mn[i] = fmin(mn[i], v);
                mx[i] = fmax(mx[i], v);
            }
            continue;
        }
        double Nv[npe];
        const double* l = topo::ip_loc<TOPO>(0, ip);
        topo::shape<TOPO>(0, l[0], l[1], l[2], Nv);
        double dx[3];
        for (int d = 0; d < 3; ++d) {
            double xi = 0.0;
            for (int n = 0; n < npe; ++n) xi += Nv[n] * xyz[3 * size_t(conn[8 * size_t(e) + n]) + d];
            dx[d] = xi - xa[d];
        }
        for (int i = 0; i < N; ++i) bj_limit(b[i], mx[i], mn[i], pa[i], g[i], dx);
    }
}

template <int N, bool MX>
__global__ void __launch_bounds__(256) k_blending(int nn, const double* xyz, const int* conn,
                                                  const int* inc_la, const int* ext_ptr,
                                                  const int* ext, const int* binc_ptr,
                                                  const int* binc, const int* bf_elem,
                                                  const int* bf_ord, const double* phi,
                                                  const double* grad, const int* gen_ptr,
                                                  const int* gen, const int* etype, double* beta) {
    const int a = xcd_block() * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    double pa[N], mn[N], mx[N], b[N], g[N][3];
    for (int i = 0; i < N; ++i) {
        pa[i] = mn[i] = mx[i] = phi[N * size_t(a) + i];
        b[i] = 1.0;
        for (int j = 0; j < 3; ++j) g[i][j] = grad[3 * N * size_t(a) + 3 * i + j];
    }
    const double xa[3] = {xyz[3 * size_t(a)], xyz[3 * size_t(a) + 1], xyz[3 * size_t(a) + 2]};
    auto incidence = [&](int la, int e, bool limit) {
        int nd[8];
        load_conn(conn, e, nd);
        for (int q = 0; q < 3; ++q) {
            const int ip = c_hex.node_ip[la][q];
            const int other = c_hex.node_sg[la][q] > 0 ? c_hex.lr[ip][1] : c_hex.lr[ip][0];
            if (!limit) {
                for (int i = 0; i < N; ++i) {
                    const double v = phi[N * size_t(nd[other]) + i];
                    mn[i] = fmin(mn[i], v);
                    mx[i] = fmax(mx[i], v);
                }
                continue;
            }
            double dx[3];
            for (int d = 0; d < 3; ++d) {
                double xi = 0.0;
                for (int n = 0; n < 8; ++n) xi += c_hex.N[0][ip][n] * xyz[3 * size_t(nd[n]) + d];
                dx[d] = xi - xa[d];
            }
            for (int i = 0; i < N; ++i) bj_limit(b[i], mx[i], mn[i], pa[i], g[i], dx);
        }
    };
    for (int pass = 0; pass < 2; ++pass) {
#pragma unroll 1
        for (int la = 0; la < 8; ++la) {
            const int e = inc_la[size_t(la) * nn + a];
            if (e >= 0) incidence(la, e, pass == 1);
        }
#pragma unroll 1
        for (int k = ext_ptr[a]; k < ext_ptr[a + 1]; ++k) incidence(ext[k] & 7, ext[k] >> 3, pass == 1);
        if constexpr (MX)
#pragma unroll 1
            for (int k = gen_ptr[a]; k < gen_ptr[a + 1]; ++k) {
                const int e = gen[k] >> 3, la = gen[k] & 7, t = etype[e];
                if (t == OAGPU_TET4) blend_generic<N, topo::kTet>(xyz, conn, phi, la, e, pass == 1, xa, pa, mn, mx, b, g);
                else if (t == OAGPU_WED6) blend_generic<N, topo::kWed>(xyz, conn, phi, la, e, pass == 1, xa, pa, mn, mx, b, g);
                else blend_generic<N, topo::kPyr>(xyz, conn, phi, la, e, pass == 1, xa, pa, mn, mx, b, g);
            }
    }
#pragma unroll 1
    for (int k = binc_ptr[a]; k < binc_ptr[a + 1]; ++k) {
        const int f = binc[k] >> 2, fk = binc[k] & 3;
        const int e = bf_elem[f], ord = bf_ord[f];
        const int* sn = c_sides[MX ? etype[e] : OAGPU_HEX8][ord];
        const int npf = sn[3] < 0 ? 3 : 4;
        double dx[3] = {0, 0, 0};
        for (int q = 0; q < npf; ++q) {
            const int n = conn[8 * size_t(e) + sn[q]];
            for (int d = 0; d < 3; ++d) dx[d] += face_shape(npf, 0, fk, q) * xyz[3 * size_t(n) + d];
        }
        for (int d = 0; d < 3; ++d) dx[d] -= xa[d];
        for (int i = 0; i < N; ++i) bj_limit(b[i], mx[i], mn[i], pa[i], g[i], dx);
    }
    for (int i = 0; i < N; ++i) beta[N * size_t(a) + i] = b[i];
}

void launch_blending(Flow& f, int which) {
    DevMesh& m = *f.m;
    const unsigned B = 256, G = xcd_grid(m.nn, B);
    const int* gp = m.mixed ? m.gen_ptr.p : nullptr;
    if (which == 0) {
        OAGPU_REQUIRE(f.beta.p, "blending factors not allocated (high resolution off)");
        (m.mixed ? k_blending<3, true> : k_blending<3, false>)<<<G, B, 0, f.stream>>>(
            m.nn, m.xyz.p, m.conn.p, m.inc_la.p, m.ext_ptr.p, m.ext.p, m.binc_ptr.p, m.binc.p,
            m.bf_elem.p, m.bf_ord.p, f.U.p, f.gradU.p, gp, m.gen.p, m.etype.p, f.beta.p);
    } else {
        OAGPU_REQUIRE((which == 2 || which == 3) && f.sst_on && f.beta_k.p, "bad blending request");
        (m.mixed ? k_blending<1, true> : k_blending<1, false>)<<<G, B, 0, f.stream>>>(
            m.nn, m.xyz.p, m.conn.p, m.inc_la.p, m.ext_ptr.p, m.ext.p, m.binc_ptr.p, m.binc.p,
            m.bf_elem.p, m.bf_ord.p, which == 2 ? f.k.p : f.om.p, which == 2 ? f.gradK.p : f.gradW.p,
            gp, m.gen.p, m.etype.p, which == 2 ? f.beta_k.p : f.beta_w.p);
    }
    OAGPU_HIP(hipGetLastError());
}

// no-slip wall nodes and the assembled symmetry area of every node
// (mesh::updateAssembledSymmetryAreaField_, meshGeometry.cpp:1787-1850: each
// symmetry face ip's exposed area added to its nearest node), summed in the
// node's face order
__global__ void __launch_bounds__(256) k_boundary_types(int nn, const double* xyz, const int* conn,
                                                        const int* binc_ptr, const int* binc,
                                                        const int* bf_elem, const int* bf_ord,
                                                        const int* bf_bc, const int* bc_type,
                                                        const int* etype, int* wnode, double* symA) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    int w = 0;
    double s[3] = {0, 0, 0};
    const double xr[3] = {xyz[3 * size_t(a)], xyz[3 * size_t(a) + 1], xyz[3 * size_t(a) + 2]};
    for (int k = binc_ptr[a]; k < binc_ptr[a + 1]; ++k) {
        const int f = binc[k] >> 2, fk = binc[k] & 3;
        const int t = bc_type[bf_bc[f]];
        if (t == OAGPU_BC_WALL) w = 1;
        if (t != OAGPU_BC_SYMMETRY || !symA) continue;
        const int e = bf_elem[f];
        const int* sn = c_sides[etype ? etype[e] : OAGPU_HEX8][bf_ord[f]];
        const int npf = sn[3] < 0 ? 3 : 4;
        double fp[4][3] = {}, A[3];
        for (int q = 0; q < npf; ++q) {
            const size_t o = 3 * size_t(conn[8 * size_t(e) + sn[q]]);
            for (int d = 0; d < 3; ++d) fp[q][d] = xyz[o + d] - xr[d];
        }
        face_areav(npf, fp, fk, A);
        for (int d = 0; d < 3; ++d) s[d] += A[d];
    }
    wnode[a] = w;
    if (symA)
        for (int d = 0; d < 3; ++d) symA[3 * size_t(a) + d] = s[d];
}

void launch_boundary_types(Flow& f) {
    DevMesh& m = *f.m;
    const unsigned B = 256;
    k_boundary_types<<<(m.nn + B - 1) / B, B, 0, f.stream>>>(
        m.nn, m.xyz.p, m.conn.p, m.binc_ptr.p, m.binc.p, m.bf_elem.p, m.bf_ord.p, m.bf_bc.p,
        f.bc_type.p, m.mixed ? m.etype.p : nullptr, f.wnode.p, f.has_sym ? f.symA.p : nullptr);
    OAGPU_HIP(hipGetLastError());
}

void launch_gradient(Flow& f, int which) {
    DevMesh& m = *f.m;
    const unsigned B = 256, G = xcd_grid(m.nn, B);
    // the symmetry correction is applied to velocity and pressure
    // (velocity.cpp:52, pressure.cpp:48), not to k / omega (default off)
    const double* sym = f.has_sym && which <= 1 ? f.symA.p : nullptr;
    const int* types = f.has_sym ? f.bc_type.p : nullptr;  // symmetry faces add no term
    const int clk = which == 0 ? kClkGradU : kClkGradP;
    const unsigned lds = occupancy_lds(B, 0, f.tune.cap_grad);
    auto go = [&](auto kern, const double* phi, double* grad) {
        clock_begin(f, clk);
        kern<<<G, B, lds, f.stream>>>(m.nn, m.nslice, m.slice_order.p, m.xyz.p, m.conn.p, m.inc_la.p,
                                    m.slots_la.p, m.ext_ptr.p, m.ext.p, m.ext_slots.p, m.binc_ptr.p,
                                    m.binc.p, m.bf_elem.p, m.bf_ord.p, m.bf_bc.p, types, sym,
                                    m.vol.p, phi, m.mixed ? m.gen_ptr.p : nullptr, m.gen.p,
                                    m.etype.p, f.prm.grad_urf, grad, nullptr, nullptr);
        clock_end(f, clk);
    };
    auto pick = [&](auto kern_hex, auto kern_mixed, const double* phi, double* grad) {
        if (m.mixed) go(kern_mixed, phi, grad);
        else go(kern_hex, phi, grad);
    };
    if (which == 0) {
        if (f.prm.u_shifted) pick(k_gradient<3, 1, false>, k_gradient<3, 1, true>, f.U.p, f.gradU.p);
        else pick(k_gradient<3, 0, false>, k_gradient<3, 0, true>, f.U.p, f.gradU.p);
    } else {
        const double* phi = f.p.p;
        double* grad = f.gradP.p;
        int sh = f.prm.p_shifted;
        if (which == 2 || which == 3) {
            OAGPU_REQUIRE(f.sst_on, "SST not enabled on this flow");
            phi = which == 2 ? f.k.p : f.om.p;
            grad = which == 2 ? f.gradK.p : f.gradW.p;
            sh = which == 2 ? f.sst.k_shifted : f.sst.w_shifted;
        }
        if (sh) pick(k_gradient<1, 1, false>, k_gradient<1, 1, true>, phi, grad);
        else pick(k_gradient<1, 0, false>, k_gradient<1, 0, true>, phi, grad);
    }
    OAGPU_HIP(hipGetLastError());
}

// grad k and grad omega (the end of an SST iteration): one launch sharing the
// element gathers and area vectors when both use the same interpolation,
// else one launch each
void launch_gradient_kw(Flow& f) {
    OAGPU_REQUIRE(f.sst_on, "SST not enabled on this flow");
    if (f.tune.scalar_pair == 0 || (f.sst.k_shifted != 0) != (f.sst.w_shifted != 0)) {
        launch_gradient(f, 2);
        launch_gradient(f, 3);
        return;
    }
    DevMesh& m = *f.m;
    const unsigned B = 256, G = xcd_grid(m.nn, B);
    const int* types = f.has_sym ? f.bc_type.p : nullptr;  // symmetry faces add no term
    const unsigned lds = occupancy_lds(B, 0, f.tune.cap_grad);
    auto go = [&](auto kern) {
        clock_begin(f, kClkGradKW);
        kern<<<G, B, lds, f.stream>>>(m.nn, m.nslice, m.slice_order.p, m.xyz.p, m.conn.p, m.inc_la.p,
                                    m.slots_la.p, m.ext_ptr.p, m.ext.p, m.ext_slots.p, m.binc_ptr.p,
                                    m.binc.p, m.bf_elem.p, m.bf_ord.p, m.bf_bc.p, types, nullptr,
                                    m.vol.p, f.k.p, m.mixed ? m.gen_ptr.p : nullptr, m.gen.p,
                                    m.etype.p, f.prm.grad_urf, f.gradK.p, f.om.p, f.gradW.p);
        clock_end(f, kClkGradKW);
    };
    if (f.sst.k_shifted) go(m.mixed ? k_gradient<2, 1, true> : k_gradient<2, 1, false>);
    else go(m.mixed ? k_gradient<2, 0, true> : k_gradient<2, 0, false>);
    OAGPU_HIP(hipGetLastError());
}

}  // namespace oagpu
