// bc.hip -- side mass flow rates and reversal flags of velocity inlets,
// static-pressure outlets and static-pressure openings.  Face-owned: one
// lane per such face ip (mass flows) or per outlet face (reversal flags),
// over the list of those faces; each value is written by exactly one lane.
// Sides of every element topology: a tri side's fourth ip lane does nothing.
#include "sides.hpp"

namespace oagpu {

struct SideArgs {
    int nio;
    const int* io_face;
    const int* bf_elem;
    const int* bf_ord;
    const int* bf_bc;
    const int* bc_type;
    const double* bc_U;
    const double* bc_p;
    const double* xyz;
    const int* conn;
    const double* U;
    const double* du;
    const double* p;
    const double* gradP;
    double rho, urf;
    int us, sp, init;
    int* rfl;
    double* msd;
    int* err;
    SideGeo sg;
};

// init: initializeMassFlowRate side part (flowModel.cpp:3802-4997): inlet
// rho U_bc . A, outlet max(0, rho U_ip . A), opening rho U_ip . A
// (:4817-4992).  Update (flowModel.cpp:6470-6598 inlet, :7323-7690 outlet,
// :8094-8380 opening): the new rate relaxed with urf against the previous
// one; a reversed outlet ip carries no flow; the outlet's Rhie-Chow rate
// takes the specified pressure on the side's nodes, the opening's the
// nodal pressure.
__global__ void __launch_bounds__(256) k_side_mdot(const SideArgs a) {
    const int i = int(xcd_block() * blockDim.x + threadIdx.x);
    if (i >= 4 * a.nio) return;
    const int f = a.io_face[i >> 2], fk = i & 3;
    const int e = a.bf_elem[f], ord = a.bf_ord[f], bc = a.bf_bc[f], t = a.bc_type[bc];
    const int* ce = a.conn + 8 * size_t(e);
    const int* sn = c_sides[a.sg.etype ? a.sg.etype[e] : OAGPU_HEX8][ord];
    if (fk == 3 && sn[3] < 0) return;  // tri side: three ips
    const size_t on = 3 * size_t(ce[sn[fk]]);
    const double xr[3] = {a.xyz[on], a.xyz[on + 1], a.xyz[on + 2]};
    int bad = 0;
    FaceIp<double> fi;
    face_ip<true>(a.sg, a.xyz, ce, e, f, ord, fk, t == OAGPU_BC_INLET ? a.us : a.sp, xr, fi, bad);
    double& md = a.msd[4 * size_t(f) + fk];
    double rip = 0;
    for (int q = 0; q < fi.npf; ++q) rip += fi.shape(a.us, fk, q) * a.rho;
    if (t == OAGPU_BC_INLET) {
        double m = 0;
        for (int j = 0; j < 3; ++j) m += (rip * a.bc_U[3 * bc + j]) * fi.A[j];
        md = a.init ? m : a.urf * m + (1.0 - a.urf) * md;
    } else if (a.init) {
        double u[3] = {0, 0, 0};
        for (int q = 0; q < fi.npf; ++q) {
            const double r = fi.shape(a.us, fk, q);
            const size_t o = 3 * size_t(ce[fi.sn[q]]);
            for (int j = 0; j < 3; ++j) u[j] += r * a.U[o + j];
        }
        double m = 0;
        for (int j = 0; j < 3; ++j) m += rip * u[j] * fi.A[j];
        md = t == OAGPU_BC_OUTLET ? fmax(0.0, m) : m;
    } else if (t == OAGPU_BC_OUTLET && a.rfl[4 * size_t(f) + fk] == 1) {
        md = 0.0;  // reversed: a slip wall
    } else {
        const double m = outlet_rhie_chow(fi, ce, fk, a.us, a.U, a.du, a.p, a.gradP, a.rho,
                                          a.bc_p[bc], t == OAGPU_BC_OUTLET);
        md = a.urf * m + (1.0 - a.urf) * md;
    }
    if (bad) *a.err = 1;
}

SideArgs side_args(Flow& f) {
    DevMesh& m = *f.m;
    SideArgs a;
    a.nio = f.nio;
    a.io_face = f.io_face.p;
    a.bf_elem = m.bf_elem.p;
    a.bf_ord = m.bf_ord.p;
    a.bf_bc = m.bf_bc.p;
    a.bc_type = f.bc_type.p;
    a.bc_U = f.bc_U.p;
    a.bc_p = f.bc_p.p;
    a.xyz = m.xyz.p;
    a.conn = m.conn.p;
    a.U = f.U.p;
    a.du = f.du.p;
    a.p = f.p.p;
    a.gradP = f.gradP.p;
    a.rho = f.prm.rho;
    a.urf = f.prm.urf_mdot;
    a.us = f.prm.u_shifted ? 1 : 0;
    a.sp = f.prm.pgrad_shifted ? 1 : 0;
    a.init = 0;
    a.rfl = f.rfl.p;
    a.msd = f.msd.p;
    a.err = f.err.p;
    a.sg = side_geo(m);
    return a;
}

void launch_side_mdot(Flow& f, bool init) {
    if (!f.has_io) return;
    SideArgs a = side_args(f);
    a.init = init ? 1 : 0;
    const unsigned B = 256;
    k_side_mdot<<<xcd_grid(size_t(4) * f.nio, B), B, 0, f.stream>>>(a);
    OAGPU_HIP(hipGetLastError());
}

// updateFlowReversalFlag, static-pressure outlets (flowModel.cpp:15226-15454):
// the face's side rates are summed and clipped at zero; an unflagged ip is
// flagged (and its rate zeroed) when the face takes in net flow, a flagged
// one released once the face-mean velocity leaves through it and the
// specified pressure does not exceed the face-mean pressure.
__global__ void __launch_bounds__(256) k_reversal(const SideArgs a) {
    const int i = int(xcd_block() * blockDim.x + threadIdx.x);
    if (i >= a.nio) return;
    const int f = a.io_face[i];
    const int e = a.bf_elem[f], ord = a.bf_ord[f], bc = a.bf_bc[f];
    if (a.bc_type[bc] != OAGPU_BC_OUTLET) return;
    const int* ce = a.conn + 8 * size_t(e);
    const int* sn = c_sides[a.sg.etype ? a.sg.etype[e] : OAGPU_HEX8][ord];
    const int npf = sn[3] < 0 ? 3 : 4;
    const double fw = 1.0 / npf;
    double* md = a.msd + 4 * size_t(f);
    int* rf = a.rfl + 4 * size_t(f);
    double side = 0;
    for (int k = 0; k < npf; ++k) {
        side += md[k];
        md[k] = fmax(md[k], 0.0);
    }
    double pavg = 0, uavg[3] = {0, 0, 0}, fp[4][3] = {};
    const size_t o0 = 3 * size_t(ce[sn[0]]);
    for (int q = 0; q < npf; ++q) {
        const size_t o = 3 * size_t(ce[sn[q]]);
        pavg += fw * a.p[ce[sn[q]]];
        for (int j = 0; j < 3; ++j) {
            uavg[j] += fw * a.U[o + j];
            fp[q][j] = a.xyz[o + j] - a.xyz[o0 + j];
        }
    }
    const double pbc = a.bc_p[bc];
    for (int k = 0; k < npf; ++k) {
        double A[3];
        face_areav(npf, fp, k, A);
        const double amag = sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
        double vn = 0;
        for (int j = 0; j < 3; ++j) vn += uavg[j] * A[j] / amag;
        if (rf[k] == 0) {
            if (side < -kSmall) {
                rf[k] = 1;
                md[k] = 0.0;
            }
        } else if (vn >= 0.0 && pbc <= pavg) {
            rf[k] = 0;
        }
    }
}

void launch_reversal(Flow& f) {
    if (!f.has_outlet) return;
    const SideArgs a = side_args(f);
    const unsigned B = 256;
    k_reversal<<<xcd_grid(size_t(f.nio), B), B, 0, f.stream>>>(a);
    OAGPU_HIP(hipGetLastError());
}

}  // namespace oagpu
