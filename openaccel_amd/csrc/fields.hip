// fields.hip -- gfx950 kernels of the element-centric and point-wise parts
// of the CVFEM hot path: static mesh geometry (SCV volumes, dual volumes,
// element-centre weights), mass flow rate initialisation and Rhie-Chow
// update, mass divergence, corrections, field scales and control residuals.
#include "device_util.hpp"
#include "topo_tables.hpp"

namespace oagpu {

using namespace hex8;

// ---------------------------------------------------------------- geometry
// SCV volumes of every hex8 element (hex_scv_det), written raw to scvw;
// other elements (etype non-null) get zeros here and their volumes from
// launch_scv_generic
__global__ void __launch_bounds__(256) k_scv_volumes(int ne, const double* xyz, const int* conn,
                                                     const int* etype, double* scvw) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    if (etype && etype[e] != OAGPU_HEX8) {
        for (int n = 0; n < 8; ++n) scvw[8 * size_t(e) + n] = 0.0;
        return;
    }
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
    k_scv_volumes<<<xcd_grid(m.ne, B), B, 0, s>>>(m.ne, m.xyz.p, m.conn.p,
                                                  m.mixed ? m.etype.p : nullptr, m.scvw.p);
    if (m.mixed) {
        launch_scv_generic(m, s);
        launch_geo_generic(m, s);
        launch_side_geo(m, s);
    }
    k_dual_volume<<<xcd_grid(m.nn, B), B, 0, s>>>(m.nn, m.inc_ptr.p, m.inc.p, m.scvw.p, m.vol.p);
    k_scv_weights<<<xcd_grid(m.ne, B), B, 0, s>>>(m.ne, m.scvw.p);
    OAGPU_HIP(hipGetLastError());
}

// updateMassDivergenceField_ interior ips (flowModel.cpp:13733-13772): divL += mdot,
// divR -= mdot; then the side mass flows of inlet, outlet and opening face
// ips whose nearest node this is (flowModel.cpp:14237-14500, 14850-14852;
// msd null without them)
__global__ void __launch_bounds__(256) k_mass_divergence(int nn, const int* inc_ptr, const int* inc,
                                                         const double* mdot, const int* binc_ptr,
                                                         const int* binc, const int* bf_bc,
                                                         const int* bc_type, const double* msd,
                                                         const int* etype, double* div) {
    const int a = xcd_block() * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    double s = 0.0;
    for (int k = inc_ptr[a]; k < inc_ptr[a + 1]; ++k) {
        const int e = inc[k] >> 3, la = inc[k] & 7;
        const int t = etype ? etype[e] : OAGPU_HEX8;
        if (t != OAGPU_HEX8) {  // the element's SCS ips with this node on either side
            const int ns = t == OAGPU_TET4 ? 6 : t == OAGPU_WED6 ? 9 : 12;
            for (int ip = 0; ip < ns; ++ip) {
                const int* lr = t == OAGPU_TET4 ? topo::c_tet_lr[ip]
                                : t == OAGPU_WED6 ? topo::c_wed_lr[ip] : topo::c_pyr_lr[ip];
                if (lr[0] == la) s += mdot[12 * size_t(e) + ip];
                else if (lr[1] == la) s -= mdot[12 * size_t(e) + ip];
            }
            continue;
        }
        for (int q = 0; q < 3; ++q) {
            const double m = mdot[12 * size_t(e) + c_node_ip[la][q]];
            if (c_node_ip_sign[la][q] > 0) s += m; else s -= m;
        }
    }
    if (msd)
        for (int k = binc_ptr[a]; k < binc_ptr[a + 1]; ++k) {
            const int f = binc[k] >> 2, t = bc_type[bf_bc[f]];
            if (t == OAGPU_BC_INLET || t == OAGPU_BC_OUTLET || t == OAGPU_BC_OPENING)
                s += msd[4 * size_t(f) + (binc[k] & 3)];
        }
    div[a] = s;
}

void launch_mass_divergence(Flow& f) {
    DevMesh& m = *f.m;
    const unsigned B = 256;
    k_mass_divergence<<<xcd_grid(m.nn, B), B, 0, f.stream>>>(
        m.nn, m.inc_ptr.p, m.inc.p, f.mdot.p, m.binc_ptr.p, m.binc.p, m.bf_bc.p, f.bc_type.p,
        f.has_io ? f.msd.p : nullptr, m.mixed ? m.etype.p : nullptr, f.div.p);
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
__global__ void __launch_bounds__(kRedBlock) k_scaled_res(int nn, const int* slice_ptr,
                                                          const int* slice_w, const int* doff,
                                                          const double* A, const double* b,
                                                          double* part) {
    __shared__ double s[N][kRedBlock];
    double acc[N];
    for (int k = 0; k < N; ++k) acc[k] = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
        const SellRow sr = sell_row<N * N>(slice_ptr, slice_w, i);
        const double* d = A + sr.base + size_t(doff[i]) * (N * N * 64);
        for (int k = 0; k < N; ++k) {
            const double r = b[N * size_t(i) + k] / d[(N + 1) * k * 64];
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
        k_scaled_res<3><<<G, kRedBlock, 0, f.stream>>>(m.nn, m.slice_ptr.p, m.slice_w.p, m.diag_off.p, f.A3.p,
                                                       f.b3.p, f.part.p);
    else
        k_scaled_res<1><<<G, kRedBlock, 0, f.stream>>>(m.nn, m.slice_ptr.p, m.slice_w.p, m.diag_off.p, f.A1.p,
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
