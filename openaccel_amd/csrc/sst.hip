// sst.hip -- nodal k-omega SST model kernels: eddy viscosity, effective
// viscosity, F1 blending and production in one pass over the nodes, and the
// clipped k / omega corrections.
//
// shearStressTransportModel.cpp (src/model/turbulence/RANS/shearStressTransport):
//   updateTurbulentDynamicViscosity :597-690, updateFOneBlending :95-174,
//   updateTurbulentProduction :176-340 (incompressible, no frame rotation;
//   no-slip wall nodes take the production the wall treatment supplies);
// turbulenceModel::updateEffectiveDynamicViscosity (turbulenceModel.cpp:121);
// equation::correctField_ with the lower clip at SMALL (equation.h:204-286).
// All three model updates read the same state (k, omega, grad U, grad k,
// grad omega, y) and the production uses the new eddy viscosity, so fusing
// them gives the reference's sequential result.
#include "device_util.hpp"
#include "flow.hpp"

#include <cfloat>
#include <vector>

namespace oagpu {

namespace sstc {  // shearStressTransportModel.h:21-37, RANSModel.h:20-22
constexpr double kA1 = 0.31, kBetaStar = 0.09, kSigW2 = 0.856;
}

__device__ __forceinline__ double strain_mag(const double g[9]) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double r = 0.5 * (g[3 * i + j] + g[3 * j + i]);
            s += r * r;
        }
    return sqrt(2.0 * s);
}

__device__ __forceinline__ double pk_by_mut(const double g[9]) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) s += g[3 * i + j] * (g[3 * i + j] + g[3 * j + i]);
    return fmax(s, 0.0);
}

struct SstNodeArgs {
    int nn;
    double rho, mu, urf_mut;
    const double *k, *w, *ymin, *gradU, *gradK, *gradW, *Pk_wall;
    const int* wnode;  // no-slip wall nodes
    double *mut, *mueff, *F1, *Pk;
};

__global__ void __launch_bounds__(256) k_sst_update(const SstNodeArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nn) return;
    using namespace sstc;
    const double rho = a.rho, mu = a.mu;
    const double k = a.k[i], w = a.w[i], y = a.ymin[i];
    double g[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) g[q] = a.gradU[9 * size_t(i) + q];
    const double minDSq = y * y;
    const double trbDiss = sqrt(k) / kBetaStar / (w + kSmall) / (y + kSmall);
    const double lamDiss = 500.0 * mu / rho / (w + kSmall) / (minDSq + kSmall);
    // eddy viscosity (F2 limiter) and effective viscosity
    const double sij = strain_mag(g);
    const double fArgTwo = fmax(2.0 * trbDiss, lamDiss);
    const double fTwo = tanh(fArgTwo * fArgTwo);
    const double mutv = kA1 * rho * k / (fmax(kA1 * w, sij * fTwo) + kSmall);
    const double mut = a.urf_mut * mutv + (1.0 - a.urf_mut) * a.mut[i];
    a.mut[i] = mut;
    a.mueff[i] = mu + mut;
    // F1 blending
    double cross = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) cross += a.gradK[3 * size_t(i) + j] * a.gradW[3 * size_t(i) + j];
    const double CDkw = fmax(2.0 * rho * kSigW2 * cross / (w + kSmall), 1.0e-10);
    const double arg = fmin(fmax(trbDiss, lamDiss),
                            4.0 * rho * kSigW2 * k / (CDkw + kSmall) / (minDSq + kSmall));
    a.F1[i] = tanh(arg * arg * arg * arg);
    // production
    a.Pk[i] = a.wnode[i] ? a.Pk_wall[i] : fmax(mut * pk_by_mut(g), 0.0);
}

void launch_sst_update(Flow& f) {
    OAGPU_REQUIRE(f.sst_on, "SST not enabled on this flow");
    DevMesh& m = *f.m;
    SstNodeArgs a;
    a.nn = m.nn;
    a.rho = f.prm.rho;
    a.mu = f.prm.mu;
    a.urf_mut = f.sst.urf_mut;
    a.k = f.k.p;
    a.w = f.om.p;
    a.ymin = f.ymin.p;
    a.gradU = f.gradU.p;
    a.gradK = f.gradK.p;
    a.gradW = f.gradW.p;
    a.Pk_wall = f.Pk_wall.p;
    a.wnode = f.wnode.p;
    a.mut = f.mut.p;
    a.mueff = f.mueff.p;
    a.F1 = f.F1.p;
    a.Pk = f.Pk.p;
    k_sst_update<<<(m.nn + 255) / 256, 256, 0, f.stream>>>(a);
    OAGPU_HIP(hipGetLastError());
}

// phi = max(phi + x, SMALL) and the block min/max of the result
__global__ void __launch_bounds__(kRedBlock) k_correct_clip(int nn, const double* x, double* phi,
                                                            double* part) {
    __shared__ double smax[kRedBlock], smin[kRedBlock];
    double vmax = -DBL_MAX, vmin = DBL_MAX;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
        double v = phi[i] + x[i];
        if (v < kSmall) v = kSmall;
        phi[i] = v;
        vmax = fmax(vmax, v);
        vmin = fmin(vmin, v);
    }
    smax[threadIdx.x] = vmax;
    smin[threadIdx.x] = vmin;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + st]);
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = smax[0];
        part[gridDim.x + blockIdx.x] = smin[0];
    }
}

void launch_correct_scalar(Flow& f, int which) {
    OAGPU_REQUIRE(f.sst_on, "SST not enabled on this flow");
    const int G = 128;
    double* phi = which == 0 ? f.k.p : f.om.p;
    k_correct_clip<<<G, kRedBlock, 0, f.stream>>>(f.m->nn, f.x1.p, phi, f.part.p);
    OAGPU_HIP(hipGetLastError());
    std::vector<double> h(2 * G);
    f.part.download(h.data(), h.size(), f.stream);
    OAGPU_HIP(hipStreamSynchronize(f.stream));
    double mx = -DBL_MAX, mn = DBL_MAX;
    for (int i = 0; i < G; ++i) {
        mx = std::max(mx, h[i]);
        mn = std::min(mn, h[G + i]);
    }
    // nodeField::updateScale (nodeField.hpp:5993-6015)
    double sc = (mx > 0 ? mx : 0.0) - (mn < 0 ? mn : 0.0);
    if (sc < 1.0e-9) sc = 1.0;
    if (mx > sc) sc = mx;
    (which == 0 ? f.k_scale : f.w_scale) = sc;
}

}  // namespace oagpu
