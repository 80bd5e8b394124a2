// graph.hip -- one outer iteration (the flow, then the k-omega SST equations
// when they are on) as a single HIP graph.
//
// Per outer iteration the reference runs segregatedFlowEquations::solve
// (src/equation/flow/segregatedFlowEquations.cpp:76-297) and
// segregatedShearStressTransportEquations::solve
// (src/equation/turbulence/RANS/shearStressTransport/segregatedShearStressTransportEquations.cpp:86-99),
// each linear solve through linearSystem<N>::solve (src/equation/linearSystem.h:416-462).
// Every data-dependent decision in that sequence -- Krylov convergence, the
// GMRES Givens rotations and least-squares update, BiCGStab's scalars and
// breakdown restarts, iterative refinement of the mixed form, the field
// scales and the scaled control residuals -- is taken here by one-workgroup
// kernels on device-resident state, so an iteration is a fixed sequence of
// launches that a graph replays with no host read inside it.
//
// Each Krylov solve is unrolled to a cap of iteration slots; every launch of
// a slot first reads the solve's flags and returns at once when its part of
// the slot has nothing to do (the solve has finished, no cycle ends here, no
// refinement step is due).  A solve that reaches its cap unfinished is
// recorded: the host then restores the state saved by the graph's first
// launch and runs the same sequence uncapped, polling the finished flag, so
// the result never depends on the caps; a cap is the solve's iteration limit
// when that is small, else the largest count seen plus a quarter (at least 4).  The same sequence run without a graph (eager) makes
// the same decisions on the same values: graph replays and eager runs are
// bitwise equal.
//
// Solves covered: Jacobi-GMRES (FP64, and FP32 under FP64 refinement) for
// the momentum, k and omega systems; for p' the smoothed-aggregation
// AMG-BiCGStab and AMG-CG (FP64, and FP32 under FP64 refinement).
#include "device_util.hpp"
#include "flow.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

namespace oagpu {
namespace {

// ---------------------------------------------------------- solver state
// ints (IterGraph::si)
enum : int {
    I_VSTOP,   // the V-cycle kernels' stop flag: nothing of the BiCGStab step runs
    I_DONE,    // the solve has finished
    I_IT,      // iterations of the running (inner) solve
    I_USED,    // iterations of all inner solves so far (refinement)
    I_J,       // GMRES: Arnoldi step of the running cycle
    I_END,     // GMRES: the running cycle has ended (least-squares update due)
    I_JEND,    // GMRES: last step of the ended cycle
    I_FRESH,   // GMRES: a cycle has started, its first basis vector is due
    I_IDONE,   // refinement: the inner solve has finished, a refinement step is due
    I_ISTART,  // refinement: a new inner solve is due
    I_MAXIT,   // iteration limit of the whole solve
    I_BUDGET,  // iteration limit of the running (inner) solve
    I_SKIP,    // BiCGStab: the rest of this iteration is skipped (breakdown restart)
    I_RESTART, // BiCGStab: r0 = r and p = r at the next direction update
    I_MRES,    // GMRES restart length
    I_MIXED,   // FP32 inner solves under FP64 refinement
    I_NSTART,  // stop flag of the V-cycle that starts an inner CG: no start due
    I_NI = 20
};
constexpr int kMaxRes = 32;  // GMRES restart length limit of the device Arnoldi step
// doubles (IterGraph::sd)
enum : int {
    D_BETA, D_TOLI, D_RN, D_TOLO, D_R0, D_R1, D_SCALE, D_IFLOOR, D_RHO, D_ALPHA, D_OMEGA, D_R0N,
    D_RHO1, D_BBETA, D_RTOL, D_ATOL, D_G = 16
};
constexpr int D_CS = D_G + kMaxRes + 1, D_SN = D_CS + kMaxRes, D_H1 = D_SN + kMaxRes,
              D_H2 = D_H1 + kMaxRes, D_H = D_H2 + kMaxRes, D_N = D_H + (kMaxRes + 1) * kMaxRes;

// record (IterGraph::rec) per iteration
enum : int {
    R_RES = 0,     // [0..2] U, [3] p', [4] k, [5] omega: raw scaled residuals (RMS of b / diag)
    R_SC0 = 8,     // [8..11] U, p, k, omega scales before the iteration
    R_IT = 12,     // [12..15] iterations of the four solves
    R_OVER = 16,   // [16..19] solve reached its cap unfinished
    R_R0 = 20,     // [20..23] initial residual norms
    R_R1 = 24,     // [24..27] final residual norms
    R_N = 32
};

enum : int { G_LIVE, G_STEP, G_END, G_REFINE, G_ISTART, G_BEGIN, G_BICG };

__device__ __forceinline__ bool gate_open(const int* si, int g) {
    if (si[I_DONE]) return false;
    switch (g) {
    case G_STEP: return !si[I_END] && !si[I_IDONE] && !si[I_ISTART];
    case G_END: return si[I_END] != 0;
    case G_REFINE: return si[I_IDONE] != 0;
    case G_ISTART: return si[I_ISTART] != 0;
    case G_BEGIN: return !si[I_IDONE] && !si[I_ISTART];
    case G_BICG: return !si[I_SKIP] && !si[I_IDONE] && !si[I_ISTART];
    default: return true;
    }
}

__device__ __forceinline__ void set_vstop(int* si) {
    si[I_VSTOP] = (si[I_DONE] || si[I_SKIP] || si[I_IDONE] || si[I_ISTART]) ? 1 : 0;
    si[I_NSTART] = (si[I_ISTART] && !si[I_DONE]) ? 0 : 1;
}

__device__ inline double g_block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// sum of the kRedGrid block partials of slot k, fixed order
__device__ inline double g_reduce(const double* part, int k, double* sh) {
    double v = 0.0;
    for (int i = threadIdx.x; i < kRedGrid; i += blockDim.x) v += part[size_t(k) * kRedGrid + i];
    return g_block_sum(v, sh);
}

// BiCGStab breakdown: an inner product with the shadow residual that vanishes
// against ||r0|| ||r|| or is not finite (Saad 2003, 7.4)
__device__ __forceinline__ bool bicg_breakdown(double d, double r0n, double rn) {
    return !isfinite(d) || fabs(d) <= 1e-14 * r0n * rn;
}

// acc = row i of the sliced-ELL block matrix A times x (s != nullptr: s .* x)
template <int N, typename T>
__device__ __forceinline__ void g_row(int i, const int* slice_ptr, const int* slice_w, const int* scols,
                                      const T* A, const T* x, const T* s, T acc[N]) {
    const int sl = i >> 6, ln = i & 63;
    const size_t s0 = size_t(slice_ptr[sl]);
    const int w = slice_w[sl];
    const int* cp = scols + s0 * 64 + ln;
    const T* ap = A + s0 * (N * N * 64) + ln;
    for (int m = 0; m < N; ++m) acc[m] = T(0);
#pragma unroll 4
    for (int q = 0; q < w; ++q) {
        const size_t c = size_t(N) * __builtin_nontemporal_load(cp + 64 * q);
        T xv[N];
        for (int r = 0; r < N; ++r) xv[r] = s ? s[c + r] * x[c + r] : x[c + r];
        const T* blk = ap + size_t(q) * (N * N * 64);
        for (int m = 0; m < N; ++m)
            for (int r = 0; r < N; ++r) acc[m] += __builtin_nontemporal_load(blk + (N * m + r) * 64) * xv[r];
    }
}

struct SellArgs {
    int nn;
    const int *slice_ptr, *slice_w, *cols;
};

// Gated launches run on a capped grid (grid-stride), so that a launch whose
// gate is closed dispatches few workgroups.
constexpr unsigned kGridCap = 2048;

// Row blocks of a capped launch with XCD-contiguous ranges: the workgroups
// on XCD x walk the x-th eighth of the nvb row blocks (the rows' gathered
// neighbours then stay in that XCD's L2).
__device__ __forceinline__ void xcd_walk(unsigned nvb, unsigned& v0, unsigned& vend, unsigned& step) {
    const unsigned per_x = (nvb + kNumXcd - 1) / kNumXcd;
    const unsigned x = blockIdx.x % kNumXcd;
    v0 = x * per_x + blockIdx.x / kNumXcd;
    vend = min(nvb, (x + 1) * per_x);
    step = gridDim.x / kNumXcd;
}

#define GRID_STRIDE(i, n) \
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < (n); i += size_t(gridDim.x) * blockDim.x)

// ------------------------------------------------------ generic kernels
// state of a new solve
__global__ void k_g_solve_init(int* si, double* sd, int max_it, int mres, int mixed, double rtol,
                               double atol, double inner_floor) {
    if (threadIdx.x != 0) return;
    for (int k = 0; k < I_NI; ++k) si[k] = 0;
    si[I_NSTART] = 1;
    si[I_MAXIT] = max_it;
    si[I_BUDGET] = max_it;
    si[I_MRES] = mres;
    si[I_MIXED] = mixed;
    sd[D_RTOL] = rtol;
    sd[D_ATOL] = atol;
    sd[D_IFLOOR] = inner_floor;
    sd[D_R0] = sd[D_R1] = 0.0;
}

// x = 0 and partials of b . b (slot 0); with r: r = b
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_g_start_vec(size_t n, const T* b, T* x, T* r,
                                                            double* part, const int* si) {
    __shared__ double sh[kRedBlock];
    if (si[I_DONE]) return;
    double bb = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const T bi = b[i];
        x[i] = T(0);
        if (r) r[i] = bi;
        bb += double(bi) * double(bi);
    }
    const double s = g_block_sum(bb, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// r = b - A x and the partials of r . r (slot 0), rows grid-strided
template <int N, typename T>
__global__ void __launch_bounds__(kRedBlock) k_g_resid(SellArgs S, const T* A, const T* x, const T* b,
                                                        T* r, double* part, const int* si, int g) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, g)) return;
    double rr = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < S.nn; i += gridDim.x * blockDim.x) {
        T acc[N];
        g_row<N, T>(i, S.slice_ptr, S.slice_w, S.cols, A, x, nullptr, acc);
        for (int m = 0; m < N; ++m) {
            const T ri = b[N * size_t(i) + m] - acc[m];
            r[N * size_t(i) + m] = ri;
            rr += double(ri) * double(ri);
        }
    }
    const double s = g_block_sum(rr, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_g_jacobi_inv1(int nn, const int* slice_ptr, const int* slice_w, const int* doff,
                                const double* A, double* dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int sl = i >> 6, ln = i & 63;
    dinv[i] = 1.0 / A[(size_t(slice_ptr[sl]) + doff[i]) * 64 + ln];
}

__global__ void k_g_jacobi_inv3(int nn, const int* slice_ptr, const int* slice_w, const int* doff,
                                const double* A, double* dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int sl = i >> 6, ln = i & 63;
    const double* d = A + (size_t(slice_ptr[sl]) + doff[i]) * (9 * 64) + ln;
    for (int k = 0; k < 3; ++k) dinv[3 * size_t(i) + k] = 1.0 / d[4 * k * 64];
}

__global__ void k_g_to_float(size_t n, const double* x, float* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = float(x[i]);
}

// ------------------------------------------------------------- refinement
// x64 += rn * d (FP32 correction)
__global__ void k_mx_refine(size_t n, const double* sd, const float* d, double* x, const int* si) {
    if (!gate_open(si, G_REFINE)) return;
    GRID_STRIDE(i, n) x[i] += sd[D_RN] * double(d[i]);
}

// outer residual norm after a refinement step: finish, or start the next
// inner solve (budget: what the earlier ones left)
__global__ void __launch_bounds__(kRedBlock) k_mx_decide(const double* part, double* sd, int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_REFINE)) return;
    const double rn = sqrt(g_reduce(part, 0, sh));
    if (threadIdx.x != 0) return;
    const int it = si[I_IT];
    si[I_USED] += it > 1 ? it : 1;
    si[I_IDONE] = 0;
    sd[D_RN] = rn;
    if (it == 0 || !(si[I_USED] < si[I_MAXIT] && rn > sd[D_TOLO])) {
        si[I_DONE] = 1;
        sd[D_R1] = rn;
    } else {
        si[I_ISTART] = 1;
    }
    set_vstop(si);
}

// start of a refined solve: r = b (x = 0), ||r||, tolerance
__global__ void __launch_bounds__(kRedBlock) k_mx_start(const double* part, double* sd, int* si) {
    __shared__ double sh[kRedBlock];
    if (si[I_DONE]) return;
    const double rn = sqrt(g_reduce(part, 0, sh));
    if (threadIdx.x != 0) return;
    sd[D_RN] = rn;
    sd[D_R0] = rn;
    sd[D_TOLO] = fmax(sd[D_RTOL] * rn, sd[D_ATOL]);
    if (si[I_USED] < si[I_MAXIT] && rn > sd[D_TOLO]) {
        si[I_ISTART] = 1;
    } else {
        si[I_DONE] = 1;
        sd[D_R1] = rn;
    }
    set_vstop(si);
}

// right-hand side of the next inner solve: out = r64 / rn in FP32, its x =
// 0, partials of out . out (slot 0)
__global__ void __launch_bounds__(kRedBlock) k_mx_tofloat(size_t n, const double* sd, const double* r64,
                                                           float* out, float* x, double* part,
                                                           const int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_ISTART)) return;
    const double s = 1.0 / sd[D_RN];
    double bb = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const float v = float(s * r64[i]);
        out[i] = v;
        x[i] = 0.f;
        bb += double(v) * double(v);
    }
    const double t = g_block_sum(bb, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ----------------------------------------------------------------- GMRES
// new Arnoldi cycle from a residual of norm beta
// (src 1: from the residual after a restart; 2: from the right-hand side at a start)
__device__ inline void gm_new_cycle(double* sd, int* si, double beta, int src) {
    si[I_J] = 0;
    si[I_END] = 0;
    si[I_FRESH] = src;
    sd[D_BETA] = beta;
    sd[D_G] = beta;
    for (int k = 1; k <= kMaxRes; ++k) sd[D_G + k] = 0.0;
    sd[D_SCALE] = 1.0 / beta;
}

// the running (inner) solve has finished
__device__ inline void finish_inner(double* sd, int* si, double rn) {
    if (si[I_MIXED]) {
        si[I_IDONE] = 1;
    } else {
        si[I_DONE] = 1;
        sd[D_R1] = rn;
    }
}

// start of an (inner) GMRES solve from ||b|| (slot 0): tolerance, first cycle
__global__ void __launch_bounds__(kRedBlock) k_gm_start(const double* part, double* sd, int* si,
                                                        int g) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, g)) return;
    const double beta = sqrt(g_reduce(part, 0, sh));
    if (threadIdx.x != 0) return;
    if (si[I_MIXED]) {  // inner: relative to its own right-hand side
        const double inner = fmax(sd[D_TOLO] / sd[D_RN], sd[D_IFLOOR]);
        sd[D_TOLI] = fmax(inner * beta, 0.0);
        si[I_BUDGET] = si[I_MAXIT] - si[I_USED];
        si[I_ISTART] = 0;
    } else {
        sd[D_R0] = beta;
        sd[D_TOLI] = fmax(sd[D_RTOL] * beta, sd[D_ATOL]);
    }
    si[I_IT] = 0;
    if (si[I_IT] < si[I_BUDGET] && beta > sd[D_TOLI]) gm_new_cycle(sd, si, beta, 2);
    else finish_inner(sd, si, beta);
    set_vstop(si);
}

// V_0 = src / beta when the cycle just started from src (which: 1 the
// residual of a restart, 2 the right-hand side of a start)
template <typename T>
__global__ void k_gm_v0(size_t n, const T* src, T* Vb, const double* sd, const int* si, int which) {
    if (si[I_DONE] || si[I_FRESH] != which) return;
    GRID_STRIDE(i, n) Vb[i] = T(sd[D_SCALE] * double(src[i]));
}

// w = A (dinv .* V_j), row per lane
template <int N, typename T>
__global__ void __launch_bounds__(256) k_gm_spmv(SellArgs S, const T* A, const T* Vb, const T* dinv,
                                                  T* w, const int* si) {
    if (!gate_open(si, G_STEP)) return;
    const T* x = Vb + size_t(si[I_J]) * (size_t(S.nn) * N);
    unsigned v, vend, step;
    xcd_walk(unsigned((S.nn + 255) / 256), v, vend, step);
    for (; v < vend; v += step) {
        const int i = int(v * 256 + threadIdx.x);
        if (i >= S.nn) break;
        T acc[N];
        g_row<N, T>(i, S.slice_ptr, S.slice_w, S.cols, A, x, dinv, acc);
        for (int m = 0; m < N; ++m) w[N * size_t(i) + m] = acc[m];
    }
}

// partials of V_{k0+k} . w for k < min(8, j + 1 - k0), to slot k0 + k
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_gm_dots(size_t n, int k0, const T* Vb, const T* w,
                                                       double* part, const int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_STEP)) return;
    const int nk = min(8, si[I_J] + 1 - k0);
    if (nk <= 0) return;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const double wi = double(w[i]);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < nk) v[k] += double(Vb[size_t(k0 + k) * n + i]) * wi;
    }
    for (int k = 0; k < nk; ++k) {
        const double s = g_block_sum(v[k], sh);
        if (threadIdx.x == 0) part[size_t(k0 + k) * kRedGrid + blockIdx.x] = s;
    }
}

// out[k] = slot k's sum, k <= j (one workgroup per k)
__global__ void __launch_bounds__(kRedBlock) k_gm_reduce(const double* part, double* out, const int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_STEP)) return;
    const int k = blockIdx.x;
    if (k > si[I_J]) return;
    const double s = g_reduce(part, k, sh);
    if (threadIdx.x == 0) out[k] = s;
}

// w -= V h fused with the partials of V_k . w for the updated w (second
// Gram-Schmidt pass); MAXK >= j + 1 (the slot's largest possible step)
template <typename T, int MAXK>
__global__ void __launch_bounds__(kRedBlock) k_gm_upd_dots(size_t n, const T* Vb, const double* h, T* w,
                                                           double* part, const int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_STEP)) return;
    const int nk = si[I_J] + 1;
    double acc[MAXK];
#pragma unroll
    for (int k = 0; k < MAXK; ++k) acc[k] = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < nk; ++k) s += h[k] * double(Vb[size_t(k) * n + i]);
        const T wt = T(double(w[i]) - s);
        w[i] = wt;
        const double wi = double(wt);
#pragma unroll
        for (int k = 0; k < MAXK; ++k)
            if (k < nk) acc[k] += double(Vb[size_t(k) * n + i]) * wi;
    }
    for (int k = 0; k < nk; ++k) {
        const double t = g_block_sum(acc[k], sh);
        if (threadIdx.x == 0) part[size_t(k) * kRedGrid + blockIdx.x] = t;
    }
}

// w -= V h fused with the partials of w . w (slot 0)
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_gm_upd_norm(size_t n, const T* Vb, const double* h, T* w,
                                                           double* part, const int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_STEP)) return;
    const int nk = si[I_J] + 1;
    double ww = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < nk; ++k) s += h[k] * double(Vb[size_t(k) * n + i]);
        const T wt = T(double(w[i]) - s);
        w[i] = wt;
        ww += double(wt) * double(wt);
    }
    const double t = g_block_sum(ww, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Hessenberg column j from the two Gram-Schmidt passes and ||w||, the
// earlier rotations applied to it, a new rotation zeroing H[j+1][j], the
// rotated right-hand side g; the cycle ends at convergence (|g[j+1]| <= tol),
// a vanishing ||w||, the restart length or the iteration budget
__global__ void __launch_bounds__(kRedBlock) k_gm_givens(const double* part, double* sd, int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_STEP)) return;
    const double ww = g_reduce(part, 0, sh);
    if (threadIdx.x != 0) return;
    const int j = si[I_J];
    const double hn = sqrt(ww);
    double* H = sd + D_H;
    double* cs = sd + D_CS;
    double* sn = sd + D_SN;
    double* g = sd + D_G;
    for (int k = 0; k <= j; ++k) H[k * kMaxRes + j] = sd[D_H1 + k] + sd[D_H2 + k];
    H[(j + 1) * kMaxRes + j] = hn;
    for (int k = 0; k < j; ++k) {
        const double a0 = H[k * kMaxRes + j], b0 = H[(k + 1) * kMaxRes + j];
        H[k * kMaxRes + j] = cs[k] * a0 + sn[k] * b0;
        H[(k + 1) * kMaxRes + j] = -sn[k] * a0 + cs[k] * b0;
    }
    const double a0 = H[j * kMaxRes + j], b0 = H[(j + 1) * kMaxRes + j];
    const double rr = sqrt(a0 * a0 + b0 * b0);
    cs[j] = rr > 0 ? a0 / rr : 1.0;
    sn[j] = rr > 0 ? b0 / rr : 0.0;
    H[j * kMaxRes + j] = rr;
    H[(j + 1) * kMaxRes + j] = 0.0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    si[I_IT] += 1;
    si[I_FRESH] = 0;
    sd[D_SCALE] = hn > 0 ? 1.0 / hn : 0.0;
    if (fabs(g[j + 1]) <= sd[D_TOLI] || hn == 0.0 || j + 1 == si[I_MRES] || si[I_IT] >= si[I_BUDGET]) {
        si[I_END] = 1;
        si[I_JEND] = j;
    } else {
        si[I_J] = j + 1;
    }
}

// V_{j+1} = w / ||w|| (j already advanced)
template <typename T>
__global__ void k_gm_next(size_t n, const T* w, T* Vb, const double* sd, const int* si) {
    if (!gate_open(si, G_STEP)) return;
    T* out = Vb + size_t(si[I_J]) * n;
    GRID_STRIDE(i, n) out[i] = T(sd[D_SCALE] * double(w[i]));
}

// end of a cycle: y = H^-1 g by back substitution (every workgroup solves the
// small triangle itself), x += dinv .* (V y)
template <typename T>
__global__ void k_gm_xupd(size_t n, const T* Vb, const T* dinv, T* x, const double* sd, const int* si) {
    __shared__ double y[kMaxRes];
    if (!gate_open(si, G_END)) return;
    const int jend = si[I_JEND];
    if (threadIdx.x == 0) {
        const double* H = sd + D_H;
        for (int k = jend; k >= 0; --k) {
            double t = sd[D_G + k];
            for (int q = k + 1; q <= jend; ++q) t -= H[k * kMaxRes + q] * y[q];
            y[k] = t / H[k * kMaxRes + k];
        }
    }
    __syncthreads();
    GRID_STRIDE(i, n) {
        double s = 0.0;
        for (int k = 0; k <= jend; ++k) s += y[k] * double(Vb[size_t(k) * n + i]);
        x[i] = T(double(x[i]) + double(dinv[i]) * s);
    }
}

// true residual norm after a cycle (slot 0): restart, or finish the (inner) solve
__global__ void __launch_bounds__(kRedBlock) k_gm_cycle(const double* part, double* sd, int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_END)) return;
    const double beta = sqrt(g_reduce(part, 0, sh));
    if (threadIdx.x != 0) return;
    si[I_END] = 0;
    if (si[I_IT] < si[I_BUDGET] && beta > sd[D_TOLI]) gm_new_cycle(sd, si, beta, 1);
    else finish_inner(sd, si, beta);
    set_vstop(si);
}

// -------------------------------------------------------------- BiCGStab
// start of an (inner) BiCGStab solve from r = b (||r|| from slot 0)
__global__ void __launch_bounds__(kRedBlock) k_bi_start(const double* part, double* sd, int* si, int g) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, g)) return;
    const double rn = sqrt(g_reduce(part, 0, sh));
    if (threadIdx.x != 0) return;
    double tol;
    if (si[I_MIXED]) {
        const double inner = fmax(sd[D_TOLO] / sd[D_RN], sd[D_IFLOOR]);
        tol = fmax(inner * rn, 0.0);
        si[I_BUDGET] = si[I_MAXIT] - si[I_USED];
        si[I_ISTART] = 0;
    } else {
        sd[D_R0] = rn;
        tol = fmax(sd[D_RTOL] * rn, sd[D_ATOL]);
    }
    sd[D_TOLI] = tol;
    si[I_IT] = 0;
    si[I_RESTART] = 1;  // r0 = r, p = v = 0
    sd[D_RHO] = sd[D_ALPHA] = sd[D_OMEGA] = 1.0;
    sd[D_R0N] = rn;
    sd[D_RHO1] = rn * rn;
    sd[D_BETA] = rn;  // running ||r||
    if (!(si[I_IT] < si[I_BUDGET] && rn > tol)) finish_inner(sd, si, rn);
    set_vstop(si);
}

// first scalar step of an iteration: breakdown of r0 . r restarts; beta
__global__ void k_bi_begin(double* sd, int* si) {
    if (!gate_open(si, G_BEGIN) || threadIdx.x != 0) return;
    si[I_SKIP] = 0;
    const double rn = sd[D_BETA];
    if (bicg_breakdown(sd[D_RHO1], sd[D_R0N], rn)) {
        si[I_RESTART] = 1;
        sd[D_RHO] = sd[D_ALPHA] = sd[D_OMEGA] = 1.0;
        sd[D_R0N] = rn;
        sd[D_RHO1] = rn * rn;
    }
    sd[D_BBETA] = (sd[D_RHO1] / sd[D_RHO]) * (sd[D_ALPHA] / sd[D_OMEGA]);
    sd[D_RHO] = sd[D_RHO1];
    set_vstop(si);
}

// p = r + beta (p - omega v); after a restart r0 = r and p = r
template <typename V>
__global__ void k_bi_p(size_t n, const V* r, const V* v, V* p, V* r0, const double* sd, const int* si) {
    if (!gate_open(si, G_BICG)) return;
    const bool restart = si[I_RESTART] != 0;
    const double beta = sd[D_BBETA], om = sd[D_OMEGA];
    GRID_STRIDE(i, n) {
        const V ri = r[i];
        if (restart) {
            r0[i] = ri;
            p[i] = V(double(ri) + beta * (0.0 - om * 0.0));
        } else {
            p[i] = V(double(ri) + beta * (double(p[i]) - om * double(v[i])));
        }
    }
}

// y = A x and the partials of a_k . y (slots 0..na-1: a0, a1; y itself when
// a_k == nullptr)
template <typename V>
__global__ void __launch_bounds__(kRedBlock) k_bi_spmv_dots(SellArgs S, const V* A, const V* x, V* y,
                                                            const V* a0, const V* a1, double* part,
                                                            const int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_BICG)) return;
    double d0 = 0.0, d1 = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < S.nn; i += gridDim.x * blockDim.x) {
        V acc[1];
        g_row<1, V>(i, S.slice_ptr, S.slice_w, S.cols, A, x, nullptr, acc);
        y[i] = acc[0];
        const double yi = double(acc[0]);
        d0 += double(a0 ? a0[i] : acc[0]) * yi;
        if (a1) d1 += double(a1[i]) * yi;
    }
    d0 = g_block_sum(d0, sh);
    if (a1) d1 = g_block_sum(d1, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = d0;
        if (a1) part[kRedGrid + blockIdx.x] = d1;
    }
}

// r0 . v: breakdown restarts (the iteration counts), else alpha = rho / r0 . v
__global__ void __launch_bounds__(kRedBlock) k_bi_alpha(const double* part, double* sd, int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_BICG)) return;
    const double r0v = g_reduce(part, 0, sh);
    if (threadIdx.x != 0) return;
    si[I_RESTART] = 0;
    const double rn = sd[D_BETA];
    if (bicg_breakdown(r0v, sd[D_R0N], rn)) {
        si[I_RESTART] = 1;
        sd[D_RHO] = sd[D_ALPHA] = sd[D_OMEGA] = 1.0;
        sd[D_R0N] = rn;
        sd[D_RHO1] = rn * rn;
        si[I_IT] += 1;
        si[I_SKIP] = 1;
        if (!(si[I_IT] < si[I_BUDGET] && rn > sd[D_TOLI])) finish_inner(sd, si, rn);
    } else {
        sd[D_ALPHA] = sd[D_RHO] / r0v;
    }
    set_vstop(si);
}

// out = x + c y with c = sign * sd[which]
template <typename V>
__global__ void k_bi_lin(size_t n, const V* x, const V* y, V* out, const double* sd, int which,
                         const int* si) {
    if (!gate_open(si, G_BICG)) return;
    const double c = sd[which];
    GRID_STRIDE(i, n) out[i] = V(1.0 * double(x[i]) + (-c) * double(y[i]));
}

// omega = (s . t) / (t . t) (0 when t . t <= 0)
__global__ void __launch_bounds__(kRedBlock) k_bi_omega(const double* part, double* sd, const int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_BICG)) return;
    const double tt = g_reduce(part, 0, sh);
    const double ts = g_reduce(part, 1, sh);
    if (threadIdx.x == 0) sd[D_OMEGA] = tt > 0 ? ts / tt : 0.0;
}

// x += alpha ph + omega sh; r = s - omega t; partials of r . r and r0 . r
template <typename V>
__global__ void __launch_bounds__(kRedBlock) k_bi_xr(size_t n, const V* ph, const V* sh_, const V* s,
                                                     const V* t, const V* r0, V* x, V* r, double* part,
                                                     const double* sd, const int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_BICG)) return;
    const double al = sd[D_ALPHA], om = sd[D_OMEGA];
    double rr = 0.0, r0r = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        x[i] = V(double(x[i]) + al * double(ph[i]) + om * double(sh_[i]));
        const V ri = V(1.0 * double(s[i]) + (-om) * double(t[i]));
        r[i] = ri;
        rr += double(ri) * double(ri);
        r0r += double(r0[i]) * double(ri);
    }
    rr = g_block_sum(rr, sh);
    r0r = g_block_sum(r0r, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = rr;
        part[kRedGrid + blockIdx.x] = r0r;
    }
}

// end of an iteration: ||r||, r0 . r; convergence, iteration limit, omega = 0
__global__ void __launch_bounds__(kRedBlock) k_bi_end(const double* part, double* sd, int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_BICG)) return;
    const double rr = g_reduce(part, 0, sh);
    const double r0r = g_reduce(part, 1, sh);
    if (threadIdx.x != 0) return;
    const double rn = sqrt(rr);
    sd[D_BETA] = rn;
    sd[D_RHO1] = r0r;
    si[I_IT] += 1;
    if (sd[D_OMEGA] == 0.0 || !(si[I_IT] < si[I_BUDGET] && rn > sd[D_TOLI])) finish_inner(sd, si, rn);
    set_vstop(si);
}

// ---------------------------------------------- CG, mixed (FP32 inner CG)
// The inner CG runs on krylov.hip's device CG kernels and their flags
// (flag[0] stop, [1] iterations, [2] converged, [3] limit; scal[0..4]).

__global__ void k_cgm_stop(int* flag) {
    if (threadIdx.x == 0) flag[0] = 1;
}

// start of an inner CG on A d = r64 / rn in FP32: r = r64 / rn, x = 0 and the
// partials of r . r (slot 1; slot 0 gets r . z after the preconditioner)
__global__ void __launch_bounds__(kRedBlock) k_cgm_tofloat(size_t n, const double* sd, const double* r64,
                                                            float* r, float* x, double* part,
                                                            const int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_ISTART)) return;
    const double s = 1.0 / sd[D_RN];
    double rr = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const float v = float(s * r64[i]);
        r[i] = v;
        x[i] = 0.f;
        rr += double(v) * double(v);
    }
    const double t = g_block_sum(rr, sh);
    if (threadIdx.x == 0) part[kRedGrid + blockIdx.x] = t;
}

// p = z (the first direction) and the partials of r . z (slot 0)
__global__ void __launch_bounds__(kRedBlock) k_cgm_pc(size_t n, const float* r, const float* z, float* p,
                                                      double* part, const int* si) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_ISTART)) return;
    double rz = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const float zi = z[i];
        p[i] = zi;
        rz += double(r[i]) * double(zi);
    }
    const double t = g_block_sum(rz, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// inner tolerance max(max(rtol, 1e-2), tol / (2 rn)) on its right-hand side
// and its iteration budget; the inner CG's scalars and flags (k_cg_start)
__global__ void __launch_bounds__(kRedBlock) k_cgm_start(const double* part, double* sd, int* si,
                                                         double* scal, int* flag) {
    __shared__ double sh[kRedBlock];
    if (!gate_open(si, G_ISTART)) return;
    const double rz = g_reduce(part, 0, sh);
    const double rr = g_reduce(part, 1, sh);
    if (threadIdx.x != 0) return;
    const double inner = fmax(fmax(sd[D_RTOL], 1.0e-2), 0.5 * sd[D_TOLO] / sd[D_RN]);
    const int budget = si[I_MAXIT] - si[I_USED];
    const double tol2 = fmax(inner * inner * rr, 0.0);
    scal[0] = rz;
    scal[2] = tol2;
    scal[3] = rr;
    scal[4] = rr;
    flag[0] = rr <= tol2 || budget <= 0 ? 1 : 0;
    flag[1] = 0;
    flag[2] = rr <= tol2 ? 1 : 0;
    flag[3] = budget;
    si[I_ISTART] = 0;
    set_vstop(si);
}

// the inner CG has stopped: a refinement step with its iteration count
__global__ void k_cgm_check(const int* flag, int* si) {
    if (threadIdx.x != 0 || si[I_DONE] || si[I_IDONE] || si[I_ISTART] || !flag[0]) return;
    si[I_IDONE] = 1;
    si[I_IT] = flag[1];
    set_vstop(si);
}

__global__ void k_bi_jacobi(size_t n, const float* dinv, const float* r, float* z, const int* si) {
    if (!gate_open(si, G_ISTART)) return;
    GRID_STRIDE(i, n) z[i] = dinv[i] * r[i];
}

// ------------------------------------------------------- iteration record
// count, initial / final norm and cap overflow of a finished or capped solve
__global__ void k_g_solve_record(const int* si, const double* sd, double* rec, int slot) {
    if (threadIdx.x != 0) return;
    rec[R_IT + slot] = si[I_MIXED] ? si[I_USED] : si[I_IT];
    rec[R_R0 + slot] = sd[D_R0];
    rec[R_R1 + slot] = sd[D_R1];
    rec[R_OVER + slot] = si[I_DONE] ? 0.0 : 1.0;
}

// the same for the device CG (flags and scalars of krylov.hip's CG)
__global__ void k_g_cg_record(const int* flag, const double* scal, double* rec, int slot) {
    if (threadIdx.x != 0) return;
    rec[R_IT + slot] = flag[1];
    rec[R_R0 + slot] = sqrt(scal[4]);
    rec[R_R1 + slot] = sqrt(scal[3]);
    rec[R_OVER + slot] = flag[0] ? 0.0 : 1.0;
}

__global__ void k_g_rec_begin(const double* sc, double* rec) {
    if (threadIdx.x != 0) return;
    for (int k = 0; k < R_N; ++k) rec[k] = 0.0;
    for (int k = 0; k < 4; ++k) rec[R_SC0 + k] = sc[k];
}

// partials of (b / diag)^2 per component (slots 0..N-1), rows grid-strided
template <int N>
__global__ void __launch_bounds__(kRedBlock) k_g_scaled_res(SellArgs S, const int* doff, const double* A,
                                                            const double* b, double* part) {
    __shared__ double sh[kRedBlock];
    double acc[N];
    for (int k = 0; k < N; ++k) acc[k] = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < S.nn; i += gridDim.x * blockDim.x) {
        const int sl = i >> 6, ln = i & 63;
        const double* d = A + (size_t(S.slice_ptr[sl]) + doff[i]) * (N * N * 64) + ln;
        for (int k = 0; k < N; ++k) {
            const double r = b[N * size_t(i) + k] / d[(N + 1) * k * 64];
            acc[k] += r * r;
        }
    }
    for (int k = 0; k < N; ++k) {
        const double t = g_block_sum(acc[k], sh);
        if (threadIdx.x == 0) part[size_t(k) * kRedGrid + blockIdx.x] = t;
    }
}

// rec[slot + k] = RMS over the rows (component k)
__global__ void __launch_bounds__(kRedBlock) k_g_scaled_res_fin(const double* part, int N, int nn,
                                                                double* rec, int slot) {
    __shared__ double sh[kRedBlock];
    for (int k = 0; k < N; ++k) {
        const double t = g_reduce(part, k, sh);
        if (threadIdx.x == 0) rec[slot + k] = sqrt(t / nn);
    }
}

// scale of a field from its extremes (nodeField::updateScale, nodeField.hpp:5993-6015)
__device__ inline double scale_of(double maxv, double minv) {
    double sc = (maxv > 0 ? maxv : 0.0) - (minv < 0 ? minv : 0.0);
    if (sc < 1.0e-9) sc = 1.0;
    if (maxv > sc) sc = maxv;
    return sc;
}

// block maximum of vmax and minimum of vmin (all lanes get them)
__device__ inline void block_minmax(double& vmax, double& vmin, double* smax, double* smin) {
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
    vmax = smax[0];
    vmin = smin[0];
    __syncthreads();
}

// extremes of the partials of slots (k, k + 1) = (max, min)
__device__ inline void reduce_minmax(const double* part, int k, double& vmax, double& vmin,
                                     double* smax, double* smin) {
    vmax = -DBL_MAX;
    vmin = DBL_MAX;
    for (int i = threadIdx.x; i < kRedGrid; i += blockDim.x) {
        vmax = fmax(vmax, part[size_t(k) * kRedGrid + i]);
        vmin = fmin(vmin, part[size_t(k + 1) * kRedGrid + i]);
    }
    block_minmax(vmax, vmin, smax, smin);
}

// partials of max / min |U| (slots 0, 1) and of max / min p (slots 2, 3)
__global__ void __launch_bounds__(kRedBlock) k_g_flow_minmax(int nn, const double* U, const double* p,
                                                             double* part) {
    __shared__ double smax[kRedBlock], smin[kRedBlock];
    double umax = -DBL_MAX, umin = DBL_MAX, pmax = -DBL_MAX, pmin = DBL_MAX;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
        const double ux = U[3 * size_t(i)], uy = U[3 * size_t(i) + 1], uz = U[3 * size_t(i) + 2];
        const double mag = sqrt(ux * ux + uy * uy + uz * uz);
        umax = fmax(umax, mag);
        umin = fmin(umin, mag);
        pmax = fmax(pmax, p[i]);
        pmin = fmin(pmin, p[i]);
    }
    block_minmax(umax, umin, smax, smin);
    block_minmax(pmax, pmin, smax, smin);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = umax;
        part[kRedGrid + blockIdx.x] = umin;
        part[2 * kRedGrid + blockIdx.x] = pmax;
        part[3 * kRedGrid + blockIdx.x] = pmin;
    }
}

// U and p scales (nodeField::updateScale; flowModel::updatePressureScale,
// flowModel.cpp:117-128, with a constant density of scale rho_s)
__global__ void __launch_bounds__(kRedBlock) k_g_flow_scales(const double* part, double rho_s, double* sc) {
    __shared__ double smax[kRedBlock], smin[kRedBlock];
    double umax, umin, pmax, pmin;
    reduce_minmax(part, 0, umax, umin, smax, smin);
    reduce_minmax(part, 2, pmax, pmin, smax, smin);
    if (threadIdx.x == 0) {
        const double us = scale_of(umax, umin);
        const double pa = pmax - pmin, pb = 0.5 * rho_s * us * us;
        sc[0] = us;
        sc[1] = pa > pb ? pa : pb;
    }
}

// phi = max(phi + x, SMALL) and the partials of its extremes (slots 0, 1)
__global__ void __launch_bounds__(kRedBlock) k_g_correct_clip(int nn, const double* x, double* phi,
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
    block_minmax(vmax, vmin, smax, smin);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = vmax;
        part[kRedGrid + blockIdx.x] = vmin;
    }
}

__global__ void __launch_bounds__(kRedBlock) k_g_scalar_scale(const double* part, double* sc) {
    __shared__ double smax[kRedBlock], smin[kRedBlock];
    double vmax, vmin;
    reduce_minmax(part, 0, vmax, vmin, smax, smin);
    if (threadIdx.x == 0) *sc = scale_of(vmax, vmin);
}

__global__ void k_g_copy(size_t n, const double* x, double* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}

// save (dir 0) or restore (dir 1) the iteration state: entry blockIdx.y of
// the table {src, dst, 32-bit words}
__global__ void k_g_state_copy(const unsigned long long* tab, int dir) {
    const unsigned long long* e = tab + 3 * blockIdx.y;
    const unsigned* a = reinterpret_cast<const unsigned*>(e[dir]);
    unsigned* b = reinterpret_cast<unsigned*>(e[1 - dir]);
    const size_t nw = size_t(e[2]);
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nw; i += size_t(gridDim.x) * blockDim.x)
        b[i] = a[i];
}


// ------------------------------------------------------------------ host
SellArgs sell_of(const DevMesh& m) { return {m.nn, m.slice_ptr.p, m.slice_w.p, m.sell_cols.p}; }

inline unsigned nblk(size_t n) { return unsigned((n + 255) / 256); }
// grid of a gated grid-stride launch
inline unsigned gblk(size_t n) { return std::min(nblk(n), kGridCap); }

bool is_mixed(const Flow& f) { return f.prm.precision == OAGPU_MIXED; }

int gm_restart(int restart) { return std::min(std::max(restart, 1), 60); }

// the solve's finished flag (poll mode: the host stops adding slots)
bool solve_done(Flow& f, bool cg) {
    int d = 0;
    const int* src = cg ? f.flag.p : f.ig.si.p + I_DONE;
    OAGPU_HIP(hipMemcpyAsync(&d, src, sizeof(int), hipMemcpyDeviceToHost, f.stream));
    OAGPU_HIP(hipStreamSynchronize(f.stream));
    return d != 0;
}

void jacobi(Flow& f, int N) {
    DevMesh& m = *f.m;
    if (N == 3)
        k_g_jacobi_inv3<<<nblk(m.nn), 256, 0, f.stream>>>(m.nn, m.slice_ptr.p, m.slice_w.p, m.diag_off.p,
                                                          f.A3.p, f.dinv.p);
    else
        k_g_jacobi_inv1<<<nblk(m.nn), 256, 0, f.stream>>>(m.nn, m.slice_ptr.p, m.slice_w.p, m.diag_off.p,
                                                          f.A1.p, f.dinv.p);
}

// Slots that also carry the end-of-cycle and refinement launches: every
// kBlockPeriod-th and the last.  Between them a finished cycle or inner solve
// waits (its step launches are gated off), so where these blocks sit changes
// no result, only how many launches a slot without them costs.
constexpr int kBlockPeriod = 4;
bool block_slot(int s, int nslots) { return s % kBlockPeriod == kBlockPeriod - 1 || s == nslots - 1; }

// one Arnoldi step of the device GMRES on vectors T and, in block slots,
// the end of a cycle and (mixed) a refinement step and the next inner start
template <int N, typename T>
void gm_slot(Flow& f, int s, int mres, bool blocks, const T* A, const T* b, const T* dinv, T* x, T* Vb,
             T* w, T* r) {
    DevMesh& m = *f.m;
    IterGraph& g = f.ig;
    hipStream_t st = f.stream;
    const SellArgs S = sell_of(m);
    const size_t n = size_t(m.nn) * N;
    int* si = g.si.p;
    double* sd = g.sd.p;
    double* part = f.part.p;
    const int jmax = std::min(s, mres - 1);
    k_gm_spmv<N, T><<<std::min(xcd_grid(size_t(m.nn), 256), kGridCap), 256, 0, st>>>(S, A, Vb, dinv, w, si);
    for (int k0 = 0; k0 <= jmax; k0 += 8)
        k_gm_dots<T><<<kRedGrid, kRedBlock, 0, st>>>(n, k0, Vb, w, part, si);
    k_gm_reduce<<<jmax + 1, kRedBlock, 0, st>>>(part, sd + D_H1, si);
    if (jmax < 8) k_gm_upd_dots<T, 8><<<kRedGrid, kRedBlock, 0, st>>>(n, Vb, sd + D_H1, w, part, si);
    else if (jmax < 16) k_gm_upd_dots<T, 16><<<kRedGrid, kRedBlock, 0, st>>>(n, Vb, sd + D_H1, w, part, si);
    else k_gm_upd_dots<T, kMaxRes><<<kRedGrid, kRedBlock, 0, st>>>(n, Vb, sd + D_H1, w, part, si);
    k_gm_reduce<<<jmax + 1, kRedBlock, 0, st>>>(part, sd + D_H2, si);
    k_gm_upd_norm<T><<<kRedGrid, kRedBlock, 0, st>>>(n, Vb, sd + D_H2, w, part, si);
    k_gm_givens<<<1, kRedBlock, 0, st>>>(part, sd, si);
    k_gm_next<T><<<gblk(n), 256, 0, st>>>(n, w, Vb, sd, si);
    if (!blocks) return;
    // end of a cycle: least-squares update, true residual, restart or finish
    k_gm_xupd<T><<<gblk(n), 256, 0, st>>>(n, Vb, dinv, x, sd, si);
    k_g_resid<N, T><<<kRedGrid, kRedBlock, 0, st>>>(S, A, x, b, r, part, si, G_END);
    k_gm_cycle<<<1, kRedBlock, 0, st>>>(part, sd, si);
    k_gm_v0<T><<<gblk(n), 256, 0, st>>>(n, r, Vb, sd, si, 1);
    if constexpr (std::is_same_v<T, float>) {  // refinement step, next inner solve
        const double* A64 = N == 3 ? f.A3.p : f.A1.p;
        const double* b64 = N == 3 ? f.b3.p : f.b1.p;
        double* x64 = N == 3 ? f.x3.p : f.x1.p;
        k_mx_refine<<<gblk(n), 256, 0, st>>>(n, sd, x, x64, si);
        k_g_resid<N, double><<<kRedGrid, kRedBlock, 0, st>>>(S, A64, x64, b64, f.r.p, part, si, G_REFINE);
        k_mx_decide<<<1, kRedBlock, 0, st>>>(part, sd, si);
        k_mx_tofloat<<<kRedGrid, kRedBlock, 0, st>>>(n, sd, f.r.p, f.fb.p, x, part, si);
        k_gm_start<<<1, kRedBlock, 0, st>>>(part, sd, si, G_ISTART);
        k_gm_v0<T><<<gblk(n), 256, 0, st>>>(n, f.fb.p, Vb, sd, si, 2);
    }
}

// GMRES(restart) on system N (momentum / k / omega), right preconditioned by
// Jacobi; mixed form: FP32 GMRES on A d = r / ||r|| under FP64 refinement
// (inner tolerance floor 1e-5), as krylov.hip's gmres / gmres_mixed
template <int N>
void enq_gmres(Flow& f, int max_it, double rtol, double atol, int restart, int cap, bool poll) {
    DevMesh& m = *f.m;
    IterGraph& g = f.ig;
    hipStream_t st = f.stream;
    const size_t n = size_t(m.nn) * N, nv = m.nsell * N * N;
    const int mres = gm_restart(restart);
    OAGPU_REQUIRE(mres <= kMaxRes, "graph iterations: GMRES restart length above 32");
    const double* A = N == 3 ? f.A3.p : f.A1.p;
    const double* b = N == 3 ? f.b3.p : f.b1.p;
    double* x = N == 3 ? f.x3.p : f.x1.p;
    int* si = g.si.p;
    double* sd = g.sd.p;
    double* part = f.part.p;
    const bool mixed = is_mixed(f);
    jacobi(f, N);
    k_g_solve_init<<<1, 1, 0, st>>>(si, sd, max_it, mres, mixed ? 1 : 0, rtol, atol, 1.0e-5);
    const int nslots = poll ? max_it + 1 : cap;
    if (!mixed) {
        k_g_start_vec<double><<<kRedGrid, kRedBlock, 0, st>>>(n, b, x, nullptr, part, si);
        k_gm_start<<<1, kRedBlock, 0, st>>>(part, sd, si, G_LIVE);
        k_gm_v0<double><<<gblk(n), 256, 0, st>>>(n, b, f.V.p, sd, si, 2);
        for (int s = 0; s < nslots; ++s) {
            if (poll && solve_done(f, false)) break;
            gm_slot<N, double>(f, s, mres, block_slot(s, nslots), A, b, f.dinv.p, x, f.V.p, f.w.p, f.r.p);
        }
    } else {
        k_g_to_float<<<nblk(nv), 256, 0, st>>>(nv, A, f.Af.p);
        k_g_to_float<<<nblk(n), 256, 0, st>>>(n, f.dinv.p, f.fdinv.p);
        k_g_start_vec<double><<<kRedGrid, kRedBlock, 0, st>>>(n, b, x, f.r.p, part, si);
        k_mx_start<<<1, kRedBlock, 0, st>>>(part, sd, si);
        k_mx_tofloat<<<kRedGrid, kRedBlock, 0, st>>>(n, sd, f.r.p, f.fb.p, f.fx.p, part, si);
        k_gm_start<<<1, kRedBlock, 0, st>>>(part, sd, si, G_ISTART);
        k_gm_v0<float><<<gblk(n), 256, 0, st>>>(n, f.fb.p, f.fV.p, sd, si, 2);
        for (int s = 0; s < nslots; ++s) {
            if (poll && solve_done(f, false)) break;
            gm_slot<N, float>(f, s, mres, block_slot(s, nslots), f.Af.p, f.fb.p, f.fdinv.p, f.fx.p, f.fV.p,
                              f.fw.p, f.fr.p);
        }
    }
    OAGPU_HIP(hipGetLastError());
}

template <typename V>
struct BiVecs {
    V *r, *r0, *p, *v, *s, *t, *ph, *sh;  // ph = M^-1 p, sh = M^-1 s
};

// one BiCGStab iteration (Saad Alg. 7.7, right preconditioned by one
// smoothed-aggregation V-cycle) and, mixed, a refinement step and the next
// inner start
template <typename V>
void bi_slot(Flow& f, bool blocks, const V* A, const V* dinv, V* x, const BiVecs<V>& w) {
    DevMesh& m = *f.m;
    IterGraph& g = f.ig;
    hipStream_t st = f.stream;
    const SellArgs S = sell_of(m);
    const size_t n = size_t(m.nn);
    int* si = g.si.p;
    double* sd = g.sd.p;
    double* part = f.part.p;
    k_bi_begin<<<1, 64, 0, st>>>(sd, si);
    k_bi_p<V><<<gblk(n), 256, 0, st>>>(n, w.r, w.v, w.p, w.r0, sd, si);
    sa_vcycle<V>(f, A, dinv, w.p, w.ph, false, si + I_VSTOP);
    k_bi_spmv_dots<V><<<kRedGrid, kRedBlock, 0, st>>>(S, A, w.ph, w.v, w.r0, nullptr, part, si);
    k_bi_alpha<<<1, kRedBlock, 0, st>>>(part, sd, si);
    k_bi_lin<V><<<gblk(n), 256, 0, st>>>(n, w.r, w.v, w.s, sd, D_ALPHA, si);
    sa_vcycle<V>(f, A, dinv, w.s, w.sh, false, si + I_VSTOP);
    k_bi_spmv_dots<V><<<kRedGrid, kRedBlock, 0, st>>>(S, A, w.sh, w.t, nullptr, w.s, part, si);
    k_bi_omega<<<1, kRedBlock, 0, st>>>(part, sd, si);
    k_bi_xr<V><<<kRedGrid, kRedBlock, 0, st>>>(n, w.ph, w.sh, w.s, w.t, w.r0, x, w.r, part, sd, si);
    k_bi_end<<<1, kRedBlock, 0, st>>>(part, sd, si);
    if (!blocks) return;
    if constexpr (std::is_same_v<V, float>) {
        k_mx_refine<<<gblk(n), 256, 0, st>>>(n, sd, x, f.x1.p, si);
        k_g_resid<1, double><<<kRedGrid, kRedBlock, 0, st>>>(S, f.A1.p, f.x1.p, f.b1.p, f.r.p, part, si,
                                                              G_REFINE);
        k_mx_decide<<<1, kRedBlock, 0, st>>>(part, sd, si);
        k_mx_tofloat<<<kRedGrid, kRedBlock, 0, st>>>(n, sd, f.r.p, w.r, x, part, si);
        k_bi_start<<<1, kRedBlock, 0, st>>>(part, sd, si, G_ISTART);
    }
}

// smoothed-aggregation AMG-BiCGStab on the p' system; mixed: FP32 inner
// solves with the FP32 hierarchy under FP64 refinement (inner tolerance
// floor 1e-4), as krylov.hip's bicgstab_amg_core / bicgstab_amg_mixed
void enq_bicgstab_sa(Flow& f, int max_it, double rtol, double atol, int cap, bool poll) {
    DevMesh& m = *f.m;
    IterGraph& g = f.ig;
    hipStream_t st = f.stream;
    const size_t n = size_t(m.nn);
    int* si = g.si.p;
    double* sd = g.sd.p;
    double* part = f.part.p;
    const bool mixed = is_mixed(f);
    jacobi(f, 1);
    k_g_solve_init<<<1, 1, 0, st>>>(si, sd, max_it, 0, mixed ? 1 : 0, rtol, atol, 1.0e-4);
    const int nslots = poll ? max_it + 1 : cap;
    if (!mixed) {
        sa_setup(f, f.A1.p, false);
        const BiVecs<double> w{f.r.p, f.r0.p, f.pk.p, f.q.p, f.s.p, f.t.p, f.z.p, f.w.p};
        k_g_start_vec<double><<<kRedGrid, kRedBlock, 0, st>>>(n, f.b1.p, f.x1.p, w.r, part, si);
        k_bi_start<<<1, kRedBlock, 0, st>>>(part, sd, si, G_LIVE);
        for (int s = 0; s < nslots; ++s) {
            if (poll && solve_done(f, false)) break;
            bi_slot<double>(f, block_slot(s, nslots), f.A1.p, f.dinv.p, f.x1.p, w);
        }
    } else {
        k_g_to_float<<<nblk(m.nsell), 256, 0, st>>>(m.nsell, f.A1.p, f.Af.p);
        k_g_to_float<<<nblk(n), 256, 0, st>>>(n, f.dinv.p, f.fdinv.p);
        sa_setup_f32(f, f.Af.p);
        const BiVecs<float> w{f.fr.p, f.fr0.p, f.fp.p, f.fq.p, f.fs.p, f.ft.p, f.fz.p, f.fsh.p};
        k_g_start_vec<double><<<kRedGrid, kRedBlock, 0, st>>>(n, f.b1.p, f.x1.p, f.r.p, part, si);
        k_mx_start<<<1, kRedBlock, 0, st>>>(part, sd, si);
        k_mx_tofloat<<<kRedGrid, kRedBlock, 0, st>>>(n, sd, f.r.p, w.r, f.fx.p, part, si);
        k_bi_start<<<1, kRedBlock, 0, st>>>(part, sd, si, G_ISTART);
        for (int s = 0; s < nslots; ++s) {
            if (poll && solve_done(f, false)) break;
            bi_slot<float>(f, block_slot(s, nslots), f.Af.p, f.fdinv.p, f.fx.p, w);
        }
    }
    OAGPU_HIP(hipGetLastError());
}

// the start of an inner FP32 CG (gated: a start is due)
void cgm_inner_start(Flow& f, bool sa) {
    IterGraph& g = f.ig;
    hipStream_t st = f.stream;
    const size_t n = size_t(f.m->nn);
    k_cgm_tofloat<<<kRedGrid, kRedBlock, 0, st>>>(n, g.sd.p, f.r.p, f.fr.p, f.fx.p, f.part.p, g.si.p);
    if (sa) {
        sa_vcycle<float>(f, f.Af.p, f.fdinv.p, f.fr.p, f.fz.p, false, g.si.p + I_NSTART);
    } else {  // no coarse level: Jacobi, z = dinv .* r
        k_bi_jacobi<<<gblk(n), 256, 0, st>>>(n, f.fdinv.p, f.fr.p, f.fz.p, g.si.p);
    }
    k_cgm_pc<<<kRedGrid, kRedBlock, 0, st>>>(n, f.fr.p, f.fz.p, f.fp.p, f.part.p, g.si.p);
    k_cgm_start<<<1, kRedBlock, 0, st>>>(f.part.p, g.sd.p, g.si.p, f.scal.p, f.flag.p);
}

// smoothed-aggregation AMG-CG (flexible CG, krylov.hip's CG kernels) on the
// p' system; mixed: FP32 inner CG with the FP32 hierarchy under FP64
// refinement, as krylov.hip's cg / cg_mixed
void enq_cg_sa(Flow& f, int max_it, double rtol, double atol, int cap, bool poll) {
    IterGraph& g = f.ig;
    hipStream_t st = f.stream;
    const size_t n = size_t(f.m->nn);
    sa_build(f);  // once per flow, from its first p' system (never inside a capture: primed first)
    const bool sa = f.sa.lv.size() > 1;
    const int nslots = poll ? max_it + 1 : cap;
    jacobi(f, 1);
    if (!is_mixed(f)) {
        if (sa) sa_setup(f, f.A1.p, false);
        cg_device_begin(f, max_it, rtol, atol, sa);
        for (int s = 0; s < nslots; ++s) {
            if (poll && solve_done(f, true)) break;
            cg_device_iteration(f, sa);
        }
        return;
    }
    int* si = g.si.p;
    double* sd = g.sd.p;
    double* part = f.part.p;
    k_g_to_float<<<nblk(f.m->nsell), 256, 0, st>>>(f.m->nsell, f.A1.p, f.Af.p);
    k_g_to_float<<<nblk(n), 256, 0, st>>>(n, f.dinv.p, f.fdinv.p);
    if (sa) sa_setup_f32(f, f.Af.p);
    k_g_solve_init<<<1, 1, 0, st>>>(si, sd, max_it, 0, 1, rtol, atol, 0.0);
    k_cgm_stop<<<1, 1, 0, st>>>(f.flag.p);
    k_g_start_vec<double><<<kRedGrid, kRedBlock, 0, st>>>(n, f.b1.p, f.x1.p, f.r.p, part, si);
    k_mx_start<<<1, kRedBlock, 0, st>>>(part, sd, si);
    cgm_inner_start(f, sa);
    for (int s = 0; s < nslots; ++s) {
        if (poll && solve_done(f, false)) break;
        cg_device_iteration_f32(f, sa);
        if (!block_slot(s, nslots)) continue;
        k_cgm_check<<<1, 1, 0, st>>>(f.flag.p, si);
        k_mx_refine<<<gblk(n), 256, 0, st>>>(n, sd, f.fx.p, f.x1.p, si);
        k_g_resid<1, double><<<kRedGrid, kRedBlock, 0, st>>>(sell_of(*f.m), f.A1.p, f.x1.p, f.b1.p, f.r.p,
                                                              part, si, G_REFINE);
        k_mx_decide<<<1, kRedBlock, 0, st>>>(part, sd, si);
        cgm_inner_start(f, sa);
    }
}

// solve `which` (0 momentum, 1 p', 2 k, 3 omega) and its record entries
void enq_solve(Flow& f, int which, int N, int method, int max_it, double rtol, int restart, int cap,
               bool poll) {
    IterGraph& g = f.ig;
    const double atol = 1e-30;
    if (method == OAGPU_GMRES) {
        if (N == 3) enq_gmres<3>(f, max_it, rtol, atol, restart, cap, poll);
        else enq_gmres<1>(f, max_it, rtol, atol, restart, cap, poll);
        k_g_solve_record<<<1, 1, 0, f.stream>>>(g.si.p, g.sd.p, g.rec.p, which);
    } else if (method == OAGPU_BICGSTAB_SA) {
        enq_bicgstab_sa(f, max_it, rtol, atol, cap, poll);
        k_g_solve_record<<<1, 1, 0, f.stream>>>(g.si.p, g.sd.p, g.rec.p, which);
    } else {  // OAGPU_CG_SA
        enq_cg_sa(f, max_it, rtol, atol, cap, poll);
        if (is_mixed(f)) k_g_solve_record<<<1, 1, 0, f.stream>>>(g.si.p, g.sd.p, g.rec.p, which);
        else k_g_cg_record<<<1, 1, 0, f.stream>>>(f.flag.p, f.scal.p, g.rec.p, which);
    }
    OAGPU_HIP(hipGetLastError());
}

void enq_scaled_res(Flow& f, int N, int slot) {
    DevMesh& m = *f.m;
    const SellArgs S = sell_of(m);
    if (N == 3)
        k_g_scaled_res<3><<<kRedGrid, kRedBlock, 0, f.stream>>>(S, m.diag_off.p, f.A3.p, f.b3.p, f.part.p);
    else
        k_g_scaled_res<1><<<kRedGrid, kRedBlock, 0, f.stream>>>(S, m.diag_off.p, f.A1.p, f.b1.p, f.part.p);
    k_g_scaled_res_fin<<<1, kRedBlock, 0, f.stream>>>(f.part.p, N, m.nn, f.ig.rec.p, slot);
}

// field_scales of fields.hip on the device (rho scale as there)
void enq_flow_scales(Flow& f) {
    double rs = (f.prm.rho > 0 ? f.prm.rho : 0.0) - (f.prm.rho < 0 ? f.prm.rho : 0.0);
    if (rs < 1.0e-9) rs = 1.0;
    if (f.prm.rho > rs) rs = f.prm.rho;
    k_g_flow_minmax<<<kRedGrid, kRedBlock, 0, f.stream>>>(f.m->nn, f.U.p, f.p.p, f.part.p);
    k_g_flow_scales<<<1, kRedBlock, 0, f.stream>>>(f.part.p, rs, f.ig.sc.p);
}

// k or omega += x, clipped at SMALL, and its scale (launch_correct_scalar of sst.hip)
void enq_correct_scalar(Flow& f, int which) {
    double* phi = which == 0 ? f.k.p : f.om.p;
    k_g_correct_clip<<<kRedGrid, kRedBlock, 0, f.stream>>>(f.m->nn, f.x1.p, phi, f.part.p);
    k_g_scalar_scale<<<1, kRedBlock, 0, f.stream>>>(f.part.p, f.ig.sc.p + 2 + which);
}

bool needs_snapshot(const Flow& f, const int* maxit) {
    for (int k = 0; k < 4; ++k)
        if (f.ig.cap[k] < maxit[k]) return true;
    return false;
}

// One outer iteration: segregatedFlowEquations::solve for one p' sub-iteration
// (segregatedFlowEquations.cpp:76-297) then, with sctl,
// segregatedShearStressTransportEquations::solve (:86-99); the order of
// capi.cpp's run_outer_iteration and oagpu_sst_iteration.  poll: no graph,
// every solve run to its end.
void enq_iteration(Flow& f, const double* ctl, const double* sctl, const int* maxit, bool poll) {
    IterGraph& g = f.ig;
    hipStream_t st = f.stream;
    const size_t nn = size_t(f.m->nn);
    if (!poll && needs_snapshot(f, maxit))
        k_g_state_copy<<<dim3(256, unsigned(g.ntab)), 256, 0, st>>>(g.tab.p, 0);
    k_g_rec_begin<<<1, 1, 0, st>>>(g.sc.p, g.rec.p);
    // momentum predictor (navierStokesEquation::solve)
    launch_assemble_momentum(f);
    enq_scaled_res(f, 3, R_RES);
    enq_solve(f, 0, 3, OAGPU_GMRES, maxit[0], ctl[1], 30, g.cap[0], poll);
    launch_correct(f, 3, 1.0);
    launch_update_mdot(f);
    enq_flow_scales(f);
    // pressure correction (pressureCorrectionEquation::solve)
    launch_assemble_pcorr(f);
    enq_scaled_res(f, 1, R_RES + 3);
    enq_solve(f, 1, 1, f.prm.p_method, maxit[1], ctl[3], 30, g.cap[1], poll);
    launch_correct(f, 1, f.prm.urf_p);
    // mass flow, divergence, velocity correction, gradients
    launch_update_mdot(f);
    if (f.has_outlet) launch_reversal(f);
    launch_mass_divergence(f);
    launch_velocity_correction(f, 1);
    launch_gradient(f, 1);
    launch_velocity_correction(f, 2);
    enq_flow_scales(f);
    launch_gradient(f, 0);
    if (f.prm.high_res) launch_blending(f, 0);
    if (sctl) {
        k_g_copy<<<nblk(nn), 256, 0, st>>>(nn, f.k.p, f.kprev.p);
        launch_sst_update(f);
        const bool pair = scalar_pair_ok(f);
        if (pair) launch_assemble_scalar_pair(f);  // omega's system to A1w / b1w
        else launch_assemble_scalar(f, 0);
        enq_scaled_res(f, 1, R_RES + 4);
        enq_solve(f, 2, 1, OAGPU_GMRES, maxit[2], sctl[1], 30, g.cap[2], poll);
        enq_correct_scalar(f, 0);
        if (pair) {  // the omega system in place of the k one while it is solved
            std::swap(f.A1, f.A1w);
            std::swap(f.b1, f.b1w);
        } else {
            launch_assemble_scalar(f, 1);
        }
        struct SwapBack {
            Flow& f;
            bool on;
            ~SwapBack() {
                if (on) {
                    std::swap(f.A1, f.A1w);
                    std::swap(f.b1, f.b1w);
                }
            }
        } back{f, pair};
        enq_scaled_res(f, 1, R_RES + 5);
        enq_solve(f, 3, 1, OAGPU_GMRES, maxit[3], sctl[1], 30, g.cap[3], poll);
        enq_correct_scalar(f, 1);
        launch_gradient_kw(f);
    }
    OAGPU_HIP(hipGetLastError());
}

// buffers of the device solves and the saved-state table
void prepare(Flow& f, const double* sctl) {
    DevMesh& m = *f.m;
    IterGraph& g = f.ig;
    const size_t nn = size_t(m.nn), n3 = 3 * nn;
    OAGPU_REQUIRE(f.prm.p_method == OAGPU_BICGSTAB_SA || f.prm.p_method == OAGPU_CG_SA,
                  "graph iterations need the smoothed-aggregation p' solver (OAGPU_BICGSTAB_SA or OAGPU_CG_SA)");
    OAGPU_REQUIRE(!sctl || f.sst_on, "SST controls given but SST is not enabled on this flow");
    OAGPU_REQUIRE(f.tscheme == 0 || f.tdt > 0, "transient flow without a time step");
    if (g.sd.n < size_t(D_N)) g.sd.alloc(D_N);
    if (g.si.n < size_t(I_NI)) g.si.alloc(I_NI);
    if (g.rec.n < size_t(R_N)) g.rec.alloc(R_N);
    if (g.sc.n < 4) g.sc.alloc(4);
    const int mres = 30;
    if (f.V_cols < mres + 1 || f.V.n < n3 * size_t(mres + 1)) {
        f.V.alloc(n3 * size_t(mres + 1));
        f.V_cols = mres + 1;
    }
    for (auto* v : {&f.r, &f.w, &f.z, &f.pk, &f.q, &f.s, &f.t, &f.r0})
        if (v->n < n3) v->alloc(n3);
    if (is_mixed(f)) {
        if (f.Af.n < m.nsell * 9) f.Af.alloc(m.nsell * 9);
        for (auto* v : {&f.fb, &f.fx, &f.fr, &f.fw, &f.fdinv, &f.fz, &f.fp, &f.fq, &f.fr0, &f.fs, &f.ft, &f.fsh})
            if (v->n < n3) v->alloc(n3);
        if (f.fV.n < n3 * size_t(mres + 1)) f.fV.alloc(n3 * size_t(mres + 1));
    }
    // the state an outer iteration reads and rewrites
    std::vector<std::pair<void*, size_t>> ent;
    auto add = [&](auto& b) {
        if (b.p && b.n) ent.push_back({(void*)b.p, b.n * sizeof(*b.p)});
    };
    add(f.U), add(f.p), add(f.gradU), add(f.gradP), add(f.du), add(f.mdot), add(f.div), add(f.mueff);
    add(f.beta), add(f.beta_k), add(f.beta_w), add(f.msd), add(f.rfl);
    add(f.k), add(f.om), add(f.gradK), add(f.gradW), add(f.mut), add(f.F1), add(f.Pk), add(f.Pk_wall);
    add(f.kprev), add(g.sc);
    size_t words = 0;
    for (auto& e : ent) words += e.second / 4;
    if (g.snap.n * 2 < words) g.snap.alloc((words + 1) / 2);
    std::vector<unsigned long long> tab;
    size_t off = 0;
    for (auto& e : ent) {
        tab.push_back((unsigned long long)e.first);
        tab.push_back((unsigned long long)((unsigned*)g.snap.p + off));
        tab.push_back(e.second / 4);
        off += e.second / 4;
    }
    if (g.tab.n < tab.size()) g.tab.alloc(tab.size());
    g.tab.upload(tab.data(), tab.size(), f.stream);
    g.ntab = int(ent.size());
}

// what the captured sequence depends on besides the device state
std::vector<double> graph_key(const Flow& f, const double* ctl, const double* sctl) {
    std::vector<double> k(ctl, ctl + 4);
    if (sctl) k.insert(k.end(), sctl, sctl + 2);
    else k.insert(k.end(), {-1.0, -1.0});
    for (int c : f.ig.cap) k.push_back(c);
    const Flow::Tuning& t = f.tune;
    for (double v : {double(f.prm.p_method), double(f.prm.precision), double(f.prm.high_res),
                     double(f.has_outlet), double(t.mom_tile), double(t.pc_occupancy),
                     double(t.cap_scalar), double(t.cap_mdot), double(t.cap_grad), double(t.scalar_pair),
                     double(t.spmv_block[0][0]), double(t.spmv_block[0][1]),
                     double(t.spmv_block[1][0]), double(t.spmv_block[1][1]),
                     double(f.sa.lv.size()), double(f.tscheme), f.tdt, f.tdt_prev,
                     double(uintptr_t(f.sa.lv.empty() ? nullptr : f.sa.lv[0].P.p)),
                     double(uintptr_t(f.ig.tab.p)), double(f.ig.ntab)})
        k.push_back(v);
    // operands captured by address
    for (const void* q : {(const void*)f.V.p, (const void*)f.fV.p, (const void*)f.Af.p, (const void*)f.r.p,
                          (const void*)f.w.p, (const void*)f.z.p, (const void*)f.pk.p, (const void*)f.q.p,
                          (const void*)f.s.p, (const void*)f.t.p, (const void*)f.r0.p, (const void*)f.fb.p,
                          (const void*)f.fx.p, (const void*)f.fr.p, (const void*)f.fw.p, (const void*)f.fdinv.p,
                          (const void*)f.fz.p, (const void*)f.fp.p, (const void*)f.fq.p, (const void*)f.fr0.p,
                          (const void*)f.fs.p, (const void*)f.ft.p, (const void*)f.fsh.p, (const void*)f.dinv.p,
                          (const void*)f.ig.snap.p, (const void*)f.ig.sd.p, (const void*)f.ig.si.p,
                          (const void*)f.ig.rec.p, (const void*)f.ig.sc.p, (const void*)f.A1.p,
                          (const void*)f.A1w.p, (const void*)f.b1.p, (const void*)f.b1w.p})
        k.push_back(double(uintptr_t(q)));
    return k;
}

// capture the iteration (the flow's kernel clocks stay off inside it)
void capture(Flow& f, const double* ctl, const double* sctl, const int* maxit) {
    IterGraph& g = f.ig;
    g.reset();
    const unsigned mask = f.prof_mask;
    f.prof_mask = 0;
    hipGraph_t gr = nullptr;
    OAGPU_HIP(hipStreamBeginCapture(f.stream, hipStreamCaptureModeThreadLocal));
    try {
        enq_iteration(f, ctl, sctl, maxit, false);
    } catch (...) {
        f.prof_mask = mask;
        (void)hipStreamEndCapture(f.stream, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        throw;
    }
    f.prof_mask = mask;
    OAGPU_HIP(hipStreamEndCapture(f.stream, &gr));
    const hipError_t e = hipGraphInstantiate(&g.exec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    OAGPU_HIP(e);
    g.key = graph_key(f, ctl, sctl);
    ++g.captures;
}

}  // namespace

int graph_iterations(Flow& f, const double* ctl, const double* sctl, int n, double tol, double* hist,
                     bool eager) {
    IterGraph& g = f.ig;
    prepare(f, sctl);
    const int maxit[4] = {int(ctl[0]), int(ctl[2]), sctl ? int(sctl[0]) : 0, sctl ? int(sctl[0]) : 0};
    for (int k = 0; k < 2; ++k) OAGPU_REQUIRE(maxit[k] >= 1, "iteration limits must be >= 1");
    OAGPU_REQUIRE(!sctl || maxit[2] >= 1, "iteration limits must be >= 1");
    const double h_sc[4] = {f.U_scale, f.p_scale, f.k_scale, f.w_scale};
    g.sc.upload(h_sc, 4, f.stream);
    const double kEps = 2.220446049250313e-16;
    int it = 0;
    double rec[R_N];
    auto read_rec = [&] {
        int e = 0;
        OAGPU_HIP(hipMemcpyAsync(rec, g.rec.p, sizeof(rec), hipMemcpyDeviceToHost, f.stream));
        OAGPU_HIP(hipMemcpyAsync(&e, f.err.p, sizeof(int), hipMemcpyDeviceToHost, f.stream));
        OAGPU_HIP(hipStreamSynchronize(f.stream));
        if (e) {
            f.err.zero(f.stream);
            throw Error(OAGPU_EGEOM, "non-positive Jacobian determinant in a kernel");
        }
    };
    // Krylov solves with a small iteration limit (the GMRES ones, 20) get it as
    // their cap and never roll back (their slots are cheap when unused); the
    // others (p', up to 200) get the largest count seen plus a quarter, at
    // least 4 slots more: a roll-back costs about one iteration and a re-capture
    auto cap_of = [&](int k, int count) {
        if (maxit[k] <= 40) return std::max(1, maxit[k]);
        return std::max(1, std::min(maxit[k], count + std::max(4, count / 4)));
    };
    auto set_caps = [&](const int* want) {
        bool changed = false;
        for (int k = 0; k < 4; ++k) {
            const int c = cap_of(k, want[k]);
            changed = changed || c != g.cap[k];
            g.cap[k] = c;
        }
        if (changed) g.reset();
        g.since = 0;
        for (int k = 0; k < 4; ++k) g.seen[k] = 0;
    };
    for (; it < n; ++it) {
        if (eager || !g.primed) {
            enq_iteration(f, ctl, sctl, maxit, true);
            read_rec();
            if (!eager) {
                int want[4];
                for (int k = 0; k < 4; ++k) want[k] = int(rec[R_IT + k]);
                set_caps(want);
                g.primed = true;
            }
        } else {
            if (!g.exec || g.key != graph_key(f, ctl, sctl)) capture(f, ctl, sctl, maxit);
            OAGPU_HIP(hipGraphLaunch(g.exec, f.stream));
            ++g.launches;
            read_rec();
            bool over = false;
            for (int k = 0; k < 4; ++k) over = over || rec[R_OVER + k] != 0.0;
            if (over) {  // a solve needed more slots: redo the iteration uncapped
                k_g_state_copy<<<dim3(256, unsigned(g.ntab)), 256, 0, f.stream>>>(g.tab.p, 1);
                enq_iteration(f, ctl, sctl, maxit, true);
                read_rec();
                ++g.replays_redone;
                int want[4];
                for (int k = 0; k < 4; ++k) want[k] = std::max(g.cap[k], int(rec[R_IT + k]));
                set_caps(want);
            } else {
                for (int k = 0; k < 4; ++k) g.seen[k] = std::max(g.seen[k], int(rec[R_IT + k]));
                if (++g.since >= 16) {  // tighten caps that the last 16 iterations left unused
                    int want[4];
                    bool shrink = false;
                    for (int k = 0; k < 4; ++k) {
                        want[k] = g.seen[k];
                        shrink = shrink || cap_of(k, want[k]) + 4 < g.cap[k];
                    }
                    if (shrink) set_caps(want);
                    else {
                        g.since = 0;
                        for (int k = 0; k < 4; ++k) g.seen[k] = 0;
                    }
                }
            }
        }
        const int nsolve = sctl ? 4 : 2;
        for (int k = 0; k < nsolve; ++k)
            if (!std::isfinite(rec[R_R0 + k]) || !std::isfinite(rec[R_R1 + k]))
                throw Error(OAGPU_ENONFINITE, "the linear solve produced a non-finite residual");
        double* h = hist ? hist + size_t(it) * 10 : nullptr;
        double hv[10];
        for (int k = 0; k < 3; ++k) hv[k] = rec[R_RES + k] / (rec[R_SC0] + kEps);
        hv[3] = rec[R_RES + 3] / (rec[R_SC0 + 1] + kEps);
        hv[4] = rec[R_IT];
        hv[5] = rec[R_IT + 1];
        hv[6] = sctl ? rec[R_RES + 4] / (rec[R_SC0 + 2] + kEps) : 0.0;
        hv[7] = sctl ? rec[R_RES + 5] / (rec[R_SC0 + 3] + kEps) : 0.0;
        hv[8] = sctl ? rec[R_IT + 2] : 0.0;
        hv[9] = sctl ? rec[R_IT + 3] : 0.0;
        if (h) std::copy(hv, hv + 10, h);
        if (tol > 0) {
            double mx = std::max({hv[0], hv[1], hv[2], hv[3]});
            if (sctl) mx = std::max({mx, hv[6], hv[7]});
            if (mx < tol) {
                ++it;
                break;
            }
        }
    }
    double sc[4];
    g.sc.download(sc, 4, f.stream);
    OAGPU_HIP(hipStreamSynchronize(f.stream));
    f.U_scale = sc[0];
    f.p_scale = sc[1];
    if (sctl) {
        f.k_scale = sc[2];
        f.w_scale = sc[3];
    }
    return it;
}

}  // namespace oagpu
