// amg.hip -- aggregation algebraic multigrid preconditioner for the scalar
// (pressure-correction) systems.
//
// The reference reaches algebraic multigrid through liblinsolve's HYPRE
// family (BoomerAMG inside GMRES/CG, linearSystem.h:309-358); the library is
// not part of the reference tree, so this is a design of its own, chosen for
// the MI355X: plain (unsmoothed) aggregation (Vanek, Mandel, Brezina,
// Computing 56 (1996), phase 1-3 greedy aggregation over the strong
// couplings) with piecewise-constant prolongation, so that the Galerkin
// product P^T A P is a plain sum of fine coefficients into their aggregate's
// row.
//
// * The hierarchy structure (aggregates, coarse patterns, the fine-slot ->
//   coarse-slot map) depends on the mesh graph and on the coupling strengths
//   of the first matrix solved, and is built once, on the host.  Every solve recomputes the coarse values on the device from the
//   current fine matrix: one lane per coarse sliced-ELL slot sums the fine
//   slots listed for it in a fixed order (no atomics, bitwise reproducible).
// * Cycle: V(1,1) with damped Jacobi (omega = 2/3) and the coarse correction
//   scaled by alpha = 2 (unsmoothed aggregation under-corrects), symmetric, so
//   it preconditions CG and BiCGStab.  The coarsest level (<= 64 rows) is
//   solved with its dense inverse, computed per solve by Gauss-Jordan in one
//   workgroup's LDS.  Under (flexible) CG the coarse levels above the
//   one-workgroup tail are K-cycle levels instead: two flexible-CG steps
//   each, preconditioned by the cycle below, with no alpha scaling; the
//   iteration count then stays nearly flat with the number of levels (on a
//   7-point Laplacian, 128^3: 22 iterations to 1e-6 against 47 for V(1,1);
//   on the 128^3 cavity's p' systems 25 against 63 in FP64).  Each K-cycle
//   launches ~35 kernels against ~12 for a V-cycle, most of them on levels
//   small enough to be latency-bound, so the gain in time is smaller than in
//   iterations (FP64 4 %, mixed form within 1 % on that cavity).
// * Values and vectors of every level are of the Krylov type V: double for
//   the FP64 form, float for the FP32 inner solve of the mixed form.
#include "device_util.hpp"
#include "flow.hpp"
#include "sell.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <numeric>
#include <thread>
#include <type_traits>

namespace oagpu {

constexpr int kDenseMax = 64;      // largest coarsest level inverted densely
constexpr int kMaxLevels = 16;
constexpr double kOmega = 2.0 / 3.0;
constexpr double kAlpha = 2.0;
// strength thresholds: the fine level's separates the wall-normal couplings of
// stretched cells; the coarse levels keep nearly all couplings (on the 128^3
// cavity a coarse threshold of 0.04 and above grew the K-cycle's coarse
// levels by half for no fewer iterations; on the 8M-cell step 0.02 gave
// 24-27 AMG-BiCGStab iterations per p' solve where the graph-only
// aggregation took 85)
constexpr double kTheta = 0.08;
constexpr double kThetaCoarse = 0.02;

struct SellView {
    int n;
    const int* slice_ptr;
    const int* slice_w;
    const int* cols;
    const int* diag_off;
};

__device__ __forceinline__ size_t sv_base(const SellView& s, int i, int& w) {
    const int sl = i >> 6;
    w = s.slice_w[sl];
    return size_t(s.slice_ptr[sl]) * 64 + (i & 63);
}

// --------------------------------------------------------------- set-up
// Galerkin coarse values: one lane per coarse sliced-ELL slot sums the fine
// slots listed for it (members ascending, then their slots ascending; an
// empty list is padding or an absent entry and gives 0), so the sum order is
// fixed and the result bitwise reproducible.
template <typename V>
__global__ void __launch_bounds__(256) k_amg_galerkin(size_t nsell, const int* gptr,
                                                      const int* glist, const V* Af, V* Ac) {
    const size_t p = size_t(xcd_block()) * blockDim.x + threadIdx.x;
    if (p >= nsell) return;
    V s = V(0);
    const int k1 = gptr[p + 1];
    for (int k = gptr[p]; k < k1; ++k) s += Af[glist[k]];
    Ac[p] = s;
}

template <typename V>
__global__ void __launch_bounds__(256) k_amg_dinv(SellView s, const V* A, V* dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    int w;
    const size_t b = sv_base(s, i, w);
    const V d = A[b + size_t(s.diag_off[i]) * 64];
    dinv[i] = d != V(0) ? V(1) / d : V(0);
}

// In-place Gauss-Jordan inverse of the coarsest matrix (n <= kDenseMax) in
// LDS, double arithmetic whatever V is.  A pivot that vanishes (a singular,
// pure-Neumann coarsest level) drops its row and column from the inverse.
template <typename V>
__global__ void __launch_bounds__(256) k_amg_dense_inverse(SellView s, const V* A, double* inv,
                                                           V* vinv) {
    __shared__ double M[kDenseMax][kDenseMax + 1];
    __shared__ double colk[kDenseMax];
    __shared__ double piv;
    __shared__ int skip;
    const int n = s.n;
    __shared__ double dmax;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) M[e / n][e % n] = 0.0;
    __syncthreads();
    if (threadIdx.x < n) {
        const int i = threadIdx.x;
        int w;
        const size_t b = sv_base(s, i, w);
        for (int q = 0; q < w; ++q) M[i][s.cols[b + size_t(q) * 64]] += double(A[b + size_t(q) * 64]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int i = 0; i < n; ++i) m = fmax(m, fabs(M[i][i]));
        dmax = m;
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        if (threadIdx.x == 0) {
            piv = M[k][k];
            skip = !(fabs(piv) > 1.0e-13 * dmax);
        }
        __syncthreads();
        if (skip) {
            for (int j = threadIdx.x; j < n; j += blockDim.x) {
                M[k][j] = 0.0;
                M[j][k] = 0.0;
            }
            __syncthreads();
            continue;
        }
        // row k of the inverse-in-progress, column k saved before elimination
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            colk[j] = M[j][k];
            M[k][j] = (j == k ? 1.0 : M[k][j]) / piv;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
            const int i = e / n, j = e % n;
            if (i != k) M[i][j] = (j == k ? 0.0 : M[i][j]) - colk[i] * M[k][j];
        }
        __syncthreads();
    }
    // stored transposed: lane i of the apply kernel reads inv[j][i], coalesced
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, j = e % n;
        inv[size_t(j) * n + i] = M[i][j];
        vinv[size_t(j) * n + i] = V(M[i][j]);
    }
}

// ----------------------------------------------------------------- cycle
// x = omega dinv .* r
template <typename V>
__global__ void __launch_bounds__(256) k_amg_jacobi0(int n, V omega, const V* dinv, const V* r, V* x,
                                                     const int* flag) {
    if (flag[0]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = omega * dinv[i] * r[i];
}

// res = r - A x
template <typename V>
__global__ void __launch_bounds__(256) k_amg_residual(SellView s, const V* A, const V* x, const V* r,
                                                      V* res, const int* flag) {
    if (flag[0]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    int w;
    const size_t b = sv_base(s, i, w);
    V acc = V(0);
    for (int q = 0; q < w; ++q) acc += A[b + size_t(q) * 64] * x[s.cols[b + size_t(q) * 64]];
    res[i] = r[i] - acc;
}

// rc[I] = sum of res over the members of I (mem == nullptr: res stored in
// member order, a contiguous range); xc = omega dinv_c rc (first smoothing
// step of the coarse level from a zero guess; skipped when dinv_c is null)
template <typename V>
__global__ void __launch_bounds__(256) k_amg_restrict(int nc, const int* mem_ptr, const int* mem,
                                                      const V* res, V omega, const V* dinvc, V* rc,
                                                      V* xc, const int* flag) {
    if (flag[0]) return;
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nc) return;
    V s = V(0);
    const int k1 = mem_ptr[I + 1];
    if (mem) {
#pragma unroll 4
        for (int k = mem_ptr[I]; k < k1; ++k) s += res[mem[k]];
    } else {
#pragma unroll 8
        for (int k = mem_ptr[I]; k < k1; ++k) s += res[k];  // res in member order
    }
    rc[I] = s;
    if (dinvc) xc[I] = omega * dinvc[I] * s;
}

// t = x + alpha P ec
template <typename V>
__global__ void __launch_bounds__(256) k_amg_prolong(int n, const int* agg, V alpha, const V* ec,
                                                     const V* x, V* t, const int* flag) {
    if (flag[0]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) t[i] = x[i] + alpha * ec[agg[i]];
}

// xo = t + omega dinv (r - A t)
template <typename V>
__global__ void __launch_bounds__(256) k_amg_smooth(SellView s, const V* A, const V* dinv, V omega,
                                                    const V* r, const V* t, V* xo, const int* flag) {
    if (flag[0]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    int w;
    const size_t b = sv_base(s, i, w);
    V acc = V(0);
    for (int q = 0; q < w; ++q) acc += A[b + size_t(q) * 64] * t[s.cols[b + size_t(q) * 64]];
    xo[i] = t[i] + omega * dinv[i] * (r[i] - acc);
}

__device__ inline double amg_block_sum(double v, double* sh) {
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

// ------------------------------------------------------------ K-cycle
// A coarse level c of the K-cycle (Notay and Vassilevski, Numer. Linear
// Algebra Appl. 15 (2008)) is solved by two flexible-CG steps, each
// preconditioned by the cycle of level c:
//   c1 = B rc, v1 = A c1, rho1 = c1.v1, a1 = c1.rc, r2 = rc - (a1/rho1) v1,
//   c2 = B r2, gam = c2.v1, bet = c2.A c2, a2 = c2.r2, rho2 = bet - gam^2/rho1,
//   e = (a1/rho1 - gam a2/(rho1 rho2)) c1 + (a2/rho2) c2.
// Dot products are block partials over a fixed grid of G blocks
// (kp[k*G+b]); every block that needs a step coefficient reduces them in the
// same fixed order (no inter-block signalling; bitwise reproducible).  r2 is
// formed inside the pre-smoothing kernel of the second visit, e inside the
// prolongation that consumes it (on the coarse levels; for the fine level by
// k_kc_combine, as every one of the fine prolongation's blocks reducing the
// partials would cost more than that pass).
struct KcRef {
    const double* kp1 = nullptr;  // rho1, a1
    const double* kp2 = nullptr;  // bet, gam, a2
    int G = 0;
};

__device__ inline double kc_reduce(const double* kp, int k, int G, double* sh) {
    double v = 0.0;
    for (int i = threadIdx.x; i < G; i += blockDim.x) v += kp[size_t(k) * G + i];
    return amg_block_sum(v, sh);
}

// a1 / rho1 (0 when c1 = 0)
__device__ inline double kc_step1(const KcRef& k, double* sh) {
    const double rho1 = kc_reduce(k.kp1, 0, k.G, sh), a1 = kc_reduce(k.kp1, 1, k.G, sh);
    return rho1 != 0.0 ? a1 / rho1 : 0.0;
}

// e = ca c1 + cb c2
__device__ inline void kc_coeffs(const KcRef& k, double* sh, double& ca, double& cb) {
    const double rho1 = kc_reduce(k.kp1, 0, k.G, sh), a1 = kc_reduce(k.kp1, 1, k.G, sh);
    const double bet = kc_reduce(k.kp2, 0, k.G, sh), gam = kc_reduce(k.kp2, 1, k.G, sh);
    const double a2 = kc_reduce(k.kp2, 2, k.G, sh);
    ca = 0.0;
    cb = 0.0;
    if (rho1 > 0.0) {
        ca = a1 / rho1;
        const double rho2 = bet - gam * gam / rho1;
        if (rho2 > 1e-12 * bet) {
            cb = a2 / rho2;
            ca -= gam * cb / rho1;
        }
    }
}

// Pre-smoothing from a zero guess fused with the residual:
//   x = omega D^-1 r,  res = r - A x  (x of the neighbours formed on the fly)
// With kc.kp1 (second visit of a K-cycle level) the right-hand side is
// r - (a1/rho1) v1, formed on the fly and written to rout.
// With spos the residual is stored at spos[i] (member order of the next
// level, for the one-workgroup tail's restriction).
// FINE only names the fine-level instance separately (profiles, timing).
// S: storage type of the matrix values (float values under double vectors:
// the fine level of the mixed form's BiCGStab preconditioner).
template <typename V, bool FINE, typename S = V>
__global__ void __launch_bounds__(256) k_amg_pre_residual(SellView s, const S* A, const V* dinv,
                                                          V omega, const V* r, V* x, V* res,
                                                          const int* flag, KcRef kc = {},
                                                          const V* v1 = nullptr,
                                                          V* rout = nullptr,
                                                          const int* spos = nullptr) {
    __shared__ double sh[256];
    if (flag[0]) return;
    const V sc = kc.kp1 ? V(kc_step1(kc, sh)) : V(0);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    int w;
    const size_t b = sv_base(s, i, w);
    V acc = V(0);
    if (kc.kp1) {
#pragma unroll 4
        for (int q = 0; q < w; ++q) {
            const int j = s.cols[b + size_t(q) * 64];
            acc += V(A[b + size_t(q) * 64]) * (dinv[j] * (r[j] - sc * v1[j]));
        }
    } else {
#pragma unroll 4
        for (int q = 0; q < w; ++q) {
            const int j = s.cols[b + size_t(q) * 64];
            acc += V(A[b + size_t(q) * 64]) * (dinv[j] * r[j]);
        }
    }
    const V ri = kc.kp1 ? r[i] - sc * v1[i] : r[i];
    if (kc.kp1) rout[i] = ri;
    x[i] = omega * dinv[i] * ri;
    res[spos ? spos[i] : i] = ri - omega * acc;
}

// Prolongation fused with post-smoothing (out of place):
//   t = x + alpha P ec (formed on the fly for the row and its neighbours),
//   out = t + omega D^-1 (r - A t).
// Grid-stride; with part != nullptr the block also writes its partial of
// r . out (a fixed partition of gridDim blocks) and, with p != nullptr,
// copies out into p (first CG iteration).
template <typename V, bool FINE, typename S = V>
__global__ void __launch_bounds__(256) k_amg_prolong_smooth(SellView s, const S* A, const V* dinv,
                                                            V omega, const V* r, const V* x,
                                                            const int* agg, V alpha, const V* ec,
                                                            V* out, double* part, V* p,
                                                            const int* flag, const V* qv = nullptr,
                                                            double* qpart = nullptr, KcRef kc = {},
                                                            const V* ec2 = nullptr) {
    __shared__ double sh[256];
    if (flag[0]) return;
    // K-cycle correction: alpha P (ca c1 + cb c2), c1 = ec, c2 = ec2
    V ca = alpha, cb = V(0);
    if (kc.kp2) {
        double a, b;
        kc_coeffs(kc, sh, a, b);
        ca = V(a);
        cb = V(b);
    }
    double rz = 0.0, qz = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
        int w;
        const size_t b = sv_base(s, i, w);
        V acc = V(0);
        if (ec2) {
#pragma unroll 4
            for (int q = 0; q < w; ++q) {
                const int j = s.cols[b + size_t(q) * 64];
                const int J = agg[j];
                acc += V(A[b + size_t(q) * 64]) * (x[j] + (ca * ec[J] + cb * ec2[J]));
            }
        } else {
#pragma unroll 4
            for (int q = 0; q < w; ++q) {
                const int j = s.cols[b + size_t(q) * 64];
                acc += V(A[b + size_t(q) * 64]) * (x[j] + ca * ec[agg[j]]);
            }
        }
        const V ri = r[i];
        const int I = agg[i];
        const V ti = x[i] + (ec2 ? ca * ec[I] + cb * ec2[I] : ca * ec[I]);
        const V oi = ti + omega * dinv[i] * (ri - acc);
        out[i] = oi;
        if (p) p[i] = oi;
        rz += double(ri) * double(oi);
        if (qv) qz += double(qv[i]) * double(oi);
    }
    if (part) {
        const double t = amg_block_sum(rz, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = t;
    }
    if (qpart) {
        const double t = amg_block_sum(qz, sh);
        if (threadIdx.x == 0) qpart[blockIdx.x] = t;
    }
}

// v = A c (v may be null); block partials of c.(A c), c.y1 and c.y2 (y2 may
// be null) into kp[k*G+b]
template <typename V>
__global__ void __launch_bounds__(256) k_kc_dots(SellView s, const V* A, const V* c, V* v,
                                                 const V* y1, const V* y2, double* kp,
                                                 const int* flag) {
    __shared__ double sh[256];
    if (flag[0]) return;
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
        int w;
        const size_t b = sv_base(s, i, w);
        V acc = V(0);
#pragma unroll 4
        for (int q = 0; q < w; ++q) acc += A[b + size_t(q) * 64] * c[s.cols[b + size_t(q) * 64]];
        if (v) v[i] = acc;
        const double ci = double(c[i]);
        d0 += ci * double(acc);
        d1 += ci * double(y1[i]);
        if (y2) d2 += ci * double(y2[i]);
    }
    const int G = gridDim.x;
    d0 = amg_block_sum(d0, sh);
    d1 = amg_block_sum(d1, sh);
    d2 = amg_block_sum(d2, sh);
    if (threadIdx.x == 0) {
        kp[blockIdx.x] = d0;
        kp[G + blockIdx.x] = d1;
        kp[2 * G + blockIdx.x] = d2;
    }
}

// c1 = ca c1 + cb c2: the correction e of a K-cycle level formed in place,
// for the fine-level prolongation (gathering both vectors there costs more
// than this pass over the coarse level)
template <typename V>
__global__ void __launch_bounds__(256) k_kc_combine(int n, KcRef kc, V* c1, const V* c2,
                                                    const int* flag) {
    __shared__ double sh[256];
    if (flag[0]) return;
    double a, b;
    kc_coeffs(kc, sh, a, b);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c1[i] = V(a) * c1[i] + V(b) * c2[i];
}

// x = Minv r on the coarsest level (one workgroup; Minv stored transposed)
template <typename V>
__global__ void __launch_bounds__(kDenseMax) k_amg_dense_apply(int n, const V* minv, const V* r, V* x,
                                                               const int* flag) {
    if (flag[0]) return;
    const int i = threadIdx.x;
    if (i >= n) return;
    V acc = V(0);
#pragma unroll 8
    for (int j = 0; j < n; ++j) acc += minv[size_t(j) * n + i] * r[j];
    x[i] = acc;
}

// ----------------------------------------------------- coarse-level tail
// The levels from `t` down, once they are small (at most kTailRows rows),
// run inside one workgroup: restriction from level t-1, pre-smoothing and
// residual, restriction, the coarsest solve, then prolongation and
// post-smoothing back up to level t, phases separated by workgroup
// barriers.  Each row is computed exactly as by the one-level kernels above
// (same loops, same order), so the result is bitwise that of the launch
// sequence it replaces — which on these levels was bound by launch latency.
constexpr int kTailRows = 2048;
constexpr int kTailThreads = 1024;

template <typename V>
struct TailLevel {
    SellView s;
    const V* A;
    const V* dinv;
    V *R, *X, *S, *T;
    const int* agg;      // level l-1 row -> level l row
    const int* mem_ptr;  // members (level l-1 rows) of each level l row
    const int* mem;
    const int* mpos;     // level l-1 row -> its position in mem (residuals are stored there)
};

template <typename V>
struct TailArgs {
    int t, last, dense_n;
    V omega, alpha;
    const V* res_in;  // residual of level t-1, stored in member order of level t
    const V* inv;     // dense inverse of the coarsest level (transposed) or null
    TailLevel<V> lv[kMaxLevels];
};

template <typename V>
// restriction from a residual stored in member order: each aggregate sums a
// contiguous range (no index gather; the same order as over mem)
__device__ __forceinline__ void tail_restrict(const TailLevel<V>& c, const V* res) {
    for (int I = threadIdx.x; I < c.s.n; I += blockDim.x) {
        V s = V(0);
        const int k1 = c.mem_ptr[I + 1];
#pragma unroll 8
        for (int k = c.mem_ptr[I]; k < k1; ++k) s += res[k];
        c.R[I] = s;
    }
}

template <typename V>
__global__ void __launch_bounds__(kTailThreads) k_amg_tail(const TailArgs<V> a, const int* flag) {
    if (flag[0]) return;
    const V om = a.omega, al = a.alpha;
    tail_restrict(a.lv[a.t], a.res_in);
    __syncthreads();
    for (int l = a.t; l < a.last; ++l) {
        const TailLevel<V>& L = a.lv[l];
        for (int i = threadIdx.x; i < L.s.n; i += blockDim.x) {  // as k_amg_pre_residual
            int w;
            const size_t b = sv_base(L.s, i, w);
            V acc = V(0);
#pragma unroll 4
            for (int q = 0; q < w; ++q) {
                const int j = L.s.cols[b + size_t(q) * 64];
                acc += L.A[b + size_t(q) * 64] * (L.dinv[j] * L.R[j]);
            }
            const V ri = L.R[i];
            L.X[i] = om * L.dinv[i] * ri;
            L.S[a.lv[l + 1].mpos[i]] = ri - om * acc;
        }
        __syncthreads();
        tail_restrict(a.lv[l + 1], L.S);
        __syncthreads();
    }
    const TailLevel<V>& C = a.lv[a.last];
    if (a.dense_n) {  // as k_amg_dense_apply
        for (int i = threadIdx.x; i < C.s.n; i += blockDim.x) {
            V acc = V(0);
#pragma unroll 8
            for (int j = 0; j < C.s.n; ++j) acc += a.inv[size_t(j) * C.s.n + i] * C.R[j];
            C.X[i] = acc;
        }
        __syncthreads();
    } else {  // as k_amg_jacobi0 and 4 x 2 k_amg_smooth
        for (int i = threadIdx.x; i < C.s.n; i += blockDim.x) C.X[i] = om * C.dinv[i] * C.R[i];
        __syncthreads();
        for (int it = 0; it < 8; ++it) {
            const V* tin = (it & 1) ? C.T : C.X;
            V* tout = (it & 1) ? C.X : C.T;
            for (int i = threadIdx.x; i < C.s.n; i += blockDim.x) {
                int w;
                const size_t b = sv_base(C.s, i, w);
                V acc = V(0);
                for (int q = 0; q < w; ++q) acc += C.A[b + size_t(q) * 64] * tin[C.s.cols[b + size_t(q) * 64]];
                tout[i] = tin[i] + om * C.dinv[i] * (C.R[i] - acc);
            }
            __syncthreads();
        }
    }
    for (int l = a.last - 1; l >= a.t; --l) {  // as k_amg_prolong_smooth<V, false>
        const TailLevel<V>& L = a.lv[l];
        const V* ec = l + 1 == a.last ? C.X : a.lv[l + 1].T;
        const int* agg = a.lv[l + 1].agg;
        for (int i = threadIdx.x; i < L.s.n; i += blockDim.x) {
            int w;
            const size_t b = sv_base(L.s, i, w);
            V acc = V(0);
#pragma unroll 4
            for (int q = 0; q < w; ++q) {
                const int j = L.s.cols[b + size_t(q) * 64];
                acc += L.A[b + size_t(q) * 64] * (L.X[j] + al * ec[agg[j]]);
            }
            const V ri = L.R[i];
            const V ti = L.X[i] + al * ec[agg[i]];
            L.T[i] = ti + om * L.dinv[i] * (ri - acc);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ host side
static SellView view_of(const DevMesh& m) {
    return {m.nn, m.slice_ptr.p, m.slice_w.p, m.sell_cols.p, m.diag_off.p};
}
static SellView view_of(const AmgLevel& l) {
    return {l.n, l.slice_ptr.p, l.slice_w.p, l.cols.p, l.diag_off.p};
}

// Strong couplings of a CRS matrix (Vanek, Mandel, Brezina 1996):
// j != i is a strong neighbour of i when |a_ij| >= theta sqrt(|a_ii a_jj|).
// On the anisotropic cells of wall-resolved meshes (aspect ratios in the
// hundreds) the wall-normal couplings of the pressure-correction operator
// exceed the tangential ones by the square of the aspect ratio; aggregating
// along the strong couplings only gives the semi-coarsening that damped
// Jacobi needs there.  On isotropic meshes every neighbour is strong.
static void strong_graph(int n, const std::vector<int>& rp, const std::vector<int>& ci,
                         const std::vector<double>& a, double theta, std::vector<int>& srp,
                         std::vector<int>& sci) {
    std::vector<double> d(n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) d[i] = std::fabs(a[k]);
    srp.assign(size_t(n) + 1, 0);
    sci.clear();
    sci.reserve(ci.size());
    for (int i = 0; i < n; ++i) {
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k];
            if (j != i && std::fabs(a[k]) >= theta * std::sqrt(d[i] * d[j])) sci.push_back(j);
        }
        srp[i + 1] = int(sci.size());
    }
}

// Greedy aggregation over the strong-coupling graph (rows list neighbours
// only).  Phase 1: a node whose strong neighbours are all free roots an
// aggregate of itself and them (a node without strong neighbours stays a
// singleton); phase 2: a node left over joins the aggregate of its strongest
// aggregated neighbour; phase 3: what is still free groups with its free
// strong neighbours.
static int aggregate(int n, const std::vector<int>& rp, const std::vector<int>& ci,
                     const std::vector<double>& w, std::vector<int>& agg) {
    agg.assign(n, -1);
    int na = 0;
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        bool free = true;
        for (int k = rp[i]; k < rp[i + 1] && free; ++k) free = agg[ci[k]] < 0;
        if (!free) continue;
        for (int k = rp[i]; k < rp[i + 1]; ++k) agg[ci[k]] = na;
        agg[i] = na;
        ++na;
    }
    std::vector<int> a2 = agg;
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        double best = -1.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (agg[ci[k]] >= 0 && w[k] > best) {
                best = w[k];
                a2[i] = agg[ci[k]];
            }
    }
    for (int i = 0; i < n; ++i) {
        if (a2[i] >= 0) continue;
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (a2[ci[k]] < 0) a2[ci[k]] = na;
        a2[i] = na;
        ++na;
    }
    agg.swap(a2);
    return na;
}

// |a_ij| of each strong coupling, in the strong graph's order
static std::vector<double> strong_weights(int n, const std::vector<int>& rp,
                                          const std::vector<int>& ci, const std::vector<double>& a,
                                          const std::vector<int>& srp, const std::vector<int>& sci) {
    std::vector<double> w(sci.size());
    for (int i = 0; i < n; ++i) {
        int q = srp[i];
        for (int k = rp[i]; k < rp[i + 1] && q < srp[i + 1]; ++k)
            if (ci[k] == sci[q]) w[q++] = std::fabs(a[k]);
    }
    return w;
}

template <typename T>
static void alloc_vec(DBuf<T>& b, size_t n) {
    if (b.n < n) b.alloc(n);
}

void amg_build(Flow& f) {
    if (f.amg.built) return;
    DevMesh& m = *f.m;
    Amg& amg = f.amg;
    amg.lv.clear();
    amg.lv.reserve(kMaxLevels);
    amg.lv.emplace_back();
    {
        AmgLevel& l0 = amg.lv[0];
        l0.n = m.nn;
        l0.h_row_ptr = m.h_row_ptr;
        l0.h_cols = m.h_cols;
    }
    std::vector<int> fine_slice_ptr = m.h_slice_ptr, fine_slice_w = m.h_slice_w;
    size_t fine_nsell = m.nsell;
    // values of the scalar system in CRS order: the strength of the couplings
    // (the hierarchy is built once, from the first matrix solved)
    std::vector<double> fvals(size_t(m.nnz));
    {
        std::vector<double> sell(m.nsell);
        f.A1.download(sell.data(), sell.size(), f.stream);
        OAGPU_HIP(hipStreamSynchronize(f.stream));
        sell_to_crs(m, 1, sell.data(), fvals.data());
    }
    double theta = kTheta;
    while (int(amg.lv.size()) < kMaxLevels && amg.lv.back().n > kDenseMax) {
        AmgLevel& fl = amg.lv.back();
        const int nf = fl.n;
        std::vector<int> agg, srp, sci;
        strong_graph(nf, fl.h_row_ptr, fl.h_cols, fvals, theta, srp, sci);
        const std::vector<double> sw = strong_weights(nf, fl.h_row_ptr, fl.h_cols, fvals, srp, sci);
        const int nc = aggregate(nf, srp, sci, sw, agg);
        theta = kThetaCoarse;
        if (nc >= nf || nc * 5 > nf * 4) break;  // coarsening has stalled
        // members of each aggregate, ascending
        std::vector<int> mem_ptr(size_t(nc) + 1, 0), mem(nf);
        for (int i = 0; i < nf; ++i) ++mem_ptr[agg[i] + 1];
        std::partial_sum(mem_ptr.begin(), mem_ptr.end(), mem_ptr.begin());
        {
            std::vector<int> pos(mem_ptr.begin(), mem_ptr.end() - 1);
            for (int i = 0; i < nf; ++i) mem[pos[agg[i]]++] = i;
        }
        // coarse pattern: aggregates of the neighbours of the members, sorted
        std::vector<int> crp(size_t(nc) + 1, 0), ccols;
        std::vector<int> mark(nc, -1), row;
        for (int I = 0; I < nc; ++I) {
            row.clear();
            for (int k = mem_ptr[I]; k < mem_ptr[I + 1]; ++k) {
                const int i = mem[k];
                for (int q = fl.h_row_ptr[i]; q < fl.h_row_ptr[i + 1]; ++q) {
                    const int J = agg[fl.h_cols[q]];
                    if (mark[J] != I) {
                        mark[J] = I;
                        row.push_back(J);
                    }
                }
            }
            std::sort(row.begin(), row.end());
            if (row.size() > 255) break;
            ccols.insert(ccols.end(), row.begin(), row.end());
            crp[I + 1] = int(ccols.size());
        }
        if (int(ccols.size()) != crp[nc] || crp[nc] == 0) break;  // a coarse row over 255 entries
        // coarse values P^T A P (sums of the members' coefficients): the next
        // level's coupling strengths
        {
            std::vector<double> cv(ccols.size(), 0.0);
            for (int I = 0; I < nc; ++I) {
                const int* cb = ccols.data() + crp[I];
                const int clen = crp[I + 1] - crp[I];
                for (int k = mem_ptr[I]; k < mem_ptr[I + 1]; ++k) {
                    const int i = mem[k];
                    for (int q = fl.h_row_ptr[i]; q < fl.h_row_ptr[i + 1]; ++q) {
                        const int g = int(std::lower_bound(cb, cb + clen, agg[fl.h_cols[q]]) - cb);
                        cv[size_t(crp[I]) + g] += fvals[q];
                    }
                }
            }
            fvals.swap(cv);
        }
        AmgLevel cl;
        cl.n = nc;
        cl.h_row_ptr = crp;
        cl.h_cols = ccols;
        cl.h_agg = agg;
        SellHost sh = make_sell(nc, crp, ccols);
        cl.nslice = sh.nslice;
        cl.nsell = sh.nsell;
        // coarse slot <- fine slots, in the order of the members and their slots
        {
            std::vector<int> gptr(sh.nsell + 1, 0), glist;
            std::vector<size_t> target;
            target.reserve(size_t(fl.h_row_ptr[nf]));
            for (int I = 0; I < nc; ++I) {
                const int* cb = ccols.data() + crp[I];
                const int clen = crp[I + 1] - crp[I];
                for (int k = mem_ptr[I]; k < mem_ptr[I + 1]; ++k) {
                    const int i = mem[k];
                    for (int q = fl.h_row_ptr[i]; q < fl.h_row_ptr[i + 1]; ++q) {
                        const int g = int(std::lower_bound(cb, cb + clen, agg[fl.h_cols[q]]) - cb);
                        target.push_back(sell_index(sh.slice_ptr, I, g, 1, 0));
                        ++gptr[target.back() + 1];
                    }
                }
            }
            std::partial_sum(gptr.begin(), gptr.end(), gptr.begin());
            glist.resize(target.size());
            std::vector<int> pos(gptr.begin(), gptr.end() - 1);
            size_t t = 0;
            for (int I = 0; I < nc; ++I)
                for (int k = mem_ptr[I]; k < mem_ptr[I + 1]; ++k) {
                    const int i = mem[k];
                    for (int q = fl.h_row_ptr[i]; q < fl.h_row_ptr[i + 1]; ++q)
                        glist[pos[target[t++]]++] =
                            int(sell_index(fine_slice_ptr, i, q - fl.h_row_ptr[i], 1, 0));
                }
            cl.gptr.upload(gptr);
            cl.glist.upload(glist);
        }
        std::vector<int> doff(nc);
        for (int I = 0; I < nc; ++I) {
            const int* b = ccols.data() + crp[I];
            doff[I] = int(std::lower_bound(b, b + (crp[I + 1] - crp[I]), I) - b);
        }
        cl.slice_ptr.upload(sh.slice_ptr);
        cl.slice_w.upload(sh.slice_w);
        cl.cols.upload(sh.cols);
        cl.diag_off.upload(doff);
        cl.agg.upload(agg);
        cl.mem_ptr.upload(mem_ptr);
        cl.mem.upload(mem);
        {
            std::vector<int> mpos(nf);
            for (int k = 0; k < nf; ++k) mpos[mem[k]] = k;
            cl.mpos.upload(mpos);
        }
        cl.A.alloc(cl.nsell);
        cl.fA.alloc(cl.nsell);
        cl.A.zero(f.stream);
        cl.fA.zero(f.stream);
        for (auto* b : {&cl.dinv, &cl.X, &cl.T, &cl.R, &cl.S, &cl.C1, &cl.C2, &cl.V1, &cl.R2})
            b->alloc(nc);
        for (auto* b : {&cl.fdinv, &cl.fX, &cl.fT, &cl.fR, &cl.fS, &cl.fC1, &cl.fC2, &cl.fV1, &cl.fR2})
            b->alloc(nc);
        fine_slice_ptr = sh.slice_ptr;
        fine_slice_w = sh.slice_w;
        fine_nsell = sh.nsell;
        amg.lv.push_back(std::move(cl));
    }
    // fine-level temporaries
    AmgLevel& l0 = amg.lv[0];
    for (auto* b : {&l0.T, &l0.S}) b->alloc(m.nn);
    for (auto* b : {&l0.fT, &l0.fS}) b->alloc(m.nn);
    const int nl = int(amg.lv.size());
    amg.dense_n = nl > 1 && amg.lv[nl - 1].n <= kDenseMax ? amg.lv[nl - 1].n : 0;
    if (amg.dense_n) {
        amg.dense.alloc(size_t(amg.dense_n) * amg.dense_n);
        amg.fdense.alloc(size_t(amg.dense_n) * amg.dense_n);
    }
    OAGPU_HIP(hipStreamSynchronize(f.stream));
    amg.built = true;
}

template <typename V>
static V* lvA(AmgLevel& l) {
    if constexpr (std::is_same_v<V, float>) return l.fA.p;
    else return l.A.p;
}
template <typename V>
static V* lvv(AmgLevel& l, int which) {  // 0 dinv, 1 X, 2 T, 3 R, 4 S, 5 C1, 6 C2, 7 V1, 8 R2
    if constexpr (std::is_same_v<V, float>) {
        DBuf<float>* b[9] = {&l.fdinv, &l.fX, &l.fT, &l.fR, &l.fS, &l.fC1, &l.fC2, &l.fV1, &l.fR2};
        return b[which]->p;
    } else {
        DBuf<double>* b[9] = {&l.dinv, &l.X, &l.T, &l.R, &l.S, &l.C1, &l.C2, &l.V1, &l.R2};
        return b[which]->p;
    }
}

static inline unsigned blocks(int n) { return unsigned((n + 255) / 256); }

template <typename V>
void amg_setup(Flow& f, const V* A0, const V* dinv0) {
    (void)dinv0;
    amg_build(f);
    Amg& amg = f.amg;
    amg.last_float = std::is_same_v<V, float>;
    const int nl = int(amg.lv.size());
    for (int l = 1; l < nl; ++l) {
        AmgLevel& c = amg.lv[l];
        const V* Af = l == 1 ? A0 : lvA<V>(amg.lv[l - 1]);
        k_amg_galerkin<V><<<xcd_grid(c.nsell, 256), 256, 0, f.stream>>>(c.nsell, c.gptr.p, c.glist.p,
                                                                        Af, lvA<V>(c));
        k_amg_dinv<V><<<blocks(c.n), 256, 0, f.stream>>>(view_of(c), lvA<V>(c), lvv<V>(c, 0));
    }
    if (amg.dense_n && nl > 1) {
        AmgLevel& c = amg.lv[nl - 1];
        V* vinv;
        if constexpr (std::is_same_v<V, float>) vinv = amg.fdense.p;
        else vinv = amg.dense.p;
        k_amg_dense_inverse<V><<<1, 256, 0, f.stream>>>(view_of(c), lvA<V>(c), amg.dense.p, vinv);
    }
    OAGPU_HIP(hipGetLastError());
}

// Host driver of one preconditioner application: the fine level and the
// levels above the tail run as grid-wide kernels, levels from the tail down
// inside one workgroup.  Every level's vectors are its own, so the two visits
// of a K-cycle level reuse the deeper levels' buffers one after the other.
template <typename V>
struct AmgCycle {
    Flow& f;
    Amg& amg;
    const int* flag;
    V om, al;
    int t, last;
    bool kcycle;
    const V* A0;
    const V* dinv0;
    bool timing = false;  // clock the tail (eager launches only)

    SellView view(int l) const { return l == 0 ? view_of(*f.m) : view_of(amg.lv[l]); }
    // a coarse level's residual is stored in the member order of the next
    // level, so that its restriction sums contiguous ranges; the fine level's
    // only when the tail restricts it (scattered stores cost more there than
    // the gather of the restriction saves)
    const int* spos(int l) const { return l > 0 || t == 1 ? amg.lv[l + 1].mpos.p : nullptr; }
    const int* rmem(int l) const { return spos(l) ? nullptr : amg.lv[l + 1].mem.p; }
    const V* A(int l) const { return l == 0 ? A0 : lvA<V>(amg.lv[l]); }
    const V* D(int l) const { return l == 0 ? dinv0 : lvv<V>(amg.lv[l], 0); }
    V* X(int l) const { return lvv<V>(amg.lv[l], l == 0 ? 2 : 1); }  // T0 on level 0

    // correction on level l + 1 of the residual S_l of level l, applied by
    // the prolongation as alpha P ec, or for a K-cycle level as
    // P (ca c1 + cb c2) with the coefficients from kc
    struct Corr {
        const V* ec;
        const V* ec2;
        V alpha;
        KcRef kc;
    };

    Corr coarse(int l) {
        const int c = l + 1;
        AmgLevel& cl = amg.lv[c];
        const V* S = lvv<V>(amg.lv[l], 4);
        if (kcycle && c < t && c < last && 16 + 3 * (2 * c + 2) <= 64) {
            V* rc = lvv<V>(cl, 3);
            V* c1 = lvv<V>(cl, 5);
            V* c2 = lvv<V>(cl, 6);
            V* v1 = lvv<V>(cl, 7);
            V* r2 = lvv<V>(cl, 8);
            const int G = int(std::min<unsigned>(blocks(cl.n), unsigned(kRedGrid)));
            double* kp1 = f.part.p + size_t(16 + 6 * c) * kRedGrid;
            double* kp2 = kp1 + 3 * size_t(kRedGrid);
            k_amg_restrict<V><<<blocks(cl.n), 256, 0, f.stream>>>(cl.n, cl.mem_ptr.p, rmem(l), S, om,
                                                                  nullptr, rc, nullptr, flag);
            level(c, rc, c1);
            k_kc_dots<V><<<G, 256, 0, f.stream>>>(view(c), A(c), c1, v1, rc, nullptr, kp1, flag);
            KcRef k;
            k.kp1 = kp1;
            k.G = G;
            level(c, rc, c2, k, v1, r2);  // right-hand side r2 = rc - (a1/rho1) v1
            k_kc_dots<V><<<G, 256, 0, f.stream>>>(view(c), A(c), c2, nullptr, v1, r2, kp2, flag);
            k.kp2 = kp2;
            if (l == 0) {
                k_kc_combine<V><<<blocks(cl.n), 256, 0, f.stream>>>(cl.n, k, c1, c2, flag);
                return {c1, nullptr, V(1), {}};
            }
            return {c1, c2, V(1), k};
        }
        if (t <= last && c == t) {  // the tail restricts S_l itself
            TailArgs<V> ta{};
            ta.t = t;
            ta.last = last;
            ta.dense_n = amg.dense_n;
            ta.omega = om;
            ta.alpha = al;
            ta.res_in = S;
            if constexpr (std::is_same_v<V, float>) ta.inv = amg.dense_n ? amg.fdense.p : nullptr;
            else ta.inv = amg.dense_n ? amg.dense.p : nullptr;
            for (int k = t; k <= last; ++k) {
                AmgLevel& L = amg.lv[k];
                ta.lv[k] = {view(k), A(k), D(k), lvv<V>(L, 3), lvv<V>(L, 1), lvv<V>(L, 4), lvv<V>(L, 2),
                            L.agg.p, L.mem_ptr.p, L.mem.p, L.mpos.p};
            }
            if (timing) clock_begin(f, kClkAmgTail);
            k_amg_tail<V><<<1, kTailThreads, 0, f.stream>>>(ta, flag);
            if (timing) clock_end(f, kClkAmgTail);
            return {t == last ? X(last) : lvv<V>(amg.lv[t], 2), nullptr, al, {}};
        }
        V* rc = lvv<V>(cl, 3);
        k_amg_restrict<V><<<blocks(cl.n), 256, 0, f.stream>>>(cl.n, cl.mem_ptr.p, rmem(l), S, om,
                                                              nullptr, rc, nullptr, flag);
        if (c < last) {  // a V-cycle level above the tail
            level(c, rc, lvv<V>(cl, 2));
            return {lvv<V>(cl, 2), nullptr, al, {}};
        }
        if (amg.dense_n) {  // coarsest level
            const V* inv;
            if constexpr (std::is_same_v<V, float>) inv = amg.fdense.p;
            else inv = amg.dense.p;
            k_amg_dense_apply<V><<<1, kDenseMax, 0, f.stream>>>(cl.n, inv, rc, X(c), flag);
        } else {  // coarsening stalled above the dense size: Jacobi steps
            k_amg_jacobi0<V><<<blocks(cl.n), 256, 0, f.stream>>>(cl.n, om, D(c), rc, X(c), flag);
            for (int it = 0; it < 4; ++it) {
                k_amg_smooth<V><<<blocks(cl.n), 256, 0, f.stream>>>(view(c), A(c), D(c), om, rc, X(c),
                                                                    lvv<V>(cl, 2), flag);
                k_amg_smooth<V><<<blocks(cl.n), 256, 0, f.stream>>>(view(c), A(c), D(c), om, rc,
                                                                    lvv<V>(cl, 2), X(c), flag);
            }
        }
        return {X(c), nullptr, al, {}};
    }

    // out ~ A_l^-1 rhs on a level 1 <= l < min(t, last): pre-smoothing and
    // residual, the coarse correction, prolongation and post-smoothing.
    // With pre.kp1 the right-hand side is rhs - (a1/rho1) v1, written to rout.
    void level(int l, const V* rhs, V* out, KcRef pre = {}, const V* v1 = nullptr,
               V* rout = nullptr) {
        AmgLevel& L = amg.lv[l];
        k_amg_pre_residual<V, false><<<blocks(L.n), 256, 0, f.stream>>>(
            view(l), A(l), D(l), om, rhs, X(l), lvv<V>(L, 4), flag, pre, v1, rout, spos(l));
        const V* r = pre.kp1 ? rout : rhs;
        const Corr cc = coarse(l);
        k_amg_prolong_smooth<V, false><<<blocks(L.n), 256, 0, f.stream>>>(
            view(l), A(l), D(l), om, r, X(l), amg.lv[l + 1].agg.p, cc.alpha, cc.ec, out, nullptr,
            nullptr, flag, nullptr, nullptr, cc.kc, cc.ec2);
    }
};

template <typename V>
void amg_vcycle(Flow& f, const V* A0, const V* dinv0, const V* r, V* z, double* rz_part, V* p_init,
                bool timing, bool kcycle, const V* q, double* qz_part, const float* A0f) {
    Amg& amg = f.amg;
    const int nl = int(amg.lv.size());
    OAGPU_REQUIRE(nl > 1, "AMG hierarchy without a coarse level");
    const int last = nl - 1;
    // first level handled by the one-workgroup tail (all deeper ones are smaller)
    int t = last + 1;
    for (int l = 1; l <= last; ++l)
        if (amg.lv[l].n <= kTailRows) {
            t = l;
            break;
        }
    // K-cycle levels only with at least two levels above the tail: with one,
    // the extra launches cost more than the iterations they save (32^3 cavity)
    kcycle = kcycle && t >= 3;
    AmgCycle<V> cy{f, amg, f.flag.p, V(kOmega), V(kAlpha), t, last, kcycle, A0, dinv0};
    AmgLevel& l0 = amg.lv[0];
    const bool fine_f32 = std::is_same_v<V, double> && A0f != nullptr;
    if (timing) clock_begin(f, kClkAmgPre);
    if (fine_f32)
        k_amg_pre_residual<V, true, float><<<blocks(l0.n), 256, 0, f.stream>>>(
            cy.view(0), A0f, dinv0, cy.om, r, cy.X(0), lvv<V>(l0, 4), cy.flag, {}, nullptr, nullptr,
            cy.spos(0));
    else
        k_amg_pre_residual<V, true><<<blocks(l0.n), 256, 0, f.stream>>>(
            cy.view(0), A0, dinv0, cy.om, r, cy.X(0), lvv<V>(l0, 4), cy.flag, {}, nullptr, nullptr,
            cy.spos(0));
    if (timing) clock_end(f, kClkAmgPre);
    cy.timing = timing;
    const auto cc = cy.coarse(0);
    if (timing) clock_begin(f, kClkAmgPost);
    if (fine_f32)
        k_amg_prolong_smooth<V, true, float><<<kRedGrid, 256, 0, f.stream>>>(
            cy.view(0), A0f, dinv0, cy.om, r, cy.X(0), amg.lv[1].agg.p, cc.alpha, cc.ec, z, rz_part,
            p_init, cy.flag, q, qz_part, cc.kc, cc.ec2);
    else
        k_amg_prolong_smooth<V, true><<<kRedGrid, 256, 0, f.stream>>>(
            cy.view(0), A0, dinv0, cy.om, r, cy.X(0), amg.lv[1].agg.p, cc.alpha, cc.ec, z, rz_part,
            p_init, cy.flag, q, qz_part, cc.kc, cc.ec2);
    if (timing) clock_end(f, kClkAmgPost);
    OAGPU_HIP(hipGetLastError());
}

// ================================================== smoothed aggregation
// Smoothed aggregation (Vanek, Mandel, Brezina, Computing 56 (1996)) with
// filtered prolongation smoothing:
//   P = (I - omega_p D_F^-1 A_F) T,
// T the piecewise-constant prolongation of the aggregates above, A_F the
// matrix with its weak couplings (below the strength threshold) dropped and
// added to the diagonal D_F, omega_p = 4 / (3 lambda) with lambda the
// spectral radius of D_F^-1 A_F (power iteration when the hierarchy is
// built).  Coarse operators are Galerkin products A_{l+1} = P_l^T A_l P_l;
// the cycle is V(1,1) with damped Jacobi (omega = 2/3) and the coarsest
// level's dense inverse.  On the wall-resolved p' systems of a
// backward-facing step (cells of aspect ratio up to ~250) this takes
// BiCGStab to a 1e-2 reduction in 3-6 iterations where the unsmoothed
// V-cycle needs 20-30, independently of the mesh size; filtering keeps the
// coarse operators about as large, in total, as the fine matrix, but their
// rows grow long (tens to a few hundred entries), so the coarse levels are
// stored CSR and every row kernel there runs G lanes per row (G from the
// level's mean row length) with a fixed-order shuffle reduction.
//
// The structure -- aggregates, the patterns of P_l, of B_l = A_l P_l and of
// A_{l+1}, and the term lists that compute their values -- is built on the
// host (in up to 16 threads) from the first matrix solved.  Every solve
// recomputes all values on the device in FP64: per entry of P_l the sum s_e
// of its row's strong couplings into that aggregate, per row of P_l the
// filtered diagonal d_F = (row sum) - sum_e s_e and the entries
// [own aggregate] (1 - omega_p) - omega_p s_e / d_F, then every entry of B_l
// and of A_{l+1} as a fixed-order sum over its listed terms (no atomics:
// bitwise reproducible).

// f(lo, hi) over [0, n) in up to 16 host threads, contiguous chunks; f must
// not throw
template <typename F>
static void par_for(int n, F&& f) {
    const int nt = std::max(1, std::min(16, int(std::thread::hardware_concurrency())));
    if (n < 8192 || nt == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    const int chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo < hi) th.emplace_back([&f, lo, hi] { f(lo, hi); });
    }
    for (auto& x : th) x.join();
}

template <int G, typename T>
__device__ __forceinline__ T sub_sum(T v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}

// lanes per row of a CSR level from its mean row length
static int sa_lanes(size_t nnz, int n) {
    const double len = n ? double(nnz) / n : 0.0;
    return len < 12 ? 4 : len < 24 ? 8 : len < 48 ? 16 : len < 96 ? 32 : 64;
}

template <int G>
static unsigned sa_grid(int n) {
    return unsigned((size_t(n) * G + 255) / 256);
}

// Dispatch a kernel template on the lane count g (4 .. 64).
#define SA_DISPATCH(g, KERNEL, ...)                    \
    switch (g) {                                       \
    case 4: KERNEL<4>(__VA_ARGS__); break;             \
    case 8: KERNEL<8>(__VA_ARGS__); break;             \
    case 16: KERNEL<16>(__VA_ARGS__); break;           \
    case 32: KERNEL<32>(__VA_ARGS__); break;           \
    default: KERNEL<64>(__VA_ARGS__); break;           \
    }

// ---------------------------------------------------------- set-up kernels
// s[e] = sum over the listed couplings of P entry e (values of A_l; V: the
// set-up's value type, FP64 or the mixed form's FP32, sums in double)
template <typename V>
__global__ void __launch_bounds__(256) k_sa_psum(int ne, const int* pt_ptr, const int* pt_idx,
                                                 const V* A, V* sum) {
    const int e = int(xcd_block() * blockDim.x + threadIdx.x);
    if (e >= ne) return;
    double s = 0.0;
    const int k1 = pt_ptr[e + 1];
    for (int k = pt_ptr[e]; k < k1; ++k) s += double(A[pt_idx[k]]);
    sum[e] = V(s);
}

// P row i from its row sum and diagonal: d_F = rowsum - sum_e s_e (a_ii if
// not positive), P_e = [own] (1 - omega_p) - omega_p s_e / d_F
template <typename V>
__device__ __forceinline__ void sa_prow(int i, double rowsum, double aii, const int* prp,
                                        const unsigned char* own, const int* ppos,
                                        const V* sum, double omega_p, V* P) {
    const int e0 = prp[i], e1 = prp[i + 1];
    double st = 0.0;
    for (int e = e0; e < e1; ++e) st += double(sum[e]);
    double dF = rowsum - st;
    if (!(dF > 0.0)) dF = aii;
    const double sc = dF != 0.0 ? omega_p / dF : 0.0;
    for (int e = e0; e < e1; ++e) P[ppos[e]] = V((own[e] ? 1.0 - omega_p : 0.0) - sc * double(sum[e]));
}

// fine level (sliced-ELL rows, short): one lane per row
template <typename V>
__global__ void __launch_bounds__(256) k_sa_prow_sell(SellView s, const V* A, const int* prp,
                                                      const unsigned char* own, const int* ppos,
                                                      const V* sum, double omega_p, V* P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    int w;
    const size_t b = sv_base(s, i, w);
    double rs = 0.0;
    for (int q = 0; q < w; ++q) rs += double(A[b + size_t(q) * 64]);
    sa_prow<V>(i, rs, double(A[b + size_t(s.diag_off[i]) * 64]), prp, own, ppos, sum, omega_p, P);
}

// coarse level (CSR rows): G lanes sum the row, its first lane writes P's row
template <int G, typename V>
__global__ void __launch_bounds__(256) k_sa_prow_csr(int n, const int* rp, const int* diag, const V* A,
                                                     const int* prp, const unsigned char* own,
                                                     const int* ppos, const V* sum, double omega_p,
                                                     V* P) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int i = int(t / G), lane = int(t % G);
    if (i >= n) return;
    double rs = 0.0;
    for (int k = rp[i] + lane; k < rp[i + 1]; k += G) rs += double(A[k]);
    rs = sub_sum<G>(rs);
    if (lane == 0) sa_prow<V>(i, rs, double(A[diag[i]]), prp, own, ppos, sum, omega_p, P);
}

// out[pos ? pos[e] : e] = sum over the terms t of entry e of X[t.x] * Y[t.y]
template <typename V>
__global__ void __launch_bounds__(256) k_sa_terms(int ne, const int* ptr, const int2* terms,
                                                  const V* X, const V* Y, V* out) {
    const int e = int(xcd_block() * blockDim.x + threadIdx.x);
    if (e >= ne) return;
    double s = 0.0;
    const int k1 = ptr[e + 1];
    for (int k = ptr[e]; k < k1; ++k) {
        const int2 t = terms[k];
        s += double(X[t.x]) * double(Y[t.y]);
    }
    out[e] = V(s);
}

// as k_sa_terms with G lanes per entry: the group's lanes read consecutive
// terms (coalesced), each sums every G-th term in double, the group adds the
// partial sums in a fixed order (bitwise reproducible)
template <typename V, int G>
__global__ void __launch_bounds__(256) k_sa_terms_g(int ne, const int* ptr, const int2* terms,
                                                    const V* X, const V* Y, V* out) {
    const size_t t = size_t(xcd_block()) * blockDim.x + threadIdx.x;
    const int e = int(t / G), lane = int(t % G);
    if (e >= ne) return;  // a group of G lanes (G divides 64) leaves together
    double s = 0.0;
    const int k1 = ptr[e + 1];
    for (int k = ptr[e] + lane; k < k1; k += G) {
        const int2 q = terms[k];
        s += double(X[q.x]) * double(Y[q.y]);
    }
    s = sub_sum<G>(s);
    if (lane == 0) out[e] = V(s);
}

// lanes per entry of a term list from its mean term count
static int sa_term_lanes(size_t nterms, size_t ne) {
    const double len = ne ? double(nterms) / double(ne) : 0.0;
    return len < 6.0 ? 1 : len < 24.0 ? 4 : 16;
}

template <typename V>
static void launch_sa_terms(hipStream_t st, int lanes, size_t ne, const int* ptr, const int2* terms,
                            const V* X, const V* Y, V* out) {
    switch (lanes) {
    case 4:
        k_sa_terms_g<V, 4><<<xcd_grid(ne * 4, 256), 256, 0, st>>>(int(ne), ptr, terms, X, Y, out);
        break;
    case 16:
        k_sa_terms_g<V, 16><<<xcd_grid(ne * 16, 256), 256, 0, st>>>(int(ne), ptr, terms, X, Y, out);
        break;
    default:
        k_sa_terms<V><<<xcd_grid(ne, 256), 256, 0, st>>>(int(ne), ptr, terms, X, Y, out);
    }
}

__global__ void __launch_bounds__(256) k_sa_dinv_f(int n, const int* diag, const float* A, float* fdinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = double(A[diag[i]]);
    fdinv[i] = float(d != 0.0 ? 1.0 / d : 0.0);
}

__global__ void __launch_bounds__(256) k_sa_dinv(int n, const int* diag, const double* A, double* dinv,
                                                 float* fdinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = A[diag[i]];
    dinv[i] = d != 0.0 ? 1.0 / d : 0.0;
    if (fdinv) fdinv[i] = float(dinv[i]);
}

__global__ void __launch_bounds__(256) k_sa_to_float(size_t n, const double* x, float* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = float(x[i]);
}

// Gauss-Jordan inverse of the coarsest matrix (CSR, n <= kDenseMax) in LDS,
// as k_amg_dense_inverse; stored transposed for k_amg_dense_apply
template <typename V, typename IN = double>
__global__ void __launch_bounds__(256) k_sa_dense_inverse(int n, const int* rp, const int* ci,
                                                          const IN* A, V* vinv) {
    __shared__ double M[kDenseMax][kDenseMax + 1];
    __shared__ double colk[kDenseMax];
    __shared__ double piv, dmax;
    __shared__ int skip;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) M[e / n][e % n] = 0.0;
    __syncthreads();
    if (int(threadIdx.x) < n) {
        const int i = threadIdx.x;
        for (int k = rp[i]; k < rp[i + 1]; ++k) M[i][ci[k]] += double(A[k]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int i = 0; i < n; ++i) m = fmax(m, fabs(M[i][i]));
        dmax = m;
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        if (threadIdx.x == 0) {
            piv = M[k][k];
            skip = !(fabs(piv) > 1.0e-13 * dmax);
        }
        __syncthreads();
        if (skip) {
            for (int j = threadIdx.x; j < n; j += blockDim.x) {
                M[k][j] = 0.0;
                M[j][k] = 0.0;
            }
            __syncthreads();
            continue;
        }
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            colk[j] = M[j][k];
            M[k][j] = (j == k ? 1.0 : M[k][j]) / piv;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
            const int i = e / n, j = e % n;
            if (i != k) M[i][j] = (j == k ? 0.0 : M[i][j]) - colk[i] * M[k][j];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, j = e % n;
        vinv[size_t(j) * n + i] = V(M[i][j]);
    }
}

// ------------------------------------------------------------ cycle kernels
// CSR rows, G lanes per row: pre-smoothing from zero fused with the residual
//   x = omega D^-1 r,  res = r - A x
template <typename V, int G>
__global__ void __launch_bounds__(256) k_csr_pre(int n, const int* rp, const int* ci, const V* A,
                                                 const V* dinv, V omega, const V* r, V* x, V* res,
                                                 const int* flag) {
    if (flag[0]) return;
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int i = int(t / G), lane = int(t % G);
    if (i >= n) return;
    V acc = V(0);
    for (int k = rp[i] + lane; k < rp[i + 1]; k += G) {
        const int j = ci[k];
        acc += A[k] * (dinv[j] * r[j]);
    }
    acc = sub_sum<G>(acc);
    if (lane == 0) {
        const V ri = r[i];
        x[i] = omega * dinv[i] * ri;
        res[i] = ri - omega * acc;
    }
}

// out = t + omega D^-1 (r - A t)
template <typename V, int G>
__global__ void __launch_bounds__(256) k_csr_smooth(int n, const int* rp, const int* ci, const V* A,
                                                    const V* dinv, V omega, const V* r, const V* tv,
                                                    V* out, const int* flag) {
    if (flag[0]) return;
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int i = int(t / G), lane = int(t % G);
    if (i >= n) return;
    V acc = V(0);
    for (int k = rp[i] + lane; k < rp[i + 1]; k += G) acc += A[k] * tv[ci[k]];
    acc = sub_sum<G>(acc);
    if (lane == 0) out[i] = tv[i] + omega * dinv[i] * (r[i] - acc);
}

// rc[I] = sum over the pairs (i, P slot) of row I of P^T of P * res[i]
template <typename V, int G>
__global__ void __launch_bounds__(256) k_csr_restrict(int nc, const int* rptr, const int2* rterms,
                                                      const V* P, const V* res, V* rc, const int* flag) {
    if (flag[0]) return;
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int I = int(t / G), lane = int(t % G);
    if (I >= nc) return;
    V s = V(0);
    for (int k = rptr[I] + lane; k < rptr[I + 1]; k += G) {
        const int2 q = rterms[k];
        s += P[q.y] * res[q.x];
    }
    s = sub_sum<G>(s);
    if (lane == 0) rc[I] = s;
}

// t = x + P ec (P sliced-ELL, one lane per row)
template <typename V>
__global__ void __launch_bounds__(256) k_sa_prolong(int n, const int* p_slice_ptr, const int* p_slice_w,
                                                    const int* p_cols, const V* P, const V* ec,
                                                    const V* x, V* t, const int* flag) {
    if (flag[0]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int pw = p_slice_w[i >> 6];
    const size_t pb = size_t(p_slice_ptr[i >> 6]) * 64 + (i & 63);
    V s = x[i];
    for (int qp = 0; qp < pw; ++qp) {
        const int col = p_cols[pb + size_t(qp) * 64];
        if (col >= 0) s += P[pb + size_t(qp) * 64] * ec[col];
    }
    t[i] = s;
}

// kernel launchers on the lane count
template <int G, typename V>
static void launch_csr_pre(hipStream_t st, int n, const int* rp, const int* ci, const V* A, const V* d,
                           V om, const V* r, V* x, V* res, const int* flag) {
    k_csr_pre<V, G><<<sa_grid<G>(n), 256, 0, st>>>(n, rp, ci, A, d, om, r, x, res, flag);
}
template <int G, typename V>
static void launch_csr_smooth(hipStream_t st, int n, const int* rp, const int* ci, const V* A,
                              const V* d, V om, const V* r, const V* t, V* out, const int* flag) {
    k_csr_smooth<V, G><<<sa_grid<G>(n), 256, 0, st>>>(n, rp, ci, A, d, om, r, t, out, flag);
}
template <int G, typename V>
static void launch_csr_restrict(hipStream_t st, int nc, const int* rptr, const int2* rt, const V* P,
                                const V* res, V* rc, const int* flag) {
    k_csr_restrict<V, G><<<sa_grid<G>(nc), 256, 0, st>>>(nc, rptr, rt, P, res, rc, flag);
}
template <int G, typename V>
static void launch_prow_csr(hipStream_t st, int n, const int* rp, const int* diag, const V* A,
                            const int* prp, const unsigned char* own, const int* ppos, const V* sum,
                            double om, V* P) {
    k_sa_prow_csr<G, V><<<sa_grid<G>(n), 256, 0, st>>>(n, rp, diag, A, prp, own, ppos, sum, om, P);
}

// ----------------------------------------------------------- host build
// spectral radius of D_F^-1 A_F (power iteration from a fixed start vector:
// 50 steps, 15 on matrices of more than 2e7 entries)
static double sa_lambda(int n, const std::vector<int>& rp, const std::vector<int>& ci,
                        const std::vector<double>& af, const std::vector<double>& dF) {
    std::vector<double> x(n), y(n);
    for (int i = 0; i < n; ++i) x[i] = 1.0 + 0.5 * std::sin(0.7 * i);
    double lam = 0.0;
    const int steps = rp[n] > 20000000 ? 15 : 50;
    for (int it = 0; it < steps; ++it) {
        double nx = 0.0;
        for (int i = 0; i < n; ++i) nx += x[i] * x[i];
        nx = std::sqrt(nx);
        if (nx == 0.0) break;
        par_for(n, [&](int lo, int hi) {
            for (int i = lo; i < hi; ++i) {
                double acc = 0.0;
                for (int k = rp[i]; k < rp[i + 1]; ++k) acc += af[k] * x[ci[k]];
                y[i] = dF[i] != 0.0 ? acc / dF[i] / nx : 0.0;
            }
        });
        double ny = 0.0;
        for (int i = 0; i < n; ++i) ny += y[i] * y[i];
        lam = std::sqrt(ny);
        x.swap(y);
    }
    return lam;
}

// Two-pass parallel fill of a CSR structure: count(i) gives row i's length,
// fill(i, out) writes it; returns row pointers (int64 offsets checked < 2^31)
template <typename T, typename C, typename F>
static std::vector<int> par_csr(int n, std::vector<T>& out, C&& count, F&& fill) {
    std::vector<long long> len(size_t(n) + 1, 0);
    par_for(n, [&](int lo, int hi) {
        for (int i = lo; i < hi; ++i) len[i + 1] = count(i);
    });
    for (int i = 0; i < n; ++i) len[i + 1] += len[i];
    OAGPU_REQUIRE(len[n] < (1LL << 31), "smoothed-aggregation list longer than 2^31 entries");
    out.resize(size_t(len[n]));
    par_for(n, [&](int lo, int hi) {
        for (int i = lo; i < hi; ++i) fill(i, out.data() + len[i]);
    });
    std::vector<int> rp(size_t(n) + 1);
    for (int i = 0; i <= n; ++i) rp[i] = int(len[i]);
    return rp;
}

void sa_build(Flow& f) {
    if (f.sa.built) return;
    DevMesh& m = *f.m;
    Sa& sa = f.sa;
    sa.lv.clear();
    sa.lv.reserve(kMaxLevels);
    // level 0: the flow's scalar system (pattern of the mesh graph, FP64 values)
    std::vector<int> rp = m.h_row_ptr, ci = m.h_cols;
    std::vector<double> av(size_t(m.nnz));
    {
        std::vector<double> sell(m.nsell);
        f.A1.download(sell.data(), sell.size(), f.stream);
        OAGPU_HIP(hipStreamSynchronize(f.stream));
        sell_to_crs(m, 1, sell.data(), av.data());
    }
    // index of each CSR entry in the level's value array (level 0: sliced-ELL)
    std::vector<int> aidx(ci.size());
    for (int i = 0; i < m.nn; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) aidx[k] = int(sell_index(m.h_slice_ptr, i, k - rp[i], 1, 0));
    int n = m.nn;
    sa.lv.emplace_back();
    sa.lv[0].n = n;
    sa.lv[0].h_rp = rp;
    sa.lv[0].h_cols = ci;
    double theta = kTheta;
    while (int(sa.lv.size()) < kMaxLevels && n > kDenseMax) {
        SaLevel& L = sa.lv.back();
        std::vector<int> srp, sci, agg;
        strong_graph(n, rp, ci, av, theta, srp, sci);
        const std::vector<double> sw = strong_weights(n, rp, ci, av, srp, sci);
        const int nc = aggregate(n, srp, sci, sw, agg);
        theta = kThetaCoarse;
        if (nc <= 0 || nc >= n || nc * 5 > n * 4) break;  // coarsening has stalled
        // strong couplings per CSR entry, filtered diagonal and values
        std::vector<char> strong(ci.size(), 0);
        std::vector<double> dF(n), af(ci.size(), 0.0);
        par_for(n, [&](int lo, int hi) {
            for (int i = lo; i < hi; ++i) {
                int q = srp[i];
                double aii = 0.0, weak = 0.0;
                for (int k = rp[i]; k < rp[i + 1]; ++k) {
                    if (q < srp[i + 1] && ci[k] == sci[q]) {
                        strong[k] = 1;
                        ++q;
                    }
                    if (ci[k] == i) aii = av[k];
                    else if (!strong[k]) weak += av[k];
                }
                dF[i] = aii + weak > 0.0 ? aii + weak : aii;
                for (int k = rp[i]; k < rp[i + 1]; ++k) af[k] = ci[k] == i ? dF[i] : (strong[k] ? av[k] : 0.0);
            }
        });
        const double lam = sa_lambda(n, rp, ci, af, dF);
        const double omega_p = lam > 0.0 ? 4.0 / (3.0 * lam) : 0.0;
        // pattern of P: the row's own aggregate and those of its strong neighbours
        auto prow = [&](int i, std::vector<int>& row) {
            row.clear();
            row.push_back(agg[i]);
            for (int k = rp[i]; k < rp[i + 1]; ++k)
                if (strong[k]) row.push_back(agg[ci[k]]);
            std::sort(row.begin(), row.end());
            row.erase(std::unique(row.begin(), row.end()), row.end());
        };
        std::vector<int> pcols;
        const std::vector<int> prp = par_csr(
            n, pcols,
            [&](int i) {
                thread_local std::vector<int> row;
                prow(i, row);
                return (long long)row.size();
            },
            [&](int i, int* out) {
                thread_local std::vector<int> row;
                prow(i, row);
                std::copy(row.begin(), row.end(), out);
            });
        const size_t np = pcols.size();
        // P sliced-ELL (padding columns -1); per P entry its slot there,
        // whether it is the own aggregate, and its couplings (strong,
        // off-diagonal, column in that aggregate) as value indices of A
        SellHost ps = make_sell(n, prp, pcols);
        for (int i = 0; i < n; ++i)
            for (int q = prp[i + 1] - prp[i]; q < ps.slice_w[i / 64]; ++q)
                ps.cols[sell_index(ps.slice_ptr, i, q, 1, 0)] = -1;
        std::vector<int> ppos(np);
        std::vector<unsigned char> own(np);
        for (int i = 0; i < n; ++i)
            for (int e = prp[i]; e < prp[i + 1]; ++e) {
                ppos[e] = int(sell_index(ps.slice_ptr, i, e - prp[i], 1, 0));
                own[e] = pcols[e] == agg[i];
            }
        std::vector<int> pt_idx;
        std::vector<int> pt_ptr(np + 1, 0);
        {
            // per row: entries grouped by P position, in CSR order
            std::vector<int> cnt(np, 0);
            par_for(n, [&](int lo, int hi) {
                for (int i = lo; i < hi; ++i) {
                    const int* pb = pcols.data() + prp[i];
                    const int plen = prp[i + 1] - prp[i];
                    for (int k = rp[i]; k < rp[i + 1]; ++k)
                        if (strong[k]) ++cnt[prp[i] + int(std::lower_bound(pb, pb + plen, agg[ci[k]]) - pb)];
                }
            });
            for (size_t e = 0; e < np; ++e) pt_ptr[e + 1] = pt_ptr[e] + cnt[e];
            pt_idx.resize(size_t(pt_ptr[np]));
            par_for(n, [&](int lo, int hi) {
                std::vector<int> at;
                for (int i = lo; i < hi; ++i) {
                    const int* pb = pcols.data() + prp[i];
                    const int plen = prp[i + 1] - prp[i];
                    at.assign(pt_ptr.begin() + prp[i], pt_ptr.begin() + prp[i + 1]);
                    for (int k = rp[i]; k < rp[i + 1]; ++k)
                        if (strong[k]) pt_idx[at[std::lower_bound(pb, pb + plen, agg[ci[k]]) - pb]++] = aidx[k];
                }
            });
        }
        // host values of P (for the next level's structure)
        std::vector<double> pv(np);
        par_for(n, [&](int lo, int hi) {
            for (int i = lo; i < hi; ++i) {
                const double sc = dF[i] != 0.0 ? omega_p / dF[i] : 0.0;
                const int* pb = pcols.data() + prp[i];
                const int plen = prp[i + 1] - prp[i];
                for (int e = prp[i]; e < prp[i + 1]; ++e) pv[e] = own[e] ? 1.0 - omega_p : 0.0;
                for (int k = rp[i]; k < rp[i + 1]; ++k)
                    if (strong[k]) pv[prp[i] + int(std::lower_bound(pb, pb + plen, agg[ci[k]]) - pb)] -= sc * av[k];
            }
        });
        // B = A P: per row the sorted union of its neighbours' P columns;
        // terms (A value index, P slot) grouped by B entry in (A entry, P
        // entry) order
        auto brow = [&](int i, std::vector<int>& row) {
            row.clear();
            for (int k = rp[i]; k < rp[i + 1]; ++k)
                row.insert(row.end(), pcols.begin() + prp[ci[k]], pcols.begin() + prp[ci[k] + 1]);
            std::sort(row.begin(), row.end());
            row.erase(std::unique(row.begin(), row.end()), row.end());
        };
        std::vector<int> bcols;
        const std::vector<int> brp = par_csr(
            n, bcols,
            [&](int i) {
                thread_local std::vector<int> row;
                brow(i, row);
                return (long long)row.size();
            },
            [&](int i, int* out) {
                thread_local std::vector<int> row;
                brow(i, row);
                std::copy(row.begin(), row.end(), out);
            });
        const int nb = int(bcols.size());
        std::vector<int2> bterms;
        std::vector<int> bptr;
        {
            // terms per B entry: count, then fill in (A entry, P entry) order
            std::vector<int> cnt(size_t(nb), 0);
            par_for(n, [&](int lo, int hi) {
                for (int i = lo; i < hi; ++i) {
                    const int* bb = bcols.data() + brp[i];
                    const int blen = brp[i + 1] - brp[i];
                    for (int k = rp[i]; k < rp[i + 1]; ++k)
                        for (int e = prp[ci[k]]; e < prp[ci[k] + 1]; ++e)
                            ++cnt[brp[i] + int(std::lower_bound(bb, bb + blen, pcols[e]) - bb)];
                }
            });
            std::vector<long long> bp(size_t(nb) + 1, 0);
            for (int e = 0; e < nb; ++e) bp[e + 1] = bp[e] + cnt[e];
            OAGPU_REQUIRE(bp[nb] < (1LL << 31), "smoothed-aggregation list longer than 2^31 entries");
            bptr.assign(bp.begin(), bp.end());
            bterms.resize(size_t(bp[nb]));
            par_for(n, [&](int lo, int hi) {
                std::vector<int> at;
                for (int i = lo; i < hi; ++i) {
                    const int* bb = bcols.data() + brp[i];
                    const int blen = brp[i + 1] - brp[i];
                    at.assign(bptr.begin() + brp[i], bptr.begin() + brp[i + 1]);
                    for (int k = rp[i]; k < rp[i + 1]; ++k)
                        for (int e = prp[ci[k]]; e < prp[ci[k] + 1]; ++e) {
                            const int b = int(std::lower_bound(bb, bb + blen, pcols[e]) - bb);
                            bterms[at[b]++] = make_int2(aidx[k], ppos[e]);
                        }
                }
            });
        }
        // host values of B, through the CSR indices of A and P
        std::vector<double> bv(size_t(nb), 0.0);
        {
            std::vector<int> a_of, p_of(ps.nsell, -1);
            if (sa.lv.size() == 1) {
                a_of.assign(m.nsell, -1);
                for (size_t k = 0; k < ci.size(); ++k) a_of[aidx[k]] = int(k);
            }
            for (size_t e = 0; e < np; ++e) p_of[ppos[e]] = int(e);
            par_for(nb, [&](int lo, int hi) {
                for (int e = lo; e < hi; ++e) {
                    double acc = 0.0;
                    for (int t = bptr[e]; t < bptr[e + 1]; ++t) {
                        const int ka = a_of.empty() ? bterms[t].x : a_of[bterms[t].x];
                        acc += av[ka] * pv[p_of[bterms[t].y]];
                    }
                    bv[e] = acc;
                }
            });
        }
        // P^T: per coarse row, (fine row, P slot) with fine rows ascending
        std::vector<int> rptr(size_t(nc) + 1, 0);
        for (size_t e = 0; e < np; ++e) ++rptr[pcols[e] + 1];
        std::partial_sum(rptr.begin(), rptr.end(), rptr.begin());
        std::vector<int2> rterms(np);
        std::vector<int> rcrs(np);  // P CSR entry of each pair
        {
            std::vector<int> pos(rptr.begin(), rptr.end() - 1);
            for (int i = 0; i < n; ++i)
                for (int e = prp[i]; e < prp[i + 1]; ++e) {
                    const int at = pos[pcols[e]]++;
                    rterms[at] = make_int2(i, ppos[e]);
                    rcrs[at] = e;
                }
        }
        // A_{l+1} = P^T B: per coarse row the sorted union of the B columns of
        // its P^T rows; terms (P slot, B entry) grouped by entry in (P^T pair,
        // B entry) order; values on the host as well
        auto crow = [&](int I, std::vector<int>& row) {
            row.clear();
            for (int k = rptr[I]; k < rptr[I + 1]; ++k) {
                const int i = rterms[k].x;
                row.insert(row.end(), bcols.begin() + brp[i], bcols.begin() + brp[i + 1]);
            }
            std::sort(row.begin(), row.end());
            row.erase(std::unique(row.begin(), row.end()), row.end());
        };
        std::vector<int> ccols;
        const std::vector<int> crp = par_csr(
            nc, ccols,
            [&](int I) {
                thread_local std::vector<int> row;
                crow(I, row);
                return (long long)row.size();
            },
            [&](int I, int* out) {
                thread_local std::vector<int> row;
                crow(I, row);
                std::copy(row.begin(), row.end(), out);
            });
        const int ncn = int(ccols.size());
        std::vector<int2> cterms;
        std::vector<int> cptr;
        std::vector<double> cv(size_t(ncn), 0.0);
        {
            std::vector<int> cnt(size_t(ncn), 0);
            par_for(nc, [&](int lo, int hi) {
                for (int I = lo; I < hi; ++I) {
                    const int* cb = ccols.data() + crp[I];
                    const int clen = crp[I + 1] - crp[I];
                    for (int k = rptr[I]; k < rptr[I + 1]; ++k) {
                        const int i = rterms[k].x;
                        for (int e = brp[i]; e < brp[i + 1]; ++e)
                            ++cnt[crp[I] + int(std::lower_bound(cb, cb + clen, bcols[e]) - cb)];
                    }
                }
            });
            std::vector<long long> cp(size_t(ncn) + 1, 0);
            for (int e = 0; e < ncn; ++e) cp[e + 1] = cp[e] + cnt[e];
            OAGPU_REQUIRE(cp[ncn] < (1LL << 31), "smoothed-aggregation list longer than 2^31 entries");
            cptr.assign(cp.begin(), cp.end());
            cterms.resize(size_t(cp[ncn]));
            par_for(nc, [&](int lo, int hi) {
                std::vector<int> at;
                for (int I = lo; I < hi; ++I) {
                    const int* cb = ccols.data() + crp[I];
                    const int clen = crp[I + 1] - crp[I];
                    at.assign(cptr.begin() + crp[I], cptr.begin() + crp[I + 1]);
                    for (int k = rptr[I]; k < rptr[I + 1]; ++k) {
                        const int i = rterms[k].x;
                        for (int e = brp[i]; e < brp[i + 1]; ++e) {
                            const int c = crp[I] + int(std::lower_bound(cb, cb + clen, bcols[e]) - cb);
                            cterms[at[c - crp[I]]++] = make_int2(rterms[k].y, e);
                            cv[c] += pv[rcrs[k]] * bv[e];
                        }
                    }
                }
            });
        }
        // device copies of this level's transfer data
        L.omega_p = omega_p;
        L.p_nsell = ps.nsell;
        L.p_slice_ptr.upload(ps.slice_ptr);
        L.p_slice_w.upload(ps.slice_w);
        L.p_cols.upload(ps.cols);
        L.prp.upload(prp);
        L.ppos.upload(ppos);
        L.own.upload(own);
        L.pt_ptr.upload(pt_ptr);
        if (pt_idx.empty()) pt_idx.push_back(0);
        L.pt_idx.upload(pt_idx);
        L.psum.alloc(std::max<size_t>(np, 1));
        L.P.alloc(ps.nsell);
        L.P.zero(f.stream);
        L.fP.alloc(ps.nsell);
        L.fP.zero(f.stream);
        L.nb = size_t(nb);
        L.b_ptr.upload(bptr);
        L.b_terms.upload(bterms);
        L.b_lanes = sa_term_lanes(bterms.size(), L.nb);
        L.B.alloc(std::max<size_t>(L.nb, 1));
        L.nac = size_t(ncn);
        L.c_ptr.upload(cptr);
        L.c_terms.upload(cterms);
        L.c_lanes = sa_term_lanes(cterms.size(), size_t(ncn));
        L.r_ptr.upload(rptr);
        L.r_terms.upload(rterms);
        L.r_lanes = sa_lanes(np, nc);
        L.h_agg = agg;
        L.h_prp = prp;
        L.h_pcols = pcols;
        for (auto* b : {&L.X, &L.S, &L.T}) b->alloc(size_t(n));
        for (auto* b : {&L.fX, &L.fS, &L.fT}) b->alloc(size_t(n));
        // the next level, CSR
        SaLevel C;
        C.n = nc;
        C.a_rp.upload(crp);
        C.a_ci.upload(ccols);
        std::vector<int> diag(nc);
        for (int I = 0; I < nc; ++I) {
            const int* b = ccols.data() + crp[I];
            diag[I] = crp[I] + int(std::lower_bound(b, b + (crp[I + 1] - crp[I]), I) - b);
        }
        C.a_diag.upload(diag);
        C.A.alloc(std::max<size_t>(size_t(ncn), 1));
        C.fA.alloc(std::max<size_t>(size_t(ncn), 1));
        C.dinv.alloc(nc);
        C.fdinv.alloc(nc);
        C.lanes = sa_lanes(size_t(ncn), nc);
        for (auto* b : {&C.R, &C.X, &C.S, &C.T, &C.Y}) b->alloc(size_t(nc));
        for (auto* b : {&C.fR, &C.fX, &C.fS, &C.fT, &C.fY}) b->alloc(size_t(nc));
        C.h_rp = crp;
        C.h_cols = ccols;
        rp = crp;
        ci = ccols;
        av.swap(cv);
        aidx.resize(ci.size());
        std::iota(aidx.begin(), aidx.end(), 0);
        n = nc;
        sa.lv.push_back(std::move(C));
    }
    const int nl = int(sa.lv.size());
    sa.dense_n = nl > 1 && sa.lv[nl - 1].n <= kDenseMax ? sa.lv[nl - 1].n : 0;
    if (sa.dense_n) {
        sa.dense.alloc(size_t(sa.dense_n) * sa.dense_n);
        sa.fdense.alloc(size_t(sa.dense_n) * sa.dense_n);
    }
    OAGPU_HIP(hipStreamSynchronize(f.stream));
    sa.built = true;
}

void sa_setup(Flow& f, const double* A0, bool want_float) {
    sa_build(f);
    Sa& sa = f.sa;
    const int nl = int(sa.lv.size());
    hipStream_t st = f.stream;
    auto tf = [&](size_t n, const double* x, float* y) {
        if (n) k_sa_to_float<<<unsigned((n + 255) / 256), 256, 0, st>>>(n, x, y);
    };
    for (int l = 0; l + 1 < nl; ++l) {
        SaLevel& L = sa.lv[l];
        SaLevel& C = sa.lv[l + 1];
        const double* Al = l == 0 ? A0 : L.A.p;
        const int np = L.h_prp.back();
        if (np)
            k_sa_psum<double><<<xcd_grid(size_t(np), 256), 256, 0, st>>>(np, L.pt_ptr.p, L.pt_idx.p, Al, L.psum.p);
        if (l == 0)
            k_sa_prow_sell<double><<<blocks(L.n), 256, 0, st>>>(view_of(*f.m), Al, L.prp.p, L.own.p, L.ppos.p,
                                                        L.psum.p, L.omega_p, L.P.p);
        else
            SA_DISPATCH(L.lanes, launch_prow_csr, st, L.n, L.a_rp.p, L.a_diag.p, Al, L.prp.p, L.own.p,
                        L.ppos.p, L.psum.p, L.omega_p, L.P.p)
        if (L.nb)
            launch_sa_terms<double>(st, L.b_lanes, L.nb, L.b_ptr.p, L.b_terms.p, Al, L.P.p, L.B.p);
        if (L.nac)
            launch_sa_terms<double>(st, L.c_lanes, L.nac, L.c_ptr.p, L.c_terms.p, L.P.p, L.B.p, C.A.p);
        k_sa_dinv<<<blocks(C.n), 256, 0, st>>>(C.n, C.a_diag.p, C.A.p, C.dinv.p,
                                               want_float ? C.fdinv.p : nullptr);
        if (want_float) {
            tf(L.p_nsell, L.P.p, L.fP.p);
            tf(L.nac, C.A.p, C.fA.p);
        }
    }
    if (sa.dense_n) {
        SaLevel& C = sa.lv[nl - 1];
        k_sa_dense_inverse<double><<<1, 256, 0, st>>>(C.n, C.a_rp.p, C.a_ci.p, C.A.p, sa.dense.p);
        if (want_float)
            k_sa_dense_inverse<float><<<1, 256, 0, st>>>(C.n, C.a_rp.p, C.a_ci.p, C.A.p, sa.fdense.p);
    }
    OAGPU_HIP(hipGetLastError());
}

// The mixed form's numeric set-up, FP32 throughout: P, the products B = A P
// and P^T B, the inverse diagonals and the coarsest dense inverse computed
// from the FP32 fine values A0f straight into the FP32 level arrays the
// FP32 V-cycle reads (sums accumulated in double); the FP64 level arrays are
// left as the last FP64 set-up wrote them.
void sa_setup_f32(Flow& f, const float* A0f) {
    sa_build(f);
    Sa& sa = f.sa;
    const int nl = int(sa.lv.size());
    hipStream_t st = f.stream;
    for (int l = 0; l + 1 < nl; ++l) {
        SaLevel& L = sa.lv[l];
        SaLevel& C = sa.lv[l + 1];
        const float* Al = l == 0 ? A0f : L.fA.p;
        float* psum = reinterpret_cast<float*>(L.psum.p);  // the FP64 arrays' space, half used
        float* B = reinterpret_cast<float*>(L.B.p);
        const int np = L.h_prp.back();
        if (np)
            k_sa_psum<float><<<xcd_grid(size_t(np), 256), 256, 0, st>>>(np, L.pt_ptr.p, L.pt_idx.p, Al, psum);
        if (l == 0)
            k_sa_prow_sell<float><<<blocks(L.n), 256, 0, st>>>(view_of(*f.m), Al, L.prp.p, L.own.p, L.ppos.p,
                                                               psum, L.omega_p, L.fP.p);
        else
            SA_DISPATCH(L.lanes, launch_prow_csr, st, L.n, L.a_rp.p, L.a_diag.p, Al, L.prp.p, L.own.p,
                        L.ppos.p, (const float*)psum, L.omega_p, L.fP.p)
        if (L.nb)
            launch_sa_terms<float>(st, L.b_lanes, L.nb, L.b_ptr.p, L.b_terms.p, Al, L.fP.p, B);
        if (L.nac)
            launch_sa_terms<float>(st, L.c_lanes, L.nac, L.c_ptr.p, L.c_terms.p, L.fP.p, B, C.fA.p);
        k_sa_dinv_f<<<blocks(C.n), 256, 0, st>>>(C.n, C.a_diag.p, C.fA.p, C.fdinv.p);
    }
    if (sa.dense_n) {
        SaLevel& C = sa.lv[nl - 1];
        k_sa_dense_inverse<float, float><<<1, 256, 0, st>>>(C.n, C.a_rp.p, C.a_ci.p, C.fA.p, sa.fdense.p);
    }
    OAGPU_HIP(hipGetLastError());
}

template <typename V>
static V* sav(SaLevel& l, int which) {  // 0 R, 1 X, 2 S, 3 T, 4 Y, 5 A, 6 dinv, 7 P
    if constexpr (std::is_same_v<V, float>) {
        DBuf<float>* b[8] = {&l.fR, &l.fX, &l.fS, &l.fT, &l.fY, &l.fA, &l.fdinv, &l.fP};
        return b[which]->p;
    } else {
        DBuf<double>* b[8] = {&l.R, &l.X, &l.S, &l.T, &l.Y, &l.A, &l.dinv, &l.P};
        return b[which]->p;
    }
}

template <typename V>
void sa_vcycle(Flow& f, const V* A0, const V* dinv0, const V* r, V* z, bool timing,
               const int* stop) {
    Sa& sa = f.sa;
    const int nl = int(sa.lv.size());
    OAGPU_REQUIRE(nl > 1, "smoothed-aggregation hierarchy without a coarse level");
    const int last = nl - 1;
    const V om = V(kOmega);
    const int* flag = stop ? stop : f.flag.p;
    hipStream_t st = f.stream;
    auto rhs = [&](int l) -> const V* { return l == 0 ? r : sav<V>(sa.lv[l], 0); };
    // down: pre-smoothing from zero with the residual, restriction
    for (int l = 0; l < last; ++l) {
        SaLevel& L = sa.lv[l];
        SaLevel& C = sa.lv[l + 1];
        if (l == 0) {
            if (timing) clock_begin(f, kClkAmgPre);
            k_amg_pre_residual<V, true><<<blocks(L.n), 256, 0, st>>>(view_of(*f.m), A0, dinv0, om, r,
                                                                     sav<V>(L, 1), sav<V>(L, 2), flag);
            if (timing) clock_end(f, kClkAmgPre);
        } else {
            SA_DISPATCH(L.lanes, launch_csr_pre, st, L.n, L.a_rp.p, L.a_ci.p, (const V*)sav<V>(L, 5),
                        (const V*)sav<V>(L, 6), om, rhs(l), sav<V>(L, 1), sav<V>(L, 2), flag)
        }
        SA_DISPATCH(L.r_lanes, launch_csr_restrict, st, C.n, L.r_ptr.p, L.r_terms.p,
                    (const V*)sav<V>(L, 7), (const V*)sav<V>(L, 2), sav<V>(C, 0), flag)
    }
    // coarsest level
    SaLevel& Lc = sa.lv[last];
    V* xc = sav<V>(Lc, 4);
    if (sa.dense_n) {
        const V* inv;
        if constexpr (std::is_same_v<V, float>) inv = sa.fdense.p;
        else inv = sa.dense.p;
        k_amg_dense_apply<V><<<1, kDenseMax, 0, st>>>(Lc.n, inv, rhs(last), xc, flag);
    } else {  // coarsening stalled above the dense size: Jacobi steps
        k_amg_jacobi0<V><<<blocks(Lc.n), 256, 0, st>>>(Lc.n, om, sav<V>(Lc, 6), rhs(last), xc, flag);
        for (int it = 0; it < 8; ++it) {
            const V* tin = (it & 1) ? sav<V>(Lc, 3) : xc;
            V* tout = (it & 1) ? xc : sav<V>(Lc, 3);
            SA_DISPATCH(Lc.lanes, launch_csr_smooth, st, Lc.n, Lc.a_rp.p, Lc.a_ci.p,
                        (const V*)sav<V>(Lc, 5), (const V*)sav<V>(Lc, 6), om, rhs(last), tin, tout, flag)
        }
    }
    // up: prolongation and post-smoothing
    const V* ec = xc;
    for (int l = last - 1; l >= 0; --l) {
        SaLevel& L = sa.lv[l];
        V* out = l == 0 ? z : sav<V>(L, 4);
        if (l == 0 && timing) clock_begin(f, kClkAmgPost);
        k_sa_prolong<V><<<blocks(L.n), 256, 0, st>>>(L.n, L.p_slice_ptr.p, L.p_slice_w.p, L.p_cols.p,
                                                     sav<V>(L, 7), ec, sav<V>(L, 1), sav<V>(L, 3), flag);
        if (l == 0)
            k_amg_smooth<V><<<blocks(L.n), 256, 0, st>>>(view_of(*f.m), A0, dinv0, om, r, sav<V>(L, 3), out,
                                                         flag);
        else
            SA_DISPATCH(L.lanes, launch_csr_smooth, st, L.n, L.a_rp.p, L.a_ci.p, (const V*)sav<V>(L, 5),
                        (const V*)sav<V>(L, 6), om, rhs(l), (const V*)sav<V>(L, 3), out, flag)
        if (l == 0 && timing) clock_end(f, kClkAmgPost);
        ec = out;
    }
    OAGPU_HIP(hipGetLastError());
}

template void amg_setup<double>(Flow&, const double*, const double*);
template void amg_setup<float>(Flow&, const float*, const float*);
template void amg_vcycle<double>(Flow&, const double*, const double*, const double*, double*,
                                 double*, double*, bool, bool, const double*, double*,
                                 const float*);
template void amg_vcycle<float>(Flow&, const float*, const float*, const float*, float*, double*,
                                float*, bool, bool, const float*, double*, const float*);
template void sa_vcycle<double>(Flow&, const double*, const double*, const double*, double*, bool,
                                const int*);
template void sa_vcycle<float>(Flow&, const float*, const float*, const float*, float*, bool,
                               const int*);

}  // namespace oagpu
