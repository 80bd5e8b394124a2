// krylov.hip -- device Krylov solvers for the block systems (N = 1, 3), whose
// values are stored sliced-ELL (64-row slices, lane-contiguous).
//
// Replaces liblinsolve's linearSolver::Base<N>::solve (reached from
// linearSystem<N>::solveSystem_, src/equation/linearSystem.h:416-462).  The
// reference delegates to PETSc / Trilinos / HYPRE; these are Jacobi
// preconditioned textbook methods (Saad, Iterative Methods for Sparse Linear
// Systems, 2nd ed.: GMRES Alg. 9.5 with right preconditioning and classical
// Gram-Schmidt applied twice, PCG Alg. 9.1, BiCGStab Alg. 7.7 right
// preconditioned), all vectors resident in HBM.
//
// Every reduction uses one fixed partition (kRedGrid blocks of kRedBlock
// lanes, grid-stride) and a fixed-order tree, so results are bitwise
// reproducible run to run.  CG keeps its scalars on the device: each
// consumer block re-reduces the kRedGrid partials itself, so an iteration is
// three launches with no host synchronisation; the host only polls the
// convergence flag every few iterations.
#include "device_util.hpp"
#include "flow.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <type_traits>

namespace oagpu {

template <int N>
__global__ void __launch_bounds__(256) k_jacobi_inv(int nn, const int* slice_ptr, const int* slice_w,
                                                    const int* doff, const double* A, double* dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const SellRow sr = sell_row<N * N>(slice_ptr, slice_w, i);
    for (int k = 0; k < N; ++k)
        dinv[N * size_t(i) + k] = 1.0 / A[sr.base + size_t(doff[i] * N * N + (N + 1) * k) * 64];
}

// acc = sum over the sliced-ELL slots of row i of A_block * xv, with
// xv = s .* x (s = nullptr: xv = x).  The lanes of a wavefront own the 64 rows
// of one slice: every column index and value load is one contiguous 256/512 B
// wavefront access.
template <int N, typename V>
__device__ __forceinline__ void sell_row_product(int i, const int* slice_ptr, const int* slice_w,
                                                 const int* scols, const V* A, const V* x,
                                                 const V* s, V acc[N]) {
    const int sl = i >> 6, ln = i & 63;
    const size_t s0 = size_t(slice_ptr[sl]);
    const int w = slice_w[sl];
    const int* cp = scols + s0 * 64 + ln;
    const V* ap = A + s0 * (N * N * 64) + ln;
    for (int m = 0; m < N; ++m) acc[m] = V(0);
#pragma unroll 2
    for (int q = 0; q < w; ++q) {
        const size_t c = size_t(N) * cp[64 * q];
        V xv[N];
        for (int r = 0; r < N; ++r) xv[r] = s ? s[c + r] * x[c + r] : x[c + r];
        const V* blk = ap + size_t(q) * (N * N * 64);
        for (int m = 0; m < N; ++m)
            for (int r = 0; r < N; ++r) acc[m] += blk[(N * m + r) * 64] * xv[r];
    }
}

// y = A (s .* x) (s = nullptr: y = A x), one row per lane, consecutive
// slices on one XCD (its L2 holds the neighbours' x); the block values and
// column indices, read once per product and larger than the caches, are
// loaded non-temporally so that they do not evict x
template <int N, typename T = double>
__global__ void __launch_bounds__(256) k_spmv(int nn, const int* slice_ptr, const int* slice_w,
                                              const int* scols, const T* A, const T* x,
                                              const T* s, T* y) {
    const int i = int(xcd_block() * blockDim.x + threadIdx.x);
    if (i >= nn) return;
    const int sl = i >> 6, ln = i & 63;
    const size_t s0 = size_t(slice_ptr[sl]);
    const int w = slice_w[sl];
    const int* cp = scols + s0 * 64 + ln;
    const T* ap = A + s0 * (N * N * 64) + ln;
    T acc[N];
    for (int m = 0; m < N; ++m) acc[m] = T(0);
#pragma unroll 4
    for (int q = 0; q < w; ++q) {
        const size_t c = size_t(N) * __builtin_nontemporal_load(cp + 64 * q);
        T xv[N];
        for (int r = 0; r < N; ++r) xv[r] = s ? s[c + r] * x[c + r] : x[c + r];
        const T* blk = ap + size_t(q) * (N * N * 64);
        for (int m = 0; m < N; ++m)
            for (int r = 0; r < N; ++r)
                acc[m] += __builtin_nontemporal_load(blk + (N * m + r) * 64) * xv[r];
    }
    for (int m = 0; m < N; ++m) y[N * size_t(i) + m] = acc[m];
}

__device__ inline double block_sum(double v, double* sh) {
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

// every block reduces the kRedGrid partials of slot `k` in the same order
__device__ inline double reduce_partials(const double* part, int k, double* sh) {
    double v = 0.0;
    for (int i = threadIdx.x; i < kRedGrid; i += blockDim.x) v += part[size_t(k) * kRedGrid + i];
    return block_sum(v, sh);
}

template <typename T>
struct DotPairsT {
    const T* x[64];
    const T* y[64];
};
using DotPairs = DotPairsT<double>;

__global__ void __launch_bounds__(kRedBlock) k_dots(size_t n, int np, DotPairs P, double* part) {
    __shared__ double sh[kRedBlock];
    for (int k = 0; k < np; ++k) {
        double v = 0.0;
        const double* x = P.x[k];
        const double* y = P.y[k];
        for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
             i += size_t(gridDim.x) * blockDim.x)
            v += x[i] * y[i];
        const double s = block_sum(v, sh);
        if (threadIdx.x == 0) part[size_t(k) * kRedGrid + blockIdx.x] = s;
    }
}

__global__ void __launch_bounds__(kRedBlock) k_reduce(int np, const double* part, double* out) {
    __shared__ double sh[kRedBlock];
    const int k = blockIdx.x;
    if (k >= np) return;
    const double s = reduce_partials(part, k, sh);
    if (threadIdx.x == 0) out[k] = s;
}

// x_k . y for up to 8 vectors x_k against one y (a Gram-Schmidt step): y is
// read once per element for all of them; partials of pair k0 + k to slot k0 + k
template <typename T = double>
__global__ void __launch_bounds__(kRedBlock) k_dots_y(size_t n, int k0, int nk, DotPairsT<T> P,
                                                      const T* y, double* part) {
    __shared__ double sh[kRedBlock];
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        const double yi = double(y[i]);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < nk) v[k] += double(P.x[k0 + k][i]) * yi;
    }
    for (int k = 0; k < nk; ++k) {
        const double s = block_sum(v[k], sh);
        if (threadIdx.x == 0) part[size_t(k0 + k) * kRedGrid + blockIdx.x] = s;
    }
}

// host-side: np dot products in one pass, results copied to the host
static void dots(Flow& f, size_t n, int np, const double* const* xs, const double* const* ys,
                 double* out) {
    DotPairs P;
    bool same_y = np > 1;
    for (int k = 0; k < np; ++k) {
        P.x[k] = xs[k];
        P.y[k] = ys[k];
        same_y = same_y && ys[k] == ys[0];
    }
    if (same_y) {
        for (int k0 = 0; k0 < np; k0 += 8)
            k_dots_y<double><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, k0, std::min(8, np - k0), P,
                                                                   ys[0], f.part.p);
    } else {
        k_dots<<<kRedGrid, kRedBlock, 0, f.stream>>>(n, np, P, f.part.p);
    }
    k_reduce<<<np, kRedBlock, 0, f.stream>>>(np, f.part.p, f.red.p);
    OAGPU_HIP(hipGetLastError());
    f.red.download(out, np, f.stream);
    OAGPU_HIP(hipStreamSynchronize(f.stream));
}

// ------------------------------------------------------------------ GMRES
// w -= sum_k h[k] V_k
template <typename T = double>
__global__ void k_gs_update(size_t n, int nk, const T* V, const double* h, T* w) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < nk; ++k) s += h[k] * double(V[size_t(k) * n + i]);
    w[i] = T(double(w[i]) - s);
}

// w -= sum_k h[k] V_k fused with the block partials of V_k . w for the
// updated w (the second Gram-Schmidt pass), nk <= MAXK
template <int MAXK, typename T = double>
__global__ void __launch_bounds__(kRedBlock) k_gs_update_dots(size_t n, int nk, const T* V,
                                                              const double* h, T* w,
                                                              double* part) {
    __shared__ double sh[kRedBlock];
    double acc[MAXK];
#pragma unroll
    for (int k = 0; k < MAXK; ++k) acc[k] = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < nk; ++k) s += h[k] * double(V[size_t(k) * n + i]);
        const T wt = T(double(w[i]) - s);
        w[i] = wt;
        const double wi = double(wt);
#pragma unroll
        for (int k = 0; k < MAXK; ++k)
            if (k < nk) acc[k] += double(V[size_t(k) * n + i]) * wi;
    }
    for (int k = 0; k < nk; ++k) {
        const double t = block_sum(acc[k], sh);
        if (threadIdx.x == 0) part[size_t(k) * kRedGrid + blockIdx.x] = t;
    }
}

// w -= sum_k h[k] V_k fused with the block partials of w . w (the norm that
// follows the second Gram-Schmidt pass)
template <typename T = double>
__global__ void __launch_bounds__(kRedBlock) k_gs_update_norm(size_t n, int nk, const T* V,
                                                              const double* h, T* w,
                                                              double* part) {
    __shared__ double sh[kRedBlock];
    double ww = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < nk; ++k) s += h[k] * double(V[size_t(k) * n + i]);
        const T wt = T(double(w[i]) - s);
        w[i] = wt;
        const double wi = double(wt);
        ww += wi * wi;
    }
    const double t = block_sum(ww, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <typename T = double>
__global__ void k_scale_to(size_t n, double a, const T* x, T* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = T(a * double(x[i]));
}

// x += dinv .* (V y)
template <typename T = double>
__global__ void k_gmres_x(size_t n, int nk, const T* V, const double* y, const T* dinv, T* x) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < nk; ++k) s += y[k] * double(V[size_t(k) * n + i]);
    x[i] = T(double(x[i]) + double(dinv[i]) * s);
}

template <typename T = double>
__global__ void k_residual_copy(size_t n, const T* b, const T* Ax, T* r) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) r[i] = b[i] - Ax[i];
}

// np dot products of T vectors (accumulated in double) to the host
template <typename T>
static void dots_t(Flow& f, size_t n, int np, const T* const* xs, const T* y, double* out) {
    DotPairsT<T> P;
    for (int k = 0; k < np; ++k) P.x[k] = xs[k];
    for (int k0 = 0; k0 < np; k0 += 8)
        k_dots_y<T><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, k0, std::min(8, np - k0), P, y, f.part.p);
    k_reduce<<<np, kRedBlock, 0, f.stream>>>(np, f.part.p, f.red.p);
    OAGPU_HIP(hipGetLastError());
    f.red.download(out, np, f.stream);
    OAGPU_HIP(hipStreamSynchronize(f.stream));
}

template <int N, typename T = double>
static void spmv(Flow& f, const T* A, const T* x, const std::type_identity_t<T>* s, T* y) {
    DevMesh& m = *f.m;
    constexpr bool f32 = std::is_same_v<T, float>;
    const int clk = N == 1 ? (f32 ? kClkSpmv1f : kClkSpmv1) : (f32 ? kClkSpmv3f : kClkSpmv3);
    const unsigned B = f.tune.spmv_block[N == 3][f32];
    clock_begin(f, clk);
    k_spmv<N, T><<<xcd_grid(size_t(m.nn), B), B, 0, f.stream>>>(m.nn, m.slice_ptr.p, m.slice_w.p,
                                                                 m.sell_cols.p, A, x, s, y);
    clock_end(f, clk);
    OAGPU_HIP(hipGetLastError());
}

// GMRES(m) right-preconditioned by Jacobi (x = D^-1 V y), restarted from the
// true residual, on values and vectors of type T (double; float for the
// inner solve of the mixed form), Gram-Schmidt coefficients and dot
// products in double.  Starts from x = 0; returns the iterations, r0 / r1
// the initial / final residual norms (the final one from b - A x).
template <int N, typename T>
static int gmres_core(Flow& f, const T* A, const T* b, const T* dinv, T* x, T* Vb, T* w, T* r,
                      int max_it, double rtol, double atol, int mres, double* r0out,
                      double* r1out) {
    DevMesh& m = *f.m;
    const size_t n = size_t(m.nn) * N;
    const unsigned B = 256, G = unsigned((n + B - 1) / B);
    OAGPU_HIP(hipMemsetAsync(x, 0, n * sizeof(T), f.stream));
    double beta;
    {
        const T* xs[1] = {b};
        dots_t<T>(f, n, 1, xs, b, &beta);
        beta = std::sqrt(beta);
    }
    const double r0 = beta;
    const double tol = std::max(rtol * r0, atol);
    *r0out = r0;
    std::vector<double> H(size_t(mres + 1) * mres), cs(mres), sn(mres), g(mres + 1), y(mres),
        h(mres + 1), h2(mres + 1);
    const T* rvec = b;
    int it = 0;
    while (it < max_it && beta > tol) {
        k_scale_to<T><<<G, B, 0, f.stream>>>(n, 1.0 / beta, rvec, Vb);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int jend = -1;
        for (int j = 0; j < mres && it < max_it; ++j) {
            spmv<N, T>(f, A, Vb + size_t(j) * n, dinv, w);
            double hn;
            std::vector<const T*> xs(j + 1);
            for (int k = 0; k <= j; ++k) xs[k] = Vb + size_t(k) * n;
            // classical Gram-Schmidt, twice
            if (j + 1 <= 16) {  // device-resident step: one host synchronisation
                DotPairsT<T> P;
                for (int k = 0; k <= j; ++k) P.x[k] = xs[k];
                for (int k0 = 0; k0 <= j; k0 += 8)
                    k_dots_y<T><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, k0, std::min(8, j + 1 - k0),
                                                                      P, w, f.part.p);
                k_reduce<<<j + 1, kRedBlock, 0, f.stream>>>(j + 1, f.part.p, f.red.p);
                k_gs_update_dots<16, T><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, j + 1, Vb, f.red.p,
                                                                               w, f.part.p);
                k_reduce<<<j + 1, kRedBlock, 0, f.stream>>>(j + 1, f.part.p, f.red.p + 64);
                k_gs_update_norm<T><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, j + 1, Vb, f.red.p + 64,
                                                                          w, f.part.p);
                k_reduce<<<1, kRedBlock, 0, f.stream>>>(1, f.part.p, f.red.p + 128);
                OAGPU_HIP(hipGetLastError());
                double hr[129];
                f.red.download(hr, 129, f.stream);
                OAGPU_HIP(hipStreamSynchronize(f.stream));
                for (int k = 0; k <= j; ++k) h[k] = hr[k] + hr[64 + k];
                hn = std::sqrt(hr[128]);
            } else {
                dots_t<T>(f, n, j + 1, xs.data(), w, h.data());
                f.red.upload(h.data(), j + 1, f.stream);
                k_gs_update<T><<<G, B, 0, f.stream>>>(n, j + 1, Vb, f.red.p, w);
                dots_t<T>(f, n, j + 1, xs.data(), w, h2.data());
                f.red.upload(h2.data(), j + 1, f.stream);
                k_gs_update_norm<T><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, j + 1, Vb, f.red.p, w,
                                                                          f.part.p);
                k_reduce<<<1, kRedBlock, 0, f.stream>>>(1, f.part.p, f.red.p);
                OAGPU_HIP(hipGetLastError());
                for (int k = 0; k <= j; ++k) h[k] += h2[k];
                f.red.download(&hn, 1, f.stream);
                OAGPU_HIP(hipStreamSynchronize(f.stream));
                hn = std::sqrt(hn);
            }
            for (int k = 0; k <= j; ++k) H[size_t(mres) * k + j] = h[k];
            H[size_t(mres) * (j + 1) + j] = hn;
            if (hn > 0) k_scale_to<T><<<G, B, 0, f.stream>>>(n, 1.0 / hn, w, Vb + size_t(j + 1) * n);
            for (int k = 0; k < j; ++k) {
                const double a0 = H[size_t(mres) * k + j], b0 = H[size_t(mres) * (k + 1) + j];
                H[size_t(mres) * k + j] = cs[k] * a0 + sn[k] * b0;
                H[size_t(mres) * (k + 1) + j] = -sn[k] * a0 + cs[k] * b0;
            }
            const double a0 = H[size_t(mres) * j + j], b0 = H[size_t(mres) * (j + 1) + j];
            const double rr = std::sqrt(a0 * a0 + b0 * b0);
            cs[j] = rr > 0 ? a0 / rr : 1.0;
            sn[j] = rr > 0 ? b0 / rr : 0.0;
            H[size_t(mres) * j + j] = rr;
            H[size_t(mres) * (j + 1) + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            ++it;
            jend = j;
            if (std::fabs(g[j + 1]) <= tol || hn == 0.0) break;
        }
        for (int k = jend; k >= 0; --k) {
            double t = g[k];
            for (int q = k + 1; q <= jend; ++q) t -= H[size_t(mres) * k + q] * y[q];
            y[k] = t / H[size_t(mres) * k + k];
        }
        f.red.upload(y.data(), jend + 1, f.stream);
        k_gmres_x<T><<<G, B, 0, f.stream>>>(n, jend + 1, Vb, f.red.p, dinv, x);
        spmv<N, T>(f, A, x, nullptr, w);
        k_residual_copy<T><<<G, B, 0, f.stream>>>(n, b, w, r);
        rvec = r;
        const T* rs[1] = {r};
        dots_t<T>(f, n, 1, rs, r, &beta);
        beta = std::sqrt(beta);
    }
    *r1out = beta;
    return it;
}

static int gmres_restart_len(int mres) { return std::min(std::max(mres, 1), 60); }

template <int N>
static int gmres(Flow& f, int max_it, double rtol, double atol, int mres, oagpu_control_data* cd) {
    DevMesh& m = *f.m;
    const size_t n = size_t(m.nn) * N;
    mres = gmres_restart_len(mres);
    if (f.V_cols < mres + 1 || f.V.n < n * size_t(mres + 1)) {
        f.V.alloc(n * size_t(mres + 1));
        f.V_cols = mres + 1;
    }
    const double* A = N == 3 ? f.A3.p : f.A1.p;
    const double* b = N == 3 ? f.b3.p : f.b1.p;
    double* x = N == 3 ? f.x3.p : f.x1.p;
    return gmres_core<N, double>(f, A, b, f.dinv.p, x, f.V.p, f.w.p, f.r.p, max_it, rtol, atol, mres,
                                 &cd->solver_initial_res, &cd->solver_final_res);
}

// -------------------------------------------------------------------- CG
// Vectors and matrix values of type V (double, or float for the inner solve
// of the mixed form); dot products are always accumulated in double.
// scal: [0..1] rz ping-pong, [2] tol^2, [3] last rr, [4] rr0
// PRE: the preconditioner is applied by separate launches (AMG V-cycle),
// which also form p = z and the r.z partials.
template <typename V, bool PRE>
__global__ void __launch_bounds__(kRedBlock) k_cg_init(size_t n, const V* b, const V* dinv, V* x,
                                                       V* r, V* z, V* p, double* part) {
    __shared__ double sh[kRedBlock];
    double rz = 0.0, rr = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        const V ri = b[i];
        x[i] = V(0);
        r[i] = ri;
        if (!PRE) {
            const V zi = dinv[i] * ri;
            z[i] = zi;
            p[i] = zi;
            rz += double(ri) * double(zi);
        }
        rr += double(ri) * double(ri);
    }
    const double s0 = block_sum(rz, sh);
    const double s1 = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s0;
        part[kRedGrid + blockIdx.x] = s1;
    }
}

// flag[0]: stop (converged or out of iterations), flag[1]: iterations done,
// flag[2]: converged, flag[3]: iteration limit
__global__ void __launch_bounds__(kRedBlock) k_cg_start(const double* part, double rtol, double atol,
                                                        int max_it, double* scal, int* flag) {
    __shared__ double sh[kRedBlock];
    const double rz = reduce_partials(part, 0, sh);
    const double rr = reduce_partials(part, 1, sh);
    if (threadIdx.x == 0) {
        const double tol2 = fmax(rtol * rtol * rr, atol * atol);
        scal[0] = rz;
        scal[2] = tol2;
        scal[3] = rr;
        scal[4] = rr;
        flag[0] = rr <= tol2 || max_it <= 0 ? 1 : 0;
        flag[1] = 0;
        flag[2] = rr <= tol2 ? 1 : 0;
        flag[3] = max_it;
    }
}

template <int N, typename V>
__global__ void __launch_bounds__(kRedBlock) k_cg_spmv(int nn, const int* slice_ptr,
                                                       const int* slice_w, const int* scols,
                                                       const V* A, const V* p, V* q, double* part,
                                                       const int* flag) {
    __shared__ double sh[kRedBlock];
    if (flag[0]) return;
    double pq = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
        V acc[N];
        sell_row_product<N, V>(i, slice_ptr, slice_w, scols, A, p, nullptr, acc);
        for (int m = 0; m < N; ++m) {
            q[N * size_t(i) + m] = acc[m];
            pq += double(p[N * size_t(i) + m]) * double(acc[m]);
        }
    }
    const double s = block_sum(pq, sh);
    if (threadIdx.x == 0) part[2 * kRedGrid + blockIdx.x] = s;
}

template <typename V, bool PRE>
__global__ void __launch_bounds__(kRedBlock) k_cg_update1(size_t n, const V* p, const V* q,
                                                          const V* dinv, V* x, V* r, V* z,
                                                          double* part, const double* scal,
                                                          const int* flag) {
    __shared__ double sh[kRedBlock];
    if (flag[0]) return;
    const int it = flag[1];
    const double pq = reduce_partials(part, 2, sh);
    const V alpha = V(scal[it & 1] / pq);
    double rz = 0.0, rr = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        x[i] += alpha * p[i];
        const V ri = r[i] - alpha * q[i];
        r[i] = ri;
        if (!PRE) {
            const V zi = dinv[i] * ri;
            z[i] = zi;
            rz += double(ri) * double(zi);
        }
        rr += double(ri) * double(ri);
    }
    const double s0 = block_sum(rz, sh);
    const double s1 = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        part[3 * kRedGrid + blockIdx.x] = s0;
        part[4 * kRedGrid + blockIdx.x] = s1;
    }
}

__global__ void k_clear_flag(int* flag) { flag[0] = 0; }

// Block partials (fixed grid of kRedGrid blocks) of r . z into rz_part and,
// with q, of q . z into qz_part; with p, p = z: what the unsmoothed cycle's
// last kernel produces on the way, for the smoothed-aggregation cycle.
template <typename V>
__global__ void __launch_bounds__(kRedBlock) k_pc_partials(size_t n, const V* r, const V* z,
                                                           const V* q, V* p, double* rz_part,
                                                           double* qz_part, const int* flag) {
    __shared__ double sh[kRedBlock];
    if (flag[0]) return;
    double rz = 0.0, qz = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        const V zi = z[i];
        rz += double(r[i]) * double(zi);
        if (q) qz += double(q[i]) * double(zi);
        if (p) p[i] = zi;
    }
    rz = block_sum(rz, sh);
    if (q) qz = block_sum(qz, sh);
    if (threadIdx.x == 0) {
        rz_part[blockIdx.x] = rz;
        if (q) qz_part[blockIdx.x] = qz;
    }
}

// FLEX: flexible CG (the preconditioner is the nonlinear K-cycle):
// beta = -(z . q) / (p . q) keeps the new direction A-orthogonal to the last
// one (Notay, SIAM J. Sci. Comput. 22 (2000)); else beta = r.z / r_old.z_old
template <typename V, bool FLEX = false>
__global__ void __launch_bounds__(kRedBlock) k_cg_update2(size_t n, const V* z, V* p,
                                                          const double* part, double* scal,
                                                          int* flag) {
    __shared__ double sh[kRedBlock];
    if (flag[0]) return;
    const int it = flag[1];
    const double rz = reduce_partials(part, 3, sh);
    const double rr = reduce_partials(part, 4, sh);
    const bool conv = rr <= scal[2];
    if (!conv) {
        V beta;
        if constexpr (FLEX) {
            const double pq = reduce_partials(part, 2, sh), qz = reduce_partials(part, 5, sh);
            beta = V(pq != 0.0 ? -qz / pq : 0.0);
        } else {
            beta = V(rz / scal[it & 1]);
        }
        for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
             i += size_t(gridDim.x) * blockDim.x)
            p[i] = z[i] + beta * p[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[(it + 1) & 1] = rz;
        scal[3] = rr;
        if (conv) flag[2] = 1;
    }
}

// end of an iteration: count it, stop on convergence or at the limit (a
// separate launch: the iteration kernels read flag[1] while they run)
__global__ void k_cg_tick(int* flag) {
    if (flag[0]) return;
    const int done = flag[1] + 1;
    flag[1] = done;
    if (flag[2] || done >= flag[3]) flag[0] = 1;
}

// Device-resident PCG on A x = b (values V); returns iterations, fills
// initial / final ||r||_2 in V arithmetic.  amg: preconditioner = one AMG
// K-cycle under flexible CG (N = 1), else Jacobi under CG.
template <int N, typename V>
static void cg_iteration(Flow& f, const V* A, const V* dinv, V* x, V* r, V* z, V* p, V* q, bool amg,
                         bool sa = false) {
    DevMesh& m = *f.m;
    const size_t n = size_t(m.nn) * N;
    k_cg_spmv<N, V><<<kRedGrid, kRedBlock, 0, f.stream>>>(m.nn, m.slice_ptr.p, m.slice_w.p,
                                                           m.sell_cols.p, A, p, q, f.part.p,
                                                           f.flag.p);
    if (amg) {
        if constexpr (N == 1) {
            k_cg_update1<V, true><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, p, q, dinv, x, r, z,
                                                                         f.part.p, f.scal.p,
                                                                         f.flag.p);
            if (sa) {
                sa_vcycle<V>(f, A, dinv, r, z);
                k_pc_partials<V><<<kRedGrid, kRedBlock, 0, f.stream>>>(
                    n, r, z, q, nullptr, f.part.p + 3 * kRedGrid, f.part.p + 5 * kRedGrid, f.flag.p);
            } else {
                amg_vcycle<V>(f, A, dinv, r, z, f.part.p + 3 * kRedGrid, nullptr, false, true, q,
                              f.part.p + 5 * kRedGrid);
            }
            k_cg_update2<V, true><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, z, p, f.part.p, f.scal.p,
                                                                         f.flag.p);
        }
    } else {
        k_cg_update1<V, false><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, p, q, dinv, x, r, z,
                                                                      f.part.p, f.scal.p, f.flag.p);
        k_cg_update2<V><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, z, p, f.part.p, f.scal.p, f.flag.p);
    }
    k_cg_tick<<<1, 1, 0, f.stream>>>(f.flag.p);
}

// Device-resident PCG on A x = b (values V); returns iterations, fills
// initial / final ||r||_2 in V arithmetic.  amg: preconditioner = one AMG
// K-cycle under flexible CG (N = 1), else Jacobi under CG.  Every iteration reads its index and the
// stop flag from device memory, so a block of kPoll iterations is the same
// launch sequence every time: with graphs on it is captured once per set of
// operands and replayed (one launch per block instead of ~15 per iteration);
// the host reads the stop flag between blocks.
constexpr int kPoll = 8;

template <int N, typename V>
static int cg_core(Flow& f, const V* A, const V* b, const V* dinv, V* x, V* r, V* z, V* p, V* q,
                   int max_it, double rtol, double atol, double* r0, double* r1, bool amg = false,
                   bool sa = false) {
    DevMesh& m = *f.m;
    const size_t n = size_t(m.nn) * N;
    if constexpr (N != 1) OAGPU_REQUIRE(!amg, "AMG preconditioning is for scalar systems");
    k_clear_flag<<<1, 1, 0, f.stream>>>(f.flag.p);  // a kernel, not a memset: capturable
    if (amg) {
        k_cg_init<V, true><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, b, dinv, x, r, z, p, f.part.p);
        if constexpr (N == 1) {
            if (sa) {
                sa_vcycle<V>(f, A, dinv, r, z);
                k_pc_partials<V><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, r, z, nullptr, p, f.part.p,
                                                                        nullptr, f.flag.p);
            } else {
                amg_vcycle<V>(f, A, dinv, r, z, f.part.p, p, true, true);  // z, p = z, r.z
            }
        }
    } else {
        k_cg_init<V, false><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, b, dinv, x, r, z, p, f.part.p);
    }
    k_cg_start<<<1, kRedBlock, 0, f.stream>>>(f.part.p, rtol, atol, max_it, f.scal.p, f.flag.p);
    OAGPU_HIP(hipGetLastError());
    int hflag[2] = {0, 0};
    const std::array<const void*, 10> key = {A, dinv, x, r, z, p, q, (const void*)(size_t)N,
                                             (const void*)(size_t)sizeof(V),
                                             (const void*)(size_t)(amg + 2 * sa)};
    const bool graphs = f.prm.use_graphs != 0;
    if (graphs && (f.cg_graph.exec == nullptr || f.cg_graph.key != key)) {
        f.cg_graph.reset();
        hipGraph_t g = nullptr;
        OAGPU_HIP(hipStreamBeginCapture(f.stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < kPoll; ++k) cg_iteration<N, V>(f, A, dinv, x, r, z, p, q, amg, sa);
        OAGPU_HIP(hipStreamEndCapture(f.stream, &g));
        const hipError_t e = hipGraphInstantiate(&f.cg_graph.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        OAGPU_HIP(e);
        f.cg_graph.key = key;
    }
    for (int done = 0; done < max_it; done += kPoll) {
        if (graphs) {
            OAGPU_HIP(hipGraphLaunch(f.cg_graph.exec, f.stream));
        } else {
            for (int k = 0; k < kPoll; ++k) cg_iteration<N, V>(f, A, dinv, x, r, z, p, q, amg, sa);
            OAGPU_HIP(hipGetLastError());
        }
        f.flag.download(hflag, 2, f.stream);
        OAGPU_HIP(hipStreamSynchronize(f.stream));
        if (hflag[0]) break;
    }
    f.flag.download(hflag, 2, f.stream);
    double sc[5];
    f.scal.download(sc, 5, f.stream);
    OAGPU_HIP(hipStreamSynchronize(f.stream));
    *r0 = std::sqrt(sc[4]);
    *r1 = std::sqrt(sc[3]);
    return hflag[1];
}

// The CG of the graph-captured iteration (FP64 scalar system, sa: one
// smoothed-aggregation V-cycle as preconditioner, else Jacobi): the launches
// of cg_core without its host polls.  f.flag[0] stops every later launch once
// the solve has converged or used max_it iterations; f.flag[1] counts them,
// f.scal[4] / [3] hold the initial / latest r . r.
void cg_device_begin(Flow& f, int max_it, double rtol, double atol, bool sa) {
    const size_t n = size_t(f.m->nn);
    k_clear_flag<<<1, 1, 0, f.stream>>>(f.flag.p);
    if (sa) {
        k_cg_init<double, true><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, f.b1.p, f.dinv.p, f.x1.p, f.r.p,
                                                                       f.z.p, f.pk.p, f.part.p);
        sa_vcycle<double>(f, f.A1.p, f.dinv.p, f.r.p, f.z.p);
        k_pc_partials<double><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, f.r.p, f.z.p, nullptr, f.pk.p,
                                                                     f.part.p, nullptr, f.flag.p);
    } else {
        k_cg_init<double, false><<<kRedGrid, kRedBlock, 0, f.stream>>>(n, f.b1.p, f.dinv.p, f.x1.p, f.r.p,
                                                                        f.z.p, f.pk.p, f.part.p);
    }
    k_cg_start<<<1, kRedBlock, 0, f.stream>>>(f.part.p, rtol, atol, max_it, f.scal.p, f.flag.p);
    OAGPU_HIP(hipGetLastError());
}

void cg_device_iteration(Flow& f, bool sa) {
    cg_iteration<1, double>(f, f.A1.p, f.dinv.p, f.x1.p, f.r.p, f.z.p, f.pk.p, f.q.p, sa, sa);
    OAGPU_HIP(hipGetLastError());
}

// one iteration of the FP32 inner CG of the mixed form (FP32 values, the
// FP32 hierarchy's V-cycle or Jacobi)
void cg_device_iteration_f32(Flow& f, bool sa) {
    cg_iteration<1, float>(f, f.Af.p, f.fdinv.p, f.fx.p, f.fr.p, f.fz.p, f.fp.p, f.fq.p, sa, sa);
    OAGPU_HIP(hipGetLastError());
}

// sa: the smoothed-aggregation V-cycle (OAGPU_CG_SA), else the unsmoothed
// hierarchy's K-cycle (OAGPU_CG_AMG)
template <int N>
static int cg(Flow& f, int max_it, double rtol, double atol, oagpu_control_data* cd, bool amg,
              bool sa = false) {
    const double* A = N == 3 ? f.A3.p : f.A1.p;
    const double* b = N == 3 ? f.b3.p : f.b1.p;
    double* x = N == 3 ? f.x3.p : f.x1.p;
    if (amg && sa) {
        sa_setup(f, A, false);
        amg = sa = f.sa.lv.size() > 1;  // no coarse level (tiny system): Jacobi-CG
    } else if (amg) {
        amg_setup<double>(f, A, f.dinv.p);
        amg = f.amg.lv.size() > 1;  // no coarse level (tiny system): Jacobi-CG
    }
    return cg_core<N, double>(f, A, b, f.dinv.p, x, f.r.p, f.z.p, f.pk.p, f.q.p, max_it, rtol, atol,
                              &cd->solver_initial_res, &cd->solver_final_res, amg, sa);
}

// ---------------------------------------------- mixed: FP32 CG + FP64 refinement
__global__ void k_to_float(size_t n, double s, const double* x, float* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = float(s * x[i]);
}

__global__ void k_refine(size_t n, double s, const float* d, double* x) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] += s * double(d[i]);
}

// Iterative refinement: r = b - A x in FP64, FP32 PCG on A d = r / ||r||,
// x += ||r|| d in FP64, until ||r|| <= max(rtol ||b||, atol) or the total
// inner iteration budget max_it is spent.  The FP32 matrix is A rounded.
template <int N>
static int cg_mixed(Flow& f, int max_it, double rtol, double atol, oagpu_control_data* cd, bool amg,
                    bool sa = false) {
    DevMesh& m = *f.m;
    const size_t n = size_t(m.nn) * N, nv = m.nsell * N * N;
    const double* A = N == 3 ? f.A3.p : f.A1.p;
    const double* b = N == 3 ? f.b3.p : f.b1.p;
    double* x = N == 3 ? f.x3.p : f.x1.p;
    if (f.Af.n < nv) f.Af.alloc(nv);
    for (auto* v : {&f.fb, &f.fx, &f.fr, &f.fz, &f.fp, &f.fq, &f.fdinv})
        if (v->n < n) v->alloc(n);
    const unsigned B = 256;
    k_to_float<<<unsigned((nv + B - 1) / B), B, 0, f.stream>>>(nv, 1.0, A, f.Af.p);
    k_to_float<<<unsigned((n + B - 1) / B), B, 0, f.stream>>>(n, 1.0, f.dinv.p, f.fdinv.p);
    if (amg && sa) {
        sa_setup_f32(f, f.Af.p);  // FP32 levels for the FP32 inner solve
        amg = sa = f.sa.lv.size() > 1;
    } else if (amg) {
        amg_setup<float>(f, f.Af.p, f.fdinv.p);
        amg = f.amg.lv.size() > 1;
    }
    OAGPU_HIP(hipMemsetAsync(x, 0, n * sizeof(double), f.stream));
    OAGPU_HIP(hipMemcpyAsync(f.r.p, b, n * sizeof(double), hipMemcpyDeviceToDevice, f.stream));
    double rn;
    {
        const double* xs[1] = {f.r.p};
        dots(f, n, 1, xs, xs, &rn);
        rn = std::sqrt(rn);
    }
    cd->solver_initial_res = rn;
    const double tol = std::max(rtol * rn, atol);
    // Inner FP32 solves stop at 1e-2 of their right-hand side.  The true
    // residual an FP32 CG attains on these Laplacians stalls near 1e-4 (its
    // recursive residual keeps falling), so a tighter inner tolerance only
    // adds iterations: on a 64^3 cavity p' system the total AMG-CG count to
    // 1e-6 is 65 with a 1e-2 floor (3 refinement steps), 82 with 1e-4, 60 for
    // FP64 CG alone.
    const double inner_rtol = std::max(rtol, 1.0e-2);
    int used = 0;
    while (used < max_it && rn > tol) {
        const double inner = std::max(inner_rtol, 0.5 * tol / rn);
        k_to_float<<<unsigned((n + B - 1) / B), B, 0, f.stream>>>(n, 1.0 / rn, f.r.p, f.fb.p);
        double i0, i1;
        const int it = cg_core<N, float>(f, f.Af.p, f.fb.p, f.fdinv.p, f.fx.p, f.fr.p, f.fz.p, f.fp.p,
                                         f.fq.p, max_it - used, inner, 0.0, &i0, &i1, amg, sa);
        used += std::max(it, 1);
        k_refine<<<unsigned((n + B - 1) / B), B, 0, f.stream>>>(n, rn, f.fx.p, x);
        spmv<N>(f, A, x, nullptr, f.w.p);
        k_residual_copy<<<unsigned((n + B - 1) / B), B, 0, f.stream>>>(n, b, f.w.p, f.r.p);
        const double* xs[1] = {f.r.p};
        dots(f, n, 1, xs, xs, &rn);
        rn = std::sqrt(rn);
        if (it == 0) break;
    }
    cd->solver_final_res = rn;
    return used;
}

// Mixed form of GMRES: FP32 GMRES(m) on A d = r / ||r|| (A rounded to FP32,
// FP32 basis), x += ||r|| d and r = b - A x in FP64, until ||r|| <= max(rtol
// ||b||, atol) or the iteration budget is spent.  The momentum and k / omega
// systems are solved to 1e-1 / 1e-2 of their right-hand side, which one FP32
// inner solve reaches; the FP64 residual confirms it.
template <int N>
static int gmres_mixed(Flow& f, int max_it, double rtol, double atol, int mres,
                       oagpu_control_data* cd) {
    DevMesh& m = *f.m;
    const size_t n = size_t(m.nn) * N, nv = m.nsell * N * N;
    mres = gmres_restart_len(mres);
    const double* A = N == 3 ? f.A3.p : f.A1.p;
    const double* b = N == 3 ? f.b3.p : f.b1.p;
    double* x = N == 3 ? f.x3.p : f.x1.p;
    if (f.Af.n < nv) f.Af.alloc(nv);
    for (auto* v : {&f.fb, &f.fx, &f.fr, &f.fw, &f.fdinv})
        if (v->n < n) v->alloc(n);
    if (f.fV.n < n * size_t(mres + 1)) f.fV.alloc(n * size_t(mres + 1));
    const unsigned B = 256;
    k_to_float<<<unsigned((nv + B - 1) / B), B, 0, f.stream>>>(nv, 1.0, A, f.Af.p);
    k_to_float<<<unsigned((n + B - 1) / B), B, 0, f.stream>>>(n, 1.0, f.dinv.p, f.fdinv.p);
    OAGPU_HIP(hipMemsetAsync(x, 0, n * sizeof(double), f.stream));
    OAGPU_HIP(hipMemcpyAsync(f.r.p, b, n * sizeof(double), hipMemcpyDeviceToDevice, f.stream));
    double rn;
    {
        const double* xs[1] = {f.r.p};
        dots(f, n, 1, xs, xs, &rn);
        rn = std::sqrt(rn);
    }
    cd->solver_initial_res = rn;
    const double tol = std::max(rtol * rn, atol);
    int used = 0;
    while (used < max_it && rn > tol) {
        const double inner = std::max(tol / rn, 1.0e-5);
        k_to_float<<<unsigned((n + B - 1) / B), B, 0, f.stream>>>(n, 1.0 / rn, f.r.p, f.fb.p);
        double i0, i1;
        const int it = gmres_core<N, float>(f, f.Af.p, f.fb.p, f.fdinv.p, f.fx.p, f.fV.p, f.fw.p,
                                            f.fr.p, max_it - used, inner, 0.0, mres, &i0, &i1);
        used += std::max(it, 1);
        k_refine<<<unsigned((n + B - 1) / B), B, 0, f.stream>>>(n, rn, f.fx.p, x);
        spmv<N>(f, A, x, nullptr, f.w.p);
        k_residual_copy<<<unsigned((n + B - 1) / B), B, 0, f.stream>>>(n, b, f.w.p, f.r.p);
        const double* xs[1] = {f.r.p};
        dots(f, n, 1, xs, xs, &rn);
        rn = std::sqrt(rn);
        if (it == 0) break;
    }
    cd->solver_final_res = rn;
    return used;
}

// --------------------------------------------------------------- BiCGStab
__global__ void k_bicg_p(size_t n, double beta, double omega, const double* r, const double* v,
                         double* p) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = r[i] + beta * (p[i] - omega * v[i]);
}

__global__ void k_lin2(size_t n, double a, const double* x, double b, const double* y, double* out) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a * x[i] + b * y[i];
}

__global__ void k_bicg_x(size_t n, double alpha, const double* p, double omega, const double* s,
                         const double* dinv, double* x) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] += dinv[i] * (alpha * p[i] + omega * s[i]);
}

// Breakdown of BiCGStab: an inner product with the shadow residual r0 that
// vanishes against ||r0|| ||r|| (or is not finite) leaves alpha or beta
// undefined; the iteration then restarts with r0 = r (Saad 2003, 7.4).
static bool bicg_breakdown(double d, double r0n, double rn) {
    return !std::isfinite(d) || std::fabs(d) <= 1e-14 * r0n * rn;
}

static void bicg_restart(Flow& f, size_t n, double* rho, double* alpha, double* omega, double* r0n,
                         double rn) {
    OAGPU_HIP(hipMemcpyAsync(f.r0.p, f.r.p, n * sizeof(double), hipMemcpyDeviceToDevice, f.stream));
    OAGPU_HIP(hipMemsetAsync(f.pk.p, 0, n * sizeof(double), f.stream));
    OAGPU_HIP(hipMemsetAsync(f.q.p, 0, n * sizeof(double), f.stream));
    *rho = *alpha = *omega = 1.0;
    *r0n = rn;
}

template <int N>
static int bicgstab(Flow& f, int max_it, double rtol, double atol, oagpu_control_data* cd) {
    DevMesh& m = *f.m;
    const size_t n = size_t(m.nn) * N;
    const double* A = N == 3 ? f.A3.p : f.A1.p;
    const double* b = N == 3 ? f.b3.p : f.b1.p;
    double* x = N == 3 ? f.x3.p : f.x1.p;
    const unsigned B = 256, G = unsigned((n + B - 1) / B);
    OAGPU_HIP(hipMemsetAsync(x, 0, n * sizeof(double), f.stream));
    OAGPU_HIP(hipMemcpyAsync(f.r.p, b, n * sizeof(double), hipMemcpyDeviceToDevice, f.stream));
    OAGPU_HIP(hipMemcpyAsync(f.r0.p, b, n * sizeof(double), hipMemcpyDeviceToDevice, f.stream));
    OAGPU_HIP(hipMemsetAsync(f.pk.p, 0, n * sizeof(double), f.stream));
    OAGPU_HIP(hipMemsetAsync(f.q.p, 0, n * sizeof(double), f.stream));  // v
    double rn;
    {
        const double* xs[1] = {f.r.p};
        dots(f, n, 1, xs, xs, &rn);
        rn = std::sqrt(rn);
    }
    cd->solver_initial_res = rn;
    const double tol = std::max(rtol * rn, atol);
    double rho = 1, alpha = 1, omega = 1, r0n = rn;
    int it = 0;
    while (it < max_it && rn > tol) {
        double rho1;
        {
            const double* xs[1] = {f.r0.p};
            const double* ys[1] = {f.r.p};
            dots(f, n, 1, xs, ys, &rho1);
        }
        if (bicg_breakdown(rho1, r0n, rn)) {  // r0 . r ~ 0: restart from r
            bicg_restart(f, n, &rho, &alpha, &omega, &r0n, rn);
            continue;
        }
        const double beta = (rho1 / rho) * (alpha / omega);
        rho = rho1;
        k_bicg_p<<<G, B, 0, f.stream>>>(n, beta, omega, f.r.p, f.q.p, f.pk.p);
        spmv<N>(f, A, f.pk.p, f.dinv.p, f.q.p);  // v = A M^-1 p
        double r0v;
        {
            const double* xs[1] = {f.r0.p};
            const double* ys[1] = {f.q.p};
            dots(f, n, 1, xs, ys, &r0v);
        }
        if (bicg_breakdown(r0v, r0n, rn)) {  // r0 . v ~ 0: alpha undefined
            bicg_restart(f, n, &rho, &alpha, &omega, &r0n, rn);
            ++it;
            continue;
        }
        alpha = rho / r0v;
        k_lin2<<<G, B, 0, f.stream>>>(n, 1.0, f.r.p, -alpha, f.q.p, f.s.p);
        spmv<N>(f, A, f.s.p, f.dinv.p, f.t.p);  // t = A M^-1 s
        double tt_ts[2];
        {
            const double* xs[2] = {f.t.p, f.t.p};
            const double* ys[2] = {f.t.p, f.s.p};
            dots(f, n, 2, xs, ys, tt_ts);
        }
        omega = tt_ts[0] > 0 ? tt_ts[1] / tt_ts[0] : 0.0;
        k_bicg_x<<<G, B, 0, f.stream>>>(n, alpha, f.pk.p, omega, f.s.p, f.dinv.p, x);
        k_lin2<<<G, B, 0, f.stream>>>(n, 1.0, f.s.p, -omega, f.t.p, f.r.p);
        ++it;
        const double* xs[1] = {f.r.p};
        dots(f, n, 1, xs, xs, &rn);
        rn = std::sqrt(rn);
        if (omega == 0.0) break;
    }
    cd->solver_final_res = rn;
    return it;
}

// ------------------------------------------------------ AMG-BiCGStab
template <typename V>
__global__ void k_axpby_add(size_t n, double a, const V* x, double b, const V* y, V* out) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = V(double(out[i]) + a * double(x[i]) + b * double(y[i]));
}

template <typename V>
__global__ void k_bicg_p_t(size_t n, double beta, double omega, const V* r, const V* v, V* p) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = V(double(r[i]) + beta * (double(p[i]) - omega * double(v[i])));
}

template <typename V>
__global__ void k_lin2_t(size_t n, double a, const V* x, double b, const V* y, V* out) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = V(a * double(x[i]) + b * double(y[i]));
}

template <typename V>
struct BicgVecs {
    V *r, *r0, *p, *v, *s, *t, *ph, *sh;  // ph = M^-1 p, sh = M^-1 s
};

// BiCGStab (Saad Alg. 7.7) right-preconditioned by one AMG V-cycle on
// values and vectors of type V, scalars and dot products in double: the p'
// systems of flows with static-pressure outlets, or on distorted meshes, are
// not symmetric, and AMG-CG stalls on them.  Af (V = double only): the fine
// level's values in FP32, read by the V-cycle's fine-level kernels.  Starts
// from x = 0; a vanishing inner product with the shadow residual restarts
// it from the current residual.
template <typename V>
static int bicgstab_amg_core(Flow& f, const V* A, const float* Af, const V* b, const V* dinv, V* x,
                             const BicgVecs<V>& w, int max_it, double rtol, double atol,
                             double* r0out, double* r1out, bool sa = false) {
    DevMesh& m = *f.m;
    const size_t n = size_t(m.nn);
    const unsigned B = 256, G = unsigned((n + B - 1) / B);
    k_clear_flag<<<1, 1, 0, f.stream>>>(f.flag.p);  // the V-cycle kernels honour the CG stop flag
    OAGPU_HIP(hipMemsetAsync(x, 0, n * sizeof(V), f.stream));
    OAGPU_HIP(hipMemcpyAsync(w.r, b, n * sizeof(V), hipMemcpyDeviceToDevice, f.stream));
    auto restart = [&](double* rho, double* alpha, double* omega, double* r0n, double rn) {
        OAGPU_HIP(hipMemcpyAsync(w.r0, w.r, n * sizeof(V), hipMemcpyDeviceToDevice, f.stream));
        OAGPU_HIP(hipMemsetAsync(w.p, 0, n * sizeof(V), f.stream));
        OAGPU_HIP(hipMemsetAsync(w.v, 0, n * sizeof(V), f.stream));
        *rho = *alpha = *omega = 1.0;
        *r0n = rn;
    };
    auto dot = [&](const V* a, const V* c) {
        double d;
        const V* xs[1] = {a};
        dots_t<V>(f, n, 1, xs, c, &d);
        return d;
    };
    auto apply = [&](const V* r, V* z) {
        if (sa)
            sa_vcycle<V>(f, A, dinv, r, z, true);
        else
            amg_vcycle<V>(f, A, dinv, r, z, f.part.p + 8 * kRedGrid, nullptr, true, false, nullptr,
                          nullptr, Af);
        OAGPU_HIP(hipGetLastError());
    };
    // r . r and r0 . r in one pass over r (r0 = r at a start or restart)
    auto norms = [&](double* rr, double* r0r) {
        double d[2];
        const V* xs[2] = {w.r, w.r0};
        dots_t<V>(f, n, 2, xs, w.r, d);
        *rr = d[0];
        *r0r = d[1];
    };
    double rn = std::sqrt(dot(w.r, w.r));
    *r0out = rn;
    const double tol = std::max(rtol * rn, atol);
    double rho = 1, alpha = 1, omega = 1, r0n = rn;
    restart(&rho, &alpha, &omega, &r0n, rn);
    double rho1 = rn * rn;  // r0 . r
    int it = 0;
    while (it < max_it && rn > tol) {
        if (bicg_breakdown(rho1, r0n, rn)) {  // r0 . r ~ 0: restart from r
            restart(&rho, &alpha, &omega, &r0n, rn);
            rho1 = rn * rn;
            continue;
        }
        const double beta = (rho1 / rho) * (alpha / omega);
        rho = rho1;
        k_bicg_p_t<V><<<G, B, 0, f.stream>>>(n, beta, omega, w.r, w.v, w.p);
        apply(w.p, w.ph);
        spmv<1, V>(f, A, w.ph, nullptr, w.v);  // v = A M^-1 p
        const double r0v = dot(w.r0, w.v);
        if (bicg_breakdown(r0v, r0n, rn)) {  // r0 . v ~ 0: alpha undefined
            restart(&rho, &alpha, &omega, &r0n, rn);
            rho1 = rn * rn;
            ++it;
            continue;
        }
        alpha = rho / r0v;
        k_lin2_t<V><<<G, B, 0, f.stream>>>(n, 1.0, w.r, -alpha, w.v, w.s);
        apply(w.s, w.sh);
        spmv<1, V>(f, A, w.sh, nullptr, w.t);  // t = A M^-1 s
        double tt_ts[2];
        {
            const V* xs[2] = {w.t, w.s};  // t . t and s . t in one pass over t
            dots_t<V>(f, n, 2, xs, w.t, tt_ts);
        }
        omega = tt_ts[0] > 0 ? tt_ts[1] / tt_ts[0] : 0.0;
        k_axpby_add<V><<<G, B, 0, f.stream>>>(n, alpha, w.ph, omega, w.sh, x);
        k_lin2_t<V><<<G, B, 0, f.stream>>>(n, 1.0, w.s, -omega, w.t, w.r);
        ++it;
        double rr;
        norms(&rr, &rho1);
        rn = std::sqrt(rr);
        if (omega == 0.0) break;
    }
    *r1out = rn;
    return it;
}

// The mixed form's p' solve: FP32 AMG-BiCGStab on A d = r / ||r|| with the
// FP32 hierarchy, x += ||r|| d and r = b - A x in FP64, until ||r|| <=
// max(rtol ||b||, atol) or the iteration budget is spent (kMixedBicgInner =
// false: FP64 BiCGStab whose V-cycle reads the fine level's values in FP32).
constexpr bool kMixedBicgInner = true;

static int bicgstab_amg_mixed(Flow& f, int max_it, double rtol, double atol, oagpu_control_data* cd,
                              bool sa) {
    DevMesh& m = *f.m;
    const size_t n = size_t(m.nn);
    const unsigned B = 256;
    const double* A = f.A1.p;
    const double* b = f.b1.p;
    double* x = f.x1.p;
    if (f.Af.n < m.nsell) f.Af.alloc(m.nsell);
    k_to_float<<<unsigned((m.nsell + B - 1) / B), B, 0, f.stream>>>(m.nsell, 1.0, A, f.Af.p);
    if (!kMixedBicgInner && !sa) {
        amg_setup<double>(f, A, f.dinv.p);
        BicgVecs<double> w{f.r.p, f.r0.p, f.pk.p, f.q.p, f.s.p, f.t.p, f.z.p, f.w.p};
        return bicgstab_amg_core<double>(f, A, f.Af.p, b, f.dinv.p, x, w, max_it, rtol, atol,
                                         &cd->solver_initial_res, &cd->solver_final_res);
    }
    for (auto* v : {&f.fb, &f.fx, &f.fr, &f.fz, &f.fp, &f.fq, &f.fdinv, &f.fr0, &f.fs, &f.ft, &f.fsh})
        if (v->n < n) v->alloc(n);
    k_to_float<<<unsigned((n + B - 1) / B), B, 0, f.stream>>>(n, 1.0, f.dinv.p, f.fdinv.p);
    if (sa)
        sa_setup_f32(f, f.Af.p);  // FP32 levels for the FP32 inner solve
    else
        amg_setup<float>(f, f.Af.p, f.fdinv.p);
    OAGPU_HIP(hipMemsetAsync(x, 0, n * sizeof(double), f.stream));
    OAGPU_HIP(hipMemcpyAsync(f.r.p, b, n * sizeof(double), hipMemcpyDeviceToDevice, f.stream));
    double rn;
    {
        const double* xs[1] = {f.r.p};
        dots(f, n, 1, xs, xs, &rn);
        rn = std::sqrt(rn);
    }
    cd->solver_initial_res = rn;
    const double tol = std::max(rtol * rn, atol);
    BicgVecs<float> w{f.fr.p, f.fr0.p, f.fp.p, f.fq.p, f.fs.p, f.ft.p, f.fz.p, f.fsh.p};
    int used = 0;
    while (used < max_it && rn > tol) {
        const double inner = std::max(tol / rn, 1.0e-4);
        k_to_float<<<unsigned((n + B - 1) / B), B, 0, f.stream>>>(n, 1.0 / rn, f.r.p, f.fb.p);
        double i0, i1;
        const int it = bicgstab_amg_core<float>(f, f.Af.p, nullptr, f.fb.p, f.fdinv.p, f.fx.p, w,
                                                max_it - used, inner, 0.0, &i0, &i1, sa);
        used += std::max(it, 1);
        k_refine<<<unsigned((n + B - 1) / B), B, 0, f.stream>>>(n, rn, f.fx.p, x);
        spmv<1>(f, A, x, nullptr, f.w.p);
        k_residual_copy<<<unsigned((n + B - 1) / B), B, 0, f.stream>>>(n, b, f.w.p, f.r.p);
        const double* xs[1] = {f.r.p};
        dots(f, n, 1, xs, xs, &rn);
        rn = std::sqrt(rn);
        if (it == 0) break;
    }
    cd->solver_final_res = rn;
    return used;
}

// sa: the smoothed-aggregation hierarchy (OAGPU_BICGSTAB_SA), else the
// unsmoothed one (OAGPU_BICGSTAB_AMG)
static int bicgstab_amg_solve(Flow& f, int max_it, double rtol, double atol, oagpu_control_data* cd,
                              bool sa) {
    if (sa) sa_build(f);
    else amg_build(f);
    if ((sa ? f.sa.lv.size() : f.amg.lv.size()) <= 1) return bicgstab<1>(f, max_it, rtol, atol, cd);  // tiny: Jacobi
    if (f.prm.precision == OAGPU_MIXED) return bicgstab_amg_mixed(f, max_it, rtol, atol, cd, sa);
    if (sa) sa_setup(f, f.A1.p, false);
    else amg_setup<double>(f, f.A1.p, f.dinv.p);
    BicgVecs<double> w{f.r.p, f.r0.p, f.pk.p, f.q.p, f.s.p, f.t.p, f.z.p, f.w.p};
    return bicgstab_amg_core<double>(f, f.A1.p, nullptr, f.b1.p, f.dinv.p, f.x1.p, w, max_it, rtol,
                                     atol, &cd->solver_initial_res, &cd->solver_final_res, sa);
}

// SpMV block-size candidates (64, 128, 256 rows per workgroup) timed on the
// flow's system N (FP32: its values rounded); the fastest is kept in
// f.tune.spmv_block.  y = A x into the Krylov work vector.
void time_spmv_variants(Flow& f, int N, bool f32, double* best_ms, double* default_ms) {
    DevMesh& m = *f.m;
    const size_t n = size_t(m.nn) * N, nv = m.nsell * N * N;
    const double* A = N == 3 ? f.A3.p : f.A1.p;
    const double* x = N == 3 ? f.x3.p : f.x1.p;
    const unsigned Bc = 256;
    if (f32) {
        if (f.Af.n < nv) f.Af.alloc(nv);
        for (auto* v : {&f.fx, &f.fw})
            if (v->n < n) v->alloc(n);
        k_to_float<<<unsigned((nv + Bc - 1) / Bc), Bc, 0, f.stream>>>(nv, 1.0, A, f.Af.p);
        k_to_float<<<unsigned((n + Bc - 1) / Bc), Bc, 0, f.stream>>>(n, 1.0, x, f.fx.p);
    }
    hipEvent_t e0, e1;
    OAGPU_HIP(hipEventCreate(&e0));
    OAGPU_HIP(hipEventCreate(&e1));
    const unsigned cand[3] = {64, 128, 256};
    unsigned best = 256;
    *best_ms = 1e30;
    for (unsigned B : cand) {
        f.tune.spmv_block[N == 3][f32] = B;
        float ms = 0.f;
        for (int rep = 0; rep < 2; ++rep) {  // first pass warms the caches
            OAGPU_HIP(hipEventRecord(e0, f.stream));
            for (int k = 0; k < 5; ++k) {
                if (f32) {
                    if (N == 3) spmv<3, float>(f, f.Af.p, f.fx.p, nullptr, f.fw.p);
                    else spmv<1, float>(f, f.Af.p, f.fx.p, nullptr, f.fw.p);
                } else {
                    if (N == 3) spmv<3>(f, A, x, nullptr, f.w.p);
                    else spmv<1>(f, A, x, nullptr, f.w.p);
                }
            }
            OAGPU_HIP(hipEventRecord(e1, f.stream));
            OAGPU_HIP(hipEventSynchronize(e1));
            OAGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
        }
        const double per = double(ms) / 5.0;
        if (B == 256) *default_ms = per;
        if (per < *best_ms) {
            *best_ms = per;
            best = B;
        }
    }
    f.tune.spmv_block[N == 3][f32] = best;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

int krylov_solve(Flow& f, int N, int method, int max_it, double rtol, double atol, int restart,
                 oagpu_control_data* cd) {
    DevMesh& m = *f.m;
    OAGPU_REQUIRE(N == 1 || N == 3, "block size must be 1 or 3");
    const unsigned B = 256;
    if (N == 3)
        k_jacobi_inv<3><<<(m.nn + B - 1) / B, B, 0, f.stream>>>(m.nn, m.slice_ptr.p, m.slice_w.p,
                                                                 m.diag_off.p, f.A3.p, f.dinv.p);
    else
        k_jacobi_inv<1><<<(m.nn + B - 1) / B, B, 0, f.stream>>>(m.nn, m.slice_ptr.p, m.slice_w.p,
                                                                 m.diag_off.p, f.A1.p, f.dinv.p);
    OAGPU_HIP(hipGetLastError());
    int it = 0;
    if (method == OAGPU_GMRES) {
        if (f.prm.precision == OAGPU_MIXED)
            it = N == 3 ? gmres_mixed<3>(f, max_it, rtol, atol, restart, cd)
                        : gmres_mixed<1>(f, max_it, rtol, atol, restart, cd);
        else
            it = N == 3 ? gmres<3>(f, max_it, rtol, atol, restart, cd)
                        : gmres<1>(f, max_it, rtol, atol, restart, cd);
    }
    else if (method == OAGPU_CG || method == OAGPU_CG_AMG || method == OAGPU_CG_SA) {
        const bool amg = method != OAGPU_CG, sa = method == OAGPU_CG_SA;
        OAGPU_REQUIRE(!amg || N == 1, "AMG preconditioning is for scalar systems (N = 1)");
        if (f.prm.precision == OAGPU_MIXED)
            it = N == 3 ? cg_mixed<3>(f, max_it, rtol, atol, cd, false)
                        : cg_mixed<1>(f, max_it, rtol, atol, cd, amg, sa);
        else
            it = N == 3 ? cg<3>(f, max_it, rtol, atol, cd, false)
                        : cg<1>(f, max_it, rtol, atol, cd, amg, sa);
    }
    else if (method == OAGPU_BICGSTAB)
        it = N == 3 ? bicgstab<3>(f, max_it, rtol, atol, cd) : bicgstab<1>(f, max_it, rtol, atol, cd);
    else if (method == OAGPU_BICGSTAB_AMG || method == OAGPU_BICGSTAB_SA) {
        OAGPU_REQUIRE(N == 1, "AMG preconditioning is for scalar systems (N = 1)");
        it = bicgstab_amg_solve(f, max_it, rtol, atol, cd, method == OAGPU_BICGSTAB_SA);
    } else
        throw Error(OAGPU_EINVAL, "unknown Krylov method");
    cd->n_iterations = it;
    if (!std::isfinite(cd->solver_final_res) || !std::isfinite(cd->solver_initial_res))
        throw Error(OAGPU_ENONFINITE, "the linear solve produced a non-finite residual");
    return it;
}

}  // namespace oagpu
