/*
 * krylov.c -- oracle Krylov solvers for the case's block-CRS systems.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference hands A x = b to liblinsolve (linearSystem.h:416-462), a
 * submodule absent from the reference tree; its PETSc/Trilinos/HYPRE
 * iterations are not restated.  These are textbook Jacobi-preconditioned
 * GMRES(m) (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed.,
 * Alg. 9.5, right preconditioning, classical Gram-Schmidt applied twice),
 * PCG (Alg. 9.1) and BiCGStab (Alg. 7.7, right preconditioned), in the same
 * form the HIP path runs, so that the two can be compared iterate by iterate.
 */
#include "case_internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

struct or_case; /* opaque; accessed through the public getters */

typedef struct {
    int n, N, nnz;
    const int *rp, *cols, *doff;
    const double *A;
} sys_t;

static void spmv(const sys_t *s, const double *x, double *y)
{
    const int N = s->N, NN = N * N;
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
    for (int i = 0; i < s->n; ++i) {
        double acc[3] = {0, 0, 0};
        for (int k = s->rp[i]; k < s->rp[i + 1]; ++k) {
            const double *blk = &s->A[(size_t)NN * k];
            const double *xv = &x[N * s->cols[k]];
            for (int m = 0; m < N; ++m)
                for (int q = 0; q < N; ++q) acc[m] += blk[N * m + q] * xv[q];
        }
        for (int m = 0; m < N; ++m) y[N * i + m] = acc[m];
    }
}

static double dot(int n, const double *a, const double *b)
{
    double s = 0;
#pragma omp parallel for num_threads(or_nthreads) reduction(+ : s) schedule(static)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static void jacobi_inv(const sys_t *s, double *dinv)
{
    const int N = s->N, NN = N * N;
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
    for (int i = 0; i < s->n; ++i) {
        const double *d = &s->A[(size_t)NN * (s->rp[i] + s->doff[i])];
        for (int m = 0; m < N; ++m) dinv[N * i + m] = 1.0 / d[(N + 1) * m];
    }
}

static int gmres(const sys_t *s, const double *b, double *x, int max_it, double rtol,
                 double atol, int m, double *res)
{
    const int n = s->n * s->N;
    double *V = (double *)calloc((size_t)n * (m + 1), sizeof(double));
    double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
    double *cs = (double *)calloc((size_t)m, sizeof(double));
    double *sn = (double *)calloc((size_t)m, sizeof(double));
    double *g = (double *)calloc((size_t)m + 1, sizeof(double));
    double *y = (double *)calloc((size_t)m, sizeof(double));
    double *h = (double *)calloc((size_t)m + 1, sizeof(double));
    double *w = (double *)calloc((size_t)n, sizeof(double));
    double *z = (double *)calloc((size_t)n, sizeof(double));
    double *dinv = (double *)calloc((size_t)n, sizeof(double));
    jacobi_inv(s, dinv);
    spmv(s, x, w);
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
    for (int i = 0; i < n; ++i) w[i] = b[i] - w[i];
    double beta = sqrt(dot(n, w, w));
    const double r0 = beta;
    const double tol = fmax(rtol * r0, atol);
    res[0] = r0;
    int it = 0;
    while (it < max_it && beta > tol) {
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
        for (int i = 0; i < n; ++i) V[i] = w[i] / beta;
        memset(g, 0, sizeof(double) * (m + 1));
        g[0] = beta;
        int jend = -1;
        for (int j = 0; j < m && it < max_it; ++j) {
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
            for (int i = 0; i < n; ++i) z[i] = dinv[i] * V[(size_t)n * j + i];
            spmv(s, z, w);
            /* classical Gram-Schmidt, twice */
            for (int k = 0; k <= j; ++k) h[k] = 0;
            for (int pass = 0; pass < 2; ++pass) {
                double hk[64];
                for (int k = 0; k <= j; ++k) hk[k] = dot(n, &V[(size_t)n * k], w);
                for (int k = 0; k <= j; ++k) {
                    const double *vk = &V[(size_t)n * k];
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
                    for (int i = 0; i < n; ++i) w[i] -= hk[k] * vk[i];
                    h[k] += hk[k];
                }
            }
            const double hn = sqrt(dot(n, w, w));
            for (int k = 0; k <= j; ++k) H[(size_t)m * k + j] = h[k];
            H[(size_t)m * (j + 1) + j] = hn;
            if (hn > 0)
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
                for (int i = 0; i < n; ++i) V[(size_t)n * (j + 1) + i] = w[i] / hn;
            /* previous rotations on column j */
            for (int k = 0; k < j; ++k) {
                const double a = H[(size_t)m * k + j], bb = H[(size_t)m * (k + 1) + j];
                H[(size_t)m * k + j] = cs[k] * a + sn[k] * bb;
                H[(size_t)m * (k + 1) + j] = -sn[k] * a + cs[k] * bb;
            }
            const double a = H[(size_t)m * j + j], bb = H[(size_t)m * (j + 1) + j];
            const double rr = sqrt(a * a + bb * bb);
            cs[j] = rr > 0 ? a / rr : 1.0;
            sn[j] = rr > 0 ? bb / rr : 0.0;
            H[(size_t)m * j + j] = rr;
            H[(size_t)m * (j + 1) + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            ++it;
            jend = j;
            if (fabs(g[j + 1]) <= tol || hn == 0.0) break;
        }
        /* back substitution and update x += M^-1 V y */
        for (int k = jend; k >= 0; --k) {
            double t = g[k];
            for (int q = k + 1; q <= jend; ++q) t -= H[(size_t)m * k + q] * y[q];
            y[k] = t / H[(size_t)m * k + k];
        }
        memset(z, 0, sizeof(double) * n);
        for (int k = 0; k <= jend; ++k) {
            const double *vk = &V[(size_t)n * k];
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
            for (int i = 0; i < n; ++i) z[i] += y[k] * vk[i];
        }
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
        for (int i = 0; i < n; ++i) x[i] += dinv[i] * z[i];
        spmv(s, x, w);
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
        for (int i = 0; i < n; ++i) w[i] = b[i] - w[i];
        beta = sqrt(dot(n, w, w));
    }
    res[1] = beta;
    free(V); free(H); free(cs); free(sn); free(g); free(y); free(h); free(w); free(z); free(dinv);
    return it;
}

static int pcg(const sys_t *s, const double *b, double *x, int max_it, double rtol,
               double atol, double *res)
{
    const int n = s->n * s->N;
    double *r = (double *)calloc((size_t)n, sizeof(double));
    double *z = (double *)calloc((size_t)n, sizeof(double));
    double *p = (double *)calloc((size_t)n, sizeof(double));
    double *q = (double *)calloc((size_t)n, sizeof(double));
    double *dinv = (double *)calloc((size_t)n, sizeof(double));
    jacobi_inv(s, dinv);
    spmv(s, x, q);
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
    for (int i = 0; i < n; ++i) {
        r[i] = b[i] - q[i];
        z[i] = dinv[i] * r[i];
        p[i] = z[i];
    }
    double rz = dot(n, r, z);
    double rn = sqrt(dot(n, r, r));
    const double tol = fmax(rtol * rn, atol);
    res[0] = rn;
    int it = 0;
    while (it < max_it && rn > tol) {
        spmv(s, p, q);
        const double alpha = rz / dot(n, p, q);
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
        }
        ++it;
        rn = sqrt(dot(n, r, r));
        if (rn <= tol) break;
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
        for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
        const double rz1 = dot(n, r, z);
        const double beta = rz1 / rz;
        rz = rz1;
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    res[1] = rn;
    free(r); free(z); free(p); free(q); free(dinv);
    return it;
}

static int bicgstab(const sys_t *s, const double *b, double *x, int max_it, double rtol,
                    double atol, double *res)
{
    const int n = s->n * s->N;
    double *r = (double *)calloc((size_t)n, sizeof(double));
    double *r0 = (double *)calloc((size_t)n, sizeof(double));
    double *p = (double *)calloc((size_t)n, sizeof(double));
    double *v = (double *)calloc((size_t)n, sizeof(double));
    double *sv = (double *)calloc((size_t)n, sizeof(double));
    double *t = (double *)calloc((size_t)n, sizeof(double));
    double *ph = (double *)calloc((size_t)n, sizeof(double));
    double *sh = (double *)calloc((size_t)n, sizeof(double));
    double *dinv = (double *)calloc((size_t)n, sizeof(double));
    jacobi_inv(s, dinv);
    spmv(s, x, v);
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
    for (int i = 0; i < n; ++i) {
        r[i] = b[i] - v[i];
        r0[i] = r[i];
        p[i] = 0;
        v[i] = 0;
    }
    double rho = 1, alpha = 1, omega = 1;
    double rn = sqrt(dot(n, r, r));
    const double tol = fmax(rtol * rn, atol);
    res[0] = rn;
    int it = 0;
    while (it < max_it && rn > tol) {
        const double rho1 = dot(n, r0, r);
        const double beta = (rho1 / rho) * (alpha / omega);
        rho = rho1;
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
        for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
        for (int i = 0; i < n; ++i) ph[i] = dinv[i] * p[i];
        spmv(s, ph, v);
        alpha = rho / dot(n, r0, v);
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
        for (int i = 0; i < n; ++i) sv[i] = r[i] - alpha * v[i];
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
        for (int i = 0; i < n; ++i) sh[i] = dinv[i] * sv[i];
        spmv(s, sh, t);
        const double tt = dot(n, t, t);
        omega = tt > 0 ? dot(n, t, sv) / tt : 0.0;
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * ph[i] + omega * sh[i];
            r[i] = sv[i] - omega * t[i];
        }
        ++it;
        rn = sqrt(dot(n, r, r));
        if (omega == 0.0) break;
    }
    res[1] = rn;
    free(r); free(r0); free(p); free(v); free(sv); free(t); free(ph); free(sh); free(dinv);
    return it;
}

int or_solve(or_case *c, int N, int kind, int max_it, double rtol, double atol,
             int restart, double *res)
{
    sys_t s;
    s.N = N;
    s.rp = or_case_row_ptr(c);
    s.cols = or_case_cols(c);
    s.doff = or_case_diag_off(c);
    s.nnz = or_case_nnz(c);
    s.A = or_case_vals(c, N);
    s.n = or_case_nn(c);
    const double *b = or_case_rhs(c, N);
    double *x = or_case_x(c, N);
    if (restart > 60) restart = 60;
    if (kind == 0) return gmres(&s, b, x, max_it, rtol, atol, restart, res);
    if (kind == 1) return pcg(&s, b, x, max_it, rtol, atol, res);
    return bicgstab(&s, b, x, max_it, rtol, atol, res);
}
