/*
 * graph.c -- oracle restatement of the CRS node graph and of the
 * applyCoeff_ scatter.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   nodeGraph::buildGraph_      src/mesh/graph/nodeGraph.cpp:29-392
 *   assembler<N>::applyCoeff_   src/assemble/assembler.h:311-453
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

void or_free(void *p) { free(p); }

static int cmp_int(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

int or_graph_build(int nn, int ne, const int *conn, const int *etype, int reduced,
                   int *row_ptr, int **cols_out)
{
    /* candidate (row, col) insertions, counted then filled; std::set per row
     * in the reference (nodeGraph.cpp:36) == sort + unique per row here.
     * Elements of any topology: conn [ne][8], npe leading entries used. */
    int *cnt = (int *)calloc((size_t)nn + 1, sizeof(int));
    if (!cnt) return -1;
    for (int e = 0; e < ne; ++e) {
        const int t = etype ? etype[e] : OR_HEX8;
        const int *en = &conn[8 * e], npe = or_topo_npe(t), nscs = or_topo_nscs(t);
        const int *lr = t == OR_HEX8 ? or_hex_lrscv() : or_topo_lrscv(t);
        if (reduced) {
            for (int ip = 0; ip < nscs; ++ip) {
                cnt[en[lr[2 * ip]]] += 2;
                cnt[en[lr[2 * ip + 1]]] += 2;
            }
        } else {
            for (int a = 0; a < npe; ++a) cnt[en[a]] += npe;
        }
    }
    size_t *start = (size_t *)malloc(((size_t)nn + 1) * sizeof(size_t));
    if (!start) { free(cnt); return -1; }
    start[0] = 0;
    for (int i = 0; i < nn; ++i) start[i + 1] = start[i] + (size_t)cnt[i];
    int *cand = (int *)malloc(start[nn] * sizeof(int) + 1);
    if (!cand) { free(cnt); free(start); return -1; }
    memset(cnt, 0, sizeof(int) * ((size_t)nn + 1));
    for (int e = 0; e < ne; ++e) {
        const int t = etype ? etype[e] : OR_HEX8;
        const int *en = &conn[8 * e], npe = or_topo_npe(t), nscs = or_topo_nscs(t);
        const int *lr = t == OR_HEX8 ? or_hex_lrscv() : or_topo_lrscv(t);
        if (reduced) {
            for (int ip = 0; ip < nscs; ++ip) {
                const int L = en[lr[2 * ip]], R = en[lr[2 * ip + 1]];
                cand[start[L] + cnt[L]++] = L;
                cand[start[L] + cnt[L]++] = R;
                cand[start[R] + cnt[R]++] = L;
                cand[start[R] + cnt[R]++] = R;
            }
        } else {
            for (int a = 0; a < npe; ++a) {
                const int r = en[a];
                for (int c = 0; c < npe; ++c) cand[start[r] + cnt[r]++] = en[c];
            }
        }
    }
    row_ptr[0] = 0;
    size_t nnz = 0;
    for (int i = 0; i < nn; ++i) {
        int *rc = &cand[start[i]];
        const int n = cnt[i];
        qsort(rc, (size_t)n, sizeof(int), cmp_int);
        int u = 0;
        for (int k = 0; k < n; ++k)
            if (u == 0 || rc[k] != rc[u - 1]) rc[u++] = rc[k];
        cnt[i] = u;
        nnz += (size_t)u;
        row_ptr[i + 1] = (int)nnz;
    }
    int *cols = (int *)malloc(nnz * sizeof(int) + 1);
    if (!cols) { free(cnt); free(start); free(cand); return -1; }
    for (int i = 0; i < nn; ++i)
        memcpy(&cols[row_ptr[i]], &cand[start[i]], (size_t)cnt[i] * sizeof(int));
    free(cnt);
    free(start);
    free(cand);
    *cols_out = cols;
    return 0;
}

int or_graph_find(const int *row_ptr, const int *cols, int row, int col)
{
    for (int k = row_ptr[row]; k < row_ptr[row + 1]; ++k)
        if (cols[k] == col) return k - row_ptr[row];
    return -1;
}

void or_apply_coeff(int N, int n, const int *nodes, const double *lhs,
                    const double *rhs, int n_owned, const int *row_ptr,
                    const int *cols, const int *diag_off, double *vals, double *b)
{
    /* local column ids sorted by matrix column id (assembler.h:356-361) */
    int perm[32];
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = 1; i < n; ++i) { /* insertion sort == std::sort on <= 8 ids */
        const int p = perm[i];
        int j = i - 1;
        while (j >= 0 && nodes[perm[j]] > nodes[p]) {
            perm[j + 1] = perm[j];
            --j;
        }
        perm[j + 1] = p;
    }
    const int BS = N, BS2 = N * N;
    for (int r = 0; r < n; ++r) {
        const int cur = perm[r];
        const int row = nodes[cur];
        const double *lhsCur = &lhs[(size_t)BS * BS * cur * n];
        const double *rhsCur = &rhs[BS * cur];
        if (row >= n_owned) continue; /* owned rows only (:380) */
        double *rowVals = &vals[(size_t)BS2 * row_ptr[row]];
        const int *rowCols = &cols[row_ptr[row]];
        const int length = row_ptr[row + 1] - row_ptr[row];
        int offset = 0, offsetOld = 0;
        for (int j = 0; j < n; ++j) {
            const int pidx = perm[j];
            const int col = nodes[pidx];
            /* sorted merge scan (:388-398) */
            while (offset < length && rowCols[offset] != col) ++offset;
            if (offset < length) {
                for (int m = 0; m < BS; ++m)
                    for (int k = 0; k < BS; ++k)
                        rowVals[BS2 * offset + BS * m + k] += lhsCur[BS * n * m + BS * pidx + k];
                offsetOld = offset;
            } else {
                /* reduced stencil: unfound column lumped into the diagonal
                 * block (:423-445) */
                double *diag = &vals[(size_t)BS2 * (row_ptr[row] + diag_off[row])];
                for (int m = 0; m < BS; ++m)
                    for (int k = 0; k < BS; ++k)
                        diag[BS * m + k] += lhsCur[BS * n * m + BS * pidx + k];
                offset = offsetOld;
            }
        }
        for (int k = 0; k < BS; ++k) b[BS * row + k] += rhsCur[k];
    }
}
