/*
 * case.c -- oracle restatement of the segregated pressure-velocity outer
 * iteration on hex8 meshes.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restated (file:line in the reference):
 *   momentum interior     assemble/flow/segregatedFlow/navierStokesAssembler/
 *                         navierStokesAssemblerElemTerms.cpp:15-441
 *   momentum node terms   navierStokesAssemblerNodeTerms.cpp:16-149
 *   no-slip wall          navierStokesAssemblerElemBoundaryConditions.cpp:2659-2884
 *                         (macro IP_ZERO_NORMAL_STRESS_FIXED_VEL__ :592-654)
 *   relaxation            assemble/phiAssembler/phiAssembler.h:743-753, 966-1152
 *                         navierStokesAssembler.cpp:234-239
 *   du coefficients       navierStokesAssembler.cpp:84-232
 *   p' interior           pressureCorrectionAssemblerElemTerms.cpp:17-616
 *   pressure reference    pressureCorrectionAssembler.cpp:19-63
 *   mass flow update      model/flow/flowModel.cpp:5272-5688
 *   mass flow init        flowModel.cpp:3303-3470
 *   mass divergence       flowModel.cpp:13541-13775
 *   Green-Gauss gradient  field/nodeField/nodeField.hpp:3253-4950
 *   field correction      equation/equation.h:204-286
 *   velocity correction   equation/flow/segregatedFlowEquations.cpp:133-279
 *   scales                nodeField.hpp:5993-6015, flowModel.cpp:117-128
 *   dual nodal volume     mesh/meshGeometry.cpp:1277
 *   exposed area vectors  mesh/meshGeometry.cpp:1441-1528
 */
#include "case_internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static void *xcalloc(size_t n, size_t s)
{
    void *p = calloc(n ? n : 1, s);
    return p;
}

int or_nthreads = 1;

void or_set_threads(int n) { or_nthreads = n > 0 ? n : 1; }

/* Greedy colouring of the elements by shared nodes (with one thread: one
 * class, element order, so single-threaded results keep their summation
 * order). */
void or_color_elements(or_case *c)
{
    const int want = or_nthreads > 1 ? 2 : 1;
    if (c->color_ptr && c->color_threads == want) return;
    free(c->color_ptr);
    free(c->color_elem);
    c->color_threads = want;
    c->color_elem = (int *)xcalloc((size_t)c->ne, sizeof(int));
    if (want == 1) {
        c->ncolor = 1;
        c->color_ptr = (int *)xcalloc(2, sizeof(int));
        c->color_ptr[1] = c->ne;
        for (int e = 0; e < c->ne; ++e) c->color_elem[e] = e;
        return;
    }
    int *col = (int *)xcalloc((size_t)c->ne, sizeof(int));
    unsigned long long *used = (unsigned long long *)xcalloc((size_t)c->nn, sizeof(unsigned long long));
    int ncol = 0, base = 0;
    /* classes of at most 64 colours per sweep; elements that find none wait
     * for the next sweep */
    int left = c->ne;
    for (int e = 0; e < c->ne; ++e) col[e] = -1;
    while (left > 0) {
        memset(used, 0, sizeof(unsigned long long) * (size_t)c->nn);
        int top = 0;
        for (int e = 0; e < c->ne; ++e) {
            if (col[e] >= 0) continue;
            const int npe = or_topo_npe(or_etype(c, e));
            unsigned long long m = 0;
            for (int a = 0; a < npe; ++a) m |= used[c->conn[8 * e + a]];
            if (m == ~0ULL) continue;
            int k = 0;
            while (m & (1ULL << k)) ++k;
            for (int a = 0; a < npe; ++a) used[c->conn[8 * e + a]] |= 1ULL << k;
            col[e] = base + k;
            if (k + 1 > top) top = k + 1;
            --left;
        }
        base += top;
        ncol = base;
    }
    c->ncolor = ncol;
    c->color_ptr = (int *)xcalloc((size_t)ncol + 1, sizeof(int));
    for (int e = 0; e < c->ne; ++e) c->color_ptr[col[e] + 1]++;
    for (int k = 0; k < ncol; ++k) c->color_ptr[k + 1] += c->color_ptr[k];
    int *fill = (int *)xcalloc((size_t)ncol, sizeof(int));
    for (int e = 0; e < c->ne; ++e) c->color_elem[c->color_ptr[col[e]] + fill[col[e]]++] = e;
    free(fill);
    free(col);
    free(used);
}

or_case *or_case_create(int nn, int ne, const double *xyz, const int *conn,
                        const int *etype, int nbf, const int *bf_elem, const int *bf_ord,
                        const int *bf_bc, int nbc, const double *bc_U, const or_params *prm)
{
    if (etype) {
        for (int e = 0; e < ne; ++e)
            if (etype[e] < OR_HEX8 || etype[e] > OR_PYR5) return NULL;
    }
    for (int f = 0; f < nbf; ++f) {
        if (bf_elem[f] < 0 || bf_elem[f] >= ne) return NULL;
        const int t = etype ? etype[bf_elem[f]] : OR_HEX8;
        if (bf_ord[f] < 0 || bf_ord[f] >= or_topo_nsides(t)) return NULL;
    }
    or_case *c = (or_case *)xcalloc(1, sizeof(or_case));
    c->nn = nn;
    c->ne = ne;
    c->nbf = nbf;
    c->nbc = nbc;
    c->prm = *prm;
    c->xyz = (double *)xcalloc((size_t)nn * 3, sizeof(double));
    memcpy(c->xyz, xyz, sizeof(double) * 3 * (size_t)nn);
    c->conn = (int *)xcalloc((size_t)ne * 8, sizeof(int));
    memcpy(c->conn, conn, sizeof(int) * 8 * (size_t)ne);
    if (etype) {
        c->etype = (int *)xcalloc((size_t)ne, sizeof(int));
        memcpy(c->etype, etype, sizeof(int) * (size_t)ne);
    }
    c->bf_elem = (int *)xcalloc((size_t)nbf, sizeof(int));
    c->bf_ord = (int *)xcalloc((size_t)nbf, sizeof(int));
    c->bf_bc = (int *)xcalloc((size_t)nbf, sizeof(int));
    if (nbf) {
        memcpy(c->bf_elem, bf_elem, sizeof(int) * (size_t)nbf);
        memcpy(c->bf_ord, bf_ord, sizeof(int) * (size_t)nbf);
        memcpy(c->bf_bc, bf_bc, sizeof(int) * (size_t)nbf);
    }
    c->bc_U = (double *)xcalloc((size_t)nbc * 3, sizeof(double));
    if (nbc) memcpy(c->bc_U, bc_U, sizeof(double) * 3 * (size_t)nbc);

    c->row_ptr = (int *)xcalloc((size_t)nn + 1, sizeof(int));
    if (or_graph_build(nn, ne, conn, etype, prm->reduced, c->row_ptr, &c->cols) != 0) {
        or_case_destroy(c);
        return NULL;
    }
    c->nnz = c->row_ptr[nn];
    c->diag_off = (int *)xcalloc((size_t)nn, sizeof(int));
    for (int i = 0; i < nn; ++i) c->diag_off[i] = or_graph_find(c->row_ptr, c->cols, i, i);

    c->A3 = (double *)xcalloc((size_t)c->nnz * 9, sizeof(double));
    c->b3 = (double *)xcalloc((size_t)nn * 3, sizeof(double));
    c->x3 = (double *)xcalloc((size_t)nn * 3, sizeof(double));
    c->A1 = (double *)xcalloc((size_t)c->nnz, sizeof(double));
    c->b1 = (double *)xcalloc((size_t)nn, sizeof(double));
    c->x1 = (double *)xcalloc((size_t)nn, sizeof(double));

    c->U = (double *)xcalloc((size_t)nn * 3, sizeof(double));
    c->p = (double *)xcalloc((size_t)nn, sizeof(double));
    c->gradU = (double *)xcalloc((size_t)nn * 9, sizeof(double));
    c->gradP = (double *)xcalloc((size_t)nn * 3, sizeof(double));
    c->du = (double *)xcalloc((size_t)nn * 3, sizeof(double));
    c->mdot = (double *)xcalloc((size_t)ne * 12, sizeof(double));
    c->div = (double *)xcalloc((size_t)nn, sizeof(double));
    c->vol = (double *)xcalloc((size_t)nn, sizeof(double));
    c->rho = (double *)xcalloc((size_t)nn, sizeof(double));
    c->mu = (double *)xcalloc((size_t)nn, sizeof(double));
    c->beta = (double *)xcalloc((size_t)nn * 3, sizeof(double));
    c->bareav = (double *)xcalloc((size_t)nbf * 12, sizeof(double));
    c->bnode = (int *)xcalloc((size_t)nn, sizeof(int));
    c->wnode = (int *)xcalloc((size_t)nn, sizeof(int));
    c->symA = (double *)xcalloc((size_t)nn * 3, sizeof(double));
    c->bc_type = (int *)xcalloc((size_t)nbc, sizeof(int));
    c->bc_p = (double *)xcalloc((size_t)nbc, sizeof(double));
    c->bc_k = (double *)xcalloc((size_t)nbc, sizeof(double));
    c->bc_w = (double *)xcalloc((size_t)nbc, sizeof(double));
    c->msd = (double *)xcalloc((size_t)nbf * 4, sizeof(double));
    c->rfl = (int *)xcalloc((size_t)nbf * 4, sizeof(int));
    for (int i = 0; i < nn; ++i) {
        c->rho[i] = prm->rho;
        c->mu[i] = prm->mu;
    }
    /* exposed area vectors from the face nodes in side-ordinal order
     * (Quad3DSCS / Tri3DSCS determinant) */
    for (int f = 0; f < nbf; ++f) {
        or_side sd;
        or_side_get(c, f, &sd);
        double fc[12];
        for (int k = 0; k < sd.npf; ++k) {
            for (int d = 0; d < 3; ++d) fc[3 * k + d] = c->xyz[3 * sd.fn[k] + d];
            c->bnode[sd.fn[k]] = 1;
        }
        if (sd.npf == 3)
            or_tri_face_areav(fc, &c->bareav[12 * f]);
        else
            or_quad_face_areav(fc, &c->bareav[12 * f]);
    }
    memcpy(c->wnode, c->bnode, sizeof(int) * (size_t)nn);
    or_dual_volume(c);
    c->U_scale = 1.0;
    c->p_scale = 1.0;
    return c;
}

void or_case_destroy(or_case *c)
{
    if (c) {
        free(c->color_ptr);
        free(c->color_elem);
    }
    if (!c) return;
    void *ptrs[] = {c->xyz, c->conn, c->etype, c->bf_elem, c->bf_ord, c->bf_bc, c->bc_U,
                    c->row_ptr, c->cols, c->diag_off, c->A3, c->b3, c->x3,
                    c->A1, c->b1, c->x1, c->U, c->p, c->gradU, c->gradP, c->du,
                    c->mdot, c->div, c->vol, c->rho, c->mu, c->beta, c->bareav,
                    c->bnode, c->wnode, c->symA, c->bc_type, c->bc_p, c->bc_k, c->bc_w,
                    c->msd, c->rfl,
                    c->k, c->w, c->gradK, c->gradW, c->ymin, c->mut, c->F1,
                    c->Pk, c->Pk_wall, c->kprev, c->beta_k, c->beta_w,
                    c->Uold, c->Uold2, c->kold, c->kold2, c->wold, c->wold2};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    free(c);
}

/* boundary physical types (mesh.cpp zone boundaries): no-slip wall nodes
 * and the assembled symmetry area (mesh::updateAssembledSymmetryAreaField_,
 * meshGeometry.cpp:1787-1850: each symmetry face ip's exposed area added to
 * its nearest node) */
int or_case_set_bc_types(or_case *c, const int *type)
{
    for (int b = 0; b < c->nbc; ++b)
        if (type[b] < OR_BC_WALL || type[b] > OR_BC_OPENING) return 1;
    memcpy(c->bc_type, type, sizeof(int) * (size_t)c->nbc);
    memset(c->wnode, 0, sizeof(int) * (size_t)c->nn);
    memset(c->symA, 0, sizeof(double) * 3 * (size_t)c->nn);
    c->has_sym = 0;
    c->has_outlet = 0;
    for (int f = 0; f < c->nbf; ++f) {  /* a pressure-specified boundary */
        const int t = c->bc_type[c->bf_bc[f]];
        if (t == OR_BC_OUTLET || t == OR_BC_OPENING) c->has_outlet = 1;
    }
    for (int f = 0; f < c->nbf; ++f) {
        or_side sd;
        or_side_get(c, f, &sd);
        const int t = c->bc_type[c->bf_bc[f]];
        for (int k = 0; k < sd.npf; ++k) {
            const int nd = sd.fn[k];
            if (t == OR_BC_WALL) c->wnode[nd] = 1;
            if (t == OR_BC_SYMMETRY) {
                for (int d = 0; d < 3; ++d) c->symA[3 * nd + d] += c->bareav[12 * f + 3 * k + d];
                c->has_sym = 1;
            }
        }
    }
    return 0;
}

void or_case_set_bc_values(or_case *c, const double *p, const double *k, const double *w)
{
    for (int b = 0; b < c->nbc; ++b) {
        c->bc_p[b] = p ? p[b] : 0.0;
        c->bc_k[b] = k ? k[b] : 0.0;
        c->bc_w[b] = w ? w[b] : 0.0;
    }
}

const int *or_case_reversal(const or_case *c) { return c->rfl; }
void or_update_reversal(or_case *c) { or_bc_reversal(c); }

static int sym_normal(const or_case *c, int i, double n[3])
{
    const double *a = &c->symA[3 * i];
    const double amag = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    if (!(amag > 0)) return 0;
    for (int d = 0; d < 3; ++d) n[d] = a[d] / amag;
    return 1;
}

int or_case_nn(const or_case *c) { return c->nn; }
int or_case_nnz(const or_case *c) { return c->nnz; }
const int *or_case_row_ptr(const or_case *c) { return c->row_ptr; }
const int *or_case_cols(const or_case *c) { return c->cols; }
const int *or_case_diag_off(const or_case *c) { return c->diag_off; }

double *or_case_field(or_case *c, const char *name)
{
    if (!strcmp(name, "U")) return c->U;
    if (!strcmp(name, "Uold")) return c->Uold;
    if (!strcmp(name, "Uold2")) return c->Uold2;
    if (!strcmp(name, "kold")) return c->kold;
    if (!strcmp(name, "kold2")) return c->kold2;
    if (!strcmp(name, "wold")) return c->wold;
    if (!strcmp(name, "wold2")) return c->wold2;
    if (!strcmp(name, "p")) return c->p;
    if (!strcmp(name, "gradU")) return c->gradU;
    if (!strcmp(name, "gradP")) return c->gradP;
    if (!strcmp(name, "du")) return c->du;
    if (!strcmp(name, "mdot")) return c->mdot;
    if (!strcmp(name, "div")) return c->div;
    if (!strcmp(name, "vol")) return c->vol;
    if (!strcmp(name, "rho")) return c->rho;
    if (!strcmp(name, "mu")) return c->mu;
    if (!strcmp(name, "beta")) return c->beta;
    if (!strcmp(name, "bareav")) return c->bareav;
    if (!strcmp(name, "xyz")) return c->xyz;
    if (!strcmp(name, "k")) return c->k;
    if (!strcmp(name, "omega")) return c->w;
    if (!strcmp(name, "gradK")) return c->gradK;
    if (!strcmp(name, "gradW")) return c->gradW;
    if (!strcmp(name, "ymin")) return c->ymin;
    if (!strcmp(name, "mut")) return c->mut;
    if (!strcmp(name, "F1")) return c->F1;
    if (!strcmp(name, "Pk")) return c->Pk;
    if (!strcmp(name, "Pk_wall")) return c->Pk_wall;
    if (!strcmp(name, "kprev")) return c->kprev;
    if (!strcmp(name, "symA")) return c->symA;
    if (!strcmp(name, "msd")) return c->msd;
    if (!strcmp(name, "beta_k")) return c->beta_k;
    if (!strcmp(name, "beta_w")) return c->beta_w;
    return NULL;
}

double *or_case_vals(or_case *c, int N) { return N == 3 ? c->A3 : c->A1; }
double *or_case_rhs(or_case *c, int N) { return N == 3 ? c->b3 : c->b1; }
double *or_case_x(or_case *c, int N) { return N == 3 ? c->x3 : c->x1; }

#define gather_coords or_gather_coords

void or_dual_volume(or_case *c)
{
    memset(c->vol, 0, sizeof(double) * (size_t)c->nn);
    double co[24], v[8];
    for (int e = 0; e < c->ne; ++e) {
        const int t = or_etype(c, e), npe = or_topo_npe(t);
        gather_coords(c, e, co);
        or_elem_scv(t, co, v);
        for (int a = 0; a < npe; ++a) c->vol[c->conn[8 * e + a]] += v[a];
    }
}

/* SCV weights of the element-centre average (pressureCorrectionAssemblerElemTerms.cpp:346-368;
 * the SCV ipNodeMap of every topology is the identity) */
static void scv_weights(int topo, const double *co, double *w)
{
    const int npe = or_topo_npe(topo);
    double v[8], V = 0.0;
    or_elem_scv(topo, co, v);
    for (int k = 0; k < npe; ++k) V += v[k];
    const double inv = 1.0 / V;
    for (int k = 0; k < npe; ++k) w[k] = 0.0;
    for (int k = 0; k < npe; ++k) w[k] += v[k] * inv;
}

/* ------------------------------------------------------------ mass flow */
static void mdot_init(or_case *c)
{
    /* initializeMassFlowRateInterior_: shape functions shifted iff the
     * velocity interpolation is linear_linear (flowModel.cpp:3310-3316) */
    or_tables tab[4];
    or_tables_init(tab);
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
    for (int e = 0; e < c->ne; ++e) {
        double co[24], av[36];
        const int t = or_etype(c, e);
        const or_tables *T = &tab[t];
        const int npe = T->npe;
        const double *sf = T->sf[c->prm.u_shifted ? 1 : 0];
        const int *en = &c->conn[8 * e];
        gather_coords(c, e, co);
        or_elem_areav(t, co, av);
        for (int ip = 0; ip < T->nscs; ++ip) {
            double u[3] = {0, 0, 0}, rip = 0;
            for (int ic = 0; ic < npe; ++ic) {
                const double r = sf[npe * ip + ic];
                rip += r * c->rho[en[ic]];
                for (int j = 0; j < 3; ++j) u[j] += r * c->U[3 * en[ic] + j];
            }
            double tt = 0;
            for (int j = 0; j < 3; ++j) tt += rip * u[j] * av[3 * ip + j];
            c->mdot[12 * e + ip] = tt;
        }
    }
    or_bc_mdot_init(c);
}

void or_update_mdot(or_case *c)
{
    or_tables tab[4];
    or_tables_init(tab);
    const double urf = c->prm.urf_mdot;
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
    for (int e = 0; e < c->ne; ++e) {
        double co[24], av[36], dndx[288], detj[12], w[8];
        const int t = or_etype(c, e);
        const or_tables *T = &tab[t];
        const int npe = T->npe, *lr = T->lr;
        const double *sfu = T->sf[c->prm.u_shifted ? 1 : 0], *sfc = T->sf[0];
        const int *en = &c->conn[8 * e];
        gather_coords(c, e, co);
        or_elem_areav(t, co, av);
        scv_weights(t, co, w);
        or_elem_gradop(t, c->prm.pgrad_shifted, co, dndx, detj);
        double *md = &c->mdot[12 * e];
        for (int ip = 0; ip < T->nscs; ++ip) {
            const int il = lr[2 * ip], ir = lr[2 * ip + 1];
            double xip[3] = {0, 0, 0}, uip[3] = {0, 0, 0}, Gp[3] = {0, 0, 0};
            double dp[3] = {0, 0, 0}, duip[3] = {0, 0, 0}, rip = 0;
            for (int ic = 0; ic < npe; ++ic) {
                const int n = en[ic];
                const double rv = sfu[npe * ip + ic], rc = sfc[npe * ip + ic], ws = w[ic];
                rip += rv * c->rho[n];
                for (int j = 0; j < 3; ++j) {
                    duip[j] += rv * c->du[3 * n + j];
                    uip[j] += rv * c->U[3 * n + j];
                    dp[j] += dndx[3 * npe * ip + 3 * ic + j] * c->p[n];
                    xip[j] += rc * co[3 * ic + j];
                    Gp[j] += ws * c->gradP[3 * n + j];
                }
            }
            const double ur = md[ip] / rip;
            /* deferred correction of rho vanishes (blending factor 0 for
             * incompressible upwind); rhoHR = upwind rho */
            const double rhoHR = ur > 0 ? c->rho[en[il]] : c->rho[en[ir]];
            double tt = 0;
            for (int j = 0; j < 3; ++j) {
                tt += rhoHR * uip[j] * av[3 * ip + j];
                tt -= rhoHR * duip[j] * (dp[j] - Gp[j]) * av[3 * ip + j];
            }
            md[ip] = urf * tt + (1.0 - urf) * md[ip];
        }
    }
    or_bc_mdot_update(c);
}

void or_mass_divergence(or_case *c)
{
    memset(c->div, 0, sizeof(double) * (size_t)c->nn);
    or_color_elements(c);
    OR_PARALLEL
    OR_FOR_COLORED(c, e) {
        const int t = or_etype(c, e);
        const int *en = &c->conn[8 * e];
        const int *lr = t == OR_HEX8 ? or_hex_lrscv() : or_topo_lrscv(t);
        for (int ip = 0; ip < or_topo_nscs(t); ++ip) {
            const double m = c->mdot[12 * e + ip];
            c->div[en[lr[2 * ip]]] += m;
            c->div[en[lr[2 * ip + 1]]] -= m;
        }
    }
    /* side flows of inlets and outlets to the ip's nearest node; static walls
     * and symmetry planes carry none */
    or_bc_divergence(c);
}

/* ------------------------------------------------------------- gradients */
/* symmetry-plane gradient corrections (velocity and pressure set
 * correctGradient_, velocity.cpp:52, pressure.cpp:48):
 * vector, nodeVectorField.cpp:46-140: G -= (G n) n^T + n (n^T G) - 2 n n^T (n^T G n);
 * scalar, nodeScalarField.cpp:45-110: g -= (g . n) n */
static void symmetry_correct(or_case *c, double *grad, int N)
{
    for (int i = 0; i < c->nn; ++i) {
        double n[3];
        if (!sym_normal(c, i, n)) continue;
        double *g = &grad[3 * N * i];
        if (N == 1) {
            double dot = 0;
            for (int j = 0; j < 3; ++j) dot += g[j] * n[j];
            for (int j = 0; j < 3; ++j) g[j] -= dot * n[j];
            continue;
        }
        double ng[3] = {0, 0, 0}, gn[3] = {0, 0, 0}, ngn = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                ng[a] += g[3 * a + b] * n[b];
                gn[b] += g[3 * a + b] * n[a];
            }
        for (int b = 0; b < 3; ++b) ngn += gn[b] * n[b];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                g[3 * a + b] -= ng[a] * n[b];
                g[3 * a + b] -= n[a] * gn[b];
                g[3 * a + b] += 2.0 * n[a] * n[b] * ngn;
            }
    }
}

void or_gradient(or_case *c, int which)
{
    if (which == 0) {
        or_gradient_field(c, c->U, c->gradU, 3, c->prm.u_shifted);
        if (c->has_sym) symmetry_correct(c, c->gradU, 3);
    } else if (which == 1) {
        or_gradient_field(c, c->p, c->gradP, 1, c->prm.p_shifted);
        if (c->has_sym) symmetry_correct(c, c->gradP, 1);
    }
    else if (which == 2 && c->k) or_gradient_field(c, c->k, c->gradK, 1, c->sst.k_shifted);
    else if (which == 3 && c->w) or_gradient_field(c, c->w, c->gradW, 1, c->sst.w_shifted);
}

void or_gradient_field(or_case *c, const double *phi, double *grad, int N, int shifted)
{
    const double gurf = c->prm.grad_urf;
    /* relax the existing gradient (nodeField.hpp:3268, relax :6362) */
    for (size_t k = 0; k < (size_t)c->nn * 3 * N; ++k) grad[k] *= (1.0 - gurf);
    or_tables tab[4];
    or_tables_init(tab);
    or_color_elements(c);
    OR_PARALLEL
    {
    double co[24], av[36];
    OR_FOR_COLORED(c, e) {
        const int t = or_etype(c, e);
        const or_tables *T = &tab[t];
        const int npe = T->npe, *lr = T->lr;
        const double *sf = T->sf[shifted ? 1 : 0];
        const int *en = &c->conn[8 * e];
        gather_coords(c, e, co);
        or_elem_areav(t, co, av);
        for (int ip = 0; ip < T->nscs; ++ip) {
            const int L = en[lr[2 * ip]], R = en[lr[2 * ip + 1]];
            double pip[3] = {0, 0, 0};
            for (int ic = 0; ic < npe; ++ic)
                for (int j = 0; j < N; ++j) pip[j] += sf[npe * ip + ic] * phi[N * en[ic] + j];
            const double ivL = 1.0 / c->vol[L], ivR = 1.0 / c->vol[R];
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < 3; ++j) {
                    grad[3 * N * L + i * N + j] +=
                        gurf * (pip[i] - phi[N * L + i]) * av[3 * ip + j] * ivL;
                    grad[3 * N * R + i * N + j] -=
                        gurf * (pip[i] - phi[N * R + i]) * av[3 * ip + j] * ivR;
                }
        }
    }
    }
    /* boundary ip contribution (walls; symmetry planes are skipped) */
    for (int f = 0; f < c->nbf; ++f) {
        if (c->bc_type[c->bf_bc[f]] == OR_BC_SYMMETRY) continue;
        or_side sd;
        or_side_get(c, f, &sd);
        const int npf = sd.npf;
        const int *fn = sd.fn;
        double fsf[16];
        or_side_shape(npf, shifted, fsf);
        const double *aw = &c->bareav[12 * f];
        for (int ip = 0; ip < npf; ++ip) {
            const int nn_ = fn[ip];
            double pip[3] = {0, 0, 0};
            for (int ic = 0; ic < npf; ++ic)
                for (int j = 0; j < N; ++j) pip[j] += fsf[npf * ip + ic] * phi[N * fn[ic] + j];
            const double iv = 1.0 / c->vol[nn_];
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < 3; ++j)
                    grad[3 * N * nn_ + i * N + j] +=
                        gurf * (pip[i] - phi[N * nn_ + i]) * aw[3 * ip + j] * iv;
        }
    }
}

/* ------------------------------------------------------------- momentum */
void or_assemble_momentum(or_case *c)
{
    const int nn = c->nn;
    const or_params *P = &c->prm;
    memset(c->A3, 0, sizeof(double) * 9 * (size_t)c->nnz);
    memset(c->b3, 0, sizeof(double) * 3 * (size_t)nn);
    double lhs[576], rhs[24];
    int nodes[8];

    /* node terms, steady (navierStokesAssemblerNodeTerms.cpp:16-149: the
     * pseudo-time term rho V / dt on the diagonal) or transient (BDF1 :151-268,
     * BDF2 :270-400: c0 rho V / dt on the diagonal, the time derivative of the
     * previous levels in the rhs; rho constant in time) */
    double cf[3];
    or_bdf_coeff(c, cf);
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
    for (int i = 0; i < nn; ++i) {
        double l9[9] = {0}, r3[3] = {0};
        const double rho = c->rho[i], vol = c->vol[i], div = c->div[i];
        const double dt = c->tscheme ? c->tdt : P->dt;
        const double lhsfac = (c->tscheme ? cf[0] : 1.0) * rho * vol / dt;
        const double f1 = cf[1] * rho * vol / dt, f2 = cf[2] * rho * vol / dt;
        for (int k = 0; k < 3; ++k) {
            if (div < 0) l9[4 * k] += -div;
            const double Ui = c->U[3 * i + k];
            r3[k] -= -div * Ui;
            r3[k] -= c->gradP[3 * i + k] * vol;
            l9[4 * k] += lhsfac;
            if (c->tscheme) {
                r3[k] -= lhsfac * Ui + f1 * c->Uold[3 * i + k];
                if (c->Uold2) r3[k] -= f2 * c->Uold2[3 * i + k];
            }
        }
        or_apply_coeff(3, 1, &i, l9, r3, nn, c->row_ptr, c->cols, c->diag_off, c->A3, c->b3);
    }

    /* interior element terms */
    or_tables tab[4];
    or_tables_init(tab);
    or_color_elements(c);
    OR_PARALLEL
    {
    double co[24], av[36], dndx[288], detj[12], lhs[576], rhs[24];
    int nodes[8];
    OR_FOR_COLORED(c, e) {
        const int t = or_etype(c, e);
        const or_tables *T = &tab[t];
        const int npe = T->npe, *lr = T->lr, S = 3 * npe; /* local row stride */
        const double *sfu = T->sf[P->u_shifted ? 1 : 0], *sfc = T->sf[0];
        const int *en = &c->conn[8 * e];
        memset(lhs, 0, sizeof(lhs));
        memset(rhs, 0, sizeof(rhs));
        double Uw[24], mu[8];
        for (int a = 0; a < npe; ++a) {
            nodes[a] = en[a];
            mu[a] = c->mu[en[a]];
            for (int d = 0; d < 3; ++d) Uw[3 * a + d] = c->U[3 * en[a] + d];
        }
        gather_coords(c, e, co);
        or_elem_areav(t, co, av);
        or_elem_gradop(t, P->ugrad_shifted, co, dndx, detj);
        const double *md = &c->mdot[12 * e];
        for (int ip = 0; ip < T->nscs; ++ip) {
            const int il = lr[2 * ip], ir = lr[2 * ip + 1];
            const double tm = md[ip];
            double muIp = 0, xip[3] = {0, 0, 0};
            for (int ic = 0; ic < npe; ++ic) {
                muIp += sfu[npe * ip + ic] * mu[ic];
                for (int i = 0; i < 3; ++i) xip[i] += sfc[npe * ip + ic] * co[3 * ic + i];
            }
            for (int i = 0; i < 3; ++i) {
                double uup, dcorr = 0;
                const int iu = tm > 0 ? il : ir;
                uup = Uw[3 * iu + i];
                for (int j = 0; j < 3; ++j)
                    dcorr += c->beta[3 * en[iu] + i] * (xip[j] - co[3 * iu + j]) *
                             c->gradU[9 * en[iu] + 3 * i + j];
                const double aflux = tm * (uup + dcorr);
                const int iL = 3 * il + i, iR = 3 * ir + i;
                const int rowL = iL * S, rowR = iR * S;
                rhs[iL] -= aflux;
                rhs[iR] += aflux;
                const double aL = 0.5 * (tm + fabs(tm)), aR = 0.5 * (tm - fabs(tm));
                lhs[rowL + 3 * il + i] += aL;
                lhs[rowR + 3 * il + i] -= aL;
                lhs[rowR + 3 * ir + i] -= aR;
                lhs[rowL + 3 * ir + i] += aR;
            }
            for (int ic = 0; ic < npe; ++ic) {
                const int icN = 3 * ic;
                for (int i = 0; i < 3; ++i) {
                    const int iL = 3 * il + i, iR = 3 * ir + i;
                    const int rowL = iL * S, rowR = iR * S;
                    const double *dn = &dndx[S * ip + icN];
                    double lhs_riC_i = 0;
                    for (int j = 0; j < 3; ++j) {
                        const double axj = av[3 * ip + j];
                        const double uj = Uw[icN + j];
                        lhs_riC_i += -muIp * dn[j] * axj;
                        const double lfj = -muIp * dn[i] * axj;
                        lhs[rowL + icN + j] += lfj;
                        lhs[rowR + icN + j] -= lfj;
                        rhs[iL] -= lfj * uj;
                        rhs[iR] += lfj * uj;
                    }
                    lhs[rowL + icN + i] += lhs_riC_i;
                    lhs[rowR + icN + i] -= lhs_riC_i;
                    const double ui = Uw[icN + i];
                    rhs[iL] -= lhs_riC_i * ui;
                    rhs[iR] += lhs_riC_i * ui;
                }
            }
        }
        or_apply_coeff(3, npe, nodes, lhs, rhs, nn, c->row_ptr, c->cols, c->diag_off, c->A3, c->b3);
    }
    }

    /* no-slip walls (navierStokesAssemblerElemBoundaryConditions.cpp:2659-2884)
     * and symmetry planes (:2082-2375, IP_FULL_STRESS__ :296-343 with the
     * incompressible divU term zero); sides of every topology (face_grad_op
     * of the parent element at the side's face ips) */
    double co[24];
    for (int f = 0; f < c->nbf; ++f) {
        const int btype = c->bc_type[c->bf_bc[f]];
        if (btype == OR_BC_SLIP) continue;  /* slip wall: no face terms (:2376-2657) */
        or_side sd;
        or_side_get(c, f, &sd);
        const int npe = sd.npe, npf = sd.npf, ld = 3 * npe;
        const int *en = sd.en;
        for (int a = 0; a < npe; ++a) nodes[a] = en[a];
        memset(lhs, 0, sizeof(lhs));
        memset(rhs, 0, sizeof(rhs));
        if (btype == OR_BC_INLET || btype == OR_BC_OUTLET || btype == OR_BC_OPENING) {
            or_bc_momentum(c, f, lhs, rhs);
            or_apply_coeff(3, npe, nodes, lhs, rhs, nn, c->row_ptr, c->cols, c->diag_off, c->A3, c->b3);
            continue;
        }
        double fsf[16];
        or_side_shape(npf, P->u_shifted, fsf);
        gather_coords(c, sd.e, co);
        double fdn[96], fdet[4];
        or_topo_side_gradop(sd.topo, P->ugrad_shifted, sd.ord, co, fdn, fdet);
        const double *aw = &c->bareav[12 * f];
        double muf[4];
        for (int k = 0; k < npf; ++k) muf[k] = c->mu[sd.fn[k]];
        if (btype == OR_BC_SYMMETRY) {
            for (int ip = 0; ip < npf; ++ip) {
                const int nearest = sd.sn[ip];
                const double *ax = &aw[3 * ip];
                double mub = 0;
                for (int ic = 0; ic < npf; ++ic) mub += fsf[npf * ip + ic] * muf[ic];
                for (int ic = 0; ic < npe; ++ic) {
                    const double *dn = &fdn[3 * npe * ip + 3 * ic];
                    const double *u = &c->U[3 * en[ic]];
                    for (int i = 0; i < 3; ++i) {
                        const int iR = 3 * nearest + i, rowR = iR * ld;
                        for (int j = 0; j < 3; ++j) {
                            double lf = -mub * dn[j] * ax[j];
                            lhs[rowR + 3 * ic + i] += lf;
                            rhs[iR] -= lf * u[i];
                            lf = -mub * dn[i] * ax[j];
                            lhs[rowR + 3 * ic + j] += lf;
                            rhs[iR] -= lf * u[j];
                        }
                    }
                }
            }
            or_apply_coeff(3, npe, nodes, lhs, rhs, nn, c->row_ptr, c->cols, c->diag_off, c->A3, c->b3);
            continue;
        }
        double Uw[24], bcm[8];
        for (int a = 0; a < npe; ++a) {
            bcm[a] = 1.0;
            for (int d = 0; d < 3; ++d) Uw[3 * a + d] = c->U[3 * en[a] + d];
        }
        const double *Ubc = &c->bc_U[3 * c->bf_bc[f]];
        for (int k = 0; k < npf; ++k) {
            const int ic = sd.sn[k];
            bcm[ic] = 0.0;
            for (int d = 0; d < 3; ++d) Uw[3 * ic + d] = Ubc[d];
        }
        for (int ip = 0; ip < npf; ++ip) {
            const int nearest = sd.sn[ip];
            const double *ax = &aw[3 * ip];
            const double amag = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
            double mub = 0;
            for (int ic = 0; ic < npf; ++ic) mub += fsf[npf * ip + ic] * muf[ic];
            double nx[3];
            for (int i = 0; i < 3; ++i) nx[i] = ax[i] / amag;
            for (int ic = 0; ic < npe; ++ic) {
                const double *dn = &fdn[3 * npe * ip + 3 * ic];
                for (int j = 0; j < 3; ++j) {
                    const double axj = ax[j], dnj = dn[j], uxj = Uw[3 * ic + j];
                    for (int i = 0; i < 3; ++i) {
                        const int iR = 3 * nearest + i, rowR = iR * ld;
                        const double dni = dn[i], uxi = Uw[3 * ic + i], nxi = nx[i];
                        const double om = 1.0 - nxi * nxi;
                        double lf = -mub * dnj * axj * om;
                        lhs[rowR + 3 * ic + i] += lf * bcm[ic];
                        rhs[iR] -= lf * uxi;
                        lf = -mub * dni * axj * om;
                        lhs[rowR + 3 * ic + j] += lf * bcm[ic];
                        rhs[iR] -= lf * uxj;
                        for (int l = 0; l < 3; ++l) {
                            if (l == i) continue;
                            const double nn_ = nxi * nx[l];
                            const double uxl = Uw[3 * ic + l], dnl = dn[l];
                            lf = mub * dnj * axj * nn_;
                            lhs[rowR + 3 * ic + l] += lf * bcm[ic];
                            rhs[iR] -= lf * uxl;
                            lf = mub * dnl * axj * nn_;
                            lhs[rowR + 3 * ic + j] += lf * bcm[ic];
                            rhs[iR] -= lf * uxj;
                        }
                    }
                }
            }
        }
        or_apply_coeff(3, npe, nodes, lhs, rhs, nn, c->row_ptr, c->cols, c->diag_off, c->A3, c->b3);
    }

    /* postAssemble_: relaxation diag /= urf (phiAssembler.h:966-1006), then
     * boundary rhs relaxation on wall nodes (navierStokesAssembler.cpp:238,
     * phiAssembler.h:1008-1152) */
    const double uinv = 1.0 / P->urf_U;
    for (int i = 0; i < nn; ++i) {
        double *d = &c->A3[9 * ((size_t)c->row_ptr[i] + c->diag_off[i])];
        for (int k = 0; k < 3; ++k) d[4 * k] *= uinv;
    }
    for (int i = 0; i < nn; ++i)
        if (c->bnode[i])
            for (int k = 0; k < 3; ++k) c->b3[3 * i + k] *= P->bnd_relax;
    /* applySymmetryConditions_ (navierStokesAssembler.cpp:387-471): the rhs
     * loses its component along the assembled symmetry normal */
    for (int i = 0; i < nn; ++i) {
        double n[3];
        if (!sym_normal(c, i, n)) continue;
        double *b = &c->b3[3 * i], dot = 0;
        for (int j = 0; j < 3; ++j) dot += b[j] * n[j];
        for (int j = 0; j < 3; ++j) b[j] -= dot * n[j];
    }
}

void or_normal_relaxation(or_case *c)
{
    const double urf = 0.75;
    for (int i = 0; i < c->nn; ++i) {
        double n[3];
        if (!sym_normal(c, i, n)) continue;
        double *d = &c->A3[9 * ((size_t)c->row_ptr[i] + c->diag_off[i])], mod[9] = {0};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int k = 0; k < 3; ++k) mod[3 * a + b] += d[3 * a + k] * n[k] * n[b];
        for (int k = 0; k < 9; ++k) d[k] += (1.0 - urf) * mod[k];
    }
}

void or_compute_du(or_case *c)
{
    for (int i = 0; i < c->nn; ++i) {
        const double *d = &c->A3[9 * ((size_t)c->row_ptr[i] + c->diag_off[i])];
        for (int k = 0; k < 3; ++k) c->du[3 * i + k] = c->vol[i] / (d[4 * k] + SMALL);
    }
}

/* -------------------------------------------------------- p correction */
void or_assemble_pcorr(or_case *c)
{
    const int nn = c->nn;
    const or_params *P = &c->prm;
    memset(c->A1, 0, sizeof(double) * (size_t)c->nnz);
    memset(c->b1, 0, sizeof(double) * (size_t)nn);
    or_tables tab[4];
    or_tables_init(tab);
    double lhs[64], rhs[8];
    int nodes[8];
    or_color_elements(c);
    OR_PARALLEL
    {
    double co[24], av[36], dndx[288], detj[12], w[8], lhs[64], rhs[8];
    int nodes[8];
    OR_FOR_COLORED(c, e) {
        const int t = or_etype(c, e);
        const or_tables *T = &tab[t];
        const int npe = T->npe, *lr = T->lr;
        const double *sfu = T->sf[P->u_shifted ? 1 : 0];
        const int *en = &c->conn[8 * e];
        memset(lhs, 0, sizeof(lhs));
        memset(rhs, 0, sizeof(rhs));
        for (int a = 0; a < npe; ++a) nodes[a] = en[a];
        gather_coords(c, e, co);
        or_elem_areav(t, co, av);
        scv_weights(t, co, w);
        or_elem_gradop(t, P->pgrad_shifted, co, dndx, detj);
        const double *md = &c->mdot[12 * e];
        for (int ip = 0; ip < T->nscs; ++ip) {
            const int il = lr[2 * ip], ir = lr[2 * ip + 1];
            const double tm = md[ip];
            const int rowL = il * npe, rowR = ir * npe;
            const double *dn = &dndx[3 * npe * ip];
            double uip[3] = {0, 0, 0}, duip[3] = {0, 0, 0}, dp[3] = {0, 0, 0}, Gp[3] = {0, 0, 0};
            for (int ic = 0; ic < npe; ++ic) {
                const int n = en[ic];
                const double rv = sfu[npe * ip + ic], ws = w[ic];
                for (int j = 0; j < 3; ++j) {
                    uip[j] += rv * c->U[3 * n + j];
                    duip[j] += rv * c->du[3 * n + j];
                    dp[j] += dn[3 * ic + j] * c->p[n];
                    Gp[j] += ws * c->gradP[3 * n + j];
                }
            }
            const double rhoHR = tm > 0 ? c->rho[en[il]] : c->rho[en[ir]];
            for (int ic = 0; ic < npe; ++ic) {
                double lf = 0;
                for (int j = 0; j < 3; ++j)
                    lf += -rhoHR * duip[j] * dn[3 * ic + j] * av[3 * ip + j];
                lhs[rowL + ic] += lf;
                lhs[rowR + ic] -= lf;
            }
            double m = 0;
            for (int j = 0; j < 3; ++j) {
                m += rhoHR * uip[j] * av[3 * ip + j];
                m -= rhoHR * duip[j] * (dp[j] - Gp[j]) * av[3 * ip + j];
            }
            rhs[il] -= m;
            rhs[ir] += m;
        }
        or_apply_coeff(1, npe, nodes, lhs, rhs, nn, c->row_ptr, c->cols, c->diag_off, c->A1, c->b1);
    }
    }
    /* no-slip static walls and symmetry planes add nothing to p' (no
     * frame/mesh motion); inlets, outlets and openings do */
    for (int f = 0; f < c->nbf; ++f) {
        const int btype = c->bc_type[c->bf_bc[f]];
        if (btype != OR_BC_INLET && btype != OR_BC_OUTLET && btype != OR_BC_OPENING) continue;
        memset(lhs, 0, sizeof(lhs));
        memset(rhs, 0, sizeof(rhs));
        const int npe = or_topo_npe(or_etype(c, c->bf_elem[f]));
        for (int a = 0; a < npe; ++a) nodes[a] = c->conn[8 * c->bf_elem[f] + a];
        or_bc_pcorr(c, f, lhs, rhs);
        or_apply_coeff(1, npe, nodes, lhs, rhs, nn, c->row_ptr, c->cols, c->diag_off, c->A1, c->b1);
    }
    /* pressure reference (only without a pressure-specified boundary,
     * domain::pressureLevelRequired): zero the row, keep the diagonal, b = 0 */
    const int r = c->has_outlet ? -1 : P->p_ref_node;
    if (r >= 0 && r < nn) {
        const double d = c->A1[c->row_ptr[r] + c->diag_off[r]];
        for (int k = c->row_ptr[r]; k < c->row_ptr[r + 1]; ++k) c->A1[k] = 0.0;
        c->A1[c->row_ptr[r] + c->diag_off[r]] = d;
        c->b1[r] = 0.0;
    }
}

/* ----------------------------------------------------------- correction */
void or_correct_U(or_case *c, double urf)
{
    for (size_t k = 0; k < (size_t)c->nn * 3; ++k) c->U[k] += urf * c->x3[k];
}

void or_correct_p(or_case *c, double urf)
{
    for (int k = 0; k < c->nn; ++k) c->p[k] += urf * c->x1[k];
}

void or_velocity_correction(or_case *c, int step)
{
    const double f = (step == 1 ? 1.0 : -1.0) / c->prm.urf_p;
    for (size_t k = 0; k < (size_t)c->nn * 3; ++k) c->U[k] += f * c->du[k] * c->gradP[k];
}

static void field_scale(const double *v, int n, int N, double *scale, double *mx, double *mn)
{
    double maxv = -DBL_MAX, minv = DBL_MAX;
    for (int i = 0; i < n; ++i) {
        double val;
        if (N == 1)
            val = v[i];
        else {
            double s = 0;
            for (int k = 0; k < N; ++k) s += v[N * i + k] * v[N * i + k];
            val = sqrt(s);
        }
        if (val > maxv) maxv = val;
        if (val < minv) minv = val;
    }
    double sc = (maxv > 0 ? maxv : 0.0) - (minv < 0 ? minv : 0.0);
    if (sc < 1.0e-9) sc = 1.0;
    if (maxv > sc) sc = maxv;
    *scale = sc;
    *mx = maxv;
    *mn = minv;
}

void or_field_scale(const double *v, int n, int N, double *scale)
{
    double mx, mn;
    field_scale(v, n, N, scale, &mx, &mn);
}

void or_update_scales(or_case *c, double *U_scale, double *p_scale)
{
    double s, mx, mn, ps, pmx, pmn, rs, rmx, rmn;
    field_scale(c->U, c->nn, 3, &s, &mx, &mn);
    field_scale(c->p, c->nn, 1, &ps, &pmx, &pmn);
    field_scale(c->rho, c->nn, 1, &rs, &rmx, &rmn);
    c->U_scale = s;
    const double a = pmx - pmn, b = 0.5 * rs * s * s;
    c->p_scale = a > b ? a : b;
    if (U_scale) *U_scale = c->U_scale;
    if (p_scale) *p_scale = c->p_scale;
}

void or_scaled_residual(or_case *c, int N, double *out)
{
    const double *A = N == 3 ? c->A3 : c->A1;
    const double *b = N == 3 ? c->b3 : c->b1;
    for (int k = 0; k < N; ++k) {
        double s = 0;
        for (int i = 0; i < c->nn; ++i) {
            const double d = A[(size_t)N * N * (c->row_ptr[i] + c->diag_off[i]) + (N + 1) * k];
            const double r = b[N * i + k] / d;
            s += r * r;
        }
        out[k] = sqrt(s / c->nn);
    }
}

/* ------------------------------------------------------ blending factor */
static double limiter_y(double mx, double mn, double phi, double rgrad)
{
    rgrad = rgrad > 0 ? rgrad + SMALL : rgrad - SMALL;
    const double betaLocal = fmax((mx - phi) / rgrad, (mn - phi) / rgrad);
    const double beta12Local = fmax(-(mx - phi) / rgrad, -(mn - phi) / rgrad);
    double y = fmin(betaLocal, beta12Local);
    y = (y * y + 2.0 * y) / (y * y + y + 2.0); /* blendingFactorMax_ = 1 < 2 */
    return y;
}

/* beta = min(beta, y) unless the projected gradient is negligible */
static void limit(double *beta, double mx, double mn, double phi, const double *g, const double *dx)
{
    double rgrad = 0;
    for (int j = 0; j < 3; ++j) rgrad += g[j] * dx[j];
    const double rg = rgrad > 0 ? rgrad + SMALL : rgrad - SMALL;
    const double y = limiter_y(mx, mn, phi, rgrad);
    if (rg > 100 * SMALL || rg < -100 * SMALL) *beta = fmin(*beta, y);
}

void or_update_blending(or_case *c, int which)
{
    const int N = which == 0 ? 3 : 1;
    const double *phi = which == 0 ? c->U : which == 2 ? c->k : c->w;
    const double *grad = which == 0 ? c->gradU : which == 2 ? c->gradK : c->gradW;
    double *beta;
    if (which == 0) {
        beta = c->beta;
    } else {
        double **bp = which == 2 ? &c->beta_k : &c->beta_w;
        if (!*bp) *bp = (double *)xcalloc((size_t)c->nn, sizeof(double));
        beta = *bp;
    }
    const size_t n = (size_t)c->nn * N;
    double *mn = (double *)xcalloc(n, sizeof(double)), *mx = (double *)xcalloc(n, sizeof(double));
    for (size_t k = 0; k < n; ++k) mn[k] = mx[k] = phi[k];
    or_tables tab[4];
    or_tables_init(tab);
    for (int e = 0; e < c->ne; ++e) {
        const or_tables *T = &tab[or_etype(c, e)];
        const int *en = &c->conn[8 * e], *lr = T->lr;
        for (int ip = 0; ip < T->nscs; ++ip) {
            const int L = en[lr[2 * ip]], R = en[lr[2 * ip + 1]];
            for (int i = 0; i < N; ++i) {
                mn[N * L + i] = fmin(mn[N * L + i], phi[N * R + i]);
                mx[N * L + i] = fmax(mx[N * L + i], phi[N * R + i]);
                mn[N * R + i] = fmin(mn[N * R + i], phi[N * L + i]);
                mx[N * R + i] = fmax(mx[N * R + i], phi[N * L + i]);
            }
        }
    }
    for (size_t k = 0; k < n; ++k) beta[k] = 1.0; /* blendingFactorMax_ */
    double co[24];
    for (int e = 0; e < c->ne; ++e) {
        const or_tables *T = &tab[or_etype(c, e)];
        const int npe = T->npe, *lr = T->lr;
        const double *sfc = T->sf[0]; /* coordinates: standard shape functions */
        const int *en = &c->conn[8 * e];
        gather_coords(c, e, co);
        for (int ip = 0; ip < T->nscs; ++ip) {
            double xip[3] = {0, 0, 0};
            for (int ic = 0; ic < npe; ++ic)
                for (int j = 0; j < 3; ++j) xip[j] += sfc[npe * ip + ic] * co[3 * ic + j];
            for (int side = 0; side < 2; ++side) {
                const int a = lr[2 * ip + side], node = en[a];
                double dx[3];
                for (int j = 0; j < 3; ++j) dx[j] = xip[j] - co[3 * a + j];
                for (int i = 0; i < N; ++i)
                    limit(&beta[N * node + i], mx[N * node + i], mn[N * node + i], phi[N * node + i],
                          &grad[3 * N * node + 3 * i], dx);
            }
        }
    }
    /* boundary faces: each face ip limits its nearest node */
    for (int f = 0; f < c->nbf; ++f) {
        or_side sd;
        or_side_get(c, f, &sd);
        const int npf = sd.npf;
        const int *fn = sd.fn;
        double fc[12], fsf[16];
        or_side_shape(npf, 0, fsf);
        for (int k = 0; k < npf; ++k)
            for (int d = 0; d < 3; ++d) fc[3 * k + d] = c->xyz[3 * fn[k] + d];
        for (int ip = 0; ip < npf; ++ip) {
            double xip[3] = {0, 0, 0};
            for (int ic = 0; ic < npf; ++ic)
                for (int j = 0; j < 3; ++j) xip[j] += fsf[npf * ip + ic] * fc[3 * ic + j];
            const int node = fn[ip];
            double dx[3];
            for (int j = 0; j < 3; ++j) dx[j] = xip[j] - fc[3 * ip + j];
            for (int i = 0; i < N; ++i)
                limit(&beta[N * node + i], mx[N * node + i], mn[N * node + i], phi[N * node + i],
                      &grad[3 * N * node + 3 * i], dx);
        }
    }
    free(mn);
    free(mx);
}

/* ------------------------------------------------------------ initialize */
void or_case_initialize(or_case *c)
{
    /* navierStokesEquation::initialize (:86-99): gradient, scale */
    memset(c->gradU, 0, sizeof(double) * 9 * (size_t)c->nn);
    memset(c->gradP, 0, sizeof(double) * 3 * (size_t)c->nn);
    or_gradient(c, 0);
    if (c->prm.high_res) or_update_blending(c, 0);
    /* pressureCorrectionEquation::initialize (:59-70) */
    or_gradient(c, 1);
    or_update_scales(c, NULL, NULL);
    /* postInitialize: mass flow rate and divergence (:101-133) */
    mdot_init(c);
    or_mass_divergence(c);
}

/* -------------------------------------------------------------- outer */
/* Start a physical time step of length dt with scheme 1 (BDF1) or 2 (BDF2):
 * the previous levels move back (old2 <- old <- current) for U and, with
 * SST, k and omega (the reference's prevTime fields). */
int or_time_step(or_case *c, int scheme, double dt)
{
    if ((scheme != 1 && scheme != 2) || !(dt > 0)) return 1;
    const size_t n3 = (size_t)c->nn * 3, n1 = (size_t)c->nn;
    if (!c->Uold) {
        c->Uold = (double *)xcalloc(n3, sizeof(double));
        c->Uold2 = (double *)xcalloc(n3, sizeof(double));
        if (c->k) {
            c->kold = (double *)xcalloc(n1, sizeof(double));
            c->kold2 = (double *)xcalloc(n1, sizeof(double));
            c->wold = (double *)xcalloc(n1, sizeof(double));
            c->wold2 = (double *)xcalloc(n1, sizeof(double));
        }
    }
    memcpy(c->Uold2, c->Uold, n3 * sizeof(double));
    memcpy(c->Uold, c->U, n3 * sizeof(double));
    if (c->kold) {
        memcpy(c->kold2, c->kold, n1 * sizeof(double));
        memcpy(c->kold, c->k, n1 * sizeof(double));
        memcpy(c->wold2, c->wold, n1 * sizeof(double));
        memcpy(c->wold, c->w, n1 * sizeof(double));
    }
    c->tdt_prev = c->tsteps > 0 ? c->tdt : 0.0;
    c->tdt = dt;
    c->tscheme = scheme;
    ++c->tsteps;
    return 0;
}

void or_outer_iteration(or_case *c, const double *ctl, double *hist)
{
    const or_params *P = &c->prm;
    double res[2], r3[3], r1;
    /* residual scales in force when each system is reported: those of the
     * previous iteration's end (linearSystem.h:475, navierStokesEquation.cpp:230,
     * pressureCorrectionEquation.cpp:165) */
    const double us0 = c->U_scale, ps0 = c->p_scale;
    /* momentum predictor (navierStokesEquation::solve :157-222) */
    or_assemble_momentum(c);
    or_compute_du(c);
    or_normal_relaxation(c);
    or_scaled_residual(c, 3, r3);
    memset(c->x3, 0, sizeof(double) * 3 * (size_t)c->nn);
    const int itU = or_solve(c, 3, 0, (int)ctl[0], ctl[1], 1e-30, 30, res);
    or_correct_U(c, 1.0);
    or_update_mdot(c);
    double us, ps;
    or_update_scales(c, &us, &ps);
    /* pressure correction (pressureCorrectionEquation::solve :89-164) */
    or_assemble_pcorr(c);
    or_scaled_residual(c, 1, &r1);
    memset(c->x1, 0, sizeof(double) * (size_t)c->nn);
    const int itP = or_solve(c, 1, P->p_method, (int)ctl[2], ctl[3], 1e-30, 30, res);
    or_correct_p(c, P->urf_p);
    /* segregatedFlowEquations.cpp:106-285 */
    or_update_mdot(c);
    or_bc_reversal(c);
    or_mass_divergence(c);
    or_velocity_correction(c, 1);
    or_gradient(c, 1);
    or_velocity_correction(c, 2);
    or_update_scales(c, &us, &ps);
    or_gradient(c, 0);
    if (P->high_res) or_update_blending(c, 0); /* segregatedFlowEquations.cpp:284-288 */
    if (hist) {
        for (int k = 0; k < 3; ++k) hist[k] = r3[k] / (us0 + SMALL);
        hist[3] = r1 / (ps0 + SMALL);
        hist[4] = itU;
        hist[5] = itP;
    }
}
