/*
 * case_internal.h -- state of an oracle case shared by case.c and sst.c.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#ifndef OA_CASE_INTERNAL_H
#define OA_CASE_INTERNAL_H
#include "oracle.h"
#include <float.h>

#define SMALL DBL_EPSILON /* src/basic/types.h:179 */

struct or_case {
    int nn, ne, nbf, nbc;
    double *xyz;
    int *conn;      /* [ne][8], unused trailing entries -1 */
    int *etype;     /* topology per element (OR_*), NULL: every element hex8 */
    int *bf_elem, *bf_ord, *bf_bc;
    double *bc_U;
    or_params prm;
    /* graph */
    int nnz;
    int *row_ptr, *cols, *diag_off;
    /* systems */
    double *A3, *b3, *x3, *A1, *b1, *x1;
    /* fields */
    double *U, *p, *gradU, *gradP, *du, *mdot, *div, *vol, *rho, *mu, *beta;
    double *bareav; /* exposed area vectors [nbf][4][3] */
    int *bnode;     /* node is on a boundary face (any type) */
    int *bc_type;   /* physical type of each boundary id (OR_BC_*) */
    int *wnode;     /* node is on a no-slip wall face */
    double *symA;   /* assembled symmetry area vector [nn][3] (0 off symmetry) */
    int has_sym, has_outlet;
    double *bc_p, *bc_k, *bc_w; /* outlet static pressure, inlet k / omega per id */
    double *msd;    /* side mass flow rate per boundary face ip [nbf][4] */
    int *rfl;       /* flow-reversal flag per boundary face ip [nbf][4] */
    double U_scale, p_scale;
    /* k-omega SST (sst.c); NULL until or_sst_enable */
    or_sst_params sst;
    double *k, *w, *gradK, *gradW, *ymin, *mut, *F1, *Pk, *Pk_wall, *kprev, *beta_k, *beta_w;
    double k_scale, w_scale;
    int scalar_terms;   /* or_set_scalar_terms mask (test hook), default all */
    /* transient (or_time_step): scheme 0 steady, 1 BDF1, 2 BDF2; physical
     * time step and the previous one (0 on the first step); fields of the
     * previous two time levels */
    int tscheme, tsteps;
    double tdt, tdt_prev;
    double *Uold, *Uold2, *kold, *kold2, *wold, *wold2;
    /* element colour classes (no two elements of a class share a node) for
     * the scattering loops when several threads run them; one class in
     * element order with one thread */
    int ncolor, color_threads;
    int *color_ptr, *color_elem;
};

/* BDF coefficients of the current time step (types.cpp:392-411: BDF1
 * {1, -1}; BDF2 with omega = dt / dt_prev, BDF1 on the first step) */
static inline void or_bdf_coeff(const or_case *c, double cf[3])
{
    cf[0] = 1.0;
    cf[1] = -1.0;
    cf[2] = 0.0;
    if (c->tscheme == 2 && c->tdt_prev > 0.0) {
        const double om = c->tdt / c->tdt_prev;
        cf[0] = (1.0 + 2.0 * om) / (1.0 + om);
        cf[1] = -(1.0 + om);
        cf[2] = om * om / (1.0 + om);
    }
}

/* threads of the oracle's loops (or_set_threads, default 1) */
extern int or_nthreads;
void or_color_elements(or_case *c);
#define OR_PARALLEL _Pragma("omp parallel num_threads(or_nthreads)")
/* inside OR_PARALLEL: every element once, one colour class after the other,
 * each class shared out among the threads */
#define OR_FOR_COLORED(c, e)                                                    \
    for (int col_ = 0; col_ < (c)->ncolor; ++col_)                              \
        _Pragma("omp for schedule(static)")                                     \
        for (int k_ = (c)->color_ptr[col_]; k_ < (c)->color_ptr[col_ + 1]; ++k_) \
            for (int e = (c)->color_elem[k_], once_ = 1; once_; once_ = 0)

static inline int or_etype(const or_case *c, int e) { return c->etype ? c->etype[e] : OR_HEX8; }

static inline void or_gather_coords(const or_case *c, int e, double *co)
{
    const int *en = &c->conn[8 * e];
    const int npe = or_topo_npe(or_etype(c, e));
    for (int a = 0; a < npe; ++a)
        for (int d = 0; d < 3; ++d) co[3 * a + d] = c->xyz[3 * en[a] + d];
}

/* one boundary side: its element and topology, the element's nodes, the
 * side nodes as local ordinals and as global nodes (side order, which is
 * the order of the side's face ips) */
typedef struct {
    int e, ord, topo, npe, npf;
    const int *en;
    int sn[4], fn[4];
} or_side;

static inline void or_side_get(const or_case *c, int f, or_side *s)
{
    s->e = c->bf_elem[f];
    s->ord = c->bf_ord[f];
    s->topo = or_etype(c, s->e);
    s->npe = or_topo_npe(s->topo);
    s->en = &c->conn[8 * s->e];
    s->npf = or_topo_side_nodes(s->topo, s->ord, s->sn);
    for (int k = 0; k < 4; ++k) s->fn[k] = k < s->npf ? s->en[s->sn[k]] : -1;
}

/* face shape functions [npf][npf] of a tri3 or quad4 side */
static inline void or_side_shape(int npf, int shifted, double *shape)
{
    if (npf == 3)
        or_tri_face_shape(shifted, shape);
    else
        or_quad_face_shape(shifted, shape);
}

/* master-element tables of one topology (elem.c); sized for hex8 */
typedef struct {
    int npe, nscs;
    const int *lr;       /* [nscs][2] */
    double sf[2][96];    /* SCS shape functions [nscs][npe], standard / shifted ips */
} or_tables;
void or_tables_init(or_tables tab[4]);
void or_elem_areav(int topo, const double *co, double *av);
int or_elem_gradop(int topo, int shifted, const double *co, double *g, double *detj);
void or_elem_scv(int topo, const double *co, double *v);

/* Green-Gauss gradient of an N-component nodal field (case.c) */
void or_gradient_field(or_case *c, const double *phi, double *grad, int N, int shifted);
/* scale of an N-component field, nodeField::updateScale (case.c) */
void or_field_scale(const double *v, int n, int N, double *scale);

/* inlet / outlet boundary conditions (bc.c) */
void or_bc_mdot_init(or_case *c);
void or_bc_mdot_update(or_case *c);
void or_bc_reversal(or_case *c);
void or_bc_divergence(or_case *c);
void or_bc_momentum(or_case *c, int f, double *lhs, double *rhs);
void or_bc_pcorr(or_case *c, int f, double *lhs, double *rhs);
void or_bc_scalar(or_case *c, int which, int f, const double *G8, double *lhs, double *rhs);


#endif
