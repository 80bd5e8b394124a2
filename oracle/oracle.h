/*
 * oracle.h -- CPU restatement of OpenAccel's CVFEM assemble-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product in openaccel_amd/ and the CPU baseline timed by bench.py.  It is
 * never linked into, loaded by, or called from the product path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * FP64 (reference `scalar = double`, src/basic/types.h:103-105) with int32
 * local ids (`label = int`).
 *
 * Layouts follow the reference's STK field data:
 *   coordinates / velocity / du / grad p : [node][3]
 *   grad U                               : [node][9], (i*3+j) = dU_i/dx_j
 *   mass flow rate                       : [elem][n_scs_ip]
 *   block-CRS values                     : [nnz][N*N] row-major N x N blocks
 *   rhs / solution                       : [node][N]
 */
#ifndef OA_ORACLE_H
#define OA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* topology codes shared with the product's C ABI */
enum { OR_HEX8 = 0, OR_TET4 = 1, OR_WED6 = 2, OR_PYR5 = 3 };

/* ---------------------------------------------------------------- geometry */
/* hex8 SCS area vectors [12][3] (MasterElementWork.F:1-231 hex_scs_det,
 * CalculateQuadArea.F quadAreaByTriangleFacets) */
void or_hex_scs_areav(const double *coords /*[8][3]*/, double *areav /*[12][3]*/);
/* hex8 SCV volumes [8] (MasterElementWork.F:233-457 hex_scv_det,
 * CalculateHexVolume.F, CalculatePolyhedralVolume.F) */
void or_hex_scv_vol(const double *coords, double *vol /*[8]*/);
/* hex8 shape functions at npts isoparametric points (MasterElementWork.F:2344) */
void or_hex_shape(int npts, const double *par /*[npts][3]*/, double *shape /*[npts][8]*/);
/* hex8 shape derivatives (MasterElementWork.F:2404) */
void or_hex_deriv(int npts, const double *par, double *deriv /*[npts][8][3]*/);
/* gradient operator (MasterElementWork.F:2623 hex_gradient_operator);
 * returns the number of ips with det_j <= 1e6*realmin (err flag) */
int or_hex_gradop(int npts, const double *deriv, const double *coords,
                  double *gradop /*[npts][8][3]*/, double *detj /*[npts]*/);
/* SCS gradient operator at the standard (shifted=0) or edge-midpoint
 * (shifted=1) ips (Hex8CVFEM.C:612-662) */
int or_hex_scs_gradop(int shifted, const double *coords, double *gradop, double *detj);
/* face gradient operator of face `ord` at its 4 ips (Hex8CVFEM.C:686-783) */
int or_hex_face_gradop(int shifted, int ord, const double *coords,
                       double *gradop /*[4][8][3]*/, double *detj /*[4]*/);
/* SCS shape functions [12][8] at standard / shifted ips (Hex8CVFEM.C:833-850) */
void or_hex_scs_shape(int shifted, double *shape);
/* quad4 exposed-face area vectors [4][3] (MasterElementWork.F:1467 quad3d_scs_det) */
void or_quad_face_areav(const double *coords /*[4][3]*/, double *areav /*[4][3]*/);
/* quad4 face shape functions [4][4] (Quad43DCVFEM.C:38-48, MasterElementWork.F:3651) */
void or_quad_face_shape(int shifted, double *shape);
/* master-element tables (Hex8CVFEM.C:264-276, 381-393, 419-426) */
const int *or_hex_lrscv(void);          /* [12][2] */
const int *or_hex_side_nodes(void);     /* [6][4]  */

/* tet4 / wedge6 / pyramid5 master elements (topo.c); topo = OR_* above.
 * Vertex order, SCS ips and L/R tables are those of Tet4CVFEM.C, Wed6CVFEM.C,
 * Pyr5CVFEM.C; OR_HEX8 routes to the hex8 functions. */
int or_topo_npe(int topo);
int or_topo_nscs(int topo);
const int *or_topo_lrscv(int topo);                       /* [nscs][2] */
const double *or_topo_scs_ips(int topo, int shifted);     /* [nscs][3] */
/* tet_scs_det / wed_scs_det / pyr_scs_det: area vectors [nscs][3] */
void or_topo_scs_areav(int topo, const double *coords, double *areav);
/* tet_scv_det / wed_scv_det / pyr_scv_det: SCV volumes [npe] */
void or_topo_scv_vol(int topo, const double *coords, double *vol);
/* tet_derivative, wedge_derivative, pyr_derivative / shifted_pyr_derivative */
void or_topo_deriv(int topo, int shifted, int npts, const double *par, double *deriv);
/* *_gradient_operator for any npe; returns the count of det_j <= 1e6*realmin */
int or_topo_gradop(int npe, int npts, const double *deriv, const double *coords,
                   double *gradop, double *detj);
/* SCS gradient operator [nscs][npe][3] at the standard or shifted ips */
int or_topo_scs_gradop(int topo, int shifted, const double *coords, double *gradop,
                       double *detj);
/* tet_shape_fcn, wedge_shape_fcn, pyr_shape_fcn / shifted_pyr_shape_fcn at
 * npts parametric points: shape [npts][npe] (elem.c) */
void or_topo_shape(int topo, int shifted, int npts, const double *par, double *shape);
/* sides of any topology: count, side-node ordinals [4] (returns the side's
 * node count, 3 or 4; unused entries -1), exposed-face ip locations
 * [npf][3], and the face gradient operator [npf][npe][3] at the standard or
 * shifted face ips (face_grad_op / shifted_face_grad_op) */
int or_topo_nsides(int topo);
int or_topo_side_nodes(int topo, int ord, int *nodes);
void or_topo_side_ips(int topo, int ord, int shifted, double *par);
int or_topo_side_gradop(int topo, int shifted, int ord, const double *coords, double *gradop,
                        double *detj);
/* tri3 exposed-face area vectors [3][3] (tri3d_scs_det) and shape
 * functions [3][3] (Tri3DSCS) */
void or_tri_face_areav(const double *coords /*[3][3]*/, double *areav /*[3][3]*/);
void or_tri_face_shape(int shifted, double *shape);

/* ------------------------------------------------------------- CRS graph */
/* nodeGraph::buildGraph_ (mesh/graph/nodeGraph.cpp:29-392): reduced stencil
 * (SCS-ip pairs, :68-96) when reduced != 0, else full element stencil
 * (:98-118).  Owned rows 0..nn-1, sorted local columns.  row_ptr[nn+1] must
 * be allocated by the caller; cols is allocated here (free with or_free). */
int or_graph_build(int nn, int ne, const int *conn, const int *etype, int reduced,
                   int *row_ptr, int **cols_out);
void or_free(void *p);
/* position of `col` in row `row`, or -1 (the sorted search of
 * assembler.h:380-398) */
int or_graph_find(const int *row_ptr, const int *cols, int row, int col);

/* --------------------------------------------------------------- scatter */
/* assembler<N>::applyCoeff_ (assemble/assembler.h:311-453): add the dense
 * local system lhs[(r*N+m)*(n*N)+c*N+k], rhs[r*N+m] over `n` nodes into the
 * block-CRS (row_ptr, cols, vals[nnz][N*N]) and b, owned rows only
 * (row < n_owned), columns missing from the row lumped into the diagonal
 * block (deductUnfound). diag_off[row] is the position of the diagonal. */
void or_apply_coeff(int N, int n, const int *nodes, const double *lhs,
                    const double *rhs, int n_owned, const int *row_ptr,
                    const int *cols, const int *diag_off, double *vals,
                    double *b);

/* -------------------------------------------------------------- the case */
typedef struct or_params {
    double rho, mu;          /* constant density / dynamic viscosity     */
    double dt;               /* physical (pseudo) timescale              */
    double urf_U, urf_p;     /* velocity (implicit) / pressure relaxation */
    double urf_mdot;         /* mass flow relaxation (relaxMass_)        */
    double grad_urf;         /* gradient relaxation, nodeField.h:152      */
    double bnd_relax;        /* boundary rhs relaxation, navierStokesAssembler.cpp:238 */
    int u_shifted;           /* velocity interpolation linear_linear      */
    int ugrad_shifted;       /* velocity gradient interpolation           */
    int p_shifted;           /* pressure interpolation linear_linear      */
    int pgrad_shifted;       /* pressure gradient interpolation           */
    int reduced;             /* bandwidth_reduction (reduced stencil)     */
    int p_ref_node;          /* pressure-level node (-1: none)            */
    int high_res;            /* advection_scheme high_resolution: limited
                                deferred correction (else upwind)         */
    int p_method;            /* p' solver of or_outer_iteration: 0 GMRES,
                                1 Jacobi-CG, 2 BiCGStab                   */
} or_params;

typedef struct or_case or_case;

/* Build the oracle state for a hex8 mesh.  Boundary faces: parent element,
 * face ordinal and boundary id; bc_U[nbc][3] wall velocities. */
/* conn [ne][8] (unused trailing entries -1), etype [ne] topology per
 * element (NULL: all hex8); boundary faces are quad sides of hex8 elements */
or_case *or_case_create(int nn, int ne, const double *xyz, const int *conn,
                        const int *etype, int nbf, const int *bf_elem, const int *bf_ord,
                        const int *bf_bc, int nbc, const double *bc_U,
                        const or_params *prm);
void or_case_destroy(or_case *c);

/* graph access */
int or_case_nn(const or_case *c);
int or_case_nnz(const or_case *c);
const int *or_case_row_ptr(const or_case *c);
const int *or_case_cols(const or_case *c);
const int *or_case_diag_off(const or_case *c);

/* field access (pointers into the case; caller may write them) */
double *or_case_field(or_case *c, const char *name);

/* flowModel::initializeMassFlowRateInterior_ (flowModel.cpp:3303-3470) and
 * the initialize()/postInitialize() sequence of navierStokesEquation.cpp:86-133
 * and pressureCorrectionEquation.cpp:59-74 */
void or_case_initialize(or_case *c);

/* mesh::updateDualNodalVolumeField_ (meshGeometry.cpp:1277) -> "vol" */
void or_dual_volume(or_case *c);

/* systems.  *_assemble fill the case's own matrix (vals/b) for N=3 or N=1. */
void or_assemble_momentum(or_case *c);   /* navierStokesEquation::solve :159-173 */
void or_compute_du(or_case *c);          /* navierStokesAssembler.cpp:84-232     */
/* assembleNormalRelaxation (navierStokesAssembler.cpp:473-696, called with
 * urf 0.75 by navierStokesEquation.cpp:169): symmetry-node diagonal blocks
 * += (1 - 0.75) D n n^T, after the du coefficients */
void or_normal_relaxation(or_case *c);
/* boundary physical types per boundary id: no-slip wall, velocity inlet
 * (the bc_U row; zero-gradient pressure), static-pressure outlet
 * (zero-gradient velocity), symmetry plane, slip wall (no face terms in
 * the momentum equations, navierStokesAssemblerElemBoundaryConditions.cpp:
 * 2376-2657; nothing in p' or the side mass flows; k, omega zero-gradient),
 * static-pressure opening (:3234-3552: the bc_U row is the entrainment
 * direction, normalised; a zero row takes the inward side normal,
 * velocity.cpp:4830-4887 normal_to_boundary_condition) */
enum { OR_BC_WALL = 0, OR_BC_INLET = 1, OR_BC_OUTLET = 2, OR_BC_SYMMETRY = 3, OR_BC_SLIP = 4,
       OR_BC_OPENING = 5 };
int or_case_set_bc_types(or_case *c, const int *type);
/* per boundary id: outlet / opening static pressure, inlet k and omega
 * (NULL: zero) */
void or_case_set_bc_values(or_case *c, const double *p, const double *k, const double *w);
/* flow-reversal flags of the boundary face ips [nbf][4] */
const int *or_case_reversal(const or_case *c);
/* flowModel::updateFlowReversalFlag (flowModel.cpp:14949-15454) for the
 * static-pressure outlets: clip the side mass flow, set / clear flags */
void or_update_reversal(or_case *c);
void or_assemble_pcorr(or_case *c);      /* pressureCorrectionEquation::solve :89-107 */
double *or_case_vals(or_case *c, int N); /* [nnz][N*N] */
double *or_case_rhs(or_case *c, int N);  /* [nn][N]     */
double *or_case_x(or_case *c, int N);    /* [nn][N]     */

/* field updates */
void or_update_mdot(or_case *c);         /* flowModel.cpp:5272-5688 */
void or_mass_divergence(or_case *c);     /* flowModel.cpp:13541-13775 */
void or_gradient(or_case *c, int which); /* nodeField.hpp:3253-4950; 0=U, 1=p */
void or_correct_U(or_case *c, double urf);   /* equation.h:204-286 (N=3) */
void or_correct_p(or_case *c, double urf);   /* equation.h:204-286 (N=1) */
void or_velocity_correction(or_case *c, int step); /* segregatedFlowEquations.cpp:133-279 */
void or_update_scales(or_case *c, double *U_scale, double *p_scale); /* nodeField.hpp:5993, flowModel.cpp:117 */

/* nodeField::updateBlendingFactorField (field/nodeField/nodeField.hpp:2069-3250)
 * with updateMinMaxFields (:1843-2068): Barth-Jespersen limiter over the SCS
 * ips (interior and boundary faces) with the smooth form for a maximum
 * blending factor of 1 (nodeField.h:150).  which: 0 U -> "beta" [nn][3],
 * 2 k -> "beta_k", 3 omega -> "beta_w". */
void or_update_blending(or_case *c, int which);

/* Krylov solves on the case's system.  kind: 0 = Jacobi-GMRES(m),
 * 1 = Jacobi-CG, 2 = Jacobi-BiCGStab.  Returns iterations; res[0]/res[1] =
 * initial / final ||b - A x||_2. */
int or_solve(or_case *c, int N, int kind, int max_it, double rtol,
             double atol, int restart, double *res);

/* control residual of a system: RMS over rows of (b/diag), per component
 * (definition used in place of liblinsolve's scaled_initial_res) */
void or_scaled_residual(or_case *c, int N, double *out /*[N]*/);

/* one steady outer iteration (simulation.cpp:314-377 ->
 * segregatedFlowEquations::solve :76-297); ctl[0..3] = momentum solver max_it,
 * rtol, p' solver max_it, rtol; hist[0..4] = U,V,W,P control residuals,
 * hist[5..6] = momentum / p' solver iterations. */
/* threads of the oracle's loops (OpenMP; default 1).  With one thread every
 * loop runs in element / row order; with more, the loops that scatter into
 * nodes run over colour classes of elements that share no node, so sums are
 * formed in another order (results agree to rounding). */
void or_set_threads(int n);
void or_outer_iteration(or_case *c, const double *ctl, double *hist);
/* start a physical time step (scheme 1 BDF1, 2 BDF2; 0 on success) */
int or_time_step(or_case *c, int scheme, double dt);

/* ------------------------------------------------------- k-omega SST */
/* Constants: shearStressTransportModel.h:21-37 (sigma_k1 0.85, sigma_k2 1.0,
 * sigma_w1 0.5, sigma_w2 0.856, beta1 0.075, beta2 0.0828, gamma1 0.5532,
 * gamma2 0.4403, a1 0.31), RANSModel.h:20-22 (beta* 0.09, production limit
 * 10), turbulenceModel.h:20 (C_mu 0.09). */
typedef struct or_sst_params {
    double urf_k, urf_w;     /* implicit relaxation of the k / omega rows   */
    double urf_mut;          /* eddy-viscosity relaxation (mutRef().urf())  */
    int k_shifted, w_shifted;         /* k / omega interpolation linear_linear */
    int kgrad_shifted, wgrad_shifted; /* gradient operators at shifted ips     */
} or_sst_params;

/* Allocate the SST fields: "k", "omega", "gradK", "gradW", "ymin" (wall
 * distance), "mut", "F1", "Pk", "Pk_wall" (production at no-slip wall nodes,
 * supplied by the wall treatment), "kprev".  "mu" becomes mu_eff. */
void or_sst_enable(or_case *c, const or_sst_params *s);
/* shearStressTransportModel::updateTurbulentDynamicViscosity (:597-690) and
 * turbulenceModel::updateEffectiveDynamicViscosity (turbulenceModel.cpp:121) */
void or_sst_update_mut(or_case *c);
/* shearStressTransportModel::updateFOneBlending (:95-174) */
void or_sst_update_F1(or_case *c);
/* shearStressTransportModel::updateTurbulentProduction (:176-340),
 * incompressible, no frame rotation; no-slip wall nodes take Pk_wall */
void or_sst_update_Pk(or_case *c);
/* k (which = 0) or omega (1) system into the scalar (N = 1) system:
 * turbulent*SSTAssemblerNodeTerms.cpp (steady), phiAssemblerElemTerms.hpp
 * :10-413 (upwind), relaxation phiAssembler.h:966-1006, boundary rhs
 * relaxation :1008-1152 (x 0.75), omega rows fixed at no-slip walls
 * (turbulentEddyFrequencySSTAssembler.h postAssemble_) */
void or_assemble_scalar(or_case *c, int which);
/* test hook: restrict or_assemble_scalar to some of its parts (bit 1 node
 * transient + divergence terms, 2 SST sources, 4 interior element terms,
 * 8 relaxation / boundary relaxation / wall rows); 15 = all (default) */
void or_set_scalar_terms(or_case *c, int mask);
/* phi += x with the lower clip at SMALL (equation.h:204-286, CLIP = 1),
 * then the field scale (nodeField.hpp:5993) */
void or_correct_scalar(or_case *c, int which);
/* one SST pass after the flow equations (simulation.cpp:314-377 order;
 * segregatedShearStressTransportEquations.cpp preSolve/solve): mut, mu_eff,
 * F1, Pk, k solve + correction, omega solve + correction, grad k, grad omega.
 * ctl = {max_it, rtol} of the Jacobi-GMRES(30) scalar solves; hist = {k, omega
 * control residuals, k, omega solver iterations}. */
void or_sst_iteration(or_case *c, const double *ctl, double *hist);

#ifdef __cplusplus
}
#endif
#endif
