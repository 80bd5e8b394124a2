/*
 * oagpu.h -- C ABI of the MI355X CVFEM assemble-and-solve path.
 *
 * The library replaces, behind OpenAccel's own interfaces, the bodies of
 *   assembler<N>::assemble            src/assemble/assembler.h:46-51
 *   linearSystem<N>::solve            src/equation/linearSystem.h:103-106, 416-462
 *                                     (-> liblinsolve linearSolver::Base<N>::solve)
 * and the field updates the segregated loop performs between them
 *   segregatedFlowEquations::solve    src/equation/flow/segregatedFlowEquations.cpp:76-297
 * for the momentum (block size 3) and pressure-correction (block size 1)
 * systems.  All entry points take plain pointers and sizes, return 0 on
 * success or a non-zero status (then oagpu_last_error() describes it), and
 * never throw.  Host arrays use the reference's STK field layouts:
 *   coordinates / velocity / du / grad p : [node][3]
 *   grad U                               : [node][9], (i*3+j) = dU_i/dx_j
 *   mass flow rate                       : [elem][12]   (hex8 SCS ips)
 *   block-CRS values                     : [nnz][N*N], row-major N x N blocks
 *   rhs / solution                       : [node][N]
 * Local ids are the node indices given at mesh creation; every row is owned.
 */
#ifndef OAGPU_H
#define OAGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OAGPU_OK 0
#define OAGPU_EINVAL 2   /* bad argument / unsupported input          */
#define OAGPU_EHIP 3     /* HIP runtime failure                        */
#define OAGPU_ENOMEM 4   /* allocation failure                         */
#define OAGPU_EGEOM 5    /* non-positive Jacobian met by a kernel      */
#define OAGPU_ENONFINITE 6 /* a linear solve produced a non-finite residual */
#define OAGPU_EINTERNAL 9

/* graph layout, mirroring linearSolver::GraphLayout (src/mesh/graph/nodeGraph.h,
 * controls.h:244 bandwidth_reduction) */
#define OAGPU_STENCIL_REDUCED 1
#define OAGPU_STENCIL_FULL 0
/* or-ed into the stencil argument of oagpu_mesh_create: keep the caller's
 * node and element numbering on the device.  By default the device holds the
 * mesh renumbered for locality (nodes along a Morton curve of their per-axis
 * layer indices, elements by their smallest node), the counterpart of the
 * reference's solver-side reordering (expert parameter bandwidth_reduction,
 * on by default, src/controls/controls.h:244).  Either way every array that
 * crosses this ABI (fields, graphs, scatter maps, systems, solutions, AMG
 * aggregates of the fine level, p_ref_node) is in the caller's numbering. */
#define OAGPU_MESH_CALLER_ORDER 0x100

/* element topologies (nalu node orders of Hex8CVFEM.C, Tet4CVFEM.C,
 * Wed6CVFEM.C, Pyr5CVFEM.C) */
#define OAGPU_HEX8 0
#define OAGPU_TET4 1
#define OAGPU_WED6 2
#define OAGPU_PYR5 3

/* arithmetic form: FP64 everywhere, or FP32 geometry/local matrices summed
 * into FP64 global coefficients */
#define OAGPU_FP64 0
#define OAGPU_MIXED 1

typedef struct oagpu_mesh oagpu_mesh;
typedef struct oagpu_flow oagpu_flow;

/* Physical and numerical parameters of a steady, laminar, constant-property
 * segregated flow case (the hot-path subset of controls.h:57-290). */
typedef struct oagpu_flow_params {
    double rho, mu;        /* density, dynamic viscosity                    */
    double dt;             /* physical_timescale (false transient)          */
    double urf_U, urf_p;   /* velocity / pressure relaxation factors        */
    double urf_mdot;       /* mass flow relaxation (relaxMass_)             */
    double grad_urf;       /* gradient relaxation (nodeField.h:152)         */
    double bnd_relax;      /* boundary rhs relaxation (navierStokesAssembler.cpp:238) */
    int u_shifted;         /* velocity_interpolation_type == linear_linear  */
    int ugrad_shifted;     /* velocity gradient interpolation linear_linear */
    int p_shifted;         /* pressure interpolation linear_linear          */
    int pgrad_shifted;     /* pressure gradient interpolation linear_linear */
    int p_ref_node;        /* pressure-level node id, -1 for none           */
    int precision;         /* OAGPU_FP64 or OAGPU_MIXED                     */
    int p_method;          /* p' solver of oagpu_outer_iteration: OAGPU_CG,
                              OAGPU_CG_AMG, OAGPU_BICGSTAB(_AMG), OAGPU_GMRES */
    int use_graphs;        /* 1: CG iterations replayed from a captured HIP
                              graph; 0: launched one by one (same results)  */
    int high_res;          /* advection_scheme high_resolution: limited
                              deferred correction (blending factors "beta",
                              "beta_k", "beta_w"); 0: upwind                */
} oagpu_flow_params;

/* k-omega SST transport (shearStressTransportModel.h:21-37 constants are
 * built in).  The wall treatment stays with the caller: the wall distance
 * "ymin", omega at no-slip wall nodes (rows fixed there, so it is kept) and
 * the production at wall nodes "Pk_wall" are fields it uploads. */
typedef struct oagpu_sst_params {
    double urf_k, urf_w;   /* implicit relaxation of the k / omega rows      */
    double urf_mut;        /* eddy-viscosity relaxation (mutRef().urf())     */
    int k_shifted, w_shifted;         /* k / omega interpolation linear_linear */
    int kgrad_shifted, wgrad_shifted; /* their gradient operators shifted      */
} oagpu_sst_params;

/* liblinsolve ControlData (linearSystem.h:475-482) as the device solve
 * reports it */
typedef struct oagpu_control_data {
    int n_iterations;
    double solver_initial_res; /* ||b - A x0||_2 */
    double solver_final_res;   /* ||b - A x||_2  */
    double scaled_initial_res[3];
} oagpu_control_data;

/* Krylov methods (linear_solver_settings family: MI355X) */
#define OAGPU_GMRES 0
#define OAGPU_CG 1
#define OAGPU_BICGSTAB 2
#define OAGPU_CG_AMG 3   /* flexible CG preconditioned by one aggregation-AMG
                            cycle: a K-cycle (two inner flexible-CG steps per
                            coarse level) when two or more levels lie above
                            the one-workgroup tail, else a V-cycle (scalar
                            systems; the role of HYPRE BoomerAMG in
                            liblinsolve's hypre family, linearSystem.h:331) */
#define OAGPU_BICGSTAB_AMG 4 /* BiCGStab right-preconditioned by one AMG
                            V-cycle (scalar systems that are not symmetric:
                            p' with static-pressure outlets or on distorted
                            meshes); FP64 in both forms */
#define OAGPU_BICGSTAB_SA 5  /* BiCGStab right-preconditioned by one V(1,1)-cycle
                            of smoothed aggregation (prolongation smoothed by
                            one filtered Jacobi step; Galerkin coarse levels
                            recomputed on the device every solve): the p'
                            systems of wall-resolved meshes, whose cell aspect
                            ratios stall the unsmoothed V-cycle; FP32 inner
                            solve under FP64 refinement in the mixed form */
#define OAGPU_CG_SA 6        /* CG preconditioned by the same smoothed-aggregation
                            V(1,1)-cycle (symmetric scalar systems); FP32 inner
                            solve under FP64 refinement in the mixed form */

/* ---------------------------------------------------------------- runtime */
int oagpu_init(int device);
const char *oagpu_last_error(void);
int oagpu_synchronize(void);

/* --------------------------------------------------------------- mesh */
/* Device mirror of a mesh: interleaved coordinates [n_nodes][3], element
 * connectivity conn [n_elems][8] in the nalu node order of each element's
 * topology (etype [n_elems]: OAGPU_HEX8 / TET4 / WED6 / PYR5, the unused
 * trailing entries of smaller elements ignored; etype NULL: all hex8, the
 * element buckets of the reference's blocks flattened into one list),
 * boundary sides (parent element, side ordinal, boundary id) of any
 * topology: the side ordinals of the reference's master elements (hex8
 * 0..5, tet4 0..3, wedge6 0..4 with tri sides 3 and 4, pyramid5 0..4 with
 * the quad base 4; sideNodeOrdinals_ of Tet4CVFEM.C:416-421,
 * Wed6CVFEM.C:494-501, Pyr5CVFEM.C:759-765), one face ip per side node;
 * per-side arrays of the flow (side mass flows, reversal flags) keep four
 * entries per side, the fourth unused on tri sides.  Builds the CRS node graph
 * (nodeGraph::buildGraph_, src/mesh/graph/nodeGraph.cpp:29-392) with the
 * given stencil, the node->element incidences and element->CRS scatter maps,
 * the dual nodal volumes (mesh::updateDualNodalVolumeField_,
 * src/mesh/meshGeometry.cpp:1277) and per-element SCV weights. */
int oagpu_mesh_create(int n_nodes, const double *xyz, int n_elems, const int *conn,
                      const int *etype, int n_bfaces, const int *bf_elem, const int *bf_ord,
                      const int *bf_bc, int stencil, oagpu_mesh **out);
int oagpu_mesh_destroy(oagpu_mesh *m);
int oagpu_mesh_sizes(const oagpu_mesh *m, int *n_nodes, int *n_elems, int *nnz,
                     int *n_inc, int *n_binc);
/* host copies of the CRS graph (CRSNodeGraph offsets/indices, diagOffsetRef) */
int oagpu_graph_copy(const oagpu_mesh *m, int *row_ptr, int *cols, int *diag_off);
/* host copies of the node->element incidences ((elem<<3)|local node) and of
 * their scatter maps (8 bytes per incidence: row position of each element
 * node, 0xFF = lumped into the diagonal block) */
int oagpu_scatter_map_copy(const oagpu_mesh *m, int *inc_ptr, int *inc,
                           unsigned long long *slots);
/* dual nodal volumes [n_nodes] */
int oagpu_mesh_volume_copy(const oagpu_mesh *m, double *vol);

/* Master-element geometry of n_elems elements of one topology on the device
 * (MasterElementWork.F *_scs_det, *_scv_det, *_gradient_operator with the
 * ip tables of the *CVFEM.C master elements): coords [n][npe][3] in,
 * areav [n][nscs][3], scv [n][npe], dndx [n][nscs][npe][3] at the standard
 * (shifted = 0) or shifted SCS ips, detj [n][nscs] out (NULL: not wanted).
 * npe / nscs: hex8 8 / 12, tet4 4 / 6, wedge6 6 / 9, pyramid5 5 / 12. */
int oagpu_element_geometry(int topology, int shifted, int n_elems, const double *coords,
                           double *areav, double *scv, double *dndx, double *detj);

/* The same geometry on caller-provided DEVICE arrays (same layouts), with no
 * device or host allocation: for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  The kernels' error word lives in a
 * fixed 256-byte device array of the library, one per GPU, held until the
 * process ends; it is cleared by a kernel at the start of each call, so
 * calls that overlap on one GPU race on it.  stream: a hipStream_t (NULL:
 * the null stream); the call synchronizes it.  Batches above
 * OAGPU_NOALLOC_MAX_ELEMS elements are rejected (OAGPU_EINVAL); a
 * non-positive Jacobian returns OAGPU_EGEOM, the outputs still written. */
#define OAGPU_NOALLOC_MAX_ELEMS (1 << 30)
int oagpu_element_geometry_noalloc(int topology, int shifted, int n_elems, const double *coords,
                                   double *areav, double *scv, double *dndx, double *detj,
                                   void *stream);

/* --------------------------------------------------------------- flow */
/* boundary physical types (the zone boundaries' types of the reference,
 * boundaryPhysicalType): no-slip wall, velocity inlet, static-pressure
 * outlet, symmetry plane, and slip wall (a wall whose velocity condition is
 * boundaryConditionType::slip: no face terms in the momentum equations,
 * navierStokesAssemblerElemBoundaryConditions.cpp:2376-2657; the symmetry
 * treatment of p', pressureCorrectionAssemblerElemBoundaryConditions.cpp:49,
 * which adds nothing on a static mesh; no side mass flow; its nodes' rhs
 * relaxed like other boundary nodes in steady runs, navierStokesAssembler.cpp:289;
 * k and omega zero-gradient), and static-pressure opening
 * (boundaryPhysicalType::opening with a specifiedDirection velocity and a
 * staticPressure pressure condition: side mass flows rho U_ip . A, not
 * clipped, updated by Rhie-Chow with the nodal pressure, flowModel.cpp:
 * 4817-4992 / 8094-8380; no reversal flags, :15575; momentum outflow ips as
 * an outlet, inflow ips entrained along the bc_U direction,
 * navierStokesAssemblerElemBoundaryConditions.cpp:3234-3552; p' the
 * outlet's terms at every ip, pressureCorrectionAssemblerElemBoundary
 * Conditions.cpp:3026-3560; k and omega implicit upwind,
 * phiAssemblerElemBoundaryConditions.hpp:1439-1760).  An outlet or opening
 * replaces the pressure-reference row. */
#define OAGPU_BC_WALL 0
#define OAGPU_BC_INLET 1
#define OAGPU_BC_OUTLET 2
#define OAGPU_BC_SYMMETRY 3
#define OAGPU_BC_SLIP 4
#define OAGPU_BC_OPENING 5

/* bc_U[n_bc][3]: velocity of each boundary id (no-slip walls: the wall
 * velocity; velocity inlets: the specified inlet velocity; openings: the
 * entrainment direction, normalised, or all zero for the inward normal of
 * each face ip, velocity.cpp:4813-4970). */
int oagpu_flow_create(oagpu_mesh *m, const oagpu_flow_params *prm, int n_bc,
                      const double *bc_U, oagpu_flow **out);
int oagpu_flow_destroy(oagpu_flow *f);
/* fields: "U" [nn][3], "p" [nn], "gradU" [nn][9], "gradP" [nn][3],
 * "du" [nn][3], "mdot" [ne][12], "div" [nn], "mu" [nn] (effective
 * viscosity), the SST fields once enabled, and with inlets, outlets or
 * openings "msd" [n_bfaces][4] (side mass flow rate per boundary face ip, outward
 * positive) */
int oagpu_field_upload(oagpu_flow *f, const char *name, const double *host, size_t count);
int oagpu_field_download(oagpu_flow *f, const char *name, double *host, size_t count);

/* navierStokesEquation::initialize/postInitialize (navierStokesEquation.cpp:86-133),
 * pressureCorrectionEquation::initialize (pressureCorrectionEquation.cpp:59-70):
 * gradients, scales, initial mass flow rate, mass divergence. */
int oagpu_initialize(oagpu_flow *f);

/* momentum system (navierStokesEquation::solve :159-173): node terms,
 * interior element terms, no-slip walls, implicit relaxation, boundary rhs
 * relaxation, then du = V / (a_ii + SMALL) (navierStokesAssembler.cpp:84-232) */
int oagpu_assemble_momentum(oagpu_flow *f);
/* pressure-correction system (pressureCorrectionEquation::solve :89-107)
 * including the pressure-reference row (pressureCorrectionAssembler.cpp:19-63) */
int oagpu_assemble_pcorr(oagpu_flow *f);
/* host copy of a system: vals [nnz][N*N], b [nn][N] */
int oagpu_system_download(oagpu_flow *f, int N, double *vals, double *b);
int oagpu_system_upload(oagpu_flow *f, int N, const double *vals, const double *b);
int oagpu_solution_download(oagpu_flow *f, int N, double *x);

/* field updates */
int oagpu_update_mdot(oagpu_flow *f);          /* flowModel.cpp:5272-5688   */
int oagpu_mass_divergence(oagpu_flow *f);      /* flowModel.cpp:13541-13775 */
int oagpu_gradient(oagpu_flow *f, int which);  /* nodeField.hpp:3253; 0 U, 1 p, 2 k, 3 omega */
int oagpu_correct(oagpu_flow *f, int N, double urf); /* equation.h:204-286  */
int oagpu_velocity_correction(oagpu_flow *f, int step); /* segregatedFlowEquations.cpp:133-279 */
int oagpu_update_scales(oagpu_flow *f, double *U_scale, double *p_scale); /* nodeField.hpp:5993 */
/* Physical type of each boundary id (default: every id a no-slip wall).
 * Symmetry planes (navierStokesAssemblerElemBoundaryConditions.cpp:2082,
 * symmetry projection and normal relaxation navierStokesAssembler.cpp:387-696,
 * gradient corrections nodeVectorField.cpp:46 / nodeScalarField.cpp:45, the
 * assembled symmetry area meshGeometry.cpp:1787) carry no mass flow and no
 * scalar flux.  Velocity inlets (the bc_U row; zero-gradient pressure) and
 * static-pressure outlets (zero-gradient velocity, k, omega) carry side mass
 * flows (flowModel.cpp:3802-4997 initial, :6470-6598 inlet and :7323-7690
 * outlet Rhie-Chow updates, :15226-15454 outlet reversal flags), and so do
 * static-pressure openings; with an outlet or opening the p' system has no
 * pressure-reference row.  Other types are rejected (OAGPU_EINVAL).  Resets
 * the side mass flows and reversal flags. */
int oagpu_flow_set_bc_types(oagpu_flow *f, int n_bc, const int *type);
/* Per boundary id: p[n_bc] outlet / opening static pressure, k[n_bc] / w[n_bc] inlet
 * turbulent kinetic energy / eddy frequency (the specified values of the
 * outlet's pressure and the inlet's k and omega boundary conditions);
 * NULL leaves that table unchanged; all zero until set. */
int oagpu_flow_set_bc_values(oagpu_flow *f, int n_bc, const double *p, const double *k,
                             const double *w);
/* updateFlowReversalFlag (flowModel.cpp:15226-15454) on its own, and the
 * flags [n_bfaces][4] (1: the outlet face ip is reversed, a slip wall) */
int oagpu_update_reversal(oagpu_flow *f);
int oagpu_reversal_download(oagpu_flow *f, int *flags, size_t count);
/* nodeField::updateBlendingFactorField (nodeField.hpp:2069): 0 U, 2 k, 3 omega */
int oagpu_update_blending(oagpu_flow *f, int which);
/* control residual source: RMS over rows of b / a_diag per component */
int oagpu_scaled_residual(oagpu_flow *f, int N, double *out);

/* Solve A x = b of system N (3 momentum, 1 pressure correction) from x = 0
 * with Krylov `method` (Jacobi-preconditioned, or AMG-preconditioned CG);
 * converged when ||r|| <= max(rtol ||r0||, atol) or after max_it iterations.
 * Replaces linearSolver::Base<N>::solve behind linearSystem<N>::solveSystem_
 * (linearSystem.h:416-462). */
int oagpu_solve(oagpu_flow *f, int N, int method, int max_it, double rtol, double atol,
                int restart, oagpu_control_data *cd);
/* AMG hierarchy of the scalar system (built by the first OAGPU_CG_AMG solve):
 * number of levels and rows per level (sizes[16]). */
int oagpu_amg_info(oagpu_flow *f, int *n_levels, int *sizes);
/* level >= 1: aggregate of every row of level-1 (agg[rows of level-1]), the
 * level's CRS pattern and its values from the last set-up, in the precision
 * of that set-up (NULL pointers are skipped). */
int oagpu_amg_level_copy(oagpu_flow *f, int level, int *agg, int *row_ptr, int *cols,
                         double *vals);

/* Smoothed-aggregation hierarchy of a flow (built by its first
 * OAGPU_BICGSTAB_SA solve): number of levels, rows, entries of each level's
 * matrix and of its prolongation P (to the level above), and the smoothing
 * weight omega_p of each P (arrays of up to 16 levels; NULL: not wanted). */
int oagpu_sa_info(oagpu_flow *f, int *n_levels, int *sizes, long long *p_nnz, long long *a_nnz,
                  double *omega_p);
/* level >= 1: the aggregates of the rows of level - 1, the prolongation P
 * from this level to level - 1 (CRS over the rows of level - 1) and this
 * level's matrix (CRS), as of the last numeric set-up (FP64).  Rows of level
 * 0 are the mesh nodes, in the caller's numbering. */
int oagpu_sa_level_copy(oagpu_flow *f, int level, int *agg, int *p_row_ptr, int *p_cols,
                        double *p_vals, int *a_row_ptr, int *a_cols, double *a_vals);

/* One steady outer iteration of segregatedFlowEquations::solve.
 * ctl = {momentum max_it, momentum rtol, p' max_it, p' rtol}; momentum by
 * GMRES, p' by the flow's p_method;
 * hist[6] = U, V, W, P control residuals, momentum / p' Krylov iterations. */
int oagpu_outer_iteration(oagpu_flow *f, const double *ctl, double *hist);

/* Outer iterations of simulation::run's loop (simulation.cpp:166-204), each
 * one HIP graph launch: segregatedFlowEquations::solve
 * (segregatedFlowEquations.cpp:76-297) and, with sctl, the k-omega SST pass
 * of segregatedShearStressTransportEquations::solve (:86-99) as
 * oagpu_sst_iteration runs it, replayed from one captured graph with no host
 * read inside the iteration.  ctl as oagpu_outer_iteration's, sctl =
 * {k / omega GMRES max_it, rtol} or NULL (flow only).  The flow's p_method
 * must be OAGPU_BICGSTAB_SA or (FP64 only) OAGPU_CG_SA; momentum, k and
 * omega are solved by GMRES(30).  Runs up to n iterations, fewer when tol > 0
 * and every reported scaled residual falls below tol; hist (NULL: not wanted)
 * [n][10] = U, V, W, P residuals, momentum / p' iterations, k, omega
 * residuals, k / omega iterations (0 without sctl); *done = iterations run.
 * The flow's first such iteration runs without a graph (it builds the AMG
 * hierarchy and sizes the graph's iteration caps); an iteration whose solve
 * needs more than its cap is rolled back and run again uncapped, so results
 * never depend on the caps.  eager != 0: the same device-controlled launch
 * sequence without graphs (bitwise-equal results).  No reference
 * counterpart beyond the loop it replaces. */
int oagpu_graph_iterations(oagpu_flow *f, const double *ctl, const double *sctl, int n,
                           double tol, int eager, double *hist, int *done);
/* oagpu_batch_outer_iterations with every case advanced by
 * oagpu_graph_iterations (each outer iteration one graph launch in the
 * case's stream); hist [n][iters][10] as oagpu_graph_iterations'. */
int oagpu_batch_graph_iterations(oagpu_flow *const *flows, int n, const double *ctl, int iters,
                                 double tol, int threads, double *hist, int *done,
                                 double *seconds);
/* Counters of a flow's graph iterations: graph launches, iterations rolled
 * back and run uncapped, graph captures; caps[4] = iteration slots of the
 * momentum, p', k and omega solves in the current graph (NULL: not wanted). */
int oagpu_graph_stats(oagpu_flow *f, long long *launches, long long *redone, long long *captures,
                      int *caps);

/* Batched independent cases (many small flows of any meshes, sizes and
 * parameters): each of the n flows is advanced by up to `iters` outer
 * iterations (oagpu_outer_iteration, in its own HIP stream) on a pool of
 * `threads` host threads (<= 0: min(n, 16)), so that the cases' kernels
 * overlap on the GPU.  A case stops early once all four scaled residuals
 * it reports fall below tol (tol <= 0: never).  Outputs (NULL: not wanted):
 * hist [n][iters][6] as oagpu_outer_iteration's, done [n] iterations run,
 * seconds [n] wall time of each case.  Every case runs; the first failing
 * case's status and message are returned.  No counterpart call in the
 * reference, which runs one case per process. */
int oagpu_batch_outer_iterations(oagpu_flow *const *flows, int n, const double *ctl, int iters,
                                 double tol, int threads, double *hist, int *done,
                                 double *seconds);

/* ------------------------------------------------------------ transient */
#define OAGPU_BDF1 1
#define OAGPU_BDF2 2
/* Start a physical time step of length dt (timestep loop of
 * simulation::run with transient_scheme first / second order backward
 * Euler): U, and k and omega with SST, move back one time level (fields
 * "Uold" <- "U", "Uold2" <- "Uold", likewise "kold", "wold"), and the
 * momentum and k / omega assemblies that follow add the BDF time terms
 * (navierStokesAssemblerNodeTerms.cpp:151-400, phiAssemblerNodeTerms.hpp:
 * 150-404) in place of the steady pseudo-time term: c0 rho V / dt on the
 * diagonal, -(c0 phi + c1 phi_old + c2 phi_old2) rho V / dt in the rhs, with
 * BDF1 {1, -1} and BDF2 {(1+2w)/(1+w), -(1+w), w^2/(1+w)}, w = dt / dt_prev
 * (BDF1 coefficients on the first step), rho constant in time.  The outer
 * iterations of the step then run as in the steady case. */
int oagpu_time_step(oagpu_flow *f, int scheme, double dt);

/* ---------------------------------------------------------- k-omega SST */
/* allocate the SST fields: "k", "omega", "gradK", "gradW", "ymin", "mut",
 * "F1", "Pk", "Pk_wall", "kprev"; momentum then uses "mu" = mu + mu_t */
int oagpu_sst_enable(oagpu_flow *f, const oagpu_sst_params *s);
/* shearStressTransportModel::updateTurbulentDynamicViscosity (:597-690),
 * turbulenceModel::updateEffectiveDynamicViscosity (turbulenceModel.cpp:121),
 * updateFOneBlending (:95-174), updateTurbulentProduction (:176-340), fused */
int oagpu_sst_update(oagpu_flow *f);
/* k (which 0) / omega (1) system into the scalar system (N = 1): node terms
 * and SST sources (turbulent*SSTAssemblerNodeTerms.cpp), interior
 * advection-diffusion (phiAssemblerElemTerms.hpp:10-413), relaxation,
 * boundary rhs relaxation, omega rows fixed at no-slip walls
 * (turbulentEddyFrequencySSTAssembler.h postAssemble_) */
int oagpu_assemble_scalar(oagpu_flow *f, int which);
/* phi += x clipped below at SMALL, then its scale (equation.h:204-286) */
int oagpu_correct_scalar(oagpu_flow *f, int which);
/* one SST pass after the flow equations (segregatedShearStressTransportEquations
 * preSolve + solve): mu_t, mu_eff, F1, P_k, k and omega assembled and solved
 * by Jacobi-GMRES(30) (ctl = {max_it, rtol}), corrected, their gradients;
 * hist[4] = k, omega control residuals, k, omega iterations */
int oagpu_sst_iteration(oagpu_flow *f, const double *ctl, double *hist);

/* Autotuner of the flow's tunable launches (no counterpart in the
 * reference): times the candidate variants on the flow's current mesh and
 * state and keeps the fastest for the flow's later launches -- momentum rows
 * with their off-diagonal blocks summed in an LDS tile (64 rows per
 * workgroup) or accumulated in HBM (256 rows per workgroup; hex-only meshes,
 * upwind), p' rows at most 2 waves per SIMD or at the compiler's occupancy,
 * and the Krylov SpMV at 64 / 128 / 256 rows per workgroup (FP64, and FP32
 * for the mixed form).  Assembles the momentum and p' systems of the current
 * state as a side effect.  report (n_report >= 18 doubles, NULL: none):
 * per tunable -- momentum, pcorr, spmv1, spmv3, spmv1_f32, spmv3_f32,
 * scalar, mdot, grad -- the chosen variant (1 tile / 0 HBM; 0 limited / 1
 * compiler; rows per workgroup; resident waves per CU, 0 uncapped), its time
 * and the default variant's time in ms. */
int oagpu_autotune(oagpu_flow *f, double *report, int n_report);
/* Set one tunable's variant directly (e.g. from an earlier oagpu_autotune
 * report): which = "momentum" (1 / 0), "pcorr" (0 / 1), "spmv1", "spmv3",
 * "spmv1_f32", "spmv3_f32" (64 / 128 / 256), "scalar", "mdot", "grad"
 * (resident waves per CU: 0, 4, 8 or 16), "scalar_pair" (1: k and omega,
 * and their gradients, assembled by one launch each pair when their
 * operators agree; 0: one launch per equation), "graph_cap" (iteration
 * slots of every solve in the next captured graph of oagpu_graph_iterations,
 * after its first iteration; a solve needing more rolls back and raises
 * them). */
int oagpu_tune_set(oagpu_flow *f, const char *which, int variant);

/* Device time of the library's main kernels, measured with HIP events on the
 * flow's stream (no counterpart in the reference, whose only instruments are
 * the per-iteration CPU time and a profiler region around each solve).
 * oagpu_profile(f, which) clocks "all", "none" or one of: "momentum",
 * "pcorr", "scalar" (k / omega rows), "mdot", "gradU", "gradP" (the scalar
 * gradients but grad k and grad omega computed together), "gradKW" (grad k
 * and grad omega by one launch), "amg_pre" / "amg_post" (the fine level's pre-smoothing +
 * residual and prolongation + post-smoothing kernels), "amg_tail" (the
 * one-workgroup coarse levels), "spmv1" / "spmv3" (Krylov products in
 * FP64), "spmv1_f32" / "spmv3_f32" (FP32 inner solves of the mixed form), and
 * resets the totals.  Launches replayed from captured HIP graphs are not
 * clocked.  Nothing synchronises until oagpu_kernel_stats, which returns the
 * total device time (ms) and the launch count since the last oagpu_profile. */
int oagpu_profile(oagpu_flow *f, const char *which);
int oagpu_kernel_stats(oagpu_flow *f, const char *which, double *total_ms, long long *launches);

#ifdef __cplusplus
}
#endif
#endif
