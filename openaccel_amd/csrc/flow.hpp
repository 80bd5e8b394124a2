// flow.hpp -- device state of a segregated pressure-velocity case.
#pragma once

#include <algorithm>

#include "../../include/oagpu.h"
#include "mesh.hpp"

#include <array>
#include <vector>

namespace oagpu {

constexpr int kRedGrid = 2048;  // fixed partition of every reduction: deterministic sums
constexpr int kRedBlock = 256;

// One coarse level of the aggregation AMG hierarchy (values sliced-ELL like
// the fine matrix).  Level l's rows are aggregates of level l-1's rows.
struct AmgLevel {
    int n = 0;
    int nslice = 0;
    size_t nsell = 0;
    DBuf<int> slice_ptr, slice_w, cols, diag_off;
    std::vector<int> h_row_ptr, h_cols, h_agg;  // host pattern and aggregation
    DBuf<int> agg;           // [rows of level l-1] -> row of level l
    DBuf<int> mem_ptr, mem;  // rows of level l-1 aggregated into each row (ascending)
    DBuf<int> mpos;          // [rows of level l-1] -> position in mem
    DBuf<int> gptr, glist;   // per sliced-ELL slot here: the level l-1 slots summed into it, in order
    DBuf<double> A, dinv, X, T, R, S;
    DBuf<float> fA, fdinv, fX, fT, fR, fS;
    DBuf<double> C1, C2, V1, R2;  // K-cycle vectors (levels solved by two inner FCG steps)
    DBuf<float> fC1, fC2, fV1, fR2;
};

// One level of the smoothed-aggregation hierarchy (amg.hip: sa_build).  Level
// l holds its matrix (levels >= 1, CSR; level 0 is the flow's scalar system),
// the prolongation P_l from level l + 1 to level l, and the lists that
// compute P_l, B_l = A_l P_l and A_{l+1} = P_l^T B_l on the device.
struct SaLevel {
    int n = 0;
    // matrix of this level (levels >= 1): CSR, the diagonal's index per row,
    // row kernels with `lanes` lanes per row
    int lanes = 4;
    DBuf<int> a_rp, a_ci, a_diag;
    DBuf<double> A, dinv;
    DBuf<float> fA, fdinv;
    // P_l (levels below the last): sliced-ELL over this level's rows (padding
    // columns -1) for the prolongation; per CSR entry of P_l its slot there,
    // whether its column is the row's own aggregate, and the couplings of
    // A_l (value indices) summed into it
    double omega_p = 0.0;
    size_t p_nsell = 0;
    DBuf<int> p_slice_ptr, p_slice_w, p_cols, prp, ppos, pt_ptr, pt_idx;
    DBuf<unsigned char> own;
    DBuf<double> psum, P;
    DBuf<float> fP;
    // B_l = A_l P_l, CSR over this level's rows: entry e = sum over its terms
    // (value index of A_l, slot of P_l)
    size_t nb = 0;
    int b_lanes = 1;  // lanes per entry of B's and A_{l+1}'s term sums (1, 4 or 16)
    int c_lanes = 1;
    DBuf<int> b_ptr;
    DBuf<int2> b_terms;
    DBuf<double> B;
    // A_{l+1} in CSR order: entry e = sum over its terms (slot of P_l, entry of B_l)
    size_t nac = 0;
    DBuf<int> c_ptr;
    DBuf<int2> c_terms;
    // restriction P_l^T: row I of level l + 1 = sum over (row of level l,
    // slot of P_l), r_lanes lanes per row
    int r_lanes = 4;
    DBuf<int> r_ptr;
    DBuf<int2> r_terms;
    // cycle vectors: right-hand side, pre-smoothed x, residual, prolongated t, result
    DBuf<double> R, X, S, T, Y;
    DBuf<float> fR, fX, fS, fT, fY;
    // host copies (tests): CSR pattern of this level's matrix; aggregates
    // and CSR pattern of P_l
    std::vector<int> h_rp, h_cols, h_agg, h_prp, h_pcols;
};

struct Sa {
    bool built = false;
    std::vector<SaLevel> lv;
    int dense_n = 0;           // coarsest level solved by its dense inverse when > 0
    DBuf<double> dense;
    DBuf<float> fdense;
};

struct Amg {
    bool built = false;
    bool last_float = false;   // precision of the last numeric set-up
    std::vector<AmgLevel> lv;  // lv[0]: fine level (pattern of the mesh graph; vectors only)
    int dense_n = 0;           // coarsest level solved by its dense inverse when > 0
    DBuf<double> dense;        // [dense_n][dense_n] inverse of the coarsest matrix
    DBuf<float> fdense;
};

// Kernels whose device time the library can measure (oagpu_profile /
// oagpu_kernel_stats): each clocked launch records a pair of HIP events on
// the flow's stream from a per-kernel pool; nothing waits on them until the
// host asks for the totals.  Launches captured into a HIP graph are not
// clocked.
enum ClockId {
    kClkMomentum, kClkPcorr, kClkScalar, kClkMdot, kClkGradU, kClkGradP, kClkAmgPre,
    kClkAmgPost, kClkAmgTail, kClkSpmv1, kClkSpmv3, kClkSpmv1f, kClkSpmv3f, kClkGradKW, kClkCount
};
inline const char* clock_name(int id) {
    static const char* names[kClkCount] = {"momentum", "pcorr", "scalar", "mdot", "gradU",
                                           "gradP", "amg_pre", "amg_post", "amg_tail", "spmv1",
                                           "spmv3", "spmv1_f32", "spmv3_f32", "gradKW"};
    return names[id];
}
struct KernelClock {
    std::vector<hipEvent_t> ev;  // pairs (start, stop)
    size_t used = 0;             // pairs recorded and not yet folded in
    double total_ms = 0.0;
    long long launches = 0;
    void fold() {
        for (size_t k = 0; k < used; ++k) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]) == hipSuccess) total_ms += ms;
        }
        launches += (long long)used;
        used = 0;
    }
    ~KernelClock() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

// An instantiated HIP graph of a block of CG iterations and the operands it
// was captured with.
struct CgGraph {
    hipGraphExec_t exec = nullptr;
    std::array<const void*, 10> key{};
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
    ~CgGraph() { reset(); }
};

// The whole outer iteration captured as one HIP graph (graph.hip): the
// instantiated graph and what it was captured with, the iteration-slot caps
// of its four Krylov solves (momentum, p', k, omega), and the device state
// its one-workgroup control kernels keep.
struct IterGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<double> key;    // controls, caps and operands of the captured sequence
    int cap[4] = {0, 0, 0, 0};  // iteration slots per solve in the graph
    int seen[4] = {0, 0, 0, 0}; // largest iteration count of each solve since the last re-capture
    int since = 0;              // iterations since the caps were last set
    bool primed = false;        // one uncaptured iteration has run (hierarchy built, caps known)
    long long launches = 0, replays_redone = 0, captures = 0;
    DBuf<double> sd;            // Krylov scalars of the running solve (GMRES H, g, rotations; BiCGStab rho, alpha, omega; tolerances)
    DBuf<int> si;               // its flags and counters
    DBuf<double> rec;           // per-iteration record: raw residuals, scales, iterations, caps exceeded
    DBuf<double> sc;            // field scales U, p, k, omega
    DBuf<double> snap;          // saved iteration state (restored when a solve exceeds its cap)
    DBuf<unsigned long long> tab;  // copy table of the saved state: {src, dst, 32-bit words}
    int ntab = 0;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
    }
    ~IterGraph() { reset(); }
};

struct Flow {
    DevMesh* m = nullptr;
    oagpu_flow_params prm{};
    int nbc = 0;
    DBuf<double> bc_U;                   // [nbc][3]
    // nodal / element fields, reference layouts
    DBuf<double> U, p, gradU, gradP, du, mdot, div;
    // systems
    DBuf<double> A3, b3, x3, A1, b1, x1;
    // Krylov workspace
    DBuf<double> dinv, r, z, pk, q, w, t, s, r0, ph, sh;
    DBuf<double> V;                      // GMRES basis [(m+1)][n]
    DBuf<float> Af, fb, fx, fr, fz, fp, fq, fdinv;  // FP32 inner solve of the mixed form
    DBuf<float> fV, fw;                  // FP32 GMRES basis and work vector
    DBuf<float> fr0, fs, ft, fsh;        // FP32 AMG-BiCGStab vectors
    DBuf<double> part;                   // reduction partials [64][kRedGrid]
    DBuf<double> red;                    // reduced values [64]
    DBuf<double> scal;                   // device-resident Krylov scalars
    DBuf<int> flag;                      // device-resident convergence flag / counters
    DBuf<int> err;                       // this flow's device error word (non-positive Jacobians)
    int V_cols = 0;
    double U_scale = 1.0, p_scale = 1.0;
    hipStream_t stream = nullptr;
    // launch variants chosen by oagpu_autotune (defaults: the hand-picked ones)
    struct Tuning {
        int mom_tile = 1;        // momentum rows: 1 LDS tile, 0 accumulate in HBM
        int pc_occupancy = 0;    // p' rows: 0 at most 2 waves per SIMD, 1 compiler's choice
        unsigned spmv_block[2][2] = {{256, 256}, {256, 256}};  // [N == 3][FP32]
        // resident waves per CU of the k / omega rows, the mass-flow update and
        // the gradients (0: as many as registers and LDS allow), capped by
        // reserving LDS per workgroup: a smaller window of rows in flight per
        // XCD keeps more of their gathered node and element data in its L2
        int cap_scalar = 0, cap_mdot = 0, cap_grad = 0;
        // k and omega assembled by one launch sharing the geometry (1), or
        // by one launch each (0)
        int scalar_pair = 1;
        bool tuned = false;
    } tune;
    // device time of the clocked kernels (profiling on only)
    unsigned prof_mask = 0;
    KernelClock clk[kClkCount];
    Amg amg;
    Sa sa;
    CgGraph cg_graph;
    IterGraph ig;
    // effective viscosity mu + mu_t per node (mu alone while laminar)
    DBuf<double> mueff;
    // blending factors of the limited deferred correction (high resolution)
    DBuf<double> beta, beta_k, beta_w;
    // boundary physical types per boundary id (OAGPU_BC_*), no-slip wall
    // nodes, and the assembled symmetry area vector per node (zero off the
    // symmetry planes; allocated only when a symmetry plane exists)
    std::vector<int> h_bc_type;
    DBuf<int> bc_type, wnode;
    DBuf<double> symA;
    bool has_sym = false;
    bool has_slip = false;     // some boundary id is a slip wall
    // velocity inlets, static-pressure outlets and openings: the boundary
    // faces of these types, the outlet / opening static pressure and inlet
    // k / omega per boundary id, and per face ip the side mass flow rate and
    // (outlets) the flow-reversal flag; allocated only when such a boundary
    // exists
    bool has_io = false, has_outlet = false, has_opening = false;
    int nio = 0, nio_row = 0;
    DBuf<int> io_face, io_row;  // those faces; nodes nearest to one of their ips
    std::vector<double> h_bc_p, h_bc_k, h_bc_w;
    DBuf<double> bc_p, bc_k, bc_w;
    DBuf<double> msd;  // [nbf][4]
    DBuf<int> rfl;     // [nbf][4]
    // k-omega SST state (oagpu_sst_enable)
    bool sst_on = false;
    oagpu_sst_params sst{};
    DBuf<double> k, om, gradK, gradW, ymin, mut, F1, Pk, Pk_wall, kprev;  // om: omega
    DBuf<double> A1w, b1w;  // the omega system while k is solved (k and omega assembled together)
    double k_scale = 1.0, w_scale = 1.0;
    // transient (oagpu_time_step): scheme 0 steady, 1 BDF1, 2 BDF2; the time
    // step and the previous one (0 on the first step); U, k and omega of the
    // previous two time levels
    int tscheme = 0, tsteps = 0;
    double tdt = 0.0, tdt_prev = 0.0;
    DBuf<double> Uold, Uold2, kold, kold2, wold, wold2;
};

// BDF coefficients of the current time step (types.cpp:392-411: BDF1 {1, -1};
// BDF2 with omega = dt / dt_prev; BDF1 on the first step of BDF2)
inline void bdf_coeff(const Flow& f, double c[3]) {
    c[0] = 1.0;
    c[1] = -1.0;
    c[2] = 0.0;
    if (f.tscheme == 2 && f.tdt_prev > 0.0) {
        const double om = f.tdt / f.tdt_prev;
        c[0] = (1.0 + 2.0 * om) / (1.0 + om);
        c[1] = -(1.0 + om);
        c[2] = om * om / (1.0 + om);
    }
}

// Dynamic LDS bytes per workgroup of `block` threads (with `static_lds` bytes
// of its own) that hold a CU's 160 KiB at `cap` resident waves (0: none).
inline unsigned occupancy_lds(unsigned block, unsigned static_lds, int cap) {
    if (cap <= 0) return 0;
    const unsigned wg_per_cu = std::max(1u, unsigned(cap) * 64u / block);
    const unsigned per = 163840u / wg_per_cu;
    return per > static_lds + 256 ? std::min(per - static_lds - 256, 64512u - static_lds) : 0;
}

// start / stop of a clocked launch on the flow's stream (no-ops unless the
// flow's profiling covers the kernel); a pool of 4096 pairs is folded in
// (after a stream synchronisation) when full.  Launches captured into a
// graph are not clocked (the flow's clocks are off while it is captured).
inline void clock_begin(Flow& f, int id) {
    if (!(f.prof_mask & (1u << id))) return;
    KernelClock& c = f.clk[id];
    if (c.used == 4096) {
        OAGPU_HIP(hipStreamSynchronize(f.stream));
        c.fold();
    }
    while (c.ev.size() < 2 * (c.used + 1)) {
        hipEvent_t e = nullptr;
        OAGPU_HIP(hipEventCreate(&e));
        c.ev.push_back(e);
    }
    OAGPU_HIP(hipEventRecord(c.ev[2 * c.used], f.stream));
}
inline void clock_end(Flow& f, int id) {
    if (!(f.prof_mask & (1u << id))) return;
    KernelClock& c = f.clk[id];
    OAGPU_HIP(hipEventRecord(c.ev[2 * c.used + 1], f.stream));
    ++c.used;
}

// kernels.hip
void launch_mesh_geometry(DevMesh& m, hipStream_t s);
void launch_mdot_init(Flow& f);
void launch_update_mdot(Flow& f);
void launch_mass_divergence(Flow& f);
void launch_gradient(Flow& f, int which);
void launch_gradient_kw(Flow& f);           // grad k and grad omega, one launch when they can share it
void launch_blending(Flow& f, int which);   // 0 U, 2 k, 3 omega
void launch_boundary_types(Flow& f);        // wall-node flags and symmetry areas
// bc.hip: inlet / outlet side mass flow rates (init: from U; update: inlet
// specified, outlet Rhie-Chow, relaxed) and the outlets' reversal flags
void launch_side_mdot(Flow& f, bool init);
void launch_reversal(Flow& f);
void launch_assemble_momentum(Flow& f);
void launch_assemble_pcorr(Flow& f);
void launch_correct(Flow& f, int N, double urf);
void launch_velocity_correction(Flow& f, int step);
void field_scales(Flow& f, double* U_scale, double* p_scale);
void scaled_residual(Flow& f, int N, double* out);
int krylov_solve(Flow& f, int N, int method, int max_it, double rtol, double atol, int restart,
                 oagpu_control_data* cd);
// k-omega SST (sst.hip, rows.hip)
void launch_sst_update(Flow& f);               // mu_t, mu_eff, F1, P_k
void launch_assemble_scalar(Flow& f, int which);  // 0 k, 1 omega -> system N = 1
// k -> system N = 1 and omega -> A1w / b1w in one launch (geometry shared);
// possible when both use the same gradient and interpolation operators
bool scalar_pair_ok(const Flow& f);
void launch_assemble_scalar_pair(Flow& f);
void launch_correct_scalar(Flow& f, int which);   // phi += x, clipped; field scale
// topo.hip: master-element geometry of a batch of elements of one topology
// (0 hex8, 1 tet4, 2 wedge6, 3 pyramid5), device arrays
int topology_npe(int topo);
void launch_scv_generic(DevMesh& m, hipStream_t s);  // SCV volumes of non-hex elements
void launch_geo_generic(DevMesh& m, hipStream_t s);  // static ip geometry of non-hex elements
void launch_side_geo(DevMesh& m, hipStream_t s);     // face-ip geometry of their boundary sides
int topology_nscs(int topo);
void element_geometry(int topo, int shifted, int n, const double* coords, double* A, double* V,
                      double* G, double* detj, int* err, hipStream_t s);
// The same on caller-provided device arrays, allocating nothing: the error
// word lives in the file-scope scratch array of topo.hip (one per GPU, shared
// by every call on that GPU: overlapping calls race).  Synchronizes stream s;
// returns the error word (non-zero: a non-positive Jacobian).
int element_geometry_noalloc(int topo, int shifted, int n, const double* coords, double* A,
                             double* V, double* G, double* detj, hipStream_t s);
// device address of the current GPU's scratch array (hipGetSymbolAddress,
// looked up once per device and cached)
int* scratch_device_ptr();
// autotuning (tune.hip): time the candidates of each tunable launch on the
// flow's current state and keep the fastest; report: per tunable kernel
// (momentum, pcorr, spmv1, spmv3, spmv1_f32, spmv3_f32) the chosen variant,
// its time and the default's time, ms
void autotune(Flow& f, double* report, int n_report);
void time_spmv_variants(Flow& f, int N, bool f32, double* best_ms, double* default_ms);

// krylov.hip: CG launches of the graph-captured iteration (no host reads):
// set-up (x = 0, r = b, first preconditioner application, tolerance, flags
// in f.flag / f.scal) and one iteration; sa: smoothed-aggregation V-cycle,
// else Jacobi
void cg_device_begin(Flow& f, int max_it, double rtol, double atol, bool sa);
void cg_device_iteration(Flow& f, bool sa);
void cg_device_iteration_f32(Flow& f, bool sa);  // the mixed form's FP32 inner CG

// graph.hip: n outer iterations (flow, then SST when sctl is given) each
// replayed from one HIP graph; hist[10] per iteration: scaled residuals U
// (3) and p', momentum / p' iterations, scaled residuals k and omega, their
// iterations.  eager: the same device-controlled launch sequence without a
// graph (every solve polled to its end).  Returns the iterations run.
int graph_iterations(Flow& f, const double* ctl, const double* sctl, int n, double tol,
                     double* hist, bool eager);

// amg.hip
void amg_build(Flow& f);                       // hierarchy structure (once per mesh)
template <typename V>
void amg_setup(Flow& f, const V* A0, const V* dinv0);  // numeric Galerkin levels for A0
// z = M^-1 r; the block partials of r . z go to rz_part (kRedGrid blocks) and,
// with p_init, p = z as well; with q, the partials of q . z go to qz_part.
// kcycle: the coarse levels above the one-workgroup tail are solved by two
// flexible-CG steps each (a nonlinear preconditioner: flexible CG only), else
// one V-cycle (a fixed linear operator).
// timing: HIP events around the fine-level kernels (eager launches only)
// A0f (V = double only): the fine matrix's values rounded to FP32, read by
// the fine-level kernels instead of A0 (vectors and coarse levels stay V).
template <typename V>
void amg_vcycle(Flow& f, const V* A0, const V* dinv0, const V* r, V* z, double* rz_part, V* p_init,
                bool timing = false, bool kcycle = false, const V* q = nullptr,
                double* qz_part = nullptr, const float* A0f = nullptr);
// Smoothed aggregation: hierarchy structure from the flow's first scalar
// system (once), numeric set-up of every level from the FP64 fine values A0
// (FP32 copies with want_float), and one V(1,1)-cycle z = M^-1 r on values
// and vectors of type V (A0 / dinv0: the fine level's, of type V).
void sa_build(Flow& f);
void sa_setup(Flow& f, const double* A0, bool want_float);
void sa_setup_f32(Flow& f, const float* A0f);  // the mixed form's: FP32 levels from FP32 fine values
template <typename V>
void sa_vcycle(Flow& f, const V* A0, const V* dinv0, const V* r, V* z, bool timing = false,
               const int* stop = nullptr);  // stop: the kernels' stop flag (default f.flag)

}  // namespace oagpu
