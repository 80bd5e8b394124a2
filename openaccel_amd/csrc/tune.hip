// tune.hip -- the project's autotuner: each tunable launch of a flow is timed
// in its candidate variants on the flow's own mesh and state, and the fastest
// variant is kept for the flow's later launches.
//
// Tunables (the reference has no counterpart; its element loops run on the
// CPU):
// * momentum rows: off-diagonal blocks summed in a per-wavefront LDS tile
//   (64 rows per workgroup) or read-modify-written in HBM (256 rows per
//   workgroup; hex-only meshes, upwind advection);
// * p' rows: at most 2 waves per SIMD or the compiler's occupancy;
// * sliced-ELL SpMV of the Krylov solvers (momentum and scalar systems, FP64
//   and the mixed form's FP32 inner solves): 64, 128 or 256 rows per
//   workgroup;
// * k / omega rows, mass-flow update and gradients: resident waves per CU
//   capped at 4, 8, 16 or not at all (LDS reserved per workgroup).  The
//   mass-flow update relaxes the stored rates in place, so its candidates run
//   on a saved copy of them that is put back afterwards.
#include "flow.hpp"

namespace oagpu {

namespace {
template <typename F>
double time_ms(Flow& f, F&& fn, int reps) {
    hipEvent_t e0, e1;
    OAGPU_HIP(hipEventCreate(&e0));
    OAGPU_HIP(hipEventCreate(&e1));
    fn();  // warm-up
    OAGPU_HIP(hipEventRecord(e0, f.stream));
    for (int k = 0; k < reps; ++k) fn();
    OAGPU_HIP(hipEventRecord(e1, f.stream));
    OAGPU_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    OAGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return double(ms) / reps;
}
}  // namespace

void autotune(Flow& f, double* report, int n_report) {
    double rep[9][3];
    for (auto& r : rep) r[0] = r[1] = r[2] = 0.0;
    const unsigned saved_mask = f.prof_mask;
    f.prof_mask = 0;  // the candidates' launches are not clocked
    // momentum rows
    {
        const bool tile_only = f.prm.high_res || f.m->mixed;
        int best = 1;
        f.tune.mom_tile = 1;
        const double t1 = time_ms(f, [&] { launch_assemble_momentum(f); }, 3);
        double tb = t1;
        if (!tile_only) {
            f.tune.mom_tile = 0;
            const double t0 = time_ms(f, [&] { launch_assemble_momentum(f); }, 3);
            if (t0 < tb) {
                tb = t0;
                best = 0;
            }
        }
        f.tune.mom_tile = best;
        launch_assemble_momentum(f);  // the system and du of the kept variant
        rep[0][0] = best;
        rep[0][1] = tb;
        rep[0][2] = t1;
    }
    // p' rows
    {
        f.tune.pc_occupancy = 0;
        const double t0 = time_ms(f, [&] { launch_assemble_pcorr(f); }, 3);
        f.tune.pc_occupancy = 1;
        const double t1 = time_ms(f, [&] { launch_assemble_pcorr(f); }, 3);
        f.tune.pc_occupancy = t1 < t0 ? 1 : 0;
        launch_assemble_pcorr(f);
        rep[1][0] = f.tune.pc_occupancy;
        rep[1][1] = std::min(t0, t1);
        rep[1][2] = t0;
    }
    // Krylov SpMV (the FP32 variants only for the mixed form)
    const bool mixed = f.prm.precision == OAGPU_MIXED;
    int row = 2;
    for (bool f32 : {false, true}) {
        for (int N : {1, 3}) {
            if (!f32 || mixed) {
                double best, dflt;
                time_spmv_variants(f, N, f32, &best, &dflt);
                rep[row][0] = f.tune.spmv_block[N == 3][f32];
                rep[row][1] = best;
                rep[row][2] = dflt;
            }
            ++row;
        }
    }
    // occupancy caps: the fastest of none / 16 / 8 / 4 waves per CU
    auto caps = [&](int& cap, int r, auto&& run) {
        double tb = 1e300, t0 = 0.0;
        int best = 0;
        for (int c : {0, 16, 8, 4}) {
            cap = c;
            const double t = time_ms(f, run, 3);
            if (c == 0) t0 = t;
            if (t < tb) {
                tb = t;
                best = c;
            }
        }
        cap = best;
        rep[r][0] = best;
        rep[r][1] = tb;
        rep[r][2] = t0;
    };
    if (f.sst_on)  // the k / omega assembly as the iterations run it (one launch for both when possible)
        caps(f.tune.cap_scalar, 6, [&] {
            if (scalar_pair_ok(f)) launch_assemble_scalar_pair(f);
            else launch_assemble_scalar(f, 0);
        });
    {
        DBuf<double> md(f.mdot.n), sd;
        OAGPU_HIP(hipMemcpyAsync(md.p, f.mdot.p, f.mdot.n * sizeof(double), hipMemcpyDeviceToDevice,
                                 f.stream));
        if (f.msd.p) {
            sd.alloc(f.msd.n);
            OAGPU_HIP(hipMemcpyAsync(sd.p, f.msd.p, f.msd.n * sizeof(double), hipMemcpyDeviceToDevice,
                                     f.stream));
        }
        caps(f.tune.cap_mdot, 7, [&] { launch_update_mdot(f); });
        OAGPU_HIP(hipMemcpyAsync(f.mdot.p, md.p, f.mdot.n * sizeof(double), hipMemcpyDeviceToDevice,
                                 f.stream));
        if (f.msd.p)
            OAGPU_HIP(hipMemcpyAsync(f.msd.p, sd.p, f.msd.n * sizeof(double), hipMemcpyDeviceToDevice,
                                     f.stream));
        OAGPU_HIP(hipStreamSynchronize(f.stream));
    }
    caps(f.tune.cap_grad, 8, [&] { launch_gradient(f, 1); });
    OAGPU_HIP(hipStreamSynchronize(f.stream));
    f.prof_mask = saved_mask;
    f.tune.tuned = true;
    if (report)
        for (int i = 0; i < 9 && 3 * i + 2 < n_report; ++i)
            for (int k = 0; k < 3; ++k) report[3 * i + k] = rep[i][k];
}

}  // namespace oagpu
