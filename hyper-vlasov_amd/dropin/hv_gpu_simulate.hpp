// hv_gpu_simulate.hpp -- gpu_simulate for the reference program: the function
// whose signature src/hyper_vlasov.cpp:120-123 leaves commented out, selected
// by the `use_gpu` input key (src/hyper_vlasov.cpp:143-145,
// include/init.hpp:322-323).  Same arguments as cpu_simulate
// (src/hyper_vlasov.cpp:33), same output: stats.txt with one line every
// `substeps` steps (:64-70) and the snapshot out-%04d.vtk of the same steps
// (:72-85), the particles left in `part` at the end of the run.  The loop runs
// in libhvamd.so through its C-ABI (include/hv_abi.h): FP64 for double (bit
// for bit the CPU restatement of the loop), the mixed FP32/FP64 form for float.
//
// Include it in src/hyper_vlasov.cpp after the reference's own headers (it
// uses TestParticles, Skyrme, Parameters and World of include/sim_structs.hpp,
// set_stats_filename / set_output_filename and STRING_SIZE of
// include/init.hpp, sor_tolerance and MAX_SOR_ITERATIONS of
// include/sim_structs.hpp), add -I<this repository>/include, and link with
// -L<this repository>/hyper-vlasov_amd -lhvamd.  The reference's SOR settings
// are kept: omega 1.5 (include/physics.hpp:103), its tolerance and iteration
// cap; the boost is the reference's (include/physics.hpp:237-246).  Errors end
// the program with a message, as the reference's own do (exit(1)).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hv_abi.h"

namespace hv_dropin {

inline void fail(hv_ctx *ctx, const char *what, int rc) {
    std::fprintf(stderr, "gpu_simulate: %s failed (%d): %s\n", what, rc, ctx ? hv_last_error(ctx) : "");
    if (ctx) hv_ctx_destroy(ctx);
    std::exit(1);
}

}  // namespace hv_dropin

template <typename T>
void gpu_simulate(const char *output_directory, TestParticles<T> *part, const Skyrme<T> &skm,
                  const Parameters<T> &param, const World<T> &world) {
    using hv_dropin::fail;
    const hv_world w = {{world.n[0], world.n[1], world.n[2]},
                        {(double)world.d_max[0], (double)world.d_max[1], (double)world.d_max[2]}};
    const int prec = sizeof(T) == sizeof(double) ? HV_PREC_F64 : HV_PREC_MIXED;
    const int N = part->protons + part->neutrons;
    hv_ctx *ctx = nullptr;
    int rc = hv_ctx_create(&w, part->protons, part->neutrons, prec, &ctx);
    if (rc) fail(ctx, "hv_ctx_create", rc);

    // host SoA (T) -> device; the library takes FP64 on the host side
    std::vector<double> b[9];
    T *arr[9] = {part->x, part->y, part->z, part->kx, part->ky, part->kz, part->fx, part->fy, part->fz};
    for (int q = 0; q < 9; q++) b[q].assign(arr[q], arr[q] + N);
    rc = hv_upload_particles(ctx, b[0].data(), b[1].data(), b[2].data(), b[3].data(), b[4].data(), b[5].data(),
                             b[6].data(), b[7].data(), b[8].data());
    if (rc) fail(ctx, "hv_upload_particles", rc);

    const hv_skyrme s = {(double)skm.A, (double)skm.B, (double)skm.C, (double)skm.gamma};
    hv_params p = {};
    p.t_f = param.t_f;
    p.steps = param.steps;
    p.substeps = param.substeps;
    p.t_exc = param.t_exc;
    p.eta_exc = param.eta_exc;
    p.boost_mode = HV_BOOST_REF_WEIGHTED;
    p.z = param.z;
    p.n = param.n;
    p.sigma_r = param.sigma_r;
    p.part_per_nucleon = param.part_per_nucleon;
    p.sor_omega = 1.5;
    p.sor_tol = (double)sor_tolerance<T>;
    p.sor_max_it = MAX_SOR_ITERATIONS;

    char stats_filename[STRING_SIZE];
    set_stats_filename(stats_filename, output_directory);
    FILE *stats = std::fopen(stats_filename, "w");
    if (stats == nullptr) {
        std::fprintf(stderr, "CANNOT OPEN STATS FILE!\n");
        std::exit(1);
    }
    int warn = hv_run_init(ctx, &s, &p, nullptr);
    if (warn < 0) fail(ctx, "hv_run_init", warn);
    const T dt = param.t_f / param.steps;
    const int sub = p.substeps > 0 ? p.substeps : 1;
    std::vector<double> st(4 * (size_t)sub);
    // one call per stats interval: the line and the snapshot of step s
    // (s % substeps == 0) hold the state at the top of step s
    for (int first = 0; first < p.steps; first += sub) {
        const int cnt = std::min(sub, p.steps - first);
        char output_filename[STRING_SIZE];
        set_output_filename(output_filename, output_directory, first / sub);
        rc = hv_write_vtk(ctx, output_filename);
        if (rc) {
            std::fprintf(stderr, "CANNOT OPEN OUTPUT FILE!\n");
            hv_ctx_destroy(ctx);
            std::exit(1);
        }
        rc = hv_run_steps(ctx, &s, &p, first, cnt, st.data(), nullptr);
        if (rc < 0) fail(ctx, "hv_run_steps", rc);
        warn |= rc;
        // <z>_p, <z>_n, <r^2>_p, <r^2>_n of the positions at the top of `first`
        std::fprintf(stats, "%e %e %e %e %e\n", first * dt, (T)std::sqrt(st[2]), (T)std::sqrt(st[3]), (T)st[0],
                     (T)st[1]);
        std::printf("Processed step: %i/%i.\n", first, param.steps);
    }
    std::fclose(stats);
    if (warn & HV_WARN_SOR_NOT_CONVERGED) std::fprintf(stderr, "SOR COULOMB DID NOT CONVERGE!\n");

    rc = hv_download_particles(ctx, b[0].data(), b[1].data(), b[2].data(), b[3].data(), b[4].data(), b[5].data(),
                               b[6].data(), b[7].data(), b[8].data());
    if (rc) fail(ctx, "hv_download_particles", rc);
    for (int q = 0; q < 9; q++) std::transform(b[q].begin(), b[q].end(), arr[q], [](double v) { return (T)v; });
    hv_ctx_destroy(ctx);
}
