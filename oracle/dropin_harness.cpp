// dropin_harness.cpp -- compiles hyper-vlasov_amd/dropin/hv_gpu_simulate.hpp
// against the reference's own headers (its TestParticles, Skyrme, Parameters,
// World, file-name helpers and SOR constants) and exposes gpu_simulate<T>
// through extern "C" so that a test can run it and compare it with the
// reference's loop.  This is the code a maintainer adds to
// src/hyper_vlasov.cpp, minus main (which needs GSL for
// initialize_particles).  Built only by oracle/Makefile (target `dropin`) into
// oracle/_ref/; linked against libhvamd.so.  Test infrastructure.

#define FIT_ALGORITHM_H
#include <cstdio>
#include <vector>

#include "init.hpp"
#include "tools.hpp"
#include "physics.hpp"

#include "../hyper-vlasov_amd/dropin/hv_gpu_simulate.hpp"

namespace {

template <typename T>
void run(double *x, double *y, double *z, double *kx, double *ky, double *kz, double *fx, double *fy, double *fz,
         int np, int nn, const int n[3], const double d_max[3], double A, double B, double C, double gamma,
         double t_f, int steps, int substeps, double t_exc, double eta, int zz, int nnuc, double sigma_r, int ppn,
         const char *outdir) {
    const int N = np + nn;
    std::vector<T> a[9];
    double *src[9] = {x, y, z, kx, ky, kz, fx, fy, fz};
    for (int q = 0; q < 9; q++) a[q].assign(src[q], src[q] + N);
    TestParticles<T> part{};
    part.protons = np;
    part.neutrons = nn;
    part.x = a[0].data(); part.y = a[1].data(); part.z = a[2].data();
    part.kx = a[3].data(); part.ky = a[4].data(); part.kz = a[5].data();
    part.fx = a[6].data(); part.fy = a[7].data(); part.fz = a[8].data();
    World<T> world;
    for (int i = 0; i < 3; i++) {
        world.n[i] = n[i];
        world.d_max[i] = (T)d_max[i];
    }
    Skyrme<T> skm{(T)A, (T)B, (T)C, (T)gamma};
    Parameters<T> param{};
    param.t_f = (T)t_f;
    param.steps = steps;
    param.substeps = substeps;
    param.t_exc = (T)t_exc;
    param.eta_exc = (T)eta;
    param.z = zz;
    param.n = nnuc;
    param.sigma_r = (T)sigma_r;
    param.part_per_nucleon = ppn;
    gpu_simulate<T>(outdir, &part, skm, param, world);
    for (int q = 0; q < 9; q++)
        for (int i = 0; i < N; i++) src[q][i] = (double)a[q][i];
}

}  // namespace

extern "C" {

// gpu_simulate<double> / <float> on the given state; the particles come back
// in the arrays (as doubles)
void dropin_gpu_simulate(double *x, double *y, double *z, double *kx, double *ky, double *kz, double *fx,
                         double *fy, double *fz, int np, int nn, const int n[3], const double d_max[3], double A,
                         double B, double C, double gamma, double t_f, int steps, int substeps, double t_exc,
                         double eta, int zz, int nnuc, double sigma_r, int ppn, const char *outdir, int single) {
    if (single)
        run<float>(x, y, z, kx, ky, kz, fx, fy, fz, np, nn, n, d_max, A, B, C, gamma, t_f, steps, substeps, t_exc,
                   eta, zz, nnuc, sigma_r, ppn, outdir);
    else
        run<double>(x, y, z, kx, ky, kz, fx, fy, fz, np, nn, n, d_max, A, B, C, gamma, t_f, steps, substeps,
                    t_exc, eta, zz, nnuc, sigma_r, ppn, outdir);
}

}  // extern "C"
