// ref_harness.cpp -- extern "C" entry points around the reference's own
// hot-path templates, so that tests can run the reference code itself on small
// inputs and pin the oracle against it.
//
// Built only by oracle/Makefile (target `ref`) into oracle/_ref/, with the
// reference headers taken from where they lie in the reference checkout (-I
// path).  Nothing of the reference is copied here.  FIT_ALGORITHM_H is
// predefined so the GSL-only header (include/fit_algorithm.hpp, used only by
// initialize_particles, which is never instantiated here) is skipped through
// its own include guard.
//
// src/hyper_vlasov.cpp itself cannot be compiled without GSL (its main()
// instantiates initialize_particles), so ref_simulate below replays
// cpu_simulate's call sequence (src/hyper_vlasov.cpp:44-104) with the
// reference's own kernels, minus the file output.
// Test infrastructure only; the product never loads this library.

#define FIT_ALGORITHM_H
#include <cstdio>
#include <cstring>
#include <vector>

#include "init.hpp"
#include "tools.hpp"
#include "physics.hpp"

namespace {

World<double> make_world(const int n[3], const double d_max[3]) {
    World<double> w;
    for (int i = 0; i < 3; i++) {
        w.n[i] = n[i];
        w.d_max[i] = d_max[i];
    }
    return w;
}

TestParticles<double> wrap(double *x, double *y, double *z, double *kx, double *ky, double *kz,
                           double *fx, double *fy, double *fz, int np, int nn) {
    TestParticles<double> p;
    p.protons = np;
    p.neutrons = nn;
    p.x = x; p.y = y; p.z = z;
    p.kx = kx; p.ky = ky; p.kz = kz;
    p.fx = fx; p.fy = fy; p.fz = fz;
    p.energy = nullptr; p.density_p = nullptr; p.density_n = nullptr;
    return p;
}

}  // namespace

extern "C" {

// include/tools.hpp:36-96 -> counts [p|n], 2G
void ref_deposit(double *x, double *y, double *z, int np, int nn, const int n[3],
                 const double d_max[3], double *counts) {
    World<double> w = make_world(n, d_max);
    TestParticles<double> p = wrap(x, y, z, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, np, nn);
    ScalarField<double> d{counts};
    distribute_volumetric_particles_cic(&d, &p, w);
}

// include/physics.hpp:203-234, in place on `density` (2G); temp is 2G scratch
void ref_smear(double *density, double *temp, const int n[3], const double d_max[3], double sigma_r,
               int ppn) {
    World<double> w = make_world(n, d_max);
    Parameters<double> prm{};
    prm.sigma_r = sigma_r;
    prm.part_per_nucleon = ppn;
    ScalarField<double> d{density}, t{temp};
    compute_volumetric_densities(&d, &t, prm, w);
}

// include/physics.hpp:273-281
void ref_skyrme(const double *rho, double *U, const int n[3], const double d_max[3], double A, double B,
                double C, double gamma) {
    World<double> w = make_world(n, d_max);
    Skyrme<double> s{A, B, C, gamma};
    ScalarField<double> u{U}, r{const_cast<double *>(rho)};
    compute_volumetric_skyrme_potentials(&u, r, s, w);
}

// include/tools.hpp:284-300
void ref_coulomb_bc(double *phi, const int n[3], const double d_max[3], int z) {
    World<double> w = make_world(n, d_max);
    ScalarField<double> c{phi};
    set_initial_coulomb_boundaries(&c, w, z);
}

// include/physics.hpp:100-157 (omega 1.5, tolerance 1e-4, 512 iterations, fixed there)
void ref_sor(double *phi, const double *rho_p, const int n[3], const double d_max[3]) {
    World<double> w = make_world(n, d_max);
    ScalarField<double> c{phi}, r{const_cast<double *>(rho_p)};
    compute_volumetric_coulomb_potentials_sor(&c, r, w);
}

// include/tools.hpp:378-384
void ref_merge(double *U, const double *phi, const int n[3], const double d_max[3]) {
    World<double> w = make_world(n, d_max);
    ScalarField<double> a{U}, b{const_cast<double *>(phi)};
    merge_volumetric_potentials(&a, b, w);
}

// include/physics.hpp:160-200
void ref_fdm(const double *U, double *Fx, double *Fy, double *Fz, const int n[3], const double d_max[3]) {
    World<double> w = make_world(n, d_max);
    VectorField<double> f{Fx, Fy, Fz};
    ScalarField<double> u{const_cast<double *>(U)};
    compute_volumetric_forces_fdm(&f, u, w);
}

// include/tools.hpp:99-168
void ref_gather(double *x, double *y, double *z, int np, int nn, const double *Fx, const double *Fy,
                const double *Fz, const int n[3], const double d_max[3], double *fx, double *fy,
                double *fz) {
    World<double> w = make_world(n, d_max);
    TestParticles<double> p = wrap(x, y, z, nullptr, nullptr, nullptr, fx, fy, fz, np, nn);
    VectorField<double> f{const_cast<double *>(Fx), const_cast<double *>(Fy), const_cast<double *>(Fz)};
    distribute_forces_to_particles_cic(&p, f, w);
}

// include/physics.hpp:249-258
void ref_kick(double *kx, double *ky, double *kz, double *fx, double *fy, double *fz, int np, int nn,
              double dt) {
    TestParticles<double> p = wrap(nullptr, nullptr, nullptr, kx, ky, kz, fx, fy, fz, np, nn);
    update_momenta_half(&p, dt);
}

// include/physics.hpp:261-270
void ref_drift(double *x, double *y, double *z, double *kx, double *ky, double *kz, int np, int nn,
               double dt) {
    TestParticles<double> p = wrap(x, y, z, kx, ky, kz, nullptr, nullptr, nullptr, np, nn);
    update_positions_full(&p, dt);
}

// include/physics.hpp:237-246
void ref_excite(double *kz, int np, int nn, double eta, int z, int n) {
    TestParticles<double> p = wrap(nullptr, nullptr, nullptr, nullptr, nullptr, kz, nullptr, nullptr,
                                   nullptr, np, nn);
    Parameters<double> prm{};
    prm.eta_exc = eta;
    prm.z = z;
    prm.n = n;
    nuclear_excitation(&p, prm);
}

// include/tools.hpp:328-338 and :303-325; type 0 = protons, 1 = neutrons
double ref_center_of_mass(double *z, int np, int nn, int type) {
    TestParticles<double> p = wrap(nullptr, nullptr, z, nullptr, nullptr, nullptr, nullptr, nullptr,
                                   nullptr, np, nn);
    return center_of_mass(p, type);
}

double ref_mean_squared_radius(double *x, double *y, double *z, int np, int nn, const int n[3],
                               const double d_max[3], int type) {
    World<double> w = make_world(n, d_max);
    TestParticles<double> p = wrap(x, y, z, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, np, nn);
    return mean_squared_radius(p, w, type);
}

// src/hyper_vlasov.cpp:33-118: cpu_simulate's sequence of reference calls.
// stats[4*s + {0..3}] = center_of_mass p/n and mean_squared_radius p/n at the
// start of step s (the reference prints them every `substeps` steps, :64-70);
// phi (G) receives the final Coulomb potential.
void ref_simulate(double *x, double *y, double *z, double *kx, double *ky, double *kz, double *fx,
                  double *fy, double *fz, int np, int nn, const int n[3], const double d_max[3],
                  double A, double B, double C, double gamma, double t_f, int steps, double t_exc,
                  double eta, int zz, int nnuc, double sigma_r, int ppn, double *stats, double *phi) {
    World<double> world = make_world(n, d_max);
    TestParticles<double> p = wrap(x, y, z, kx, ky, kz, fx, fy, fz, np, nn);
    TestParticles<double> *part = &p;
    Skyrme<double> skm{A, B, C, gamma};
    Parameters<double> param{};
    param.t_f = t_f;
    param.steps = steps;
    param.t_exc = t_exc;
    param.eta_exc = eta;
    param.z = zz;
    param.n = nnuc;
    param.sigma_r = sigma_r;
    param.part_per_nucleon = ppn;

    bool excited_nucleus = false;
    double dt = param.t_f / param.steps;
    VectorField<double> forces;
    create_vector_field_double(&forces, world);
    ScalarField<double> potentials, coulomb, temp_density, density;
    create_scalar_field_single(&coulomb, world);
    create_scalar_field_double(&density, world);
    create_scalar_field_double(&potentials, world);
    create_scalar_field_double(&temp_density, world);

    distribute_volumetric_particles_cic(&density, part, world);
    compute_volumetric_densities(&density, &temp_density, param, world);
    set_initial_coulomb_boundaries(&coulomb, world, param.z);
    compute_volumetric_skyrme_potentials(&potentials, density, skm, world);
    compute_volumetric_coulomb_potentials_sor(&coulomb, density, world);
    merge_volumetric_potentials(&potentials, coulomb, world);
    compute_volumetric_forces_fdm(&forces, potentials, world);
    distribute_forces_to_particles_cic(part, forces, world);
    for (int step = 0; step < param.steps; step++) {
        stats[4 * step + 0] = center_of_mass(*part, PROTONS);
        stats[4 * step + 1] = center_of_mass(*part, NEUTRONS);
        stats[4 * step + 2] = mean_squared_radius(*part, world, PROTONS);
        stats[4 * step + 3] = mean_squared_radius(*part, world, NEUTRONS);
        if (step * dt >= param.t_exc && !excited_nucleus) {
            excited_nucleus = true;
            nuclear_excitation(part, param);
        }
        update_momenta_half(part, dt);
        update_positions_full(part, dt);
        distribute_volumetric_particles_cic(&density, part, world);
        compute_volumetric_densities(&density, &temp_density, param, world);
        compute_volumetric_skyrme_potentials(&potentials, density, skm, world);
        compute_volumetric_coulomb_potentials_sor(&coulomb, density, world);
        merge_volumetric_potentials(&potentials, coulomb, world);
        compute_volumetric_forces_fdm(&forces, potentials, world);
        distribute_forces_to_particles_cic(part, forces, world);
        update_momenta_half(part, dt);
    }
    int G = n[0] * n[1] * n[2];
    for (int i = 0; i < G; i++) phi[i] = coulomb.v[i];
    free_vector_field(&forces);
    free_scalar_field(&density);
    free_scalar_field(&coulomb);
    free_scalar_field(&potentials);
    free_scalar_field(&temp_density);
}

// The snapshot cpu_simulate writes every `substeps` steps
// (src/hyper_vlasov.cpp:78-82, include/init.hpp:160-254), from given fields
// (forces 3 x 2G, density 2G, potentials 2G); single != 0 runs the float
// instantiation on the fields rounded to float.  Returns 0, or 1 if the file
// cannot be opened.
int ref_write_vtk(const char *path, const int n[3], const double d_max[3], const double *fx, const double *fy,
                  const double *fz, const double *density, const double *potentials, int single) {
    FILE *out = std::fopen(path, "wb");
    if (!out) return 1;
    const int G2 = 2 * n[0] * n[1] * n[2];
    if (single) {
        World<float> w;
        for (int i = 0; i < 3; i++) {
            w.n[i] = n[i];
            w.d_max[i] = (float)d_max[i];
        }
        std::vector<float> f[5];
        const double *src[5] = {fx, fy, fz, density, potentials};
        for (int q = 0; q < 5; q++) {
            f[q].resize(G2);
            for (int e = 0; e < G2; e++) f[q][e] = (float)src[q][e];
        }
        output_vtk_header_start(out, w);
        VectorField<float> fv{f[0].data(), f[1].data(), f[2].data()};
        output_vector_field(out, fv, w, "forces");
        ScalarField<float> d{f[3].data()}, u{f[4].data()};
        output_scalar_field(out, d, w, "density");
        output_scalar_field(out, u, w, "potentials");
    } else {
        World<double> w = make_world(n, d_max);
        std::vector<double> f[5];
        const double *src[5] = {fx, fy, fz, density, potentials};
        for (int q = 0; q < 5; q++) f[q].assign(src[q], src[q] + G2);
        output_vtk_header_start(out, w);
        VectorField<double> fv{f[0].data(), f[1].data(), f[2].data()};
        output_vector_field(out, fv, w, "forces");
        ScalarField<double> d{f[3].data()}, u{f[4].data()};
        output_scalar_field(out, d, w, "density");
        output_scalar_field(out, u, w, "potentials");
    }
    std::fclose(out);
    return 0;
}

}  // extern "C"
