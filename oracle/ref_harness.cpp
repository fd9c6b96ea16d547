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
#include <unistd.h>

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

// ---- initialize_particles (include/physics.hpp:35-97), without its GSL fit ----
// One sampling round with the reference's own templates: the body of
// initialize_particles' loop up to the densities (include/physics.hpp:42-72,
// replayed: the function itself instantiates minim_woods_saxon, which needs
// GSL).  rand() is not reseeded here.  ws = {V0_p, R12_p, a_p, V0_n, R12_n, a_n}.
void ref_ws_sample_round(int z, int n, int ppn, double sigma_k, const double ws_in[6], double fermi[2],
                         double delta[2], int counts[6], double *x, double *y, double *zz, double *kx, double *ky,
                         double *kz) {
    Parameters<double> param{};
    set_parameters(&param, z, n, ppn, 1, 1, false, sigma_k, 0.6, 1.0, 0.0, 0.0);
    WoodsSaxon<double> ws[2];
    set_woods_saxon(&ws[0], ws_in[0], ws_in[1], ws_in[2]);
    set_woods_saxon(&ws[1], ws_in[3], ws_in[4], ws_in[5]);
    Fermi<double> fermi_levels;
    set_fermi_levels(&fermi_levels, fermi[0], fermi[1]);
    const int max_part = param.max_test_part, total_p = z * ppn, total_n = n * ppn;
    TestParticles<double> part, temp_part;
    create_particles(&part, total_p, total_n);
    create_particles(&temp_part, max_part, max_part);
    generate_random_particles(&temp_part, param.r_max);
    compute_particle_energies(&temp_part, ws, param);
    int less_p = 0, equal_p = 0, more_p = 0, less_n = 0, equal_n = 0, more_n = 0;
    for (int i = 0; i < max_part; i++) {
        if (temp_part.energy[i] < fermi_levels.epsilon_p) equal_p += 2;
        if (temp_part.energy[i] < fermi_levels.epsilon_p + 0.5) more_p += 2;
        if (temp_part.energy[i] < fermi_levels.epsilon_p - 0.5) less_p += 2;
        if (temp_part.energy[i + max_part] < fermi_levels.epsilon_n) equal_n += 2;
        if (temp_part.energy[i + max_part] < fermi_levels.epsilon_n + 0.5) more_n += 2;
        if (temp_part.energy[i + max_part] < fermi_levels.epsilon_n - 0.5) less_n += 2;
    }
    double delta_part_n = total_n - equal_n, delta_part_p = total_p - equal_p;
    double delta_epsilon_n = 0.5 * delta_part_n / (more_n - less_n);
    double delta_epsilon_p = 0.5 * delta_part_p / (more_p - less_p);
    if (std::abs(delta_epsilon_p) > 0.5) delta_epsilon_p *= 0.6;
    if (std::abs(delta_epsilon_n) > 0.5) delta_epsilon_n *= 0.6;
    fermi_levels.epsilon_p += delta_epsilon_p;
    fermi_levels.epsilon_n += delta_epsilon_n;
    generate_checking_particles(&part, ws, param, &fermi_levels);
    fermi[0] = fermi_levels.epsilon_p;
    fermi[1] = fermi_levels.epsilon_n;
    delta[0] = delta_epsilon_p;
    delta[1] = delta_epsilon_n;
    const int c[6] = {equal_p, more_p, less_p, equal_n, more_n, less_n};
    for (int q = 0; q < 6; q++) counts[q] = c[q];
    const int N = total_p + total_n;
    for (int i = 0; i < N; i++) {
        x[i] = part.x[i]; y[i] = part.y[i]; zz[i] = part.z[i];
        kx[i] = part.kx[i]; ky[i] = part.ky[i]; kz[i] = part.kz[i];
    }
    free_particles(&part);
    free_particles(&temp_part);
}

// max_particles (include/physics_formulas.hpp:36-42) at the reference's r_max and k_max
int ref_ws_max_particles(int z, int n, int ppn) {
    Parameters<double> param{};
    set_parameters(&param, z, n, ppn, 1, 1, false, 0.1, 0.6, 1.0, 0.0, 0.0);
    return param.max_test_part;
}

// compute_particle_densities (include/tools.hpp:171-203)
void ref_ws_particle_densities(double *x, double *y, double *zz, int np, int nn, double sigma_r, int ppn,
                               double *rho_p, double *rho_n) {
    TestParticles<double> part = wrap(x, y, zz, x, y, zz, x, y, zz, np, nn);
    part.density_p = rho_p;
    part.density_n = rho_n;
    Parameters<double> param{};
    param.sigma_r = sigma_r;
    param.part_per_nucleon = ppn;
    compute_particle_densities(&part, param);
}

// chi_squared (include/tools.hpp:341-366): the two figures it prints, read
// back from its own output ("%0.2lf")
int ref_ws_chi_squared(double *x, double *y, double *zz, double *rho_p, double *rho_n, int np, int nn, double A,
                       double B, double C, double gamma, const double ws_in[6], int ppn, double out[2]) {
    TestParticles<double> part = wrap(x, y, zz, x, y, zz, x, y, zz, np, nn);
    part.density_p = rho_p;
    part.density_n = rho_n;
    WoodsSaxon<double> ws[2];
    set_woods_saxon(&ws[0], ws_in[0], ws_in[1], ws_in[2]);
    set_woods_saxon(&ws[1], ws_in[3], ws_in[4], ws_in[5]);
    Skyrme<double> skm{A, B, C, gamma};
    std::fflush(stdout);
    FILE *tmp = std::tmpfile();
    if (!tmp) return 1;
    const int saved = dup(1);
    dup2(fileno(tmp), 1);
    chi_squared(part, ws, skm, ppn);
    std::fflush(stdout);
    dup2(saved, 1);
    close(saved);
    std::rewind(tmp);
    const int got = std::fscanf(tmp, "CHI SQUARED P %lf CHI SQUARED N %lf", &out[0], &out[1]);
    std::fclose(tmp);
    return got == 2 ? 0 : 1;
}

}  // extern "C"
