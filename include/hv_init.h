/*
 * hv_init.h -- initial states for the step loop (off the hot path).
 *
 * Two samplers:
 *   - the reference's own: the Woods-Saxon / Fermi-level self-consistent loop
 *     of initialize_particles (include/physics.hpp:35-97), with a
 *     Levenberg-Marquardt fit of the wells in place of GSL's trust-region
 *     solver (include/fit_algorithm.hpp:96-138), and its pieces;
 *   - the benchmark configurations' simple nuclei (positions uniform in a
 *     ball, momenta uniform in a Fermi ball, reproducible from a seed), which
 *     have no counterpart in the reference.
 */
#ifndef HV_INIT_H
#define HV_INIT_H

#include "hv_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- the reference's initial state ---------------------------------------- */

/* WoodsSaxon<T> (include/sim_structs.hpp:78-81): V(r) = V0 / (1 + exp((r - R12) / a)) */
typedef struct hv_woods_saxon {
    double V0, R12, a;
} hv_woods_saxon;

/* nuclear_radius (include/physics_formulas.hpp:31-34): 1.5 A^(1/3) [fm] */
double hv_ws_nuclear_radius(int mass_number);
/* max_particles (include/physics_formulas.hpp:36-42): the test particles of
 * each species initialize_particles draws per round (phase-space count of the
 * r_max, k_max ball, times part_per_nucleon) */
int hv_ws_max_particles(double r_max, double k_max, int part_per_nucleon);

/* One sampling round of initialize_particles (include/physics.hpp:42-75) on
 * the host, in the reference's order and with the C library's rand() (not
 * reseeded here: srand is the caller's, as the reference's main calls
 * srand(128)): max_particles() particles of each species uniform in the balls
 * r_max = 1.5 A^(1/3), k_max = 1.5 fm^-1 (generate_random_particles,
 * include/tools.hpp:235-254), their energies in the wells ws[2]
 * (compute_energy, include/tools.hpp:206-224), the counts below the Fermi
 * levels and 0.5 MeV either side (counts[6] = equal_p, more_p, less_p,
 * equal_n, more_n, less_n; may be NULL), the move of each level
 * (delta[2], may be NULL; fermi[2] = {epsilon_p, epsilon_n} in and out), then
 * the z*ppn + n*ppn particles below the new levels, each followed by its
 * mirror image (generate_checking_particles, include/tools.hpp:257-281) into
 * x..kz.  An empty +-0.5 MeV shell moves a level by the relaxed maximum 0.6
 * MeV instead of the reference's division by zero. */
int hv_ws_sample_round(int z, int n, int part_per_nucleon, double sigma_k, const hv_woods_saxon ws[2],
                       double fermi[2], double delta[2], int counts[6], double *x, double *y, double *zz,
                       double *kx, double *ky, double *kz);
/* compute_particle_densities (include/tools.hpp:171-203) on the current GPU:
 * each particle's smeared proton and neutron density, the Gaussian
 * exp(-d^2 / 2 sigma_r^2) of every particle within 4 sigma_r summed in index
 * order, over ppn (2 pi sigma_r^2)^(3/2).  Allocates and frees its device
 * buffers. */
int hv_ws_particle_densities(const double *x, const double *y, const double *z, int n_protons, int n_neutrons,
                             double sigma_r, int part_per_nucleon, double *rho_p, double *rho_n);
/* minim_woods_saxon (include/fit_algorithm.hpp:96-138): per species, the
 * well (V0, R12, a) closest in least squares to the Skyrme potential
 * (include/physics_formulas.hpp:71-77) at the particles' densities, started
 * from ws[] (in/out); Levenberg-Marquardt with More's scaling, to a relative
 * step of 1e-10 (GSL stops its trust-region iteration at 1e-4).
 * iterations[2] (may be NULL): the steps taken per species. */
int hv_ws_fit(const double *x, const double *y, const double *z, const double *rho_p, const double *rho_n,
              int n_protons, int n_neutrons, const hv_skyrme *skyrme, hv_woods_saxon ws[2], int iterations[2]);
/* chi_squared (include/tools.hpp:341-366): sum over each species' particles
 * of (V_ws(r) - U_skyrme(rho))^2 / ppn, out[2] = {protons, neutrons} (the
 * figures the reference prints after initialisation) */
int hv_ws_chi_squared(const double *x, const double *y, const double *z, const double *rho_p, const double *rho_n,
                      int n_protons, int n_neutrons, const hv_skyrme *skyrme, const hv_woods_saxon ws[2],
                      int part_per_nucleon, double out[2]);
/* initialize_particles (include/physics.hpp:35-97): srand(seed), then rounds
 * of sampling (hv_ws_sample_round), densities (GPU), well fit and relaxation
 * -- the rounds draw the sequence rand() gives after srand(seed) from a
 * private generator state (glibc random_r, the generator behind rand()), so
 * other threads' rand() calls cannot shift it and the process's own rand()
 * state is left as it was
 * (0.6 new + 0.4 old, include/tools.hpp:369-375) until the Fermi levels move
 * by less than 0.1 MeV together (at most 32 rounds); wells start at V0,
 * R12 = 0.8 r_max, a (include/init.hpp:333-334).  Outputs the particles (z*ppn
 * protons first, n*ppn neutrons), the Fermi levels (fermi[2] in/out), the
 * wells ws[2], the rounds run and the final chi squared (both may be NULL).
 * Returns HV_WARN_INIT_NOT_CONVERGED when 32 rounds did not settle the levels
 * (the reference prints "INITIALIZATION DID NOT CONVERGE!"). */
int hv_init_woods_saxon(int z, int n, int part_per_nucleon, const hv_skyrme *skyrme, double sigma_r, double sigma_k,
                        double V0, double a, double fermi[2], unsigned seed, double *x, double *y, double *zz,
                        double *kx, double *ky, double *kz, hv_woods_saxon ws[2], int *iterations, double chi2[2]);

/* ---- the benchmark configurations' nuclei ---------------------------------- */
/* Fermi momentum of symmetric matter at density rho [fm^-3]:
 * k_F = (3 pi^2 rho / 2)^(1/3) [fm^-1]. */
double hv_fermi_momentum(double rho);

/* Positions uniform in the ball |r| < r_max and momenta uniform in |k| < k_max,
 * each by rejection from the enclosing cube (like generate_random_particles,
 * include/tools.hpp:235-254, but with a counter-based generator: particle i
 * draws from its own stream keyed by (seed, i), so the result does not depend
 * on how many threads draw it).  Protons are particles [0, n_protons).
 * Returns 0, or HV_ERR_ARG (-1) on bad sizes or radii. */
int hv_sample_uniform_spheres(unsigned long long seed, int n_protons, int n_neutrons, double r_max,
                              double k_max, double *x, double *y, double *z, double *kx, double *ky,
                              double *kz);

/* splitmix64 finaliser, exposed so callers can derive per-instance seeds
 * (e.g. instance i of an ensemble seeded by hv_mix64(seed + i)). */
unsigned long long hv_mix64(unsigned long long v);

#ifdef __cplusplus
}
#endif

#endif
