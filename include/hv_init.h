/*
 * hv_init.h -- seeded initial states for the step loop (off the hot path).
 *
 * The reference builds its initial state with a Woods-Saxon / Fermi-level
 * self-consistent loop and a GSL fit (initialize_particles,
 * include/physics.hpp:35-97), which cannot be built without GSL.  These entry
 * points have no counterpart in the reference: they give the simple sampled
 * nuclei the benchmark configurations ask for (positions uniform in a ball,
 * momenta uniform in a Fermi ball), reproducibly from a seed, so the CPU
 * restatement and the GPU path start from identical bits.
 */
#ifndef HV_INIT_H
#define HV_INIT_H

#ifdef __cplusplus
extern "C" {
#endif

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
