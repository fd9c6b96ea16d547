/*
 * hv_oracle.h -- CPU restatement of the hyper-vlasov per-timestep test-particle
 * loop, used ONLY as test infrastructure: by tests/, by __graft_entry__.smoke()
 * and by bench.py's CPU-baseline leg, always as the checker and never as the
 * thing measured or shipped.  The product (libhvamd.so) never links it.
 *
 * Each function names the reference code it restates (file:line, relative to
 * the reference repository).  Two deliberate, documented reformulations make
 * the restatement deterministic so that the FP64 GPU path can match it bit for
 * bit:
 *   - CIC deposit accumulates weights as 2^-40 fixed point in int64 (exact,
 *     summation-order free) instead of an OpenMP array reduction of doubles
 *     (include/tools.hpp:45), whose order is thread-count dependent;
 *   - the Gaussian smearing sums the same spherical tap set with the factorised
 *     weight gx(a) gy(b) gz(c) in a nested (c, then b, then a) order instead of
 *     one exp per pair in a flat order (include/physics.hpp:209-229);
 *     hvo_smear_literal keeps the flat reference order for pinning;
 *   - the Skyrme t^gamma uses hv_det_pow (hv_detmath.h) in place of std::pow;
 *   - the centre-of-mass / radius sums are exact integer sums of values
 *     quantised to 2^-30 (hv_stat_fix), independent of summation order, instead
 *     of OpenMP reductions of doubles (include/tools.hpp:310, 334).
 * Everything else (SOR, FDM, gather, kick, drift, boost, Coulomb faces) keeps
 * the reference's expressions and operation order exactly.
 *
 * Parity status of each function (what pins it) is listed in the test files
 * tests/test_oracle_*.py: compiled reference templates (oracle/_ref), committed
 * golden vectors (tests/golden/), and closed forms.
 */
#ifndef HV_ORACLE_H
#define HV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hvo_world {
    int n[3];
    double d_max[3];
} hvo_world;

typedef struct hvo_skyrme {
    double A, B, C, gamma;
} hvo_skyrme;

/* boost modes */
#define HVO_BOOST_REF_WEIGHTED 0 /* protons +eta N/A, neutrons -eta Z/A (include/physics.hpp:237-246) */
#define HVO_BOOST_SYMMETRIC 1    /* protons +eta, neutrons -eta */

typedef struct hvo_params {
    double dt;        /* t_f / steps (src/hyper_vlasov.cpp:35) */
    int steps;
    double t_exc;     /* boost once when step*dt >= t_exc (src/hyper_vlasov.cpp:86) */
    double eta;       /* boost strength [fm^-1] */
    int boost_mode;
    int z, n;         /* nucleus Z, N (boost weights) */
    double sigma_r;   /* smearing width [fm] */
    int part_per_nucleon;
    double omega, sor_tol;
    int sor_max_it;
} hvo_params;

/* --- single kernels (one per reference entry point) --- */
/* returns the number of test particles skipped as out of grid or non-finite */
int hvo_deposit(const double *x, const double *y, const double *z, int np, int nn,
                const hvo_world *w, int64_t *acc, double *counts);
void hvo_counts_from_fixed(const int64_t *acc, double *counts, long long len);
void hvo_smear(const double *counts, double *rho, const hvo_world *w, double sigma_r, int ppn);
void hvo_smear_literal(const double *counts, double *rho, const hvo_world *w, double sigma_r, int ppn);
void hvo_skyrme_potentials(const double *rho, double *U, const hvo_world *w, const hvo_skyrme *s);
void hvo_coulomb_bc(double *phi, const hvo_world *w, int z);
int hvo_sor(double *phi, const double *rho_p, const hvo_world *w, double omega, double tol,
            int max_it, int *iters, double *max_diff);
void hvo_merge(double *U, const double *phi, const hvo_world *w);
void hvo_fdm(const double *U, double *Fx, double *Fy, double *Fz, const hvo_world *w);
void hvo_gather(const double *x, const double *y, const double *z, int np, int nn,
                const double *Fx, const double *Fy, const double *Fz, const hvo_world *w,
                double *fx, double *fy, double *fz);
void hvo_kick(double *kx, double *ky, double *kz, const double *fx, const double *fy,
              const double *fz, int total, double dt);
void hvo_drift(double *x, double *y, double *z, const double *kx, const double *ky,
               const double *kz, int total, double dt);
void hvo_excite(double *kz, int np, int nn, double eta, int z, int n, int mode);
/* out[0..3] = sum z_p, sum z_n, sum r2_p (in box), sum r2_n (in box) as
 * exact 2^-30 fixed-point sums; cnt[0..1] in-box counts */
void hvo_stats_sums128(const double *x, const double *y, const double *z, int np, int nn,
                       const hvo_world *w, __int128 out[4], long long cnt[2]);
void hvo_stats_sums(const double *x, const double *y, const double *z, int np, int nn,
                    const hvo_world *w, double out[4], long long cnt[2]);
/* st = com_p, com_n, msr_p, msr_n (src/hyper_vlasov.cpp:65-68) */
void hvo_stats(const double *x, const double *y, const double *z, int np, int nn,
               const hvo_world *w, double st[4]);
double hvo_center_of_mass(const double *z, int np, int nn, int type);
double hvo_mean_squared_radius(const double *x, const double *y, const double *z, int np, int nn,
                               const hvo_world *w, int type);

/* --- the step loop (src/hyper_vlasov.cpp:33-118, without file I/O) --- */
/* stats[4*s + {0,1,2,3}] = com_p, com_n, msr_p, msr_n of the positions at the
 * start of step s (the reference prints these every `substeps` steps);
 * sor_iters[0] = pre-loop solve, sor_iters[s+1] = step s.  phi (G values) is the
 * Coulomb potential; if phi_inout is non-NULL it is used as the initial state
 * (warm start) and returns the final one, else the faces are set and the
 * interior starts at zero like the reference.  Returns a status: 0 ok, >0
 * number of SOR solves that hit max_it without converging. */
int hvo_simulate(double *x, double *y, double *z, double *kx, double *ky, double *kz,
                 double *fx, double *fy, double *fz, int np, int nn, const hvo_world *w,
                 const hvo_skyrme *s, const hvo_params *p, double *phi_inout, double *stats,
                 int *sor_iters);

/* The loop only, from a mid-run state: forces fx..fz of the current positions
 * and the Coulomb potential phi_inout (required) are taken as given. */
int hvo_simulate_loop(double *x, double *y, double *z, double *kx, double *ky, double *kz,
                      double *fx, double *fy, double *fz, int np, int nn, const hvo_world *w,
                      const hvo_skyrme *s, const hvo_params *p, double *phi_inout, double *stats,
                      int *sor_iters);

/* Threads of the oracle's OpenMP parallel regions in this process (what a
 * timing of the oracle ran on; no reference counterpart). */
int hvo_threads(void);

/* Total energy of a state (no reference counterpart; hv_energy_density in
 * hv_detmath.h): out = {E_kin + E_sky + E_coul, E_kin, E_sky, E_coul} [MeV],
 * E_kin = (hbar c)^2/(2 m c^2) sum_i |k_i|^2 / ppn, the grid parts
 * dV sum_cells e_sky / e_coul over the smeared density rho (2G) and phi (G);
 * every per-particle |k|^2 and per-cell dV*e is quantised to 2^-30 and summed
 * exactly, like the diagnostics sums. */
void hvo_total_energy(const double *kx, const double *ky, const double *kz, int total, const double *rho,
                      const double *phi, const hvo_world *w, const hvo_skyrme *s, int ppn, double out[4]);

/* mismatches of the fixed-divisor division (hv_detmath.h) against a / b */
long long hvo_div_fixed_mismatches(double b, double lo, double hi, long long n, unsigned long long seed);

#ifdef __cplusplus
}
#endif

#endif
