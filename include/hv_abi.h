/*
 * hv_abi.h -- C-ABI of libhvamd.so, the MI355X (gfx950) implementation of the
 * hyper-vlasov per-timestep test-particle loop.
 *
 * The reference runs this loop on the CPU as free C++ templates over raw host
 * SoA pointers, called from cpu_simulate (src/hyper_vlasov.cpp:33-118), with a
 * commented-out gpu_simulate stub of the same signature
 * (src/hyper_vlasov.cpp:120-123) and a `use_gpu` input key
 * (include/init.hpp:322-323) as its hook for a device path.  This library fills
 * that hook: one entry point per reference kernel, with the same meaning and
 * the same in-place semantics, plus a fused whole-loop fast path.
 *
 * Conventions
 *   - Every call returns int: 0 ok, < 0 error (see HV_ERR_*), > 0 warning bit
 *     mask (HV_WARN_*).  hv_last_error() gives a message.  Nothing calls exit().
 *     (The reference returns void, prints "SOR COULOMB DID NOT CONVERGE!" on
 *     non-convergence, include/physics.hpp:155-156, and exit(1)s on allocation
 *     failure, include/init.hpp:94-95.)
 *   - Device buffers belong to the context.  Host arrays stay caller-owned and
 *     are touched only inside hv_upload_* / hv_download_* and the calls that
 *     return host scalars.  (Reference: caller allocates with create_* and
 *     frees with free_*, include/init.hpp:90-157, 257-273.)
 *   - Particles are stored protons first, then neutrons; species is implied by
 *     index, as in the reference (include/tools.hpp:62, 122).
 *   - Grid arrays are row-major, z fastest: node (i,j,k) at (i*ny + j)*nz + k
 *     (IDX, include/sim_structs.hpp:38).  "Double" fields hold [proton G |
 *     neutron G] (include/init.hpp:102-111).
 *   - One HIP stream per context; a context is not thread-safe (use one per
 *     host thread).  Calls are stream-ordered; those returning host scalars
 *     synchronise.
 *   - Precision HV_PREC_F64 reproduces the CPU restatement bit for bit.
 *     HV_PREC_MIXED stores particles and the smeared grid fields in FP32 and
 *     keeps the Coulomb potential, its final SOR sweeps and all diagnostics
 *     reductions in FP64.
 */
#ifndef HV_ABI_H
#define HV_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define HV_OK 0
#define HV_ERR_ARG (-1)     /* invalid argument or size */
#define HV_ERR_HIP (-2)     /* HIP runtime error */
#define HV_ERR_ALLOC (-3)   /* device allocation failed */
#define HV_ERR_STATE (-4)   /* call not valid in the current context state */
#define HV_ERR_IO (-5)      /* a file could not be opened or written */
#define HV_WARN_SOR_NOT_CONVERGED 1  /* SOR hit max_it with max|dphi| > tol */
#define HV_WARN_PARTICLE_OUTSIDE 2   /* some particles were outside the grid or non-finite */
#define HV_WARN_SMEAR_MARGIN 4       /* a smearing tap lies within 1e-9 of the cutoff sphere */
#define HV_WARN_INIT_NOT_CONVERGED 8 /* hv_init_woods_saxon: 32 rounds did not settle the Fermi levels */

/* ---- precision ----------------------------------------------------------- */
#define HV_PREC_F64 0
#define HV_PREC_MIXED 1

/* ---- species selectors (include/sim_structs.hpp:28-30) -------------------- */
#define HV_PROTONS 0
#define HV_NEUTRONS 1

/* ---- boost modes --------------------------------------------------------- */
#define HV_BOOST_REF_WEIGHTED 0 /* protons kz += eta N/A, neutrons kz -= eta Z/A (include/physics.hpp:237-246) */
#define HV_BOOST_SYMMETRIC 1    /* protons kz += eta, neutrons kz -= eta */

/* ---- device fields (for hv_upload_field / hv_download_field) -------------- */
#define HV_FIELD_DENSITY 0    /* 2G: CIC counts after hv_cic_deposit, density after hv_smear_density */
#define HV_FIELD_POTENTIALS 1 /* 2G: U_p | U_n */
#define HV_FIELD_COULOMB 2    /* G : phi (persistent SOR warm-start state) */
#define HV_FIELD_FORCES_X 3   /* 2G */
#define HV_FIELD_FORCES_Y 4   /* 2G */
#define HV_FIELD_FORCES_Z 5   /* 2G */

/* World<T> (include/sim_structs.hpp:73-76): n nodes and half-width d_max [fm]
 * per axis; spacing 2 d_max / n; node i at d_max (2 i / n - 1). */
typedef struct hv_world {
    int n[3];
    double d_max[3];
} hv_world;

/* Skyrme<T> (include/sim_structs.hpp:84-86): U = A t + B t^gamma + 2 tau C (rho_n - rho_p)/rho0 */
typedef struct hv_skyrme {
    double A, B, C, gamma;
} hv_skyrme;

/* The loop parameters of Parameters<T> (include/sim_structs.hpp:49-53) plus
 * the knobs the reference hard-codes (SOR omega/tolerance/cap,
 * include/physics.hpp:103-104, include/sim_structs.hpp:33,44) and the boost
 * convention. */
typedef struct hv_params {
    double t_f;           /* final time [fm/c]; dt = t_f / steps (src/hyper_vlasov.cpp:35) */
    int steps;
    int substeps;         /* stats cadence of the stats.txt writer */
    double t_exc;         /* boost once at the first step with step*dt >= t_exc */
    double eta_exc;       /* boost strength [fm^-1] */
    int boost_mode;       /* HV_BOOST_* */
    int z, n;             /* nucleus Z and N */
    double sigma_r;       /* smearing width [fm] (calc_sigma(r_fwhm)) */
    int part_per_nucleon; /* test particles per nucleon (density normalisation) */
    double sor_omega;     /* reference: 1.5 */
    double sor_tol;       /* reference: 1e-4, on max |dphi| of an iteration */
    int sor_max_it;       /* reference: 512 */
} hv_params;

typedef struct hv_ctx hv_ctx;

/* ---- context ------------------------------------------------------------- */
/* Allocates every device buffer the loop needs (replaces create_* in
 * include/init.hpp:90-157 and the allocations of src/hyper_vlasov.cpp:44-51).
 * Each species may hold at most 2^23 - 1 test particles. */
int hv_ctx_create(const hv_world *world, int n_protons, int n_neutrons, int precision, hv_ctx **out);
int hv_ctx_destroy(hv_ctx *ctx);
const char *hv_last_error(const hv_ctx *ctx);
int hv_sync(hv_ctx *ctx);
/* the HIP stream of the context (hipStream_t), for callers that interleave work */
void *hv_stream(hv_ctx *ctx);
/* number of particles skipped as out-of-grid or non-finite by the last deposit */
int hv_last_outside(hv_ctx *ctx, int *count);

/* ---- particle / field transfer (host FP64 <-> device) -------------------- */
/* TestParticles<T> (include/sim_structs.hpp:55-60): x,y,z [fm], kx,ky,kz
 * [fm^-1], fx,fy,fz [MeV/fm].  Any pointer may be NULL to skip that array. */
int hv_upload_particles(hv_ctx *ctx, const double *x, const double *y, const double *z,
                        const double *kx, const double *ky, const double *kz,
                        const double *fx, const double *fy, const double *fz);
int hv_download_particles(hv_ctx *ctx, double *x, double *y, double *z, double *kx, double *ky,
                          double *kz, double *fx, double *fy, double *fz);
int hv_upload_field(hv_ctx *ctx, int which, const double *host);
int hv_download_field(hv_ctx *ctx, int which, double *host);

/* ---- one call per reference kernel ---------------------------------------- */
/* distribute_volumetric_particles_cic (include/tools.hpp:36-96): particles ->
 * HV_FIELD_DENSITY as CIC counts (deterministic fixed-point accumulation). */
int hv_cic_deposit(hv_ctx *ctx);
/* compute_volumetric_densities (include/physics.hpp:203-234): counts ->
 * smeared density, in place on HV_FIELD_DENSITY (no caller scratch needed). */
int hv_smear_density(hv_ctx *ctx, double sigma_r, int part_per_nucleon);
/* compute_volumetric_skyrme_potentials (include/physics.hpp:273-281) */
int hv_skyrme_potentials(hv_ctx *ctx, const hv_skyrme *skyrme);
/* set_initial_coulomb_boundaries (include/tools.hpp:284-300) */
int hv_set_coulomb_bc(hv_ctx *ctx, int z);
/* compute_volumetric_coulomb_potentials_sor (include/physics.hpp:100-157):
 * red-black SOR on HV_FIELD_COULOMB with source = proton density.  iters and
 * max_diff (may be NULL) receive the iterations done and the last max|dphi|.
 * tol >= 0; tol 0 runs exactly max_it iterations. */
int hv_coulomb_sor(hv_ctx *ctx, double omega, double tol, int max_it, int *iters, double *max_diff);
/* merge_volumetric_potentials (include/tools.hpp:378-384): U_p += phi */
int hv_merge_potentials(hv_ctx *ctx);
/* compute_volumetric_forces_fdm (include/physics.hpp:160-200): U -> forces */
int hv_forces_fdm(hv_ctx *ctx);
/* distribute_forces_to_particles_cic (include/tools.hpp:99-168): forces -> particle f */
int hv_cic_gather(hv_ctx *ctx);
/* update_momenta_half (include/physics.hpp:249-258) */
int hv_kick_half(hv_ctx *ctx, double dt);
/* update_positions_full (include/physics.hpp:261-270) */
int hv_drift(hv_ctx *ctx, double dt);
/* nuclear_excitation (include/physics.hpp:237-246), mode HV_BOOST_* */
int hv_excite(hv_ctx *ctx, double eta, int z, int n, int mode);
/* center_of_mass (include/tools.hpp:328-338): mean z of a species */
int hv_center_of_mass(hv_ctx *ctx, int type, double *out);
/* mean_squared_radius (include/tools.hpp:303-325): <r^2> of in-box particles of a species */
int hv_mean_squared_radius(hv_ctx *ctx, int type, double *out);

/* ---- total energy (no counterpart in the reference) ---------------------- */
/* Total energy of the current state [MeV], the conserved quantity whose drift
 * measures the integration error:
 *   E_kin  = (hbar c)^2 / (2 m c^2) sum_i |k_i|^2 / ppn   (the prefactor is
 *            kinetic_energy(), include/physics_formulas.hpp:45-49)
 *   E_sky  = dV sum_cells [A rho^2/(2 rho0) + B rho0 t^(gamma+1)/(gamma+1)
 *            + C (rho_n - rho_p)^2 / rho0],   t = (rho_p + rho_n) / rho0
 *   E_coul = dV sum_cells rho_p phi / 2
 * whose derivative in rho_q is skyrme_potential (include/physics_formulas.hpp:71-77)
 * plus phi for protons.  Evaluated on the particles' momenta and on the
 * smeared density and phi held by the context, which hv_run_init and
 * hv_run_steps leave consistent with the positions (momenta synchronous after
 * the closing half kick).  Each |k_i|^2 and dV*e is quantised to 2^-30 and
 * summed exactly (order-free).  out = {E_kin + E_sky + E_coul, E_kin, E_sky, E_coul}. */
int hv_total_energy(hv_ctx *ctx, const hv_skyrme *skyrme, int part_per_nucleon, double out[4]);

/* ---- VTK snapshots (include/init.hpp:160-254) ------------------------------ */
/* The reference's snapshot of the current fields, written where
 * src/hyper_vlasov.cpp:72-85 writes out-%04d.vtk every `substeps` steps:
 * legacy binary STRUCTURED_POINTS, VECTORS forces_p / forces_n (the FDM force
 * of the current potentials, output_vector_field), SCALARS density_{p,n,t}
 * and potentials_{p,n,t} (output_scalar_field), big-endian float32, x fastest.
 * Replaces the output_vtk_header_start / output_vector_field /
 * output_scalar_field sequence of cpu_simulate. */
int hv_write_vtk(hv_ctx *ctx, const char *path);
/* The same file from host fields in the reference layout ((i ny + j) nz + k,
 * protons then neutrons; 2G values each); single != 0 writes what the float
 * instantiation writes (d_max and the p + n sum in float).  No GPU needed. */
int hv_vtk_write_fields(const char *path, const hv_world *world, const double *density, const double *potentials,
                        const double *fx, const double *fy, const double *fz, int single);

/* ---- fused step loop (the body of cpu_simulate, src/hyper_vlasov.cpp:53-104) */
/* hv_run_init: Coulomb faces + pre-loop field solve + gather (:53-62).
 * hv_run_steps: steps [first_step, first_step + nsteps) of the loop (:63-104);
 * the boost fires once, at the first step with step*dt >= t_exc.
 * stats (may be NULL) receives 4 doubles per step: centre of mass z of protons
 * and neutrons and <r^2> of protons and neutrons for the positions at the
 * start of the step (the quantities of the stats.txt line, :65-70);
 * sor_iters (may be NULL) the SOR iterations of each step's solve. */
int hv_run_init(hv_ctx *ctx, const hv_skyrme *skyrme, const hv_params *params, int *sor_iters);
int hv_run_steps(hv_ctx *ctx, const hv_skyrme *skyrme, const hv_params *params, int first_step,
                 int nsteps, double *stats, int *sor_iters);

/* gpu_simulate's body without file output: hv_run_init + hv_run_steps(0, steps). */
int hv_simulate(hv_ctx *ctx, const hv_skyrme *skyrme, const hv_params *params, double *stats,
                int *sor_iters);

/* ---- tuning options --------------------------------------------------------- */
/* HV_OPT_SORT_EVERY: steps between re-sorts of the particle arrays by
 * (species, 4^3-cell brick) (default 8; 0 disables sorting).  Sorting never
 * changes results -- deposits and diagnostics are exact integer sums and
 * hv_download_particles returns the caller's order -- only speed. */
#define HV_OPT_SORT_EVERY 0
/* HV_OPT_SOR_RESIDENT: 1 runs each SOR solve as one cooperative launch with
 * the potential held in LDS (when the grid fits the CUs' LDS); 0 (default)
 * launches one kernel per red or black sweep.  Same bits either way.  Reading
 * it back gives 0 when the resident form is off or unavailable. */
#define HV_OPT_SOR_RESIDENT 1
/* HV_OPT_SOR_FUSED: 1 runs one launch per SOR iteration (red and black
 * sweeps fused, out of place, the red ring recomputed per block); 0 (default)
 * one launch per sweep.  Same bits. */
#define HV_OPT_SOR_FUSED 2
/* HV_OPT_SOR_RESIDENT_F32: 1 (default) runs the mixed form's FP32 correction
 * iteration as one cooperative launch with the correction and its source held
 * in LDS, when the grid fits; 0 uses one launch per sweep.  The FP64 sweeps
 * that end every mixed solve apply the reference's stop rule either way. */
#define HV_OPT_SOR_RESIDENT_F32 3
/* HV_OPT_SOR_F32_LAG: the LDS-resident FP32 iteration stops after iteration
 * it once the max |update| of iteration it - lag is below tol (default 2,
 * range 1..64), so that its tiles never wait on a whole-grid reduction of the
 * current one. */
#define HV_OPT_SOR_F32_LAG 4
/* HV_OPT_SMEAR_TILE: widest column tile (8, 16, 32 or 64, default 64) the
 * smearing's LDS tiles may use; narrower tiles trade halo re-reads for more
 * blocks.  Same bits for every width. */
#define HV_OPT_SMEAR_TILE 5
/* HV_OPT_BATCH_GROUP_KIB (ensembles only): the SOR of a batch iterates its
 * instances in groups whose sweep working set fits this many KiB (default 0:
 * all instances together).  Same bits for every value. */
#define HV_OPT_BATCH_GROUP_KIB 6
/* HV_OPT_SMEAR_FUSED: smear a single grid with the fused plane-walking kernel
 * (the D channels stay in LDS) instead of the two-pass form (default 0: the
 * two-pass form, faster on one grid: after a deposit it visits only the count
 * box that the deposit records, the nodes that can hold a count, and writes
 * exact zeros beyond that box grown by the smearing radius; ensembles default
 * to the fused kernel).  Same bits either way. */
#define HV_OPT_SMEAR_FUSED 7
/* HV_OPT_BATCH_GRAPH (ensembles only): from the second step of a
 * hv_batch_run_steps call on, replay one captured hipGraph of the step's
 * device part (particle pass, diagnostics, smearing, source split and, in the
 * mixed form, the defect, the persistent LDS-resident FP32 SOR launch and the
 * fold; with HV_OPT_BATCH_DEVICE_STEER the rest of the step as well) instead
 * of launching its kernels one by one (default 1; off while stages are
 * timed).  Same bits either way. */
#define HV_OPT_BATCH_GRAPH 8
/* HV_OPT_SOR_F32_TILES_PER_CU: LDS-resident FP32 correction tiles per CU of a
 * single grid: 1 (1024-thread blocks) or 2 (512-thread blocks, half-size
 * tiles; a tile's wait for its neighbours' faces overlaps the other tile's
 * sweeps).  Same bits either way. */
#define HV_OPT_SOR_F32_TILES_PER_CU 9
/* HV_OPT_SOR_F32_OWNED: the body of the LDS-resident FP32 correction, the
 * highest that fits the tiling (for a batch: every instance's):
 *   2 (default) deep-halo lane-owned: one neighbour exchange per iteration
 *     (two face rows and the corners; the red ring around the tile is
 *     recomputed from them),
 *   1 lane-owned: every node of a tile updated by a fixed lane from
 *     precomputed descriptors, tiles with a one-row border in LDS, two
 *     exchanges per iteration,
 *   0 the body that derives each node's indexes in every sweep.
 * Same bits for every value.  hv_get_option / hv_batch_get_option report the
 * body in use once the form has solved. */
#define HV_OPT_SOR_F32_OWNED 10
/* HV_OPT_SOR_EXTRAPOLATE (mixed form only; 0 for HV_PREC_F64, which starts
 * every solve from the last potential as the reference does): the start of
 * each Coulomb solve of the fused loop (hv_run_init / hv_run_steps) is
 * extrapolated in time from the potentials of the last converged loop solves
 * by a polynomial predictor: 0 none (the last potential), 1 line through 2
 * points, 2 quadratic through 3, 3 least-squares quadratic through 5
 * (default), 4 least-squares quadratic through 4, 5 least-squares line
 * through 3.  The FP32 correction and the FP64 finishing sweeps then run from
 * that start under the same stop rule (max |dphi| of an FP64 iteration below
 * tol); the solution differs from the FP64 form's by the stop rule's own
 * error, not by precision.  Until the ring holds a predictor's points a line
 * (or no extrapolation) is used.  Only loop solves that converged count as
 * history; run_init, uploads of particles or of the potential and
 * hv_coulomb_sor clear it. */
#define HV_OPT_SOR_EXTRAPOLATE 11
/* HV_OPT_SOR_F32_TOL_SHIFT (mixed form): the FP32 correction stops once its
 * max |update| is below tol / 2^shift (0..20, default 0); the FP64 finishing
 * sweeps apply the reference's rule with tol itself either way. */
#define HV_OPT_SOR_F32_TOL_SHIFT 12
/* HV_OPT_BATCH_DEVICE_STEER (ensembles, mixed form with the LDS-resident FP32
 * correction): 1 (default) runs the whole step on the device: the FP64
 * finishing sweeps of every instance in one cooperative launch that applies
 * the stop rule itself, then the potentials and a per-step record that keeps
 * each instance's extrapolation history; the step graph then holds the
 * entire step and the host reads back only every 16 steps (to re-plan the
 * resident launch) and at the end of the call.  0: the host reads each
 * step's FP32 iteration counts and steers the FP64 sweeps.  Same bits either
 * way. */
#define HV_OPT_BATCH_DEVICE_STEER 13
/* HV_OPT_SOR_F32_OMEGA (mixed form): the relaxation factor of the FP32
 * correction.  The FP64 finishing sweeps apply the reference's omega and stop
 * rule whatever the correction used, so the correction may relax at any
 * convergent factor: 0 (default) takes 2 / (1 + 2.2 s) from the grid, with s
 * = sqrt(1 - mu^2) and mu the Jacobi spectral radius of the grid's Laplacian
 * (s = sin(pi / (n - 1)) on an n^3 grid: 1.80 at 64^3, 1.86 at 96^3, 1.90 at
 * 128^3, the factors that took the fewest iterations on those grids), but
 * never less than the run's own sor_omega; 1 takes the run's sor_omega;
 * 1001..1999 take value / 1000. */
#define HV_OPT_SOR_F32_OMEGA 14
int hv_set_option(hv_ctx *ctx, int option, int value);
int hv_get_option(hv_ctx *ctx, int option, int *value);

/* ---- timing / introspection ---------------------------------------------- */
/* Milliseconds of device time between the last pair of hv_timer_start /
 * hv_timer_stop calls (hipEvents on the context stream). */
int hv_timer_start(hv_ctx *ctx);
int hv_timer_stop(hv_ctx *ctx, double *ms);
/* HBM bandwidth of a plain device copy on the current device: `bytes` (a
 * multiple of 16) read and written by this library's copy kernel
 * (k_copy_calibration, 16-byte words), reps timed runs after one untimed;
 * best_gbs = 2 bytes / best time [GB/s], best_us = best time.  Allocates and
 * frees its two buffers.  The same kernel is the known-size copy that scales
 * HBM byte counters (FETCH_SIZE, WRITE_SIZE) of a profiled run. */
int hv_copy_bandwidth(long long bytes, int reps, double *best_gbs, double *best_us);
/* Library build string (arch, flags). */
const char *hv_build_info(void);

/* ---- kernel-class profiling (roofline accounting) -------------------------- */
/* When enabled, the loop times kernels with hipEvent pairs on the context
 * stream and accumulates device milliseconds and timed launches per class:
 * every particle pass, smearing and potential launch, every LDS-resident SOR
 * launch, and one SOR sweep launch in eight (issued with start/stop events
 * that the runtime stamps at that dispatch's own begin and end, so launch gaps
 * stay out; the sampled average is the kernel's average duration). */
#define HV_KCLASS_SOR 0       /* one red or black SOR sweep */
#define HV_KCLASS_PARTICLE 1  /* fused gather + kick + drift + deposit + diagnostics pass */
#define HV_KCLASS_SMEAR 2     /* the smearing passes of one step (one launch = one step's set) */
#define HV_KCLASS_POTENTIAL 3 /* Skyrme + Coulomb merge */
#define HV_KCLASS_SOR_F32 4   /* one sweep of the mixed form's FP32 SOR correction */
#define HV_KCLASS_SOR_RES_F32 5 /* one LDS-resident launch of the FP32 correction (a whole solve) */
#define HV_KCLASS_SOR_STAGE 6   /* ensembles: the whole SOR solve of a step (launches and read-backs) */
#define HV_KCLASS_COUNT 7
int hv_profile_enable(hv_ctx *ctx, int on); /* also resets the counters */
int hv_profile_read(hv_ctx *ctx, int kclass, double *ms, long long *launches);
/* Launches of a class issued since the last hv_profile_enable, timed or not
 * (for the sampled SOR sweep classes: the sampled time per launch times this
 * count estimates the class's whole device time). */
int hv_profile_issued(hv_ctx *ctx, int kclass, long long *issued);
/* SOR work since the last hv_profile_enable: solves, FP32 correction
 * iterations (mixed form) and FP64 iterations. */
int hv_sor_counts(hv_ctx *ctx, long long *solves, long long *iters_f32, long long *iters_f64);

/* ---- ensembles (no counterpart in the reference) ------------------------ */
/* For callers that run many nuclei (Monte Carlo ensembles, parameter scans,
 * the dipole response of a chart of nuclei): n instances, each with its own
 * grid, particle counts, Skyrme set and loop parameters, stepped together with
 * one launch per loop stage for the whole ensemble.  Instance i gives the bits
 * of a solo context run of it with the same options (the LDS-resident and
 * fused SOR forms off: a batch uses the per-sweep SOR launches).  All
 * instances share the precision and sor_max_it. */
typedef struct hv_batch hv_batch;
int hv_batch_create(int n, const hv_world *worlds, const int *n_protons, const int *n_neutrons, int precision,
                    hv_batch **out);
int hv_batch_destroy(hv_batch *batch);
const char *hv_batch_last_error(const hv_batch *batch);
int hv_batch_size(const hv_batch *batch);
/* instance i's context: particle/field transfer, diagnostics, total energy */
hv_ctx *hv_batch_instance(hv_batch *batch, int i);
/* HV_OPT_SORT_EVERY or HV_OPT_SMEAR_TILE (applied to every instance), or
 * HV_OPT_BATCH_GROUP_KIB, HV_OPT_SOR_RESIDENT_F32, HV_OPT_SOR_F32_LAG,
 * HV_OPT_SOR_F32_OWNED, HV_OPT_SMEAR_FUSED (default 1 for a batch),
 * HV_OPT_BATCH_GRAPH, HV_OPT_BATCH_DEVICE_STEER, HV_OPT_SOR_EXTRAPOLATE and
 * HV_OPT_SOR_F32_OMEGA (every instance) */
int hv_batch_set_option(hv_batch *batch, int option, int value);
/* the value of a batch option (HV_OPT_SOR_RESIDENT_F32 and
 * HV_OPT_SOR_F32_OWNED: whether that form runs, once the batch has solved) */
int hv_batch_get_option(hv_batch *batch, int option, int *value);
/* skyrme[n], params[n]; sor_iters[n] (may be NULL): the pre-loop solves */
int hv_batch_run_init(hv_batch *batch, const hv_skyrme *skyrme, const hv_params *params, int *sor_iters);
/* steps [first_step, first_step + nsteps) of every instance; stats (may be
 * NULL) receives n * nsteps * 4 doubles, instance-major, as hv_run_steps;
 * sor_iters (may be NULL) n * nsteps ints, instance-major. */
int hv_batch_run_steps(hv_batch *batch, const hv_skyrme *skyrme, const hv_params *params, int first_step,
                       int nsteps, double *stats, int *sor_iters);
/* time per stage of the batched loop (HV_KCLASS_PARTICLE, _SMEAR,
 * _SOR_STAGE, _POTENTIAL; enabling synchronises after every stage) and of one
 * SOR sweep launch in eight (HV_KCLASS_SOR, _SOR_F32: the sweep kernel's own
 * dispatch time) */
int hv_batch_profile_enable(hv_batch *batch, int on);
int hv_batch_profile_read(hv_batch *batch, int kclass, double *ms, long long *launches);
/* SOR work since hv_batch_profile_enable: sweep launches (each covers every
 * instance still iterating) of the FP32 correction and of FP64, and the
 * interior-node iterations summed over instances (sum_i iterations_i x
 * (nx-2)(ny-2)(nz-2)_i) */
int hv_batch_sor_counts(hv_batch *batch, long long *sweeps_f32, long long *sweeps_f64, long long *node_iters_f32,
                        long long *node_iters_f64);
/* steps run by replaying a captured step graph, and graphs captured, since
 * the batch was created (HV_OPT_BATCH_GRAPH) */
int hv_batch_graph_info(hv_batch *batch, long long *graph_steps, long long *captures);

/* ---- diagnostics without allocation (no counterpart in the reference) ----- */
/* For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the four
 * values of hv_center_of_mass / hv_mean_squared_radius (include/tools.hpp:
 * 303-338; protons then neutrons: mean z, mean z, <r^2> of the in-box
 * particles, <r^2> of the in-box particles) for positions the caller already
 * holds in device memory, protons first (double arrays for HV_PREC_F64, float
 * for HV_PREC_MIXED).  `out` is device memory for 4 doubles, written on
 * `stream` (a hipStream_t; NULL = the default stream); nothing is allocated,
 * copied to the host or synchronised.  The sums are the exact fixed-point
 * sums of the context entry points, so the values are the same bits.  Scratch
 * is one file-scope device array per GPU (about 3 MiB, held until the process
 * ends) shared by every call on that GPU: calls that overlap on one GPU (other
 * streams, other threads) race.  At most HV_NOALLOC_MAX_PARTICLES particles;
 * n_protons, n_neutrons >= 1; larger sizes return HV_ERR_ARG.  A species with
 * no in-box particle gets a NaN <r^2>. */
#define HV_NOALLOC_MAX_PARTICLES (1 << 24)
int hv_stats_noalloc(const hv_world *world, int precision, const void *x, const void *y, const void *z,
                     int n_protons, int n_neutrons, double *out, void *stream);

/* ---- adjoint (no counterpart in the reference) ---------------------------- */
/* For fitting the Skyrme coefficients to a dipole response (the use of the
 * reference's fit scripts): starting from the particles the context holds
 * (uploaded, caller's order), runs hv_run_init and nsteps steps (FP64 context
 * only), and returns
 *   loss = sum_{s < nsteps} (D(t_s) - d_target[s])^2,  D = N Z / A (<z>_p - <z>_n)
 * of the positions at the start of step s (scripts/programs.py:99), the
 * series D in d_series[nsteps] (may be NULL) and, when grad is not NULL, the
 * gradient dloss/d(A, B, C, gamma) (hv_skyrme order) by the discrete adjoint of
 * the loop, with trajectory checkpoints every checkpoint_every steps.  The
 * adjoint treats each SOR solve as exact.  The particles are restored to the
 * initial state on return (fields and phi are not), so calls can be repeated.
 * Returns warning bits as hv_run_steps. */
int hv_dipole_loss_grad(hv_ctx *ctx, const hv_skyrme *skyrme, const hv_params *params, int nsteps,
                        const double *d_target, int checkpoint_every, double *loss, double *d_series, double *grad);

#ifdef __cplusplus
}
#endif

#endif /* HV_ABI_H */
