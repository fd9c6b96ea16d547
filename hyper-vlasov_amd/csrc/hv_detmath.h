/*
 * hv_detmath.h -- deterministic arithmetic shared by the HIP library and its
 * CPU oracle.
 *
 * Everything here is built from IEEE-754 basic operations (+, -, *, /, rint)
 * and exact bit manipulation only, so the same input gives the same bits on
 * gfx950 and on x86-64, provided both sides compile with -ffp-contract=off
 * (no silent a*b+c -> fma fusion).  That is what lets the FP64 GPU step loop
 * reproduce the CPU restatement bit for bit over long trajectories.
 *
 * Contents
 *   - physical constants of the reference (include/sim_structs.hpp:40-46,
 *     include/tools.hpp:295, include/physics.hpp:118)
 *   - hv_det_exp / hv_det_log / hv_det_pow: textbook Cody-Waite range
 *     reduction + Taylor / atanh series, written from the mathematics.
 *     Used for the Skyrme t^gamma term (include/physics_formulas.hpp:75,
 *     which calls std::pow); agreement with libm pow is tested to a few ulp.
 *   - fixed-point deposit weights (2^-40 quanta) so that CIC deposition is an
 *     integer sum, exact and independent of summation order.
 *   - host-only builders for the Gaussian smearing tap table and the Coulomb
 *     Dirichlet faces (these call libm exp/pow/sqrt on the host, once).
 *
 * The header is valid C99 (oracle, gcc) and HIP C++ (library, hipcc).
 */
#ifndef HV_DETMATH_H
#define HV_DETMATH_H

#include <stdint.h>
#include <string.h>
#include <math.h>

#if defined(__HIPCC__)
#define HV_HD __host__ __device__
#else
#define HV_HD
#endif

#ifdef __cplusplus
#define HV_INLINE static inline
#else
#define HV_INLINE static inline
#endif

/* ---- constants (include/sim_structs.hpp:40-46) ---------------------------- */
#define HV_MC2 935.0          /* nucleon rest energy [MeV] */
#define HV_HBARC 197.33       /* hbar*c [MeV fm] */
#define HV_RHO0 0.16          /* saturation density [fm^-3] */
#define HV_E2 1.44            /* e^2 [MeV fm], include/physics.hpp:118, tools.hpp:295 */
#define HV_PI 3.141592653589793238462643383279502884
#define HV_SOR_TOL_REF 1e-4   /* include/sim_structs.hpp:44 */
#define HV_SOR_MAX_IT_REF 512 /* include/sim_structs.hpp:33 */
#define HV_SOR_OMEGA_REF 1.5  /* include/physics.hpp:103 */

/* ---- fixed-point deposit -------------------------------------------------- */
/* A CIC weight w in [0,1] is stored as rint(w * 2^40) in an int64 accumulator.
 * A node can take 2^23 full-weight contributions before the accumulator could
 * overflow, so the library caps each species at 2^23 - 1 test particles. */
#define HV_FIX_SHIFT 40
#define HV_FIX_SCALE 1099511627776.0           /* 2^40 */
#define HV_FIX_INV 9.094947017729282379150390625e-13 /* 2^-40, exact */
#define HV_MAX_SPECIES_PARTICLES ((1 << 23) - 1)

HV_HD HV_INLINE long long hv_fix_weight(double w) {
    return (long long)rint(w * HV_FIX_SCALE);
}

HV_HD HV_INLINE double hv_fix_to_double(long long v) {
    return (double)v * HV_FIX_INV;
}

/* ---- fixed-point diagnostics sums ----------------------------------------- */
/* Centre-of-mass and <r^2> sums (include/tools.hpp:303-338) are taken as exact
 * integer sums of rint(v * 2^30) per particle, so they do not depend on the
 * order particles are visited in.  |v| is clamped to 2^24 (a particle that far
 * out is already meaningless); the quantum is 9.3e-10 fm (fm^2 for r^2). */
#define HV_STAT_SCALE 1073741824.0                 /* 2^30 */
#define HV_STAT_INV 9.31322574615478515625e-10     /* 2^-30, exact */
#define HV_STAT_CLAMP 16777216.0                   /* 2^24 */

HV_HD HV_INLINE long long hv_stat_fix(double v) {
    if (!(v < HV_STAT_CLAMP)) v = (v == v) ? HV_STAT_CLAMP : 0.0;
    if (!(v > -HV_STAT_CLAMP)) v = -HV_STAT_CLAMP;
    return (long long)rint(v * HV_STAT_SCALE);
}

/* mean = (S * 2^-30) / count with S the exact 128-bit sum (host only) */
HV_INLINE double hv_stat_mean(__int128 S, long long count) {
    return ((double)S * HV_STAT_INV) / (double)count;
}

/* ---- exact bit helpers ---------------------------------------------------- */
HV_HD HV_INLINE uint64_t hv_dbits(double x) {
    uint64_t u;
    memcpy(&u, &x, sizeof u);
    return u;
}

HV_HD HV_INLINE double hv_bitsd(uint64_t u) {
    double x;
    memcpy(&x, &u, sizeof x);
    return x;
}

/* 2^k for -1022 <= k <= 1023, exact */
HV_HD HV_INLINE double hv_pow2i(int k) {
    return hv_bitsd((uint64_t)(k + 1023) << 52);
}

/* ---- deterministic exp ---------------------------------------------------- */
/* exp(x) = 2^k * exp(r), k = rint(x / ln2), r = x - k ln2 with ln2 split in a
 * high part that has 21 trailing zero bits (k*ln2_hi is exact for |k| < 2^21)
 * and a low part; exp(r), |r| <= 0.347, by its Taylor series to degree 13
 * (truncation below 5e-18 relative). */
HV_HD HV_INLINE double hv_det_exp(double x) {
    if (x != x) return x;
    if (x > 709.0) return hv_bitsd(0x7ff0000000000000ULL); /* +inf */
    if (x < -745.0) return 0.0;
    const double inv_ln2 = 1.44269504088896338700e+00;
    const double ln2_hi = 6.93147180369123816490e-01; /* 0x3fe62e42fee00000 */
    const double ln2_lo = 1.90821492927058770002e-10; /* 0x3dea39ef35793c76 */
    double kd = rint(x * inv_ln2);
    double r = (x - kd * ln2_hi) - kd * ln2_lo;
    double p = 1.0 / 6227020800.0;       /* 1/13! */
    p = p * r + 1.0 / 479001600.0;       /* 1/12! */
    p = p * r + 1.0 / 39916800.0;        /* 1/11! */
    p = p * r + 1.0 / 3628800.0;         /* 1/10! */
    p = p * r + 1.0 / 362880.0;          /* 1/9!  */
    p = p * r + 1.0 / 40320.0;           /* 1/8!  */
    p = p * r + 1.0 / 5040.0;            /* 1/7!  */
    p = p * r + 1.0 / 720.0;             /* 1/6!  */
    p = p * r + 1.0 / 120.0;             /* 1/5!  */
    p = p * r + 1.0 / 24.0;              /* 1/4!  */
    p = p * r + 1.0 / 6.0;               /* 1/3!  */
    p = p * r + 0.5;                     /* 1/2!  */
    p = p * r + 1.0;
    p = p * r + 1.0;
    int k = (int)kd;
    if (k >= -1022) return p * hv_pow2i(k);
    /* subnormal result: scale in two exact steps */
    return (p * hv_pow2i(k + 600)) * hv_pow2i(-600);
}

/* ---- deterministic log ---------------------------------------------------- */
/* x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(f), f = (m-1)/(m+1),
 * |f| <= 0.1716, by the odd series f (1 + s/3 + s^2/5 + ... + s^12/25), s = f^2
 * (truncation below 2e-20 relative). */
HV_HD HV_INLINE double hv_det_log(double x) {
    if (x != x || x < 0.0) return hv_bitsd(0x7ff8000000000000ULL); /* NaN */
    if (x == 0.0) return -hv_bitsd(0x7ff0000000000000ULL);         /* -inf */
    uint64_t u = hv_dbits(x);
    if ((u >> 52) == 0x7ffULL) return x; /* +inf */
    int e = 0;
    if ((u >> 52) == 0) { /* subnormal: bring into the normal range, exactly */
        x = x * 18014398509481984.0; /* 2^54 */
        u = hv_dbits(x);
        e = -54;
    }
    e += (int)((u >> 52) & 0x7ffULL) - 1023;
    double m = hv_bitsd((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > 1.41421356237309504880) {
        m = m * 0.5;
        e += 1;
    }
    double f = (m - 1.0) / (m + 1.0);
    double s = f * f;
    double p = 1.0 / 25.0;
    p = p * s + 1.0 / 23.0;
    p = p * s + 1.0 / 21.0;
    p = p * s + 1.0 / 19.0;
    p = p * s + 1.0 / 17.0;
    p = p * s + 1.0 / 15.0;
    p = p * s + 1.0 / 13.0;
    p = p * s + 1.0 / 11.0;
    p = p * s + 1.0 / 9.0;
    p = p * s + 1.0 / 7.0;
    p = p * s + 1.0 / 5.0;
    p = p * s + 1.0 / 3.0;
    p = p * s + 1.0;
    double lm = (2.0 * f) * p;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double ed = (double)e;
    return ed * ln2_hi + (ed * ln2_lo + lm);
}

/* x^y for x >= 0 (the Skyrme density ratio is never negative). */
HV_HD HV_INLINE double hv_det_pow(double x, double y) {
    if (x == 0.0) return (y > 0.0) ? 0.0 : ((y == 0.0) ? 1.0 : hv_bitsd(0x7ff0000000000000ULL));
    if (x == 1.0) return 1.0;
    return hv_det_exp(y * hv_det_log(x));
}

/* ---- correctly rounded division by a fixed divisor ------------------------- */
/* a / b for a divisor b known in advance, with y = 1.0 / b (correctly rounded,
 * computed once): q = RN(a y) is within one ulp of a / b, r = a - q b is exact
 * (one fma), and RN(q + r y) is the correctly rounded quotient (Markstein's
 * theorem, for results away from overflow and underflow).  Three
 * instructions instead of a full IEEE division sequence; the same bits as
 * a / b, which tests/test_oracle_closed_forms.py checks on random operands for
 * the grid spacings used here.  Needs a fused multiply-add (fma()). */
HV_HD HV_INLINE double hv_div_fixed(double a, double b, double y) {
    double q = a * y;
    double r = fma(-q, b, a);
    return fma(r, y, q);
}

/* ---- Skyrme mean field (include/physics_formulas.hpp:71-77) --------------- */
/* U_q = A t + B t^gamma + 2 tau_q C (rho_n - rho_p)/rho0, t = (rho_p+rho_n)/rho0,
 * tau_p = -1, tau_n = +1.  Same operation order as the reference expression;
 * only pow is replaced by hv_det_pow. */
HV_HD HV_INLINE double hv_skyrme_f64(double A, double B, double C, double gamma,
                                    double rho_p, double rho_n, int neutron) {
    double tau = neutron ? 1.0 : -1.0;
    double rho = rho_p + rho_n;
    double t = rho / HV_RHO0;
    return A * t + B * hv_det_pow(t, gamma) + 2.0 * tau * C * ((rho_n - rho_p) / HV_RHO0);
}

/* Both species' potentials at one node, with the parts they share (t, t^gamma,
 * A t + B t^gamma and the isovector density) evaluated once: each result is the
 * same operations in the same order as hv_skyrme_f64 for its species, so the
 * same bits, for half the work. */
HV_HD HV_INLINE void hv_skyrme_pair_f64(double A, double B, double C, double gamma, double rho_p, double rho_n,
                                        double *u_p, double *u_n) {
    double rho = rho_p + rho_n;
    double t = rho / HV_RHO0;
    double base = A * t + B * hv_det_pow(t, gamma);
    double d = (rho_n - rho_p) / HV_RHO0;
    *u_p = base + 2.0 * -1.0 * C * d;
    *u_n = base + 2.0 * 1.0 * C * d;
}

/* ---- total energy (no counterpart in the reference) ----------------------- */
/* Energy density [MeV fm^-3] of the mean field, split into its Skyrme and
 * Coulomb parts:
 *   e_sky  = A rho^2 / (2 rho0) + B rho0 t^(gamma+1) / (gamma+1) + C (rho_n - rho_p)^2 / rho0
 *   e_coul = rho_p phi / 2
 * with rho = rho_p + rho_n and t = rho / rho0.  Its functional derivative in
 * rho_q is skyrme_potential (include/physics_formulas.hpp:71-77) plus, for
 * protons, phi (phi linear in rho_p through the symmetric discrete Laplacian).
 * t^(gamma+1) is taken as t * t^gamma with the same hv_det_pow the potential
 * uses. */
HV_HD HV_INLINE void hv_energy_density(double A, double B, double C, double gamma, double rho_p, double rho_n,
                                       double phi, double *e_sky, double *e_coul) {
    double rho = rho_p + rho_n;
    double t = rho / HV_RHO0;
    double d = rho_n - rho_p;
    *e_sky = A * rho * rho / (2.0 * HV_RHO0) + B * HV_RHO0 * (t * hv_det_pow(t, gamma)) / (gamma + 1.0) +
             C * d * d / HV_RHO0;
    *e_coul = 0.5 * rho_p * phi;
}

/* kinetic_energy() (include/physics_formulas.hpp:45-49): (hbar c)^2 / (2 m c^2) [MeV fm^2] */
#define HV_KIN_FAC (HV_HBARC * HV_HBARC / (2.0 * HV_MC2))

/* ---- grid geometry (include/tools.hpp:395-401, include/physics.hpp:102) ---- */
/* node i of n along an axis of half-width d_max sits at d_max (2 i / n - 1) */
HV_HD HV_INLINE double hv_node_pos(double d_max, int n, int i) {
    return d_max * (2.0 * i / n - 1.0);
}

HV_HD HV_INLINE double hv_spacing(double d_max, int n) {
    return 2.0 * d_max / n;
}

/* ---- Gaussian smearing tap table (host only) ------------------------------ */
/* The reference smears the CIC counts with exp(-d^2 / 2 sigma^2) over every
 * node pair with d^2 <= 16 sigma^2 (include/physics.hpp:203-234).  The kernel
 * factorises: exp(-(a^2 hx^2 + b^2 hy^2 + c^2 hz^2)/2s^2) = gx(a) gy(b) gz(c),
 * and the spherical cut is kept exactly by recording, for each (|a|,|b|), the
 * largest |c| inside the sphere.  Tables are indexed by |offset|. */
#define HV_SMEAR_MAXR 31

typedef struct hv_smear_table {
    int rx, ry, rz;                                   /* max |a|, |b|, |c| in the sphere */
    int kb[HV_SMEAR_MAXR + 1];                        /* |a| -> max |b| (-1: none) */
    int kc[HV_SMEAR_MAXR + 1][HV_SMEAR_MAXR + 1];     /* |a|,|b| -> max |c| (-1: none) */
    double gx[HV_SMEAR_MAXR + 1];
    double gy[HV_SMEAR_MAXR + 1];
    double gz[HV_SMEAR_MAXR + 1];
    double term;        /* (1/ppn) (2 pi sigma^2)^-3/2, include/physics.hpp:230 */
    double exp_term;    /* 1/(2 sigma^2), include/physics.hpp:205 */
    double cut2;        /* 16 sigma^2, include/physics.hpp:206 */
    int taps;           /* number of (a,b,c) offsets inside the sphere */
    double min_margin;  /* min |d^2 - cut2| / cut2 over the bounding box */
} hv_smear_table;

/* sigma from the full width at half maximum (include/physics_formulas.hpp:58-62) */
HV_INLINE double hv_sigma_from_fwhm(double fwhm) {
    double t = 2.0 * sqrt(2.0 * log(2.0));
    return fwhm / t;
}

/* Returns 0 on success, -1 if the sphere does not fit HV_SMEAR_MAXR. */
HV_INLINE int hv_smear_table_build(hv_smear_table *tb, const int n[3], const double d_max[3],
                                   double sigma_r, int part_per_nucleon) {
    double h[3];
    int r[3];
    memset(tb, 0, sizeof *tb);
    tb->exp_term = 1.0 / (2.0 * sigma_r * sigma_r);
    tb->cut2 = 16.0 * sigma_r * sigma_r;
    tb->term = (1.0 / part_per_nucleon) * (1.0 / pow(2.0 * HV_PI * sigma_r * sigma_r, 1.5));
    for (int ax = 0; ax < 3; ax++) {
        h[ax] = hv_spacing(d_max[ax], n[ax]);
        r[ax] = 0;
        while ((r[ax] + 1) < n[ax] && ((r[ax] + 1) * h[ax]) * ((r[ax] + 1) * h[ax]) <= tb->cut2)
            r[ax]++;
        if (r[ax] > HV_SMEAR_MAXR) return -1;
    }
    tb->rx = r[0];
    tb->ry = r[1];
    tb->rz = r[2];
    for (int a = 0; a <= HV_SMEAR_MAXR; a++) {
        tb->kb[a] = -1;
        for (int b = 0; b <= HV_SMEAR_MAXR; b++) tb->kc[a][b] = -1;
    }
    tb->min_margin = 1e300;
    tb->taps = 0;
    for (int a = 0; a <= r[0]; a++) {
        double ox = a * h[0];
        tb->gx[a] = exp(-(ox * ox) * tb->exp_term);
        for (int b = 0; b <= r[1]; b++) {
            double oy = b * h[1];
            for (int c = 0; c <= r[2]; c++) {
                double oz = c * h[2];
                double d2 = ox * ox + oy * oy + oz * oz;
                double mg = fabs(d2 - tb->cut2) / tb->cut2;
                if (mg < tb->min_margin) tb->min_margin = mg;
                if (d2 <= tb->cut2) {
                    if (c > tb->kc[a][b]) tb->kc[a][b] = c;
                    if (b > tb->kb[a]) tb->kb[a] = b;
                    tb->taps += (a ? 2 : 1) * (b ? 2 : 1) * (c ? 2 : 1);
                }
            }
        }
    }
    for (int b = 0; b <= r[1]; b++) {
        double oy = b * h[1];
        tb->gy[b] = exp(-(oy * oy) * tb->exp_term);
    }
    for (int c = 0; c <= r[2]; c++) {
        double oz = c * h[2];
        tb->gz[c] = exp(-(oz * oz) * tb->exp_term);
    }
    return 0;
}

/* ---- Coulomb Dirichlet faces (include/tools.hpp:284-300), host only ------- */
/* phi = 1.44 Z / r on every face node; interior untouched. */
HV_INLINE void hv_coulomb_faces(double *phi, const int n[3], const double d_max[3], int z) {
    int nx = n[0], ny = n[1], nz = n[2];
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) {
                    double vx = hv_node_pos(d_max[0], nx, i);
                    double vy = hv_node_pos(d_max[1], ny, j);
                    double vz = hv_node_pos(d_max[2], nz, k);
                    double rr = sqrt(vx * vx + vy * vy + vz * vz);
                    phi[((size_t)i * ny + j) * nz + k] = HV_E2 * z / rr;
                }
            }
}

#endif /* HV_DETMATH_H */
