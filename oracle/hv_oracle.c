/*
 * hv_oracle.c -- CPU restatement of the hyper-vlasov step loop.
 *
 * TEST INFRASTRUCTURE ONLY (see hv_oracle.h): callers are tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg.  Compile with
 * -ffp-contract=off so that a*b+c is never fused (the GPU library is compiled
 * the same way; together with hv_detmath.h this makes the FP64 paths agree
 * bit for bit).
 */
#include "hv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hv_detmath.h"

#include <omp.h>

int hvo_threads(void) {
    int n = 1;
#pragma omp parallel
    {
#pragma omp single
        n = omp_get_num_threads();
    }
    return n;
}

#define IDX3(i, j, k, ny, nz) ((((long long)(i)) * (ny) + (j)) * (nz) + (k))

static long long world_size(const hvo_world *w) {
    return (long long)w->n[0] * w->n[1] * w->n[2];
}

/* ------------------------------------------------------------------------- */
/* CIC deposit: include/tools.hpp:36-96.  Cell coordinate c = (n/2)(x/d_max+1)
 * (:50-52); particles with c outside [0, n) are skipped (:54-55) -- written as
 * !(0 <= c < n) so that NaN is skipped too instead of reaching (int)NaN; the
 * eight corner weights are the reference's (t*t*t products, :57-60) and
 * corners with index n are dropped (:67-94).  Weights accumulate as 2^-40
 * fixed point (see hv_oracle.h). */
int hvo_deposit(const double *x, const double *y, const double *z, int np, int nn,
                const hvo_world *w, int64_t *acc, double *counts) {
    int nx = w->n[0], ny = w->n[1], nz = w->n[2];
    long long G = world_size(w);
    int total = np + nn, skipped = 0;
    memset(acc, 0, sizeof(int64_t) * 2 * G);
    for (int i = 0; i < total; i++) {
        double cx = (nx / 2.0) * (x[i] / w->d_max[0] + 1.0);
        double cy = (ny / 2.0) * (y[i] / w->d_max[1] + 1.0);
        double cz = (nz / 2.0) * (z[i] / w->d_max[2] + 1.0);
        if (!(cx >= 0.0 && cy >= 0.0 && cz >= 0.0 && cx < nx && cy < ny && cz < nz)) {
            skipped++;
            continue;
        }
        int x0 = (int)cx, y0 = (int)cy, z0 = (int)cz;
        int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
        double d_x = cx - x0, d_y = cy - y0, d_z = cz - z0;
        double t_x = 1.0 - d_x, t_y = 1.0 - d_y, t_z = 1.0 - d_z;
        long long off = (i < np) ? 0 : G;
        acc[off + IDX3(x0, y0, z0, ny, nz)] += hv_fix_weight(t_x * t_y * t_z);
        if (x1 < nx) acc[off + IDX3(x1, y0, z0, ny, nz)] += hv_fix_weight(d_x * t_y * t_z);
        if (y1 < ny) acc[off + IDX3(x0, y1, z0, ny, nz)] += hv_fix_weight(t_x * d_y * t_z);
        if (x1 < nx && y1 < ny) acc[off + IDX3(x1, y1, z0, ny, nz)] += hv_fix_weight(d_x * d_y * t_z);
        if (z1 < nz) acc[off + IDX3(x0, y0, z1, ny, nz)] += hv_fix_weight(t_x * t_y * d_z);
        if (x1 < nx && z1 < nz) acc[off + IDX3(x1, y0, z1, ny, nz)] += hv_fix_weight(d_x * t_y * d_z);
        if (y1 < ny && z1 < nz) acc[off + IDX3(x0, y1, z1, ny, nz)] += hv_fix_weight(t_x * d_y * d_z);
        if (x1 < nx && y1 < ny && z1 < nz)
            acc[off + IDX3(x1, y1, z1, ny, nz)] += hv_fix_weight(d_x * d_y * d_z);
    }
    if (counts) hvo_counts_from_fixed(acc, counts, 2 * G);
    return skipped;
}

void hvo_counts_from_fixed(const int64_t *acc, double *counts, long long len) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < len; i++) counts[i] = hv_fix_to_double(acc[i]);
}

/* ------------------------------------------------------------------------- */
/* Gaussian smearing: include/physics.hpp:203-234 (+ copy_scalar_field,
 * include/tools.hpp:387-392, made unnecessary by writing out of place).
 * rho_q(p) = term * sum over offsets (a,b,c) with |offset|^2 <= 16 sigma^2 of
 * count_q(p + offset) exp(-|offset|^2 / 2 sigma^2), neighbours outside the grid
 * omitted, same species only.  Factorised evaluation, identical tap set:
 *   H_K(p) = sum_{|c|<=K} gz(|c|) count(p + c z)          (K = 0..rz)
 *           built as H_0 = gz(0) cnt(p), H_K = (H_{K-1} + gz(K) cnt(p-K)) + gz(K) cnt(p+K)
 *   D_a(p) = sum_{b=-kb(a)..kb(a)} gy(|b|) H_{kc(a,|b|)}(p + b y)   (b ascending)
 *   rho(p) = (sum_{a=-rx..rx, kb(|a|)>=0} gx(|a|) D_{|a|}(p + a x)) * term
 * Out-of-grid neighbours contribute exact zeros, so clipping is exact. */
void hvo_smear(const double *counts, double *rho, const hvo_world *w, double sigma_r, int ppn) {
    int nx = w->n[0], ny = w->n[1], nz = w->n[2];
    long long G = world_size(w);
    hv_smear_table tb;
    if (hv_smear_table_build(&tb, w->n, w->d_max, sigma_r, ppn) != 0) abort();
    int nh = tb.rz + 1, nd = tb.rx + 1;
    double *H = (double *)malloc(sizeof(double) * (size_t)nh * 2 * G);
    double *D = (double *)malloc(sizeof(double) * (size_t)nd * 2 * G);
    if (!H || !D) abort();
    /* z pass */
#pragma omp parallel for schedule(static)
    for (long long q = 0; q < 2 * G; q++) {
        long long node = q % G;
        int k = (int)(node % nz);
        double h = tb.gz[0] * counts[q];
        H[q] = h;
        for (int K = 1; K <= tb.rz; K++) {
            double lo = (k - K >= 0) ? counts[q - K] : 0.0;
            double hi = (k + K < nz) ? counts[q + K] : 0.0;
            h = (h + tb.gz[K] * lo) + tb.gz[K] * hi;
            H[(long long)K * 2 * G + q] = h;
        }
    }
    /* y pass */
#pragma omp parallel for schedule(static)
    for (long long q = 0; q < 2 * G; q++) {
        long long node = q % G;
        int j = (int)((node / nz) % ny);
        for (int a = 0; a <= tb.rx; a++) {
            double acc = 0.0;
            int kb = tb.kb[a];
            for (int b = -kb; b <= kb; b++) {
                int jb = j + b;
                int ab = b < 0 ? -b : b;
                double hv = (jb >= 0 && jb < ny) ? H[(long long)tb.kc[a][ab] * 2 * G + q + (long long)b * nz] : 0.0;
                acc = acc + tb.gy[ab] * hv;
            }
            D[(long long)a * 2 * G + q] = acc;
        }
    }
    /* x pass */
#pragma omp parallel for schedule(static)
    for (long long q = 0; q < 2 * G; q++) {
        long long node = q % G;
        int i = (int)(node / ((long long)ny * nz));
        double acc = 0.0;
        for (int a = -tb.rx; a <= tb.rx; a++) {
            int aa = a < 0 ? -a : a;
            if (tb.kb[aa] < 0) continue;
            int ia = i + a;
            double dv = (ia >= 0 && ia < nx) ? D[(long long)aa * 2 * G + q + (long long)a * ny * nz] : 0.0;
            acc = acc + tb.gx[aa] * dv;
        }
        rho[q] = acc * tb.term;
    }
    free(H);
    free(D);
}

/* Flat-order form of the same smearing, one exp per tap in the reference's
 * ascending-neighbour order (include/physics.hpp:215-228); O(G * taps).  Used
 * only to pin hvo_smear and to compare with the compiled reference loop. */
void hvo_smear_literal(const double *counts, double *rho, const hvo_world *w, double sigma_r, int ppn) {
    int nx = w->n[0], ny = w->n[1], nz = w->n[2];
    long long G = world_size(w);
    hv_smear_table tb;
    if (hv_smear_table_build(&tb, w->n, w->d_max, sigma_r, ppn) != 0) abort();
    double hx = hv_spacing(w->d_max[0], nx), hy = hv_spacing(w->d_max[1], ny), hz = hv_spacing(w->d_max[2], nz);
#pragma omp parallel for schedule(static)
    for (long long q = 0; q < 2 * G; q++) {
        long long node = q % G;
        int i = (int)(node / ((long long)ny * nz)), j = (int)((node / nz) % ny), k = (int)(node % nz);
        double acc = 0.0;
        for (int a = -tb.rx; a <= tb.rx; a++) {
            if (i + a < 0 || i + a >= nx) continue;
            double ox = a * hx;
            for (int b = -tb.ry; b <= tb.ry; b++) {
                if (j + b < 0 || j + b >= ny) continue;
                double oy = b * hy;
                for (int c = -tb.rz; c <= tb.rz; c++) {
                    if (k + c < 0 || k + c >= nz) continue;
                    double oz = c * hz;
                    double d2 = ox * ox + oy * oy + oz * oz;
                    if (d2 > tb.cut2) continue;
                    double fact = exp(-d2 * tb.exp_term);
                    acc += counts[q + ((long long)a * ny + b) * nz + c] * fact;
                }
            }
        }
        rho[q] = acc * tb.term;
    }
}

/* ------------------------------------------------------------------------- */
/* Skyrme potentials: include/physics.hpp:273-281 with skyrme_potential,
 * include/physics_formulas.hpp:71-77.  U = [U_p | U_n]. */
void hvo_skyrme_potentials(const double *rho, double *U, const hvo_world *w, const hvo_skyrme *s) {
    long long G = world_size(w);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < G; i++) {
        U[i] = hv_skyrme_f64(s->A, s->B, s->C, s->gamma, rho[i], rho[i + G], 0);
        U[i + G] = hv_skyrme_f64(s->A, s->B, s->C, s->gamma, rho[i], rho[i + G], 1);
    }
}

/* Coulomb Dirichlet faces: include/tools.hpp:284-300 */
void hvo_coulomb_bc(double *phi, const hvo_world *w, int z) {
    hv_coulomb_faces(phi, w->n, w->d_max, z);
}

/* ------------------------------------------------------------------------- */
/* Red-black SOR: include/physics.hpp:100-157.  Red = (i+j+k) even, interior
 * only, in place; then black.  omega, tolerance and the iteration cap are
 * parameters (the reference hard-codes 1.5, 1e-4 and 512).  Stop when the
 * largest |update| of an iteration is < tol.  Returns 1 when the cap was hit
 * with max_diff > tol (the reference prints a warning, :155-156). */
static double sor_half(double *phi, const double *rho, const hvo_world *w, int colour, double omega) {
    int nx = w->n[0], ny = w->n[1], nz = w->n[2];
    double dx = 2.0 * w->d_max[0] / nx, dy = 2.0 * w->d_max[1] / ny, dz = 2.0 * w->d_max[2] / nz;
    double inv_dx2 = 1.0 / (2.0 / (dx * dx) + 2.0 / (dy * dy) + 2.0 / (dz * dz));
    double max_diff = 0.0;
#pragma omp parallel for schedule(static) reduction(max : max_diff)
    for (int i = 1; i < nx - 1; i++) {
        for (int j = 1; j < ny - 1; j++) {
            int k0 = 1 + ((i + j + 1 + colour) & 1);
            for (int k = k0; k < nz - 1; k += 2) {
                long long idx = IDX3(i, j, k, ny, nz);
                double r = rho[idx];
                double phi_x = (phi[IDX3(i + 1, j, k, ny, nz)] + phi[IDX3(i - 1, j, k, ny, nz)]) / (dx * dx);
                double phi_y = (phi[IDX3(i, j + 1, k, ny, nz)] + phi[IDX3(i, j - 1, k, ny, nz)]) / (dy * dy);
                double phi_z = (phi[IDX3(i, j, k + 1, ny, nz)] + phi[IDX3(i, j, k - 1, ny, nz)]) / (dz * dz);
                double phi_star = (phi_x + phi_y + phi_z + 4.0 * HV_PI * HV_E2 * r) * inv_dx2;
                double phi_old = phi[idx];
                phi[idx] = (1.0 - omega) * phi_old + omega * phi_star;
                double diff = fabs(phi[idx] - phi_old);
                if (diff > max_diff) max_diff = diff;
            }
        }
    }
    return max_diff;
}

int hvo_sor(double *phi, const double *rho_p, const hvo_world *w, double omega, double tol,
            int max_it, int *iters, double *max_diff_out) {
    double max_diff = 0.0;
    int it;
    for (it = 0; it < max_it; it++) {
        double m0 = sor_half(phi, rho_p, w, 0, omega);
        double m1 = sor_half(phi, rho_p, w, 1, omega);
        max_diff = m0 > m1 ? m0 : m1;
        if (max_diff < tol) {
            it++;
            break;
        }
    }
    if (iters) *iters = it;
    if (max_diff_out) *max_diff_out = max_diff;
    return (max_diff > tol) ? 1 : 0;
}

/* U_p += phi: include/tools.hpp:378-384 */
void hvo_merge(double *U, const double *phi, const hvo_world *w) {
    long long G = world_size(w);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < G; i++) U[i] += phi[i];
}

/* F = -grad U per species: include/physics.hpp:160-200.  Central differences
 * inside, one-sided first-order differences on the faces (:172-191). */
void hvo_fdm(const double *U, double *Fx, double *Fy, double *Fz, const hvo_world *w) {
    int nx = w->n[0], ny = w->n[1], nz = w->n[2];
    long long G = world_size(w);
    double dx = 2.0 * w->d_max[0] / nx, dy = 2.0 * w->d_max[1] / ny, dz = 2.0 * w->d_max[2] / nz;
    for (int s = 0; s < 2; s++) {
        const double *u = U + s * G;
        long long off = s * G;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                for (int k = 0; k < nz; k++) {
                    long long idx = IDX3(i, j, k, ny, nz);
                    double gx, gy, gz;
                    if (i == 0) gx = (u[IDX3(1, j, k, ny, nz)] - u[idx]) / dx;
                    else if (i == nx - 1) gx = (u[idx] - u[IDX3(nx - 2, j, k, ny, nz)]) / dx;
                    else gx = (u[IDX3(i + 1, j, k, ny, nz)] - u[IDX3(i - 1, j, k, ny, nz)]) / (2.0 * dx);
                    if (j == 0) gy = (u[IDX3(i, 1, k, ny, nz)] - u[idx]) / dy;
                    else if (j == ny - 1) gy = (u[idx] - u[IDX3(i, ny - 2, k, ny, nz)]) / dy;
                    else gy = (u[IDX3(i, j + 1, k, ny, nz)] - u[IDX3(i, j - 1, k, ny, nz)]) / (2.0 * dy);
                    if (k == 0) gz = (u[IDX3(i, j, 1, ny, nz)] - u[idx]) / dz;
                    else if (k == nz - 1) gz = (u[idx] - u[IDX3(i, j, nz - 2, ny, nz)]) / dz;
                    else gz = (u[IDX3(i, j, k + 1, ny, nz)] - u[IDX3(i, j, k - 1, ny, nz)]) / (2.0 * dz);
                    Fx[off + idx] = -gx;
                    Fy[off + idx] = -gy;
                    Fz[off + idx] = -gz;
                }
    }
}

/* CIC force gather: include/tools.hpp:99-168.  Same cell/weight rules as the
 * deposit; particles outside the grid get F = 0 (:112-115). */
void hvo_gather(const double *x, const double *y, const double *z, int np, int nn,
                const double *Fx, const double *Fy, const double *Fz, const hvo_world *w,
                double *fx, double *fy, double *fz) {
    int nx = w->n[0], ny = w->n[1], nz = w->n[2];
    long long G = world_size(w);
    int total = np + nn;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < total; i++) {
        double cx = (nx / 2.0) * (x[i] / w->d_max[0] + 1.0);
        double cy = (ny / 2.0) * (y[i] / w->d_max[1] + 1.0);
        double cz = (nz / 2.0) * (z[i] / w->d_max[2] + 1.0);
        if (!(cx >= 0.0 && cy >= 0.0 && cz >= 0.0 && cx < nx && cy < ny && cz < nz)) {
            fx[i] = 0.0;
            fy[i] = 0.0;
            fz[i] = 0.0;
            continue;
        }
        int x0 = (int)cx, y0 = (int)cy, z0 = (int)cz;
        int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
        double d_x = cx - x0, d_y = cy - y0, d_z = cz - z0;
        double t_x = 1.0 - d_x, t_y = 1.0 - d_y, t_z = 1.0 - d_z;
        long long off = (i < np) ? 0 : G;
        double ax = 0.0, ay = 0.0, az = 0.0, wgt;
        long long idx = off + IDX3(x0, y0, z0, ny, nz);
        wgt = t_x * t_y * t_z;
        ax += wgt * Fx[idx]; ay += wgt * Fy[idx]; az += wgt * Fz[idx];
        if (x1 < nx) {
            idx = off + IDX3(x1, y0, z0, ny, nz); wgt = d_x * t_y * t_z;
            ax += wgt * Fx[idx]; ay += wgt * Fy[idx]; az += wgt * Fz[idx];
        }
        if (y1 < ny) {
            idx = off + IDX3(x0, y1, z0, ny, nz); wgt = t_x * d_y * t_z;
            ax += wgt * Fx[idx]; ay += wgt * Fy[idx]; az += wgt * Fz[idx];
        }
        if (x1 < nx && y1 < ny) {
            idx = off + IDX3(x1, y1, z0, ny, nz); wgt = d_x * d_y * t_z;
            ax += wgt * Fx[idx]; ay += wgt * Fy[idx]; az += wgt * Fz[idx];
        }
        if (z1 < nz) {
            idx = off + IDX3(x0, y0, z1, ny, nz); wgt = t_x * t_y * d_z;
            ax += wgt * Fx[idx]; ay += wgt * Fy[idx]; az += wgt * Fz[idx];
        }
        if (x1 < nx && z1 < nz) {
            idx = off + IDX3(x1, y0, z1, ny, nz); wgt = d_x * t_y * d_z;
            ax += wgt * Fx[idx]; ay += wgt * Fy[idx]; az += wgt * Fz[idx];
        }
        if (y1 < ny && z1 < nz) {
            idx = off + IDX3(x0, y1, z1, ny, nz); wgt = t_x * d_y * d_z;
            ax += wgt * Fx[idx]; ay += wgt * Fy[idx]; az += wgt * Fz[idx];
        }
        if (x1 < nx && y1 < ny && z1 < nz) {
            idx = off + IDX3(x1, y1, z1, ny, nz); wgt = d_x * d_y * d_z;
            ax += wgt * Fx[idx]; ay += wgt * Fy[idx]; az += wgt * Fz[idx];
        }
        fx[i] = ax;
        fy[i] = ay;
        fz[i] = az;
    }
}

/* k += F dt / (2 hbar c): include/physics.hpp:249-258 */
void hvo_kick(double *kx, double *ky, double *kz, const double *fx, const double *fy,
              const double *fz, int total, double dt) {
    double fact = dt / (2.0 * HV_HBARC);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < total; i++) {
        kx[i] += fact * fx[i];
        ky[i] += fact * fy[i];
        kz[i] += fact * fz[i];
    }
}

/* x += k dt hbar c / mc^2: include/physics.hpp:261-270 */
void hvo_drift(double *x, double *y, double *z, const double *kx, const double *ky,
               const double *kz, int total, double dt) {
    double fact = (dt * HV_HBARC) / HV_MC2;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < total; i++) {
        x[i] += fact * kx[i];
        y[i] += fact * ky[i];
        z[i] += fact * kz[i];
    }
}

/* Dipole boost: include/physics.hpp:237-246 (ref_weighted) or the symmetric
 * +eta / -eta form. */
void hvo_excite(double *kz, int np, int nn, double eta, int z, int n, int mode) {
    double zz = z, nnuc = n;
    double dp, dn;
    if (mode == HVO_BOOST_SYMMETRIC) {
        dp = eta;
        dn = eta;
    } else {
        dp = eta * nnuc / (zz + nnuc);
        dn = eta * zz / (zz + nnuc);
    }
    for (int i = 0; i < np; i++) kz[i] += dp;
    for (int i = np; i < np + nn; i++) kz[i] -= dn;
}

/* ------------------------------------------------------------------------- */
/* Diagnostics sums as exact fixed-point integers (hv_stat_fix, hv_detmath.h):
 * order free, so the GPU's block/atomic structure cannot change the result.
 * r^2 = x*x + y*y + z*z as dot() (include/math_functions.hpp:42-45); the box
 * test of mean_squared_radius (include/tools.hpp:315-318) includes the faces. */
void hvo_stats_sums128(const double *x, const double *y, const double *z, int np, int nn,
                       const hvo_world *w, __int128 out[4], long long cnt[2]) {
    int total = np + nn;
    __int128 s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    long long cp = 0, cn = 0;
    for (int i = 0; i < total; i++) {
        int inbox = !(x[i] < -w->d_max[0] || x[i] > w->d_max[0] || y[i] < -w->d_max[1] ||
                      y[i] > w->d_max[1] || z[i] < -w->d_max[2] || z[i] > w->d_max[2]);
        if (x[i] != x[i] || y[i] != y[i] || z[i] != z[i]) inbox = 0;
        double r2 = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
        long long qz = hv_stat_fix(z[i]);
        long long qr = inbox ? hv_stat_fix(r2) : 0;
        if (i < np) {
            s0 += qz;
            s2 += qr;
            cp += inbox;
        } else {
            s1 += qz;
            s3 += qr;
            cn += inbox;
        }
    }
    out[0] = s0; out[1] = s1; out[2] = s2; out[3] = s3;
    cnt[0] = cp;
    cnt[1] = cn;
}

void hvo_stats_sums(const double *x, const double *y, const double *z, int np, int nn,
                    const hvo_world *w, double out[4], long long cnt[2]) {
    __int128 s[4];
    hvo_stats_sums128(x, y, z, np, nn, w, s, cnt);
    for (int c = 0; c < 4; c++) out[c] = (double)s[c] * HV_STAT_INV;
}

/* the four stats of the stats.txt line (src/hyper_vlasov.cpp:65-70) */
void hvo_stats(const double *x, const double *y, const double *z, int np, int nn,
               const hvo_world *w, double st[4]) {
    __int128 s[4];
    long long c[2];
    hvo_stats_sums128(x, y, z, np, nn, w, s, c);
    st[0] = hv_stat_mean(s[0], np);
    st[1] = hv_stat_mean(s[1], nn);
    st[2] = hv_stat_mean(s[2], c[0]);
    st[3] = hv_stat_mean(s[3], c[1]);
}

/* centre of mass along z per species over all particles of the species:
 * include/tools.hpp:328-338 */
double hvo_center_of_mass(const double *z, int np, int nn, int type) {
    /* x, y only matter for the radius; pass z for both */
    hvo_world w = {{1, 1, 1}, {1e300, 1e300, 1e300}};
    double st[4];
    hvo_stats(z, z, z, np, nn, &w, st);
    return type == 0 ? st[0] : st[1];
}

/* <r^2> per species over in-box particles: include/tools.hpp:303-325 */
double hvo_mean_squared_radius(const double *x, const double *y, const double *z, int np, int nn,
                               const hvo_world *w, int type) {
    double st[4];
    hvo_stats(x, y, z, np, nn, w, st);
    return type == 0 ? st[2] : st[3];
}

/* ------------------------------------------------------------------------- */
/* The step loop: src/hyper_vlasov.cpp:33-118 (file output left to the caller) */
static void field_solve(const double *x, const double *y, const double *z, double *fx, double *fy,
                        double *fz, int np, int nn, const hvo_world *w, const hvo_skyrme *s,
                        const hvo_params *p, int64_t *acc, double *counts, double *rho, double *U,
                        double *phi, double *Fx, double *Fy, double *Fz, int *iters, int *warn) {
    double md;
    hvo_deposit(x, y, z, np, nn, w, acc, counts);                       /* :93 */
    hvo_smear(counts, rho, w, p->sigma_r, p->part_per_nucleon);          /* :94 */
    hvo_skyrme_potentials(rho, U, w, s);                                          /* :96 */
    *warn += hvo_sor(phi, rho, w, p->omega, p->sor_tol, p->sor_max_it, iters, &md); /* :97 */
    hvo_merge(U, phi, w);                                                /* :98 */
    hvo_fdm(U, Fx, Fy, Fz, w);                                           /* :99 */
    hvo_gather(x, y, z, np, nn, Fx, Fy, Fz, w, fx, fy, fz);              /* :101 */
}

static int simulate_impl(double *x, double *y, double *z, double *kx, double *ky, double *kz,
                         double *fx, double *fy, double *fz, int np, int nn, const hvo_world *w,
                         const hvo_skyrme *s, const hvo_params *p, double *phi_inout, double *stats,
                         int *sor_iters, int preloop) {
    long long G = world_size(w);
    int total = np + nn, warn = 0, excited = 0, it = 0;
    int64_t *acc = (int64_t *)malloc(sizeof(int64_t) * 2 * G);
    double *counts = (double *)malloc(sizeof(double) * 2 * G);
    double *rho = (double *)malloc(sizeof(double) * 2 * G);
    double *U = (double *)malloc(sizeof(double) * 2 * G);
    double *F = (double *)malloc(sizeof(double) * 6 * G);
    double *phi = (double *)calloc((size_t)G, sizeof(double));
    if (!acc || !counts || !rho || !U || !F || !phi) abort();
    if (phi_inout) memcpy(phi, phi_inout, sizeof(double) * G);
    else hvo_coulomb_bc(phi, w, p->z);                                     /* :56 */
    /* pre-loop field solve and gather (:53-62) */
    if (preloop)
        field_solve(x, y, z, fx, fy, fz, np, nn, w, s, p, acc, counts, rho, U, phi, F, F + 2 * G,
                    F + 4 * G, &it, &warn);
    if (sor_iters) sor_iters[0] = it;
    for (int step = 0; step < p->steps; step++) {
        if (stats) hvo_stats(x, y, z, np, nn, w, stats + 4 * step);     /* :65-68 */
        if (step * p->dt >= p->t_exc && !excited) {                        /* :86-89 */
            excited = 1;
            hvo_excite(kz, np, nn, p->eta, p->z, p->n, p->boost_mode);
        }
        hvo_kick(kx, ky, kz, fx, fy, fz, total, p->dt);                    /* :90 */
        hvo_drift(x, y, z, kx, ky, kz, total, p->dt);                      /* :91 */
        field_solve(x, y, z, fx, fy, fz, np, nn, w, s, p, acc, counts, rho, U, phi, F, F + 2 * G,
                    F + 4 * G, &it, &warn);                                /* :93-101 */
        if (sor_iters) sor_iters[step + 1] = it;
        hvo_kick(kx, ky, kz, fx, fy, fz, total, p->dt);                    /* :103 */
    }
    if (phi_inout) memcpy(phi_inout, phi, sizeof(double) * G);
    free(acc);
    free(counts);
    free(rho);
    free(U);
    free(F);
    free(phi);
    return warn;
}

int hvo_simulate(double *x, double *y, double *z, double *kx, double *ky, double *kz,
                 double *fx, double *fy, double *fz, int np, int nn, const hvo_world *w,
                 const hvo_skyrme *s, const hvo_params *p, double *phi_inout, double *stats,
                 int *sor_iters) {
    return simulate_impl(x, y, z, kx, ky, kz, fx, fy, fz, np, nn, w, s, p, phi_inout, stats, sor_iters, 1);
}

/* The loop only (src/hyper_vlasov.cpp:63-104): the caller supplies the forces
 * of the current positions and the Coulomb state (phi_inout, required), e.g.
 * a state taken mid-run. */
int hvo_simulate_loop(double *x, double *y, double *z, double *kx, double *ky, double *kz,
                      double *fx, double *fy, double *fz, int np, int nn, const hvo_world *w,
                      const hvo_skyrme *s, const hvo_params *p, double *phi_inout, double *stats,
                      int *sor_iters) {
    if (!phi_inout) return -1;
    return simulate_impl(x, y, z, kx, ky, kz, fx, fy, fz, np, nn, w, s, p, phi_inout, stats, sor_iters, 0);
}

/* Counts operands a (drawn uniformly in [lo, hi) from a xorshift stream) for
 * which hv_div_fixed(a, b, 1/b) differs from the IEEE quotient a / b. */
long long hvo_div_fixed_mismatches(double b, double lo, double hi, long long n, unsigned long long seed) {
    double y = 1.0 / b;
    long long bad = 0;
    unsigned long long s = seed | 1ULL;
    for (long long i = 0; i < n; i++) {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        double u = (double)(s >> 11) * (1.0 / 9007199254740992.0);
        double a = lo + (hi - lo) * u;
        if (hv_div_fixed(a, b, y) != a / b) bad++;
    }
    return bad;
}

/* ---- total energy (no reference counterpart) ------------------------------ */
void hvo_total_energy(const double *kx, const double *ky, const double *kz, int total, const double *rho,
                      const double *phi, const hvo_world *w, const hvo_skyrme *s, int ppn, double out[4]) {
    __int128 sk = 0, ss = 0, sc = 0;
    for (int i = 0; i < total; i++) {
        double k2 = kx[i] * kx[i] + ky[i] * ky[i] + kz[i] * kz[i];
        sk += hv_stat_fix(k2);
    }
    const long long G = (long long)w->n[0] * w->n[1] * w->n[2];
    const double dV = hv_spacing(w->d_max[0], w->n[0]) * hv_spacing(w->d_max[1], w->n[1]) *
                      hv_spacing(w->d_max[2], w->n[2]);
    for (long long q = 0; q < G; q++) {
        double es, ec;
        hv_energy_density(s->A, s->B, s->C, s->gamma, rho[q], rho[q + G], phi[q], &es, &ec);
        ss += hv_stat_fix(es * dV);
        sc += hv_stat_fix(ec * dV);
    }
    const double ekin = HV_KIN_FAC * ((double)sk * HV_STAT_INV) / (double)ppn;
    const double esky = (double)ss * HV_STAT_INV, ecoul = (double)sc * HV_STAT_INV;
    out[1] = ekin;
    out[2] = esky;
    out[3] = ecoul;
    out[0] = ekin + esky + ecoul;
}
