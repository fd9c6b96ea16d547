// hv_init_ws_impl.h -- the reference's Woods-Saxon / Fermi-level initial
// state (initialize_particles, include/physics.hpp:35-97), off the hot path:
// host sampling and Fermi-level updates in the reference's order with the C
// library's rand(), the particles' smeared densities (O(N^2)) on the GPU, and
// a Levenberg-Marquardt fit of the Woods-Saxon wells in place of GSL's trust
// region solver (include/fit_algorithm.hpp:96-138).  Included once, at the end
// of hv_abi.hip.

#include <cstdint>
#include <cstdlib>
#include <cstring>

#include "hv_init.h"

namespace {

constexpr double kWsKMax = 1.5;              // k_max (include/sim_structs.hpp:41)
constexpr double kWsRelax = 0.6;             // relax_coef (include/physics.hpp:36)
constexpr double kWsDeltaTol = 0.1;          // delta_epsilon_tolerance (include/sim_structs.hpp:46)
constexpr int kWsMaxIterations = 32;         // MAX_INIT_ITERATIONS (include/sim_structs.hpp:32)

// The generator of a whole initialisation (hv_init_woods_saxon): the C
// library's rand() sequence from a private state (glibc's random_r on a
// 128-byte state, the generator and state size behind rand(), seeded as
// srand(seed) seeds it), so that rand() calls from other threads of the
// process (the GPU runtime's, which starts during the first round) cannot
// shift it.  Unset, the sampling takes the process's rand() itself.
thread_local random_data *tl_ws_rng = nullptr;

struct WsRngScope {
    random_data d;
    char state[128];
    explicit WsRngScope(unsigned seed) {
        std::memset(&d, 0, sizeof d);
        std::memset(state, 0, sizeof state);
        initstate_r(seed, state, sizeof state, &d);
        tl_ws_rng = &d;
    }
    ~WsRngScope() { tl_ws_rng = nullptr; }
    WsRngScope(const WsRngScope &) = delete;
    WsRngScope &operator=(const WsRngScope &) = delete;
};

// rand_val / random_vec (include/math_functions.hpp:30-39)
inline double ws_rand_val(double lo, double hi) {
    int r;
    if (tl_ws_rng) {
        int32_t v = 0;
        random_r(tl_ws_rng, &v);
        r = v;
    } else {
        r = std::rand();
    }
    const double s = r / double(RAND_MAX);
    return lo + s * (hi - lo);
}

inline void ws_random_vec(double v[3], double m) {
    for (int i = 0; i < 3; i++) v[i] = ws_rand_val(-m, m);
}

inline double ws_dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

inline double ws_magnitude(const double *a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// kinetic_energy, fluctuation_energy, woods_saxon_potential, coulomb_potential
// (include/physics_formulas.hpp:45-56, 65-69, 80-85)
inline double ws_kinetic() { return (HV_HBARC * HV_HBARC) / (2.0 * HV_MC2); }

inline double ws_potential(const hv_woods_saxon &w, double r) { return w.V0 / (1.0 + std::exp((r - w.R12) / w.a)); }

inline double ws_coulomb(const hv_woods_saxon &w, double z, double r) {
    if (r <= w.R12) return 1.44 * (z - 1.0) / w.R12 * (1.5 - 0.5 * (r / w.R12) * (r / w.R12));
    return 1.44 * (z - 1.0) / r;
}

// compute_energy (include/tools.hpp:206-224): the well of the particle's
// species, its kinetic energy, the Coulomb term for protons and the
// fluctuation energy of sigma_k
inline double ws_energy(const double r_vec[3], const double k_vec[3], bool proton, const hv_woods_saxon ws[2],
                        double z, double sigma_k) {
    const double r = ws_magnitude(r_vec);
    const double k = ws_magnitude(k_vec);
    const hv_woods_saxon &w = proton ? ws[0] : ws[1];
    double e = 0.0;
    e += ws_potential(w, r);
    e += (k * k) * ws_kinetic();
    if (proton) e += ws_coulomb(w, z, r);
    e += 3.0 * ws_kinetic() * sigma_k * sigma_k;
    return e;
}

// skyrme_potential (include/physics_formulas.hpp:71-77) with libm pow
inline double ws_skyrme(const hv_skyrme &s, double rp, double rn, bool proton) {
    const double tau = proton ? -1.0 : 1.0;
    const double t = (rp + rn) / HV_RHO0;
    return s.A * t + s.B * std::pow(t, s.gamma) + 2.0 * tau * s.C * ((rn - rp) / HV_RHO0);
}

// compute_particle_densities (include/tools.hpp:171-203): for each particle
// the Gaussian-weighted count of all particles within 4 sigma_r, summed over
// j in index order (as the reference's inner loop), per species, then the
// normalisation.  One thread per particle i; the positions stream through LDS
// in tiles of the block size.
constexpr int kWsThreads = 256;

__global__ void __launch_bounds__(kWsThreads)
k_ws_particle_densities(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
                        int np, int total, double exp_term, double cut2, double norm, double *__restrict__ rho_p,
                        double *__restrict__ rho_n) {
    __shared__ double tx[kWsThreads], ty[kWsThreads], tz[kWsThreads];
    const int i = blockIdx.x * kWsThreads + threadIdx.x;
    const double xi = i < total ? x[i] : 0.0, yi = i < total ? y[i] : 0.0, zi = i < total ? z[i] : 0.0;
    double dp = 0.0, dn = 0.0;
    for (int base = 0; base < total; base += kWsThreads) {
        const int j0 = base + threadIdx.x;
        __syncthreads();
        tx[threadIdx.x] = j0 < total ? x[j0] : 0.0;
        ty[threadIdx.x] = j0 < total ? y[j0] : 0.0;
        tz[threadIdx.x] = j0 < total ? z[j0] : 0.0;
        __syncthreads();
        const int cnt = total - base < kWsThreads ? total - base : kWsThreads;
        for (int jj = 0; jj < cnt; jj++) {
            const double d0 = xi - tx[jj], d1 = yi - ty[jj], d2 = zi - tz[jj];
            const double dist2 = d0 * d0 + d1 * d1 + d2 * d2;
            if (dist2 > cut2) continue;
            const double f = exp(-dist2 * exp_term);
            if (base + jj < np)
                dp += f;
            else
                dn += f;
        }
    }
    if (i < total) {
        rho_p[i] = dp * norm;
        rho_n[i] = dn * norm;
    }
}

// Levenberg-Marquardt fit of one species' well: minimise
// sum_i (v_sky_i - V0 / (1 + exp((r_i - R12) / a)))^2 over (V0, R12, a)
// (woods_saxon_f / woods_saxon_df, include/fit_algorithm.hpp:50-93), with
// More's diagonal scaling and Nielsen's damping update; stops when a step
// changes every parameter by less than 1e-10 relative, or after 200 steps.
struct WsFitData {
    std::vector<double> r, v;
};

double ws_cost(const WsFitData &d, const double p[3]) {
    double c = 0.0;
    for (size_t i = 0; i < d.r.size(); i++) {
        const double f = d.v[i] - p[0] / (1.0 + std::exp((d.r[i] - p[1]) / p[2]));
        c += f * f;
    }
    return 0.5 * c;
}

// solves the 3x3 system M x = b (Gaussian elimination, partial pivoting);
// false when singular
bool ws_solve3(double M[3][3], double b[3], double x[3]) {
    int idx[3] = {0, 1, 2};
    double A[3][4];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) A[i][j] = M[i][j];
        A[i][3] = b[i];
    }
    for (int c = 0; c < 3; c++) {
        int piv = c;
        for (int r = c + 1; r < 3; r++)
            if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
        if (!(std::fabs(A[piv][c]) > 0.0)) return false;
        if (piv != c)
            for (int j = 0; j < 4; j++) std::swap(A[c][j], A[piv][j]);
        for (int r = c + 1; r < 3; r++) {
            const double f = A[r][c] / A[c][c];
            for (int j = c; j < 4; j++) A[r][j] -= f * A[c][j];
        }
    }
    for (int c = 2; c >= 0; c--) {
        double s = A[c][3];
        for (int j = c + 1; j < 3; j++) s -= A[c][j] * x[j];
        x[c] = s / A[c][c];
    }
    (void)idx;
    return true;
}

int ws_fit_species(const WsFitData &d, hv_woods_saxon *w) {
    double p[3] = {w->V0, w->R12, w->a};
    if (d.r.empty() || !(p[2] > 0.0)) return 0;
    double cost = ws_cost(d, p);
    double mu = -1.0, nu = 2.0, D[3] = {0.0, 0.0, 0.0};
    int it = 0;
    for (; it < 200; it++) {
        double JJ[3][3] = {}, g[3] = {};
        for (size_t i = 0; i < d.r.size(); i++) {
            const double e = std::exp((d.r[i] - p[1]) / p[2]);
            const double den = 1.0 + e, den2 = den * den;
            const double f = d.v[i] - p[0] / den;
            const double J[3] = {-1.0 / den, -(p[0] * e) / (p[2] * den2), -(p[0] * (d.r[i] - p[1]) * e) / (p[2] * p[2] * den2)};
            for (int a = 0; a < 3; a++) {
                g[a] += J[a] * f;
                for (int b = 0; b < 3; b++) JJ[a][b] += J[a] * J[b];
            }
        }
        double gmax = 0.0;
        for (int a = 0; a < 3; a++) {
            D[a] = std::max(D[a], std::sqrt(JJ[a][a]));
            gmax = std::max(gmax, std::fabs(g[a]));
        }
        if (mu < 0.0) mu = 1e-3 * std::max(std::max(JJ[0][0], JJ[1][1]), JJ[2][2]);
        if (!(gmax > 0.0)) break;
        bool moved = false, done = false;
        for (int tries = 0; tries < 60 && !moved; tries++) {
            double M[3][3], rhs[3], dx[3];
            for (int a = 0; a < 3; a++) {
                for (int b = 0; b < 3; b++) M[a][b] = JJ[a][b];
                M[a][a] += mu * D[a] * D[a];
                rhs[a] = -g[a];
            }
            if (!ws_solve3(M, rhs, dx)) {
                mu *= nu;
                nu *= 2.0;
                continue;
            }
            const double q[3] = {p[0] + dx[0], p[1] + dx[1], p[2] + dx[2]};
            // predicted decrease of the Gauss-Newton model
            const double pred = 0.5 * (mu * (D[0] * D[0] * dx[0] * dx[0] + D[1] * D[1] * dx[1] * dx[1] +
                                             D[2] * D[2] * dx[2] * dx[2]) -
                                       (g[0] * dx[0] + g[1] * dx[1] + g[2] * dx[2]));
            const double nc = q[2] > 0.0 ? ws_cost(d, q) : HUGE_VAL;
            if (nc < cost && pred > 0.0) {
                const double rho = (cost - nc) / pred;
                const double t = 2.0 * rho - 1.0;
                mu *= std::max(1.0 / 3.0, 1.0 - t * t * t);
                nu = 2.0;
                done = true;
                for (int a = 0; a < 3; a++)
                    if (std::fabs(dx[a]) > 1e-10 * (std::fabs(p[a]) + 1e-10)) done = false;
                for (int a = 0; a < 3; a++) p[a] = q[a];
                cost = nc;
                moved = true;
            } else {
                mu *= nu;
                nu *= 2.0;
            }
        }
        if (!moved || done) break;
    }
    w->V0 = p[0];
    w->R12 = p[1];
    w->a = p[2];
    return it;
}

int ws_check(int z, int n, int ppn) {
    if (z < 1 || n < 1 || ppn < 1) return HV_ERR_ARG;
    if ((long long)z * ppn > HV_MAX_SPECIES_PARTICLES || (long long)n * ppn > HV_MAX_SPECIES_PARTICLES)
        return HV_ERR_ARG;
    return HV_OK;
}

}  // namespace

extern "C" {

double hv_ws_nuclear_radius(int mass_number) { return 1.5 * std::pow((double)mass_number, 1.0 / 3.0); }

int hv_ws_max_particles(double r_max, double k_max, int part_per_nucleon) {
    const double t = r_max * k_max, ct = 2.0 * HV_PI;
    const double phase_space_volume = 16.0 / 9.0 * HV_PI * HV_PI * (t * t * t);
    return part_per_nucleon * (int)std::floor(phase_space_volume / (ct * ct * ct) + 0.5);
}

int hv_ws_sample_round(int z, int n, int part_per_nucleon, double sigma_k, const hv_woods_saxon ws[2],
                       double fermi[2], double delta[2], int counts[6], double *x, double *y, double *zz,
                       double *kx, double *ky, double *kz) {
    if (!ws || !fermi || !x || !y || !zz || !kx || !ky || !kz) return HV_ERR_ARG;
    int rc = ws_check(z, n, part_per_nucleon);
    if (rc) return rc;
    const double r_max = hv_ws_nuclear_radius(z + n);
    const int max_part = hv_ws_max_particles(r_max, kWsKMax, part_per_nucleon);
    const int total_p = z * part_per_nucleon, total_n = n * part_per_nucleon;
    // generate_random_particles (include/tools.hpp:235-254) on max_part
    // particles of each species, then compute_particle_energies and the
    // counts of include/physics.hpp:46-62
    std::vector<double> tr(6 * (size_t)max_part), tk(6 * (size_t)max_part);
    const int tot = 2 * max_part;
    for (int i = 0; i < tot;) {
        double v[3];
        ws_random_vec(v, r_max);
        if (ws_dot(v, v) < r_max * r_max) {
            for (int a = 0; a < 3; a++) tr[3 * (size_t)i + a] = v[a];
            i++;
        }
    }
    for (int i = 0; i < tot;) {
        double v[3];
        ws_random_vec(v, kWsKMax);
        if (ws_dot(v, v) < kWsKMax * kWsKMax) {
            for (int a = 0; a < 3; a++) tk[3 * (size_t)i + a] = v[a];
            i++;
        }
    }
    int less_p = 0, equal_p = 0, more_p = 0, less_n = 0, equal_n = 0, more_n = 0;
    for (int i = 0; i < max_part; i++) {
        const double ep = ws_energy(&tr[3 * (size_t)i], &tk[3 * (size_t)i], true, ws, z, sigma_k);
        const double en = ws_energy(&tr[3 * (size_t)(i + max_part)], &tk[3 * (size_t)(i + max_part)], false, ws, z,
                                    sigma_k);
        if (ep < fermi[0]) equal_p += 2;
        if (ep < fermi[0] + 0.5) more_p += 2;
        if (ep < fermi[0] - 0.5) less_p += 2;
        if (en < fermi[1]) equal_n += 2;
        if (en < fermi[1] + 0.5) more_n += 2;
        if (en < fermi[1] - 0.5) less_n += 2;
    }
    const double delta_part_n = total_n - equal_n, delta_part_p = total_p - equal_p;
    // the reference divides by (more - less) as is; an empty +-0.5 MeV shell
    // (no particle near the level) takes the largest relaxed step instead of
    // an infinite one
    double dn = more_n != less_n ? 0.5 * delta_part_n / (more_n - less_n) : (delta_part_n > 0 ? 1.0 : -1.0);
    double dp = more_p != less_p ? 0.5 * delta_part_p / (more_p - less_p) : (delta_part_p > 0 ? 1.0 : -1.0);
    if (std::fabs(dp) > 0.5) dp *= kWsRelax;
    if (std::fabs(dn) > 0.5) dn *= kWsRelax;
    fermi[0] += dp;
    fermi[1] += dn;
    if (delta) {
        delta[0] = dp;
        delta[1] = dn;
    }
    if (counts) {
        const int c[6] = {equal_p, more_p, less_p, equal_n, more_n, less_n};
        for (int q = 0; q < 6; q++) counts[q] = c[q];
    }
    // generate_checking_particles (include/tools.hpp:257-281): positions and
    // momenta from the cubes, kept below the species' Fermi level, each kept
    // one followed by its mirror image (-r, -k)
    const int total = total_p + total_n;
    double *P[3] = {x, y, zz}, *K[3] = {kx, ky, kz};
    for (int i = 0; i < total;) {
        const bool proton = i < total_p;
        const double eps = proton ? fermi[0] : fermi[1];
        double r[3], k[3];
        ws_random_vec(r, r_max);
        ws_random_vec(k, kWsKMax);
        for (int a = 0; a < 3; a++) {
            P[a][i] = r[a];
            K[a][i] = k[a];
        }
        if (ws_energy(r, k, proton, ws, z, sigma_k) < eps) {
            if (i + 1 < total)
                for (int a = 0; a < 3; a++) {
                    P[a][i + 1] = r[a] * -1.0;
                    K[a][i + 1] = k[a] * -1.0;
                }
            i += 2;
        }
    }
    return HV_OK;
}

int hv_ws_particle_densities(const double *x, const double *y, const double *z, int n_protons, int n_neutrons,
                             double sigma_r, int part_per_nucleon, double *rho_p, double *rho_n) {
    if (!x || !y || !z || !rho_p || !rho_n || n_protons < 0 || n_neutrons < 0 || !(sigma_r > 0.0) ||
        part_per_nucleon < 1)
        return HV_ERR_ARG;
    const int total = n_protons + n_neutrons;
    if (total == 0) return HV_OK;
    double *d = nullptr;
    const size_t N = (size_t)total;
    if (hipMalloc((void **)&d, sizeof(double) * 5 * N) != hipSuccess) return HV_ERR_ALLOC;
    int rc = HV_OK;
    const double *src[3] = {x, y, z};
    for (int a = 0; a < 3 && !rc; a++)
        if (hipMemcpy(d + a * N, src[a], sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess) rc = HV_ERR_HIP;
    if (!rc) {
        const double exp_term = 1.0 / (2.0 * sigma_r * sigma_r), cut2 = 16.0 * sigma_r * sigma_r;
        const double norm = (1.0 / part_per_nucleon) * (1.0 / std::pow(2.0 * HV_PI * sigma_r * sigma_r, 1.5));
        k_ws_particle_densities<<<(unsigned)((N + kWsThreads - 1) / kWsThreads), kWsThreads>>>(
            d, d + N, d + 2 * N, n_protons, total, exp_term, cut2, norm, d + 3 * N, d + 4 * N);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = HV_ERR_HIP;
    }
    if (!rc && (hipMemcpy(rho_p, d + 3 * N, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(rho_n, d + 4 * N, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess))
        rc = HV_ERR_HIP;
    (void)hipFree(d);
    return rc;
}

int hv_ws_fit(const double *x, const double *y, const double *z, const double *rho_p, const double *rho_n,
              int n_protons, int n_neutrons, const hv_skyrme *skyrme, hv_woods_saxon ws[2], int iterations[2]) {
    if (!x || !y || !z || !rho_p || !rho_n || !skyrme || !ws || n_protons < 1 || n_neutrons < 1) return HV_ERR_ARG;
    for (int type = 0; type < 2; type++) {
        const int start = type == 0 ? 0 : n_protons, cnt = type == 0 ? n_protons : n_neutrons;
        WsFitData d;
        d.r.resize(cnt);
        d.v.resize(cnt);
        for (int i = 0; i < cnt; i++) {
            const int idx = start + i;
            const double rv[3] = {x[idx], y[idx], z[idx]};
            d.r[i] = ws_magnitude(rv);
            d.v[i] = ws_skyrme(*skyrme, rho_p[idx], rho_n[idx], type == 0);
        }
        const int it = ws_fit_species(d, &ws[type]);
        if (iterations) iterations[type] = it;
    }
    return HV_OK;
}

int hv_ws_chi_squared(const double *x, const double *y, const double *z, const double *rho_p, const double *rho_n,
                      int n_protons, int n_neutrons, const hv_skyrme *skyrme, const hv_woods_saxon ws[2],
                      int part_per_nucleon, double out[2]) {
    if (!x || !y || !z || !rho_p || !rho_n || !skyrme || !ws || !out || part_per_nucleon < 1) return HV_ERR_ARG;
    double c[2] = {0.0, 0.0};
    const int total = n_protons + n_neutrons;
    for (int i = 0; i < total; i++) {
        const bool proton = i < n_protons;
        const double rv[3] = {x[i], y[i], z[i]};
        const double diff = ws_potential(ws[proton ? 0 : 1], ws_magnitude(rv)) - ws_skyrme(*skyrme, rho_p[i], rho_n[i], proton);
        c[proton ? 0 : 1] += diff * diff;
    }
    out[0] = c[0] / part_per_nucleon;
    out[1] = c[1] / part_per_nucleon;
    return HV_OK;
}

int hv_init_woods_saxon(int z, int n, int part_per_nucleon, const hv_skyrme *skyrme, double sigma_r, double sigma_k,
                        double V0, double a, double fermi[2], unsigned seed, double *x, double *y, double *zz,
                        double *kx, double *ky, double *kz, hv_woods_saxon ws[2], int *iterations, double chi2[2]) {
    if (!skyrme || !fermi || !ws || !x || !y || !zz || !kx || !ky || !kz || !(sigma_r > 0.0) || !(a > 0.0))
        return HV_ERR_ARG;
    int rc = ws_check(z, n, part_per_nucleon);
    if (rc) return rc;
    const double r_max = hv_ws_nuclear_radius(z + n);
    ws[0].V0 = ws[1].V0 = V0;  // set_woods_saxon (include/init.hpp:333-334)
    ws[0].R12 = ws[1].R12 = 0.8 * r_max;
    ws[0].a = ws[1].a = a;
    const int np = z * part_per_nucleon, nn = n * part_per_nucleon, total = np + nn;
    std::vector<double> rp((size_t)total), rn((size_t)total);
    const WsRngScope rng(seed);
    double total_delta = 0.0;
    int it = 0;
    do {
        double delta[2];
        rc = hv_ws_sample_round(z, n, part_per_nucleon, sigma_k, ws, fermi, delta, nullptr, x, y, zz, kx, ky, kz);
        if (rc) return rc;
        rc = hv_ws_particle_densities(x, y, zz, np, nn, sigma_r, part_per_nucleon, rp.data(), rn.data());
        if (rc) return rc;
        hv_woods_saxon old[2] = {ws[0], ws[1]};
        rc = hv_ws_fit(x, y, zz, rp.data(), rn.data(), np, nn, skyrme, ws, nullptr);
        if (rc) return rc;
        for (int q = 0; q < 2; q++) {  // relax_woods_saxon (include/tools.hpp:369-375)
            ws[q].V0 = kWsRelax * ws[q].V0 + (1.0 - kWsRelax) * old[q].V0;
            ws[q].R12 = kWsRelax * ws[q].R12 + (1.0 - kWsRelax) * old[q].R12;
            ws[q].a = kWsRelax * ws[q].a + (1.0 - kWsRelax) * old[q].a;
        }
        total_delta = std::fabs(delta[1]) + std::fabs(delta[0]);
        it++;
    } while (total_delta > kWsDeltaTol && it < kWsMaxIterations);
    if (iterations) *iterations = it;
    if (chi2) {
        rc = hv_ws_chi_squared(x, y, zz, rp.data(), rn.data(), np, nn, skyrme, ws, part_per_nucleon, chi2);
        if (rc) return rc;
    }
    return it == kWsMaxIterations && total_delta > kWsDeltaTol ? HV_WARN_INIT_NOT_CONVERGED : HV_OK;
}

}  // extern "C"
