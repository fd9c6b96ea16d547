// hv_vlasov -- command-line driver of libhvamd.so with the reference program's
// interface: the same arguments (src/hyper_vlasov.cpp:153-170), the same
// input file keys (read_input_file, include/init.hpp:276-337) and the same
// stats.txt lines (src/hyper_vlasov.cpp:64-70), so the reference's scripts
// (scripts/programs.py:42-78, 82-117) can run it unchanged.
//
//   hv_vlasov [--float] <input file> <output directory>
//   hv_vlasov --check <input file>       (parse, print the run as JSON, no GPU)
//
// --float selects the mixed FP32/FP64 form, otherwise FP64.  The initial
// nucleus is the reference's (initialize_particles, include/physics.hpp:35-97,
// through hv_init_woods_saxon): Woods-Saxon wells from V0 and a, Fermi levels
// from epsilon_p / epsilon_n, momentum width from k_fwhm, srand(seed) with the
// reference's seed 128, the wells fitted by Levenberg-Marquardt where the
// reference uses GSL; it prints the reference's "MAX TEST PART" and "CHI
// SQUARED" lines.  With the optional key `init 1` it is instead the seeded
// uniform-sphere sample of include/hv_init.h (radius r_init A^(1/3), momenta
// uniform in the Fermi ball of rho0), the benchmark configurations' nucleus.
// Every `substeps` steps it writes the stats line and, as the reference does,
// the snapshot output_dir/out-%04d.vtk (hv_write_vtk: forces, density,
// potentials).
//
// Optional keys beyond the reference's 21: sor_omega, sor_tol, sor_max_it
// (default: the reference's 1.5, 1e-4, 512), boost (0 = the reference's
// weighted split, 1 = protons +eta, neutrons -eta), init (0 = the reference's
// Woods-Saxon nucleus, 1 = uniform sphere), seed (128), r_init (1.2),
// sort_every, vtk (1; 0 writes no snapshots).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "hv_abi.h"
#include "hv_init.h"

namespace {

struct Run {
    double V0 = 0, a = 0, A = 0, B = 0, C = 0, gamma = 0, epsilon_p = 0, epsilon_n = 0, k_fwhm = 0, r_fwhm = 0;
    double t_f = 0, t_exc = 0, eta_exc = 0, d_max_scale = 0;
    int nx = 0, num_test_part = 0, steps = 0, n = 0, z = 0, substeps = 0, use_gpu = 0;
    double sor_omega = 1.5, sor_tol = 1e-4;
    int sor_max_it = 512, boost = HV_BOOST_REF_WEIGHTED, sort_every = 8, vtk = 1, init = 0;
    unsigned long long seed = 128;
    double r_init = 1.2;
};

const char *kRequired[] = {"V0",    "a",     "A",      "B",       "C",           "gamma",         "epsilon_p",
                           "epsilon_n", "k_fwhm", "r_fwhm", "t_f",   "t_exc",       "eta_exc",       "d_max_scale",
                           "nx",    "num_test_part", "steps", "n",  "z",           "substeps",      "use_gpu"};

// Keys and values separated by white space, in any order; unknown tokens are
// skipped (as the reference's reader does).  Each of the 21 keys must appear
// exactly once with a value that parses.
bool read_input(const char *path, Run &r, std::string &why) {
    std::ifstream in(path);
    if (!in) {
        why = "CANNOT OPEN INPUT FILE!";
        return false;
    }
    std::map<std::string, std::string> kv;
    std::map<std::string, int> seen;
    std::string key;
    while (in >> key) {
        bool known = false;
        for (const char *k : kRequired) known = known || key == k;
        const char *opt[] = {"sor_omega", "sor_tol", "sor_max_it", "boost", "seed", "r_init", "sort_every", "vtk",
                             "init"};
        for (const char *k : opt) known = known || key == k;
        if (!known) continue;
        std::string v;
        if (!(in >> v)) break;
        kv[key] = v;
        seen[key]++;
    }
    for (const char *k : kRequired) {
        if (seen[k] != 1) {
            why = std::string("key ") + k + (seen[k] ? " given more than once" : " missing");
            return false;
        }
    }
    auto d = [&](const char *k, double &out) {
        if (!kv.count(k)) return true;
        char *end = nullptr;
        out = std::strtod(kv[k].c_str(), &end);
        return end && *end == 0;
    };
    auto i = [&](const char *k, int &out) {
        if (!kv.count(k)) return true;
        char *end = nullptr;
        long v = std::strtol(kv[k].c_str(), &end, 10);
        out = (int)v;
        return end && *end == 0;
    };
    double seed = (double)r.seed;
    bool ok = d("V0", r.V0) && d("a", r.a) && d("A", r.A) && d("B", r.B) && d("C", r.C) && d("gamma", r.gamma) &&
              d("epsilon_p", r.epsilon_p) && d("epsilon_n", r.epsilon_n) && d("k_fwhm", r.k_fwhm) &&
              d("r_fwhm", r.r_fwhm) && d("t_f", r.t_f) && d("t_exc", r.t_exc) && d("eta_exc", r.eta_exc) &&
              d("d_max_scale", r.d_max_scale) && i("nx", r.nx) && i("num_test_part", r.num_test_part) &&
              i("steps", r.steps) && i("n", r.n) && i("z", r.z) && i("substeps", r.substeps) &&
              i("use_gpu", r.use_gpu) && d("sor_omega", r.sor_omega) && d("sor_tol", r.sor_tol) &&
              i("sor_max_it", r.sor_max_it) && i("boost", r.boost) && d("seed", seed) && d("r_init", r.r_init) &&
              i("sort_every", r.sort_every) && i("vtk", r.vtk) && i("init", r.init);
    r.seed = (unsigned long long)seed;
    if (!ok) {
        why = "a value does not parse";
        return false;
    }
    if (r.steps <= 0 || r.substeps <= 0 || r.nx < 4 || r.num_test_part <= 0 || r.z < 0 || r.n < 0 ||
        r.z + r.n <= 0 || !(r.t_f > 0) || !(r.d_max_scale > 0) || !(r.r_fwhm > 0) || (r.init != 0 && r.init != 1) ||
        (r.init == 0 && (!(r.a > 0) || !(r.k_fwhm >= 0) || r.z < 1 || r.n < 1))) {
        why = "a value is out of range";
        return false;
    }
    return true;
}

double sigma_from_fwhm(double fwhm) { return fwhm / (2.0 * std::sqrt(2.0 * std::log(2.0))); }  // calc_sigma
double nuclear_radius(int a) { return 1.5 * std::pow((double)a, 1.0 / 3.0); }                 // physics_formulas.hpp:31-34

void fill_params(const Run &r, hv_world &w, hv_skyrme &s, hv_params &p) {
    const double d = r.d_max_scale * nuclear_radius(r.z + r.n);  // include/init.hpp:329
    for (int a = 0; a < 3; a++) {
        w.n[a] = r.nx;
        w.d_max[a] = d;
    }
    s = {r.A, r.B, r.C, r.gamma};
    p = {};
    p.t_f = r.t_f;
    p.steps = r.steps;
    p.substeps = r.substeps;
    p.t_exc = r.t_exc;
    p.eta_exc = r.eta_exc;
    p.boost_mode = r.boost;
    p.z = r.z;
    p.n = r.n;
    p.sigma_r = sigma_from_fwhm(r.r_fwhm);
    p.part_per_nucleon = r.num_test_part;
    p.sor_omega = r.sor_omega;
    p.sor_tol = r.sor_tol;
    p.sor_max_it = r.sor_max_it;
}

int fail(hv_ctx *c, const char *what, int rc) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, c ? hv_last_error(c) : "");
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "BAD ARGUMENTS!\n");
        return 1;
    }
    const bool check = !std::strcmp(argv[1], "--check");
    const bool mixed = !std::strcmp(argv[1], "--float");
    if (mixed && argc < 4) {
        std::fprintf(stderr, "BAD ARGUMENTS!\n");
        return 1;
    }
    const char *input = (check || mixed) ? argv[2] : argv[1];
    Run r;
    std::string why;
    if (!read_input(input, r, why)) {
        std::fprintf(stderr, "Error: Invalid input file. (%s)\n", why.c_str());
        return 1;
    }
    hv_world w;
    hv_skyrme s;
    hv_params p;
    fill_params(r, w, s, p);
    const int np = r.z * r.num_test_part, nn = r.n * r.num_test_part;
    const double rmax = r.r_init * std::pow((double)(r.z + r.n), 1.0 / 3.0);
    const double kmax = hv_fermi_momentum(0.16);
    if (check) {
        std::printf("{\"n\": %d, \"d_max\": %.17g, \"n_protons\": %d, \"n_neutrons\": %d, \"sigma_r\": %.17g, "
                    "\"dt\": %.17g, \"steps\": %d, \"substeps\": %d, \"t_exc\": %.17g, \"eta_exc\": %.17g, "
                    "\"skyrme\": [%.17g, %.17g, %.17g, %.17g], \"sor\": [%.17g, %.17g, %d], \"boost\": %d, "
                    "\"seed\": %llu, \"r_init\": %.17g, \"k_fermi\": %.17g, \"vtk\": %d, \"init\": %d}\n",
                    w.n[0], w.d_max[0], np, nn, p.sigma_r, p.t_f / p.steps, p.steps, p.substeps, p.t_exc, p.eta_exc,
                    s.A, s.B, s.C, s.gamma, p.sor_omega, p.sor_tol, p.sor_max_it, p.boost_mode, r.seed, rmax, kmax,
                    r.vtk, r.init);
        return 0;
    }
    const char *outdir = mixed ? argv[3] : argv[2];
    std::printf("Simulation started.\n");
    const auto t0 = std::chrono::steady_clock::now();
    const int tot = np + nn;
    std::vector<double> x(tot), y(tot), z(tot), kx(tot), ky(tot), kz(tot);
    if (r.init == 0) {
        // the reference's initialize_particles (src/hyper_vlasov.cpp:139-141)
        std::printf("MAX TEST PART %i\n", hv_ws_max_particles(nuclear_radius(r.z + r.n), 1.5, r.num_test_part));
        double fermi[2] = {r.epsilon_p, r.epsilon_n}, chi2[2] = {0.0, 0.0};
        hv_woods_saxon ws[2];
        int rounds = 0;
        const int irc = hv_init_woods_saxon(r.z, r.n, r.num_test_part, &s, p.sigma_r, sigma_from_fwhm(r.k_fwhm),
                                            r.V0, r.a, fermi, (unsigned)r.seed, x.data(), y.data(), z.data(),
                                            kx.data(), ky.data(), kz.data(), ws, &rounds, chi2);
        if (irc < 0) {
            std::fprintf(stderr, "initial state failed (%d)\n", irc);
            return 1;
        }
        if (irc & HV_WARN_INIT_NOT_CONVERGED) std::fprintf(stderr, "INITIALIZATION DID NOT CONVERGE!\n");
        std::printf("V0 %0.2lf R12 %0.2lf a %0.2lf\n", ws[0].V0, ws[0].R12, ws[0].a);
        std::printf("V0 %0.2lf R12 %0.2lf a %0.2lf\n", ws[1].V0, ws[1].R12, ws[1].a);
        std::printf("FERMI P %0.2lf FERMI N %0.2lf\nITERATIONS %i\n", fermi[0], fermi[1], rounds);
        std::printf("CHI SQUARED P %0.2lf CHI SQUARED N %0.2lf\n", chi2[0], chi2[1]);
    } else if (hv_sample_uniform_spheres(r.seed, np, nn, rmax, kmax, x.data(), y.data(), z.data(), kx.data(),
                                         ky.data(), kz.data()) != 0) {
        std::fprintf(stderr, "initial state: bad sizes\n");
        return 1;
    }
    hv_ctx *c = nullptr;
    int rc = hv_ctx_create(&w, np, nn, mixed ? HV_PREC_MIXED : HV_PREC_F64, &c);
    if (rc) return fail(c, "hv_ctx_create", rc);
    hv_set_option(c, HV_OPT_SORT_EVERY, r.sort_every);
    rc = hv_upload_particles(c, x.data(), y.data(), z.data(), kx.data(), ky.data(), kz.data(), nullptr, nullptr,
                             nullptr);
    if (rc) return fail(c, "hv_upload_particles", rc);
    const std::string stats_path = std::string(outdir) + "/stats.txt";  // set_stats_filename
    FILE *stats = std::fopen(stats_path.c_str(), "w");
    if (!stats) {
        std::fprintf(stderr, "CANNOT OPEN STATS FILE!\n");
        hv_ctx_destroy(c);
        return 1;
    }
    int warn = hv_run_init(c, &s, &p, nullptr);
    if (warn < 0) return fail(c, "hv_run_init", warn);
    const double dt = p.t_f / p.steps;
    // with snapshots, one chunk of steps per stats interval: the snapshot of
    // step s (s % substeps == 0) holds the fields at the top of step s
    const int chunk = r.vtk ? p.substeps : 256;
    std::vector<double> st(4 * (size_t)chunk);
    for (int first = 0; first < p.steps; first += chunk) {
        const int cnt = p.steps - first < chunk ? p.steps - first : chunk;
        if (r.vtk) {
            char name[64];
            std::snprintf(name, sizeof name, "/out-%04d.vtk", first / p.substeps);  // set_output_filename
            rc = hv_write_vtk(c, (std::string(outdir) + name).c_str());
            if (rc) return fail(c, "hv_write_vtk", rc);
            std::printf("Processed step: %i/%i.\n", first, p.steps);
        }
        rc = hv_run_steps(c, &s, &p, first, cnt, st.data(), nullptr);
        if (rc < 0) return fail(c, "hv_run_steps", rc);
        warn |= rc;
        for (int q = 0; q < cnt; q++) {
            const int step = first + q;
            if (step % p.substeps) continue;
            const double *v = st.data() + 4 * (size_t)q;  // <z>_p, <z>_n, <r^2>_p, <r^2>_n
            std::fprintf(stats, "%e %e %e %e %e\n", step * dt, std::sqrt(v[2]), std::sqrt(v[3]), v[0], v[1]);
        }
    }
    std::fclose(stats);
    if (warn & HV_WARN_SOR_NOT_CONVERGED) std::fprintf(stderr, "SOR COULOMB DID NOT CONVERGE!\n");
    hv_ctx_destroy(c);
    std::printf("Simulation ended.\n");
    std::printf("Time taken: %0.3lfs\n",
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    return 0;
}
