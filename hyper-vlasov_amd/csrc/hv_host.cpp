// hv_host.cpp -- host-only parts of libhvamd.so: the seeded initial-state
// sampler (include/hv_init.h).
#include <cmath>
#include <cstdint>

#include "hv_abi.h"
#include "hv_init.h"

namespace {

// splitmix64 finaliser (Steele, Lea, Flood 2014)
inline uint64_t mix(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// uniform double in [0, 1) from 53 random bits
inline double unit(uint64_t key) { return (double)(mix(key) >> 11) * (1.0 / 9007199254740992.0); }

// rejection sample of a point uniform in the ball of radius r, stream `key`
void ball(uint64_t key, double r, double *out) {
    for (uint64_t d = 0;; d += 3) {
        double v[3];
        for (int a = 0; a < 3; a++) v[a] = r * (2.0 * unit(key + d + (uint64_t)a) - 1.0);
        if (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < r * r) {
            out[0] = v[0];
            out[1] = v[1];
            out[2] = v[2];
            return;
        }
    }
}

}  // namespace

extern "C" {

unsigned long long hv_mix64(unsigned long long v) { return mix(v); }

double hv_fermi_momentum(double rho) { return std::cbrt(3.0 * M_PI * M_PI * rho / 2.0); }

int hv_sample_uniform_spheres(unsigned long long seed, int n_protons, int n_neutrons, double r_max,
                              double k_max, double *x, double *y, double *z, double *kx, double *ky,
                              double *kz) {
    if (n_protons < 0 || n_neutrons < 0 || !(r_max > 0.0) || !(k_max > 0.0)) return HV_ERR_ARG;
    if (!x || !y || !z || !kx || !ky || !kz) return HV_ERR_ARG;
    const int total = n_protons + n_neutrons;
    const uint64_t base = mix(seed);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < total; i++) {
        double r[3], k[3];
        // two independent streams per particle, keyed by (seed, 2i) and (seed, 2i+1)
        ball(mix(base ^ ((uint64_t)i << 1)) , r_max, r);
        ball(mix(base ^ (((uint64_t)i << 1) | 1ULL)), k_max, k);
        x[i] = r[0];
        y[i] = r[1];
        z[i] = r[2];
        kx[i] = k[0];
        ky[i] = k[1];
        kz[i] = k[2];
    }
    return HV_OK;
}

}  // extern "C"
