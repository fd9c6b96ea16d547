// hv_host.cpp -- host-only parts of libhvamd.so: the seeded initial-state
// sampler (include/hv_init.h) and the VTK snapshot writer.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

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

// ---- VTK snapshot (include/init.hpp:160-254) -----------------------------------
namespace {

// big-endian bits of a float (include/init.hpp:37-41)
inline uint32_t be_float(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    return __builtin_bswap32(u);
}

}  // namespace

extern "C" {

// The reference's snapshot file: header (output_vtk_header_start), the force
// field as VECTORS forces_p / forces_n (output_vector_field), then density
// and potentials as SCALARS *_p / *_n / *_t (output_scalar_field), every
// value a big-endian float32 with x fastest.  Fields come in the reference
// layout ((i ny + j) nz + k, protons then neutrons); single != 0 reproduces
// the float instantiation (d_max in float, the p + n sum in float).
int hv_vtk_write_fields(const char *path, const hv_world *world, const double *density, const double *potentials,
                        const double *fx, const double *fy, const double *fz, int single) {
    if (!path || !world || !density || !potentials || !fx || !fy || !fz) return HV_ERR_ARG;
    const int nx = world->n[0], ny = world->n[1], nz = world->n[2];
    if (nx <= 0 || ny <= 0 || nz <= 0) return HV_ERR_ARG;
    const long long G = (long long)nx * ny * nz;
    FILE *out = std::fopen(path, "wb");
    if (!out) return HV_ERR_IO;
    double dm[3];
    for (int a = 0; a < 3; a++) dm[a] = single ? (double)(float)world->d_max[a] : world->d_max[a];
    std::fprintf(out, "# vtk DataFile Version 3.0\n");
    std::fprintf(out, "Volumetric data\n");
    std::fprintf(out, "BINARY\n");
    std::fprintf(out, "DATASET STRUCTURED_POINTS\n");
    std::fprintf(out, "DIMENSIONS %d %d %d\n", nx, ny, nz);
    std::fprintf(out, "ORIGIN %lf %lf %lf\n", -dm[0], -dm[1], -dm[2]);
    std::fprintf(out, "SPACING %lf %lf %lf\n", 2.0 * dm[0] / nx, 2.0 * dm[1] / ny, 2.0 * dm[2] / nz);
    std::fprintf(out, "POINT_DATA %lld\n", G);
    std::vector<uint32_t> buf((size_t)(3 * G));
    // write position of node (i, j, k): x fastest
    auto wpos = [&](long long q) {
        const long long k = q % nz, j = (q / nz) % ny, i = q / ((long long)ny * nz);
        return (k * ny + j) * nx + i;
    };
    bool ok = true;
    const char tags[2] = {'p', 'n'};
    for (int sp = 0; sp < 2 && ok; sp++) {
        const long long off = sp * G;
        for (long long q = 0; q < G; q++) {
            const long long w = wpos(q);
            buf[3 * w] = be_float((float)fx[off + q]);
            buf[3 * w + 1] = be_float((float)fy[off + q]);
            buf[3 * w + 2] = be_float((float)fz[off + q]);
        }
        std::fprintf(out, "VECTORS forces_%c float\n", tags[sp]);
        ok = std::fwrite(buf.data(), sizeof(uint32_t), (size_t)(3 * G), out) == (size_t)(3 * G);
    }
    const double *fields[2] = {density, potentials};
    const char *names[2] = {"density", "potentials"};
    for (int f = 0; f < 2 && ok; f++) {
        const double *v = fields[f];
        for (int part = 0; part < 3 && ok; part++) {  // p, n, t
            for (long long q = 0; q < G; q++) {
                float x;
                if (part < 2)
                    x = (float)v[part * G + q];
                else if (single)
                    x = (float)v[q] + (float)v[G + q];
                else
                    x = (float)(v[q] + v[G + q]);
                buf[wpos(q)] = be_float(x);
            }
            std::fprintf(out, "SCALARS %s_%c float 1\n", names[f], part == 0 ? 'p' : (part == 1 ? 'n' : 't'));
            std::fprintf(out, "LOOKUP_TABLE default\n");
            ok = std::fwrite(buf.data(), sizeof(uint32_t), (size_t)G, out) == (size_t)G;
        }
    }
    if (std::fclose(out) != 0) ok = false;
    return ok ? HV_OK : HV_ERR_IO;
}

}  // extern "C"
