// hv_kernels.h -- gfx950 kernels of the hyper-vlasov step loop.
//
// Precision templates: T is the particle / grid precision (double for the
// plain FP64 path, float for the mixed path); the Coulomb potential, the SOR
// update and every diagnostics reduction stay in double.
#pragma once

#include <type_traits>

#include "hv_device.h"

namespace hv {

template <typename T>
struct Particles {
    T *x, *y, *z, *kx, *ky, *kz, *fx, *fy, *fz;
};

// Per-step diagnostics accumulator: four exact 128-bit sums (z_p, z_n, r2_p,
// r2_n as 2^-30 fixed point) and two in-box counts.
struct StatSlot {
    unsigned long long s[4][2];
    unsigned long long inbox[2];
};

// status bits written by kernels (mirrors HV_WARN_* in hv_abi.h)
constexpr unsigned kStatusOutside = 2u;

// ---- diagnostics of one particle (include/tools.hpp:303-338) ------------------
template <typename T>
__device__ __forceinline__ void stat_values(T x, T y, T z, bool proton, const Geom<T> &g,
                                            long long (&v)[6]) {
    double xd = x, yd = y, zd = z;
    double dmx = g.dmx, dmy = g.dmy, dmz = g.dmz;
    bool inbox = !(xd < -dmx || xd > dmx || yd < -dmy || yd > dmy || zd < -dmz || zd > dmz);
    if (xd != xd || yd != yd || zd != zd) inbox = false;
    double r2 = xd * xd + yd * yd + zd * zd;
    long long qz = hv_stat_fix(zd);
    long long qr = inbox ? hv_stat_fix(r2) : 0;
    v[0] = proton ? qz : 0;
    v[1] = proton ? 0 : qz;
    v[2] = proton ? qr : 0;
    v[3] = proton ? 0 : qr;
    v[4] = (proton && inbox) ? 1 : 0;
    v[5] = (!proton && inbox) ? 1 : 0;
}

__device__ __forceinline__ void stat_commit(long long (&v)[6], StatSlot *slot) {
    block_sum_i64<6>(v);
    if (threadIdx.x == 0) {
        for (int c = 0; c < 4; c++)
            if (v[c]) atomic_add_i128(slot->s[c], v[c]);
        if (v[4]) atomicAdd(&slot->inbox[0], (unsigned long long)v[4]);
        if (v[5]) atomicAdd(&slot->inbox[1], (unsigned long long)v[5]);
    }
}

// ---- CIC deposit of one particle (include/tools.hpp:62-94) ------------------
template <typename T>
__device__ __forceinline__ void deposit_particle(const Cell<T> &c, long long off, const Geom<T> &g,
                                                 unsigned long long *acc) {
#pragma unroll
    for (int corner = 0; corner < 8; corner++) {
        int i = c.x0 + (corner & 1), j = c.y0 + ((corner >> 1) & 1), k = c.z0 + ((corner >> 2) & 1);
        if (i >= g.nx || j >= g.ny || k >= g.nz) continue;
        long long q = hv_fix_weight((double)corner_weight(c, corner));
        atomicAdd(&acc[off + idx3(i, j, k, g.ny, g.nz)], (unsigned long long)q);
    }
}

// ---- fused particle pass --------------------------------------------------
// One step of the particle side of the loop, for positions x_s and the field U
// built from x_s (src/hyper_vlasov.cpp:86-93 and :101-103):
//   F = gather(-grad U)(x_s)
//   if end_kick: k += F dt/(2 hbar c)          (end of step s-1, :103)
//   if boost:    kz += boost_p / kz -= boost_n  (:86-89)
//   k += F dt/(2 hbar c)                        (start of step s, :90)
//   x += k dt hbar c / mc^2                     (:91)
//   deposit x_{s+1} into acc, diagnostics of x_{s+1} into *slot.
// The intermediate k of the two half kicks is rounded exactly as in the
// reference's two separate updates.
template <typename T, typename UT>
__global__ void __launch_bounds__(kBlock)
k_particle_step(Particles<T> p, int np, int total, Geom<T> g, const UT *__restrict__ U,
                unsigned long long *__restrict__ acc, T kick_fac, T drift_fac, int end_kick,
                int boost, T boost_p, T boost_n, StatSlot *slot, unsigned *status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    long long v[6] = {0, 0, 0, 0, 0, 0};
    if (i < total) {
        const bool proton = i < np;
        const long long off = proton ? 0 : g.G;
        T x = p.x[i], y = p.y[i], z = p.z[i];
        T kx = p.kx[i], ky = p.ky[i], kz = p.kz[i];
        Cell<T> c;
        T fx = T(0.0), fy = T(0.0), fz = T(0.0);
        if (cic_cell(x, y, z, g, c)) gather_force_from_u<T, UT>(U + off, c, g, fx, fy, fz);
        if (end_kick) {
            kx += kick_fac * fx;
            ky += kick_fac * fy;
            kz += kick_fac * fz;
        }
        if (boost) {
            if (proton) kz += boost_p;
            else kz -= boost_n;
        }
        kx += kick_fac * fx;
        ky += kick_fac * fy;
        kz += kick_fac * fz;
        x += drift_fac * kx;
        y += drift_fac * ky;
        z += drift_fac * kz;
        p.x[i] = x; p.y[i] = y; p.z[i] = z;
        p.kx[i] = kx; p.ky[i] = ky; p.kz[i] = kz;
        if (cic_cell(x, y, z, g, c)) deposit_particle(c, off, g, acc);
        else atomicOr(status, kStatusOutside);
        stat_values(x, y, z, proton, g, v);
    }
    stat_commit(v, slot);
}

// deposit (and optionally diagnostics) of the current positions, no motion
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_deposit(Particles<T> p, int np, int total, Geom<T> g, unsigned long long *__restrict__ acc,
          StatSlot *slot, unsigned *status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    long long v[6] = {0, 0, 0, 0, 0, 0};
    if (i < total) {
        const bool proton = i < np;
        T x = p.x[i], y = p.y[i], z = p.z[i];
        Cell<T> c;
        if (acc) {
            if (cic_cell(x, y, z, g, c)) deposit_particle(c, proton ? 0 : g.G, g, acc);
            else atomicOr(status, kStatusOutside);
        }
        stat_values(x, y, z, proton, g, v);
    }
    if (slot) stat_commit(v, slot);
}

// gather F(x) from U into p.f, optionally followed by the half kick
template <typename T, typename UT>
__global__ void __launch_bounds__(kBlock)
k_gather_u(Particles<T> p, int np, int total, Geom<T> g, const UT *__restrict__ U, T kick_fac, int kick) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long long off = (i < np) ? 0 : g.G;
    Cell<T> c;
    T fx = T(0.0), fy = T(0.0), fz = T(0.0);
    if (cic_cell(p.x[i], p.y[i], p.z[i], g, c)) gather_force_from_u<T, UT>(U + off, c, g, fx, fy, fz);
    p.fx[i] = fx;
    p.fy[i] = fy;
    p.fz[i] = fz;
    if (kick) {
        p.kx[i] += kick_fac * fx;
        p.ky[i] += kick_fac * fy;
        p.kz[i] += kick_fac * fz;
    }
}

// fixed-point counts -> T (per-reference-kernel path)
template <typename T>
__global__ void k_counts_to_field(const unsigned long long *__restrict__ acc, T *__restrict__ out,
                                  long long n) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = (T)hv_fix_to_double((long long)acc[q]);
}

template <typename V>
__global__ void k_fill(V *__restrict__ p, V v, long long n) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; q < n; q += stride) p[q] = v;
}

// ---- Gaussian smearing (include/physics.hpp:203-234), three factor passes ----
// Device copy of the tap table, in T.
template <typename T>
struct SmearTab {
    int rx, ry, rz;
    int kb[HV_SMEAR_MAXR + 1];
    int kc[HV_SMEAR_MAXR + 1][HV_SMEAR_MAXR + 1];
    T gx[HV_SMEAR_MAXR + 1], gy[HV_SMEAR_MAXR + 1], gz[HV_SMEAR_MAXR + 1];
    T term;
};

// z pass: H_K(q) = (H_{K-1} + gz(K) c(q-K)) + gz(K) c(q+K), H_0 = gz(0) c(q).
// CT is the count type: unsigned long long (fixed point) or T.
template <typename T, typename CT>
__device__ __forceinline__ T count_value(const CT *c, long long q) {
    if constexpr (sizeof(CT) == 8 && !std::is_floating_point<CT>::value)
        return (T)hv_fix_to_double((long long)c[q]);
    else
        return (T)c[q];
}

template <typename T, typename CT>
__global__ void __launch_bounds__(kBlock)
k_smear_z(const CT *__restrict__ cnt, T *__restrict__ H, const SmearTab<T> *__restrict__ tb, int nz,
          long long n2) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n2) return;
    const int k = (int)(q % nz);
    const int rz = tb->rz;
    T h = tb->gz[0] * count_value<T, CT>(cnt, q);
    H[q] = h;
    for (int K = 1; K <= rz; K++) {
        T lo = (k - K >= 0) ? count_value<T, CT>(cnt, q - K) : T(0.0);
        T hi = (k + K < nz) ? count_value<T, CT>(cnt, q + K) : T(0.0);
        h = (h + tb->gz[K] * lo) + tb->gz[K] * hi;
        H[(long long)K * n2 + q] = h;
    }
}

// y pass: D_a(q) = sum_{b=-kb(a)..kb(a)} gy(|b|) H_{kc(a,|b|)}(q + b nz)
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_smear_y(const T *__restrict__ H, T *__restrict__ D, const SmearTab<T> *__restrict__ tb, int ny, int nz,
          long long n2) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n2) return;
    const int j = (int)((q / nz) % ny);
    for (int a = 0; a <= tb->rx; a++) {
        const int kb = tb->kb[a];
        T acc = T(0.0);
        for (int b = -kb; b <= kb; b++) {
            const int ab = b < 0 ? -b : b;
            const int jb = j + b;
            T hv = (jb >= 0 && jb < ny) ? H[(long long)tb->kc[a][ab] * n2 + q + (long long)b * nz] : T(0.0);
            acc = acc + tb->gy[ab] * hv;
        }
        D[(long long)a * n2 + q] = acc;
    }
}

// x pass: rho(q) = (sum_a gx(|a|) D_{|a|}(q + a ny nz)) * term
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_smear_x(const T *__restrict__ D, T *__restrict__ rho, const SmearTab<T> *__restrict__ tb, int nx, int ny,
          int nz, long long n2) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n2) return;
    const long long plane = (long long)ny * nz;
    const int i = (int)((q / plane) % nx);
    T acc = T(0.0);
    for (int a = -tb->rx; a <= tb->rx; a++) {
        const int aa = a < 0 ? -a : a;
        if (tb->kb[aa] < 0) continue;
        const int ia = i + a;
        T dv = (ia >= 0 && ia < nx) ? D[(long long)aa * n2 + q + (long long)a * plane] : T(0.0);
        acc = acc + tb->gx[aa] * dv;
    }
    rho[q] = acc * tb->term;
}

// ---- Skyrme potentials (include/physics.hpp:273-281), optionally + phi ------
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_skyrme(const T *__restrict__ rho, T *__restrict__ U, const double *__restrict__ phi, double A, double B,
         double C, double gamma, long long G) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) return;
    double rp = (double)rho[i], rn = (double)rho[i + G];
    double up = hv_skyrme_f64(A, B, C, gamma, rp, rn, 0);
    double un = hv_skyrme_f64(A, B, C, gamma, rp, rn, 1);
    if (phi) {
        // merge_volumetric_potentials (include/tools.hpp:378-384), in the
        // grid precision like the separate pass
        U[i] = (T)up + (T)phi[i];
    } else {
        U[i] = (T)up;
    }
    U[i + G] = (T)un;
}

// U_p += phi (include/tools.hpp:378-384)
template <typename T>
__global__ void k_merge(T *__restrict__ U, const double *__restrict__ phi, long long G) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < G) U[i] += (T)phi[i];
}

// F = -grad U on the grid (include/physics.hpp:160-200)
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_fdm(const T *__restrict__ U, T *__restrict__ Fx, T *__restrict__ Fy, T *__restrict__ Fz, Geom<T> g) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 2 * g.G) return;
    const long long off = q >= g.G ? g.G : 0;
    const long long node = q - off;
    const int k = (int)(node % g.nz), j = (int)((node / g.nz) % g.ny), i = (int)(node / ((long long)g.ny * g.nz));
    T fx, fy, fz;
    fdm_force<T, T>(U + off, i, j, k, g, fx, fy, fz);
    Fx[q] = fx;
    Fy[q] = fy;
    Fz[q] = fz;
}

// CIC gather of grid forces (include/tools.hpp:99-168)
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_gather_f(Particles<T> p, int np, int total, Geom<T> g, const T *__restrict__ Fx, const T *__restrict__ Fy,
           const T *__restrict__ Fz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long long off = (i < np) ? 0 : g.G;
    Cell<T> c;
    T ax = T(0.0), ay = T(0.0), az = T(0.0);
    if (cic_cell(p.x[i], p.y[i], p.z[i], g, c)) {
#pragma unroll
        for (int corner = 0; corner < 8; corner++) {
            int ii = c.x0 + (corner & 1), jj = c.y0 + ((corner >> 1) & 1), kk = c.z0 + ((corner >> 2) & 1);
            if (ii >= g.nx || jj >= g.ny || kk >= g.nz) continue;
            T w = corner_weight(c, corner);
            long long idx = off + idx3(ii, jj, kk, g.ny, g.nz);
            ax += w * Fx[idx];
            ay += w * Fy[idx];
            az += w * Fz[idx];
        }
    }
    p.fx[i] = ax;
    p.fy[i] = ay;
    p.fz[i] = az;
}

// k += F fac (include/physics.hpp:249-258)
template <typename T>
__global__ void k_kick(Particles<T> p, int total, T fac) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    p.kx[i] += fac * p.fx[i];
    p.ky[i] += fac * p.fy[i];
    p.kz[i] += fac * p.fz[i];
}

// x += k fac (include/physics.hpp:261-270)
template <typename T>
__global__ void k_drift(Particles<T> p, int total, T fac) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    p.x[i] += fac * p.kx[i];
    p.y[i] += fac * p.ky[i];
    p.z[i] += fac * p.kz[i];
}

// dipole boost (include/physics.hpp:237-246)
template <typename T>
__global__ void k_excite(T *__restrict__ kz, int np, int total, T dp, T dn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    if (i < np) kz[i] += dp;
    else kz[i] -= dn;
}

// ---- red-black SOR (include/physics.hpp:100-157) ---------------------------
struct SorCoef {
    double dx2, dy2, dz2;   // dx*dx etc.
    double inv_dx2;         // 1 / (2/dx^2 + 2/dy^2 + 2/dz^2)
    double src;             // 4 pi 1.44
    double omega, one_minus_omega;
    double tol;
};

struct SorCtl {
    int done;
    int iters;
};

// One colour of one iteration, in place.  Block (64, 4): x runs along k.
// Iteration `it` first looks at the previous iteration's max |dphi|: once it
// is below tol the solve is over and every later launch returns at once.
template <typename RT>
__global__ void __launch_bounds__(256)
k_sor_colour(double *__restrict__ phi, const RT *__restrict__ rho, int nx, int ny, int nz, SorCoef cf,
             int colour, int it, unsigned long long *__restrict__ maxd, SorCtl *__restrict__ ctl) {
    if (ctl->done) return;
    if (colour == 0 && it > 0) {
        double prev = __longlong_as_double((long long)maxd[it - 1]);
        if (prev < cf.tol) {
            if (threadIdx.x == 0 && threadIdx.y == 0 && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0) {
                ctl->done = 1;
                ctl->iters = it;
            }
            return;
        }
    }
    const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 1 + blockIdx.z;
    double diff = 0.0;
    if (k < nz - 1 && j < ny - 1 && ((i + j + k) & 1) == colour) {
        const long long idx = idx3(i, j, k, ny, nz);
        const long long sx = (long long)ny * nz;
        double r = (double)rho[idx];
        double phi_x = (phi[idx + sx] + phi[idx - sx]) / cf.dx2;
        double phi_y = (phi[idx + nz] + phi[idx - nz]) / cf.dy2;
        double phi_z = (phi[idx + 1] + phi[idx - 1]) / cf.dz2;
        double phi_star = (phi_x + phi_y + phi_z + cf.src * r) * cf.inv_dx2;
        double phi_old = phi[idx];
        double nv = cf.one_minus_omega * phi_old + cf.omega * phi_star;
        phi[idx] = nv;
        diff = fabs(nv - phi_old);
    }
    block_max_atomic(diff, &maxd[it]);
}

}  // namespace hv
