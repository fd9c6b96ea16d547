// hv_adjoint.h -- kernels of the discrete adjoint of the step loop (FP64).
//
// The force on a test particle is F(x) = gather(f, x) with the grid force
// f = -grad_h U, U_q = S_q(rho; A, B, C, gamma) (+ phi for protons),
// phi = (-L_h)^-1 (4 pi e^2 rho_p) on the interior with fixed faces, and
// rho = smear(deposit(x)).  The vector-Jacobian product of F with a
// cotangent c_i per particle runs the chain backwards, each stage the exact
// transpose of its forward kernel:
//   gather        -> deposit of c_i with the same CIC weights, summed exactly in
//                    fixed point (k_adj_deposit_vec)
//   f = -grad_h U -> minus the transposed difference operator (k_adj_fdm_t)
//   U = S(rho)    -> pointwise derivative; its theta part reduced (k_adj_skyrme)
//   phi(rho_p)    -> the same SOR on (-L_h) with zero faces (L_h is symmetric)
//   smear         -> the same smearing (its matrix is symmetric)
//   deposit and the weights of the gather -> derivative of the CIC weights
//                   (k_adj_particle)
// No counterpart in the reference.
#pragma once

#include "hv_kernels.h"

namespace hv {

// Fixed-point scale of the gathered cotangent sums: 2^k with
// N max|cot| 2^k < 2^61, so every node's sum of at most N terms
// |rint(w cot 2^k)| <= w |cot| 2^k + 1/2 stays below 2^62 in int64.  One
// thread; amax holds max |cot| over the 3N components.
__global__ void k_adj_fix_scale(const double *__restrict__ amax, int n, double *__restrict__ scale) {
    const double m = *amax * (double)n;
    int e = 0;
    if (m > 0.0) (void)frexp(m, &e);  // m < 2^e
    int k = 61 - e;
    k = k > 1000 ? 1000 : (k < -1000 ? -1000 : k);
    scale[0] = ldexp(1.0, k);
    scale[1] = ldexp(1.0, -k);
}

// transpose of the CIC gather: fbar[a][species][node] += w * cot_a, summed
// exactly as int64 multiples of 1/scale (integer adds: the sum, and so the
// gradient, does not depend on the order the particles are visited in)
__global__ void __launch_bounds__(kBlock)
k_adj_deposit_vec(Particles<double> p, int np, int total, Geom<double> g, const double *__restrict__ cot,
                  const double *__restrict__ scale, unsigned long long *__restrict__ fbar) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    Cell<double> c;
    if (!cic_cell(p.x[i], p.y[i], p.z[i], g, c)) return;
    const long long G2 = 2 * g.G, off = (i < np) ? 0 : g.G;
    const double sc = scale[0];
    const double cx = cot[i], cy = cot[total + i], cz = cot[2LL * total + i];
#pragma unroll
    for (int corner = 0; corner < 8; corner++) {
        const int ii = c.x0 + (corner & 1), jj = c.y0 + ((corner >> 1) & 1), kk = c.z0 + ((corner >> 2) & 1);
        if (ii >= g.nx || jj >= g.ny || kk >= g.nz) continue;
        const double w = corner_weight(c, corner);
        const long long q = off + idx3(ii, jj, kk, g.ny, g.nz);
        atomicAdd(&fbar[q], (unsigned long long)(long long)rint((w * cx) * sc));
        atomicAdd(&fbar[G2 + q], (unsigned long long)(long long)rint((w * cy) * sc));
        atomicAdd(&fbar[2 * G2 + q], (unsigned long long)(long long)rint((w * cz) * sc));
    }
}

// d g(c') / d U(i) along one axis of n nodes for the one-axis gradient g of
// fdm_force (central inside, one-sided on the faces), summed against gbar:
// returns sum_{c'} gbar(c') dg(c')/dU(i); gbar(m) is the cotangent at index m.
template <typename Get>
__device__ __forceinline__ double fdm_axis_t(int i, int n, double inv_h, double inv_2h, Get gbar) {
    double s = 0.0;
    if (i - 1 >= 1 && i - 1 <= n - 2) s += gbar(i - 1) * inv_2h;   // g(i-1) = (U(i) - U(i-2)) / 2h
    if (i + 1 >= 1 && i + 1 <= n - 2) s -= gbar(i + 1) * inv_2h;   // g(i+1) = (U(i+2) - U(i)) / 2h
    if (i == 1) s += gbar(0) * inv_h;                              // g(0) = (U(1) - U(0)) / h
    if (i == 0) s -= gbar(0) * inv_h;
    if (i == n - 1) s += gbar(n - 1) * inv_h;                      // g(n-1) = (U(n-1) - U(n-2)) / h
    if (i == n - 2) s -= gbar(n - 1) * inv_h;
    return s;
}

// Ubar = -D^T fbar (f = -grad_h U), both species; fbar in int64 multiples of
// 1/scale (k_adj_deposit_vec), read back exactly (scale is a power of two)
__global__ void __launch_bounds__(kBlock)
k_adj_fdm_t(const unsigned long long *__restrict__ fbar, const double *__restrict__ scale, double *__restrict__ Ubar,
            Geom<double> g) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long G = g.G, G2 = 2 * G;
    if (q >= G2) return;
    const long long off = q >= G ? G : 0;
    const long long node = q - off;
    const double inv = scale[1];
    const int k = (int)(node % g.nz), j = (int)((node / g.nz) % g.ny), i = (int)(node / ((long long)g.ny * g.nz));
    const unsigned long long *fx = fbar + off, *fy = fbar + G2 + off, *fz = fbar + 2 * G2 + off;
    const double sx = fdm_axis_t(i, g.nx, 1.0 / g.dx, 1.0 / g.two_dx,
                                 [&](int m) { return (double)(long long)fx[idx3(m, j, k, g.ny, g.nz)] * inv; });
    const double sy = fdm_axis_t(j, g.ny, 1.0 / g.dy, 1.0 / g.two_dy,
                                 [&](int m) { return (double)(long long)fy[idx3(i, m, k, g.ny, g.nz)] * inv; });
    const double sz = fdm_axis_t(k, g.nz, 1.0 / g.dz, 1.0 / g.two_dz,
                                 [&](int m) { return (double)(long long)fz[idx3(i, j, m, g.ny, g.nz)] * inv; });
    Ubar[q] = -(sx + sy + sz);
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int s = kWave / 2; s > 0; s >>= 1) v += __shfl_xor(v, s, kWave);
    return v;
}

// Skyrme stage: rhobar = (dS/drho)^T Ubar per node and the theta cotangent
// partial sums per block (fixed reduction order):
//   dS_q/dA = t, dS_q/dB = t^gamma, dS_q/dC = 2 tau_q (rho_n - rho_p)/rho0,
//   dS_q/dgamma = B t^gamma ln t,   dS_q/drho_{p,n} = (A + B gamma t^(gamma-1))/rho0 -/+ 2 tau_q C/rho0.
// partials[4 * block + {0,1,2,3}] = theta cotangents in hv_skyrme order (A, B, C, gamma).
__global__ void __launch_bounds__(kBlock)
k_adj_skyrme(const double *__restrict__ rho, const double *__restrict__ Ubar, double *__restrict__ rhobar,
             double A, double B, double C, double gamma, long long G, double *__restrict__ partials) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < G) {
        const double rp = rho[i], rn = rho[i + G];
        const double up = Ubar[i], un = Ubar[i + G];
        const double t = (rp + rn) / HV_RHO0;
        const double tg = t > 0.0 ? hv_det_pow(t, gamma) : 0.0;
        const double lt = t > 0.0 ? hv_det_log(t) : 0.0;
        const double s = up + un;
        v[0] = s * t;
        v[1] = s * tg;
        v[2] = (un - up) * (2.0 * (rn - rp) / HV_RHO0);
        v[3] = s * (B * tg * lt);
        const double common = (A + (t > 0.0 ? B * gamma * (tg / t) : 0.0)) / HV_RHO0;
        const double cc = 2.0 * C / HV_RHO0;
        rhobar[i] = up * (common + cc) + un * (common - cc);
        rhobar[i + G] = up * (common - cc) + un * (common + cc);
    }
    __shared__ double red[4][16];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
#pragma unroll
    for (int c = 0; c < 4; c++) v[c] = wave_sum_f64(v[c]);
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < 4; c++) red[c][wave] = v[c];
    __syncthreads();
    if (threadIdx.x < 4) {
        double sacc = 0.0;
        for (int w = 0; w < nw; w++) sacc += red[threadIdx.x][w];
        partials[4LL * blockIdx.x + threadIdx.x] = sacc;
    }
}

// acc[c] += sum over blocks of partials[4 b + c], in block order
__global__ void k_adj_theta_sum(const double *__restrict__ partials, int nblocks, double *__restrict__ acc) {
    const int c = threadIdx.x;
    if (c >= 4) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; b++) s += partials[4LL * b + c];
    acc[c] += s;
}

// max |v| over n values into out[block] (for the adjoint SOR tolerance)
__global__ void __launch_bounds__(kBlock)
k_adj_absmax(const double *__restrict__ v, long long n, double *__restrict__ out) {
    double m = 0.0;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        const double a = fabs(v[q]);
        m = a > m ? a : m;
    }
#pragma unroll
    for (int s = kWave / 2; s > 0; s >>= 1) {
        const double o = __shfl_xor(m, s, kWave);
        m = o > m ? o : m;
    }
    __shared__ double red[16];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    if (lane == 0) red[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / kWave); w++) m = red[w] > m ? red[w] : m;
        out[blockIdx.x] = m;
    }
}

// rhobar_p[i] += psi[split(i)] (the adjoint Poisson solution, natural order)
__global__ void k_adj_add_split(const double *__restrict__ psi, double *__restrict__ rhobar, long long G, Split sp) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) return;
    const int k = (int)(i % sp.nz), j = (int)((i / sp.nz) % sp.ny), ii = (int)(i / ((long long)sp.ny * sp.nz));
    rhobar[i] += psi[split_index(sp, ii, j, k)];
}

// Kbar' = b X + Kbar (in place) and the force cotangent cot = am Kbar'
__global__ void k_adj_kbar(double *__restrict__ Kb, const double *__restrict__ X, double *__restrict__ cot, double b,
                           double am, long long n3) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n3) return;
    const double k = b * X[q] + Kb[q];
    Kb[q] = k;
    cot[q] = am * k;
}

// Position cotangent of one step: X_i += sum_corners dW/dx (cot_i . f(corner)
// + cbar(corner)) + the loss gradient dl/dz_i (gz_p for protons, gz_n for
// neutrons), with f = -grad_h U computed like the forward gather.
__global__ void __launch_bounds__(kBlock)
k_adj_particle(Particles<double> p, int np, int total, Geom<double> g, const double *__restrict__ U,
               const double *__restrict__ cot, const double *__restrict__ cbar, double *__restrict__ X, double gz_p,
               double gz_n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const bool proton = i < np;
    X[2LL * total + i] += proton ? gz_p : gz_n;
    Cell<double> c;
    if (!cic_cell(p.x[i], p.y[i], p.z[i], g, c)) return;
    const long long off = proton ? 0 : g.G;
    const double cx = cot[i], cy = cot[total + i], cz = cot[2LL * total + i];
    const double ix = 1.0 / g.dx, iy = 1.0 / g.dy, iz = 1.0 / g.dz;
    double dX = 0.0, dY = 0.0, dZ = 0.0;
#pragma unroll
    for (int corner = 0; corner < 8; corner++) {
        const int ii = c.x0 + (corner & 1), jj = c.y0 + ((corner >> 1) & 1), kk = c.z0 + ((corner >> 2) & 1);
        if (ii >= g.nx || jj >= g.ny || kk >= g.nz) continue;
        const double wx = (corner & 1) ? c.d_x : c.t_x, dwx = (corner & 1) ? ix : -ix;
        const double wy = (corner & 2) ? c.d_y : c.t_y, dwy = (corner & 2) ? iy : -iy;
        const double wz = (corner & 4) ? c.d_z : c.t_z, dwz = (corner & 4) ? iz : -iz;
        double fx, fy, fz;
        fdm_force<double, double>(U + off, ii, jj, kk, g, fx, fy, fz);
        const double val = cx * fx + cy * fy + cz * fz + cbar[off + idx3(ii, jj, kk, g.ny, g.nz)];
        dX += dwx * wy * wz * val;
        dY += wx * dwy * wz * val;
        dZ += wx * wy * dwz * val;
    }
    X[i] += dX;
    X[(long long)total + i] += dY;
    X[2LL * total + i] += dZ;
}

}  // namespace hv
