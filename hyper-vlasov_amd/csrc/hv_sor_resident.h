// hv_sor_resident.h -- red-black SOR with the whole potential held in LDS.
//
// One persistent, cooperatively launched block per CU owns a tile of
// TI x TJ grid rows (all k) of the Coulomb potential and of its source term,
// both kept in LDS for the whole solve (a 128^3 FP64 potential + FP64 source
// is 33.5 MB, 131 KB per CU over 256 CUs).  An iteration is
//   red sweep -> publish the tile's red face values -> wait for the 4
//   neighbour tiles' red faces -> black sweep -> publish black faces and the
//   tile's max |dphi| -> wait for every tile's max of this iteration ->
//   stop when it is below tol (include/physics.hpp:104-154).
// Only tile faces travel through memory, with device-scope (sc1) stores and
// loads that bypass the per-XCD L2s, ordered by s_waitcnt + s_barrier before
// a flag store.  No kernel launch and no atomics per iteration.
//
// Every node update is the reference's expression on the same operand values
// (the red sweep reads the black values of the previous iteration, the black
// sweep the red values of this one), so the iterate and the iteration count
// equal the sweep-per-launch path bit for bit.
//
// Every wait is bounded: a spin that exceeds kSpinLimit polls raises the error
// word and all blocks leave the loop, so a scheduling fault cannot hang the GPU.
#pragma once

#include "hv_device.h"

namespace hv {

constexpr int kResThreads = 512;
constexpr int kSlotPoll = 4;  // slots per lane: at most 4 * 64 = 256 tiles
constexpr int kResRows = 4;   // rows a wave updates at once
constexpr long long kSpinLimit = 1LL << 24;

struct ResTiles {
    int TI, TJ;    // tile extent along i and j (rows of all nz nodes)
    int PI, PJ;    // tiles along i and j
    int nzh;       // k slots per colour row
    int side_i;    // TJ * nzh: values on one i-face, per colour
    int side_j;    // TI * nzh
    int xstride;   // 2 * side_i + 2 * side_j: exchange values per tile and colour
};

#ifdef HV_RES_TRACE
#define HV_TRACE(slot)                                                                     \
    do {                                                                                   \
        if (t == 0 && tid == 0 && it < 16) sy.trace[it * 16 + (slot)] = wall_clock64();   \
    } while (0)
#else
#define HV_TRACE(slot) \
    do {                \
    } while (0)
#endif

struct ResSync {
#ifdef HV_RES_TRACE
    unsigned long long *trace;
#endif
    double *xr, *xb;          // exchange buffers (red, black): nT * xstride
    int *flag_red;            // nT: last iteration + 1 whose red faces are published
    double *slots;            // max_it * nT per-tile max |dphi|, preset to -1
    int *err;                 // set on a spin timeout
    double *result;           // [0] iterations done, [1] last max |dphi|
};

__device__ __forceinline__ double ld_dev(const double *p) {
    return __longlong_as_double(
        (long long)__hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__device__ __forceinline__ void st_dev(double *p, double v) {
    __hip_atomic_store((unsigned long long *)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int ld_dev_i(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_dev_i(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename RT>
__global__ void __launch_bounds__(kResThreads)
k_sor_resident(double *__restrict__ phi, const RT *__restrict__ rho, Split sp, ResTiles tl, SorCoef cf, int max_it,
               ResSync sy) {
    extern __shared__ unsigned char smem[];
    const int TI = tl.TI, TJ = tl.TJ, nzh = tl.nzh;
    const int rows = TI * TJ, tsz = rows * nzh;
    double *pc[2];
    pc[0] = (double *)smem;          // red nodes of the tile, (li, lj, h)
    pc[1] = pc[0] + tsz;             // black nodes
    double *halo = pc[1] + tsz;      // other-colour face values: im, ip (TJ*nzh each), jm, jp (TI*nzh each)
    double *h_im = halo, *h_ip = halo + tl.side_i, *h_jm = halo + 2 * tl.side_i, *h_jp = h_jm + tl.side_j;
    double *red = h_jp + tl.side_j;  // 16 wave partials + broadcast
    RT *rc[2];
    rc[0] = (RT *)(red + 32);
    rc[1] = rc[0] + tsz;

    const int nx = sp.nx, ny = sp.ny, nz = sp.nz;
    const int t = blockIdx.x;
    const int nT = tl.PI * tl.PJ;
    const int pi = t / tl.PJ, pj = t - pi * tl.PJ;
    const int i0 = pi * TI, j0 = pj * TJ;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave, nwaves = blockDim.x / kWave;
    const bool has_im = pi > 0, has_ip = pi < tl.PI - 1, has_jm = pj > 0, has_jp = pj < tl.PJ - 1;
    const long long sx = (long long)ny * nzh;
    const float invTJ = 1.0f / (float)TJ;

    // ---- load the tile (both colours) and its source term
    for (int e = tid; e < tsz; e += blockDim.x) {
        const int r = e / nzh, h = e - r * nzh;
        const int li = r / TJ, lj = r - li * TJ;
        const int i = i0 + li, j = j0 + lj;
        double v0 = 0.0, v1 = 0.0;
        RT s0 = RT(0), s1 = RT(0);
        if (i < nx && j < ny) {
            const long long q = ((long long)i * ny + j) * nzh + h;
            v0 = phi[q];
            v1 = phi[sp.Gh + q];
            s0 = rho[q];
            s1 = rho[sp.Gh + q];
        }
        pc[0][e] = v0;
        pc[1][e] = v1;
        rc[0][e] = s0;
        rc[1][e] = s1;
    }
    // initial halo: black face values of the neighbour rows, from memory
    auto load_halo_global = [&](int colour) {
        const double *src = phi + (colour ? sp.Gh : 0);
        for (int e = tid; e < tl.side_i; e += blockDim.x) {
            const int lj = e / nzh, h = e - lj * nzh, j = j0 + lj;
            if (j >= ny) continue;
            if (has_im) h_im[e] = src[((long long)(i0 - 1) * ny + j) * nzh + h];
            if (has_ip) h_ip[e] = src[((long long)(i0 + TI) * ny + j) * nzh + h];
        }
        for (int e = tid; e < tl.side_j; e += blockDim.x) {
            const int li = e / nzh, h = e - li * nzh, i = i0 + li;
            if (i >= nx) continue;
            if (has_jm) h_jm[e] = src[((long long)i * ny + (j0 - 1)) * nzh + h];
            if (has_jp) h_jp[e] = src[((long long)i * ny + (j0 + TJ)) * nzh + h];
        }
    };
    load_halo_global(1);
    __syncthreads();

    // one colour of one iteration over the tile; returns the thread's max
    // |dphi|.  Each wave updates kResRows rows at once (independent update
    // chains in flight together); the division by dx^2 is hv_div_fixed, the
    // correctly rounded quotient.
    auto sweep = [&](int colour) -> double {
        double *dst = pc[colour];
        const double *oth = pc[colour ^ 1];
        const RT *src = rc[colour];
        double m = 0.0;
        for (int r0 = wave; r0 < rows; r0 += nwaves * kResRows) {
            for (int h = lane; h < nzh; h += kWave) {
                double nv[kResRows], po[kResRows];
                int ee[kResRows];
#pragma unroll
                for (int u = 0; u < kResRows; u++) {
                    ee[u] = -1;
                    nv[u] = 0.0;
                    po[u] = 0.0;
                    const int r = r0 + u * nwaves;
                    if (r >= rows) continue;
                    const int li = (int)(((float)r + 0.5f) * invTJ), lj = r - li * TJ;
                    const int i = i0 + li, j = j0 + lj;
                    if (i < 1 || i > nx - 2 || j < 1 || j > ny - 2) continue;
                    const int off = (i + j + colour) & 1;
                    const int k = 2 * h + off;
                    if (k < 1 || k > nz - 2) continue;
                    const int e = r * nzh + h;
                    const double xm = li > 0 ? oth[e - TJ * nzh] : h_im[lj * nzh + h];
                    const double xp = li < TI - 1 ? oth[e + TJ * nzh] : h_ip[lj * nzh + h];
                    const double ym = lj > 0 ? oth[e - nzh] : h_jm[li * nzh + h];
                    const double yp = lj < TJ - 1 ? oth[e + nzh] : h_jp[li * nzh + h];
                    const double zm = oth[e + off - 1], zp = oth[e + off];
                    const double rr = (double)src[e];
                    const double phi_x = hv_div_fixed(xp + xm, cf.dx2, cf.ydx2);
                    const double phi_y = hv_div_fixed(yp + ym, cf.dy2, cf.ydy2);
                    const double phi_z = hv_div_fixed(zp + zm, cf.dz2, cf.ydz2);
                    const double phi_star = (phi_x + phi_y + phi_z + cf.src * rr) * cf.inv_dx2;
                    po[u] = dst[e];
                    nv[u] = cf.one_minus_omega * po[u] + cf.omega * phi_star;
                    ee[u] = e;
                }
#pragma unroll
                for (int u = 0; u < kResRows; u++) {
                    if (ee[u] < 0) continue;
                    dst[ee[u]] = nv[u];
                    const double d = fabs(nv[u] - po[u]);
                    m = d > m ? d : m;
                }
            }
        }
        return m;
    };

    // publish the tile's colour-c face values into its exchange slice
    auto publish = [&](int colour) {
        double *x = (colour ? sy.xb : sy.xr) + (long long)t * tl.xstride;
        const double *v = pc[colour];
        for (int e = tid; e < tl.side_i; e += blockDim.x) {
            const int lj = e / nzh, h = e - lj * nzh;
            st_dev(x + e, v[(0 * TJ + lj) * nzh + h]);
            st_dev(x + tl.side_i + e, v[((TI - 1) * TJ + lj) * nzh + h]);
        }
        for (int e = tid; e < tl.side_j; e += blockDim.x) {
            const int li = e / nzh, h = e - li * nzh;
            st_dev(x + 2 * tl.side_i + e, v[(li * TJ + 0) * nzh + h]);
            st_dev(x + 2 * tl.side_i + tl.side_j + e, v[(li * TJ + TJ - 1) * nzh + h]);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    };

    // copy the neighbours' colour-c faces into the halo
    auto fetch = [&](int colour) {
        const double *xs = colour ? sy.xb : sy.xr;
        for (int e = tid; e < tl.side_i; e += blockDim.x) {
            // plane i0-1 is the i-high face of tile (pi-1, pj); plane i0+TI the i-low face of (pi+1, pj)
            if (has_im) h_im[e] = ld_dev(xs + (long long)(t - tl.PJ) * tl.xstride + tl.side_i + e);
            if (has_ip) h_ip[e] = ld_dev(xs + (long long)(t + tl.PJ) * tl.xstride + e);
        }
        for (int e = tid; e < tl.side_j; e += blockDim.x) {
            if (has_jm) h_jm[e] = ld_dev(xs + (long long)(t - 1) * tl.xstride + 2 * tl.side_i + tl.side_j + e);
            if (has_jp) h_jp[e] = ld_dev(xs + (long long)(t + 1) * tl.xstride + 2 * tl.side_i + e);
        }
        __syncthreads();
    };

    __shared__ int s_stop;
    __shared__ double s_max;
    int iters = max_it;
    double gmax = 0.0;
    for (int it = 0; it < max_it; it++) {
        // red
        HV_TRACE(0);
        double m = sweep(0);
        __syncthreads();
        HV_TRACE(1);
        publish(0);
        HV_TRACE(2);
        if (tid == 0) st_dev_i(sy.flag_red + t, it + 1);
        if (wave == 0) {
            // lanes 0..3 poll the four neighbour flags at once
            int nbq = -1;
            if (lane == 0 && has_im) nbq = t - tl.PJ;
            if (lane == 1 && has_ip) nbq = t + tl.PJ;
            if (lane == 2 && has_jm) nbq = t - 1;
            if (lane == 3 && has_jp) nbq = t + 1;
            int ok = nbq < 0, bad = 0;
            long long spin = 0;
            while (!__all(ok) && !bad) {
                if (!ok) ok = ld_dev_i(sy.flag_red + nbq) >= it + 1;
                if (++spin > kSpinLimit || ld_dev_i(sy.err)) bad = 1;
            }
            if (lane == 0) {
                if (bad) st_dev_i(sy.err, 1);
                s_stop = bad;
            }
        }
        __syncthreads();
        if (s_stop) break;
        HV_TRACE(3);
        fetch(0);
        HV_TRACE(4);
        // black
        double m2 = sweep(1);
        m = m2 > m ? m2 : m;
        m = wave_max_f64(m);
        if (lane == 0) red[wave] = m;
        __syncthreads();
        HV_TRACE(5);
        publish(1);
        HV_TRACE(6);
        if (tid == 0) {
            double bm = 0.0;
            for (int w = 0; w < nwaves; w++) bm = red[w] > bm ? red[w] : bm;
            st_dev(sy.slots + (long long)it * nT + t, bm);
        }
        // wait for every tile's max of this iteration (wave 0; each lane polls
        // up to kSlotPoll slots with all its loads in flight together)
        if (wave == 0) {
            double v[kSlotPoll];
            int got = 0, bad = 0;
            const double *row = sy.slots + (long long)it * nT;
            for (int m = 0; m < kSlotPoll; m++) {
                v[m] = 0.0;
                if (lane + m * kWave >= nT) got |= 1 << m;
            }
            long long spin = 0;
            while (!__all(got == (1 << kSlotPoll) - 1) && !bad) {
#pragma unroll
                for (int m = 0; m < kSlotPoll; m++)
                    if (!(got & (1 << m))) v[m] = ld_dev(row + lane + m * kWave);
#pragma unroll
                for (int m = 0; m < kSlotPoll; m++)
                    if (!(got & (1 << m)) && v[m] >= 0.0) got |= 1 << m;
                if (++spin > kSpinLimit || ld_dev_i(sy.err)) bad = 1;
            }
            double g = 0.0;
#pragma unroll
            for (int m = 0; m < kSlotPoll; m++) g = v[m] > g ? v[m] : g;
            g = wave_max_f64(g);
            bad = __any(bad);
            if (lane == 0) {
                if (bad) st_dev_i(sy.err, 1);
                s_max = g;
                s_stop = bad;
            }
        }
        __syncthreads();
        if (s_stop) break;
        HV_TRACE(7);
        gmax = s_max;
        fetch(1);
        HV_TRACE(8);
        if (gmax < cf.tol) {
            iters = it + 1;
            break;
        }
    }
    // ---- write the tile back
    __syncthreads();
    for (int e = tid; e < tsz; e += blockDim.x) {
        const int r = e / nzh, h = e - r * nzh;
        const int li = r / TJ, lj = r - li * TJ;
        const int i = i0 + li, j = j0 + lj;
        if (i < nx && j < ny) {
            const long long q = ((long long)i * ny + j) * nzh + h;
            phi[q] = pc[0][e];
            phi[sp.Gh + q] = pc[1][e];
        }
    }
    if (t == 0 && tid == 0) {
        sy.result[0] = (double)iters;
        sy.result[1] = gmax;
    }
    (void)sx;
}

}  // namespace hv
