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

// ---- FP32 correction iteration of the mixed form, LDS-resident -----------------
// The mixed form solves L delta = R for the correction of a warm FP64
// potential in FP32 (the same red-black iteration on delta), then finishes with
// FP64 sweeps under the exact stop rule.  The FP32 stage only has to bring
// delta close enough for those sweeps, so it tests convergence with a lag: at
// iteration `it` every tile reads all tiles' max |d delta| of iteration
// it - lag (written `lag` iterations earlier, normally present already) and
// all tiles stop after the same iteration once that max is below tol.
//
// Tiles otherwise synchronise only with their four neighbours, through tagged
// face words: each exchanged value travels as one 64-bit word (iteration tag
// << 32 | float bits), so a reader polls the words themselves and there is no
// separate flag and no store-completion wait on the critical path.  A tile
// overwrites its face slice of a colour only after its neighbours have read
// the previous one (each neighbour's next face, which this tile waits for, is
// published after that read), so one slice per colour suffices and a word's
// tag equals the awaited iteration when it is accepted.
constexpr int kRes32Threads = 1024;
// slots per lane of the FP32 kernel's convergence poll: 4 (at most 256 tiles)
// for 1024-thread blocks, 8 (512 tiles) for the two-per-CU 512-thread blocks
template <int NT>
constexpr int slot_poll32() { return NT == 512 ? 8 : 4; }
constexpr int kFetchMax = 4;  // halo words per thread and colour

struct ResSync32 {
#ifdef HV_RES_TRACE
    unsigned long long *trace;
#endif
    unsigned long long *xr, *xb;  // tagged exchange words (red, black): nT * xstride, tag 0 before the launch
    float *slots;                 // max_it * nT per-tile max |d delta|, preset to -1
    int *err;                     // set on a spin timeout
    float *result;                // [0] iterations done, [1] max |d delta| of the tested iteration
};

__device__ __forceinline__ float ld_dev_f(const float *p) {
    return __int_as_float(__hip_atomic_load((const int *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__device__ __forceinline__ void st_dev_f(float *p, float v) {
    __hip_atomic_store((int *)p, __float_as_int(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_tagged(unsigned long long *p, float v, unsigned tag) {
    __hip_atomic_store(p, ((unsigned long long)tag << 32) | (unsigned long long)__float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long ld_tagged(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// CTI, CTJ, CNZH: the tile shape at compile time (0: taken from `tl` at run
// time).  With a fixed shape every tile index folds into constants and LDS
// immediate offsets; the host launches a fixed-shape instance when the chosen
// tiling matches one.
//
// res32_solve_tile is the whole solve as seen by the block that owns tile t;
// k_sor_resident_f32 runs it with one block per tile, the ensemble kernel
// runs it once per queued (instance, tile) job.  A block that leaves it has
// written its tile of delta back; the LDS is free for the next job after a
// barrier.
template <int CTI, int CTJ, int CNZH, int NT = kRes32Threads>
__device__ __forceinline__ void res32_solve_tile(float *__restrict__ delta, const float *__restrict__ R, Split sp,
                                                 ResTiles tl, SorCoef cf, int max_it, int lag, ResSync32 sy,
                                                 const int t) {
    extern __shared__ unsigned char smem[];
    const int TI = CTI ? CTI : tl.TI, TJ = CTJ ? CTJ : tl.TJ, nzh = CNZH ? CNZH : tl.nzh;
    const int side_i = TJ * nzh, side_j = TI * nzh, xstride = 2 * side_i + 2 * side_j;
    constexpr int nwaves = NT / kWave;
    // rows a wave updates at once (fewer for the large fixed tile: register budget of 1024-thread blocks)
    constexpr int RR = (CTI * CTJ * CNZH >= 4096) ? 2 : kResRows;
    // rows of fewer (or other than a multiple of 64) k slots: lanes take
    // consecutive (row, slot) elements of the row set, so no lane idles
    constexpr bool FLAT = CNZH == 0 || CNZH % kWave != 0;
    constexpr int RF = 2;  // elements a lane updates at once in the FLAT sweep
    const int rows = TI * TJ, tsz = rows * nzh;
    float *pc[2], *rc[2];
    pc[0] = (float *)smem;
    pc[1] = pc[0] + tsz;
    rc[0] = pc[1] + tsz;
    rc[1] = rc[0] + tsz;
    // halo of other-colour face values, in exchange-slice order: im, ip (TJ*nzh each), jm, jp (TI*nzh each)
    float *halo = rc[1] + tsz;
    float *h_im = halo, *h_ip = halo + side_i, *h_jm = halo + 2 * side_i, *h_jp = h_jm + side_j;
    float *red = halo + xstride;  // wave partials

    const int nx = sp.nx, ny = sp.ny, nz = sp.nz;
    const int nT = tl.PI * tl.PJ;
    const int pi = t / tl.PJ, pj = t - pi * tl.PJ;
    const int i0 = pi * TI, j0 = pj * TJ;
    const int tid = threadIdx.x, lane = tid & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);  // wave-uniform: row indexes and branches stay scalar
    const bool has_im = pi > 0, has_ip = pi < tl.PI - 1, has_jm = pj > 0, has_jp = pj < tl.PJ - 1;
    const float yx = (float)cf.ydx2, yy = (float)cf.ydy2, yz = (float)cf.ydz2, inv = (float)cf.inv_dx2;
    const float om = (float)cf.omega, om1 = (float)cf.one_minus_omega, tol = (float)cf.tol;
    __shared__ int s_bad;
    __shared__ float s_max[2];  // by iteration parity: written in iteration it, read at its end
    if (tid == 0) s_bad = 0;

    for (int e = tid; e < tsz; e += NT) {
        const int r = e / nzh, h = e - r * nzh;
        const int li = r / TJ, lj = r - li * TJ;
        const int i = i0 + li, j = j0 + lj;
        float v0 = 0.0f, v1 = 0.0f, s0 = 0.0f, s1 = 0.0f;
        if (i < nx && j < ny) {
            const long long q = ((long long)i * ny + j) * nzh + h;
            v0 = delta[q];
            v1 = delta[sp.Gh + q];
            s0 = R[q];
            s1 = R[sp.Gh + q];
        }
        pc[0][e] = v0;
        pc[1][e] = v1;
        rc[0][e] = s0;
        rc[1][e] = s1;
    }
    {
        const float *src = delta + sp.Gh;  // black values around the tile for the first red sweep
        for (int e = tid; e < side_i; e += NT) {
            const int lj = e / nzh, h = e - lj * nzh, j = j0 + lj;
            if (j >= ny) continue;
            if (has_im) h_im[e] = src[((long long)(i0 - 1) * ny + j) * nzh + h];
            if (has_ip) h_ip[e] = src[((long long)(i0 + TI) * ny + j) * nzh + h];
        }
        for (int e = tid; e < side_j; e += NT) {
            const int li = e / nzh, h = e - li * nzh, i = i0 + li;
            if (i >= nx) continue;
            if (has_jm) h_jm[e] = src[((long long)i * ny + (j0 - 1)) * nzh + h];
            if (has_jp) h_jp[e] = src[((long long)i * ny + (j0 + TJ)) * nzh + h];
        }
    }
    // this thread's halo words: exchange offset in the neighbour's slice and
    // halo index (at most kFetchMax; the host checks xstride fits)
    // (offsets fit in int: nT * xstride <= 256 * kFetchMax * NT)
    int src_off[kFetchMax];
    int nfetch = 0;
    for (int x = tid; x < xstride && nfetch < kFetchMax; x += NT) {
        int off = -1;
        if (x < side_i) {
            if (has_im) off = (t - tl.PJ) * xstride + side_i + x;  // i-high face of (pi-1, pj)
        } else if (x < 2 * side_i) {
            if (has_ip) off = (t + tl.PJ) * xstride + (x - side_i);  // i-low face of (pi+1, pj)
        } else if (x < 2 * side_i + side_j) {
            if (has_jm) off = (t - 1) * xstride + 2 * side_i + side_j + (x - 2 * side_i);
        } else {
            if (has_jp) off = (t + 1) * xstride + 2 * side_i + (x - 2 * side_i - side_j);
        }
        src_off[nfetch++] = off;
    }
    __syncthreads();

    // Row sets: the ring of rows on the tile's faces (the only rows that read
    // halo values) and the interior rows.  Tiles thinner than 3 rows in i or
    // j have no interior; their ring is every row.
    const bool split_rows = TI >= 3 && TJ >= 3;
    const int n_inner = split_rows ? (TI - 2) * (TJ - 2) : 0;
    const int n_ring = rows - n_inner;
    auto row_at = [&](bool inner, int q, int &li, int &lj) {
        if (inner) {
            li = 1 + q / (TJ - 2);
            lj = 1 + q - (li - 1) * (TJ - 2);
        } else if (!split_rows) {
            li = q / TJ;
            lj = q - li * TJ;
        } else if (q < TJ) {
            li = 0;
            lj = q;
        } else if (q < 2 * TJ) {
            li = TI - 1;
            lj = q - TJ;
        } else {
            li = 1 + ((q - 2 * TJ) >> 1);
            lj = ((q - 2 * TJ) & 1) ? TJ - 1 : 0;
        }
    };

    // the update of k_sor_colour_f32, on LDS operands, over one row set.
    // Branch-free over the RR rows a wave takes at once: every lane loads its
    // operands (indexes clamped into the tile; values of nodes that are not
    // updated are discarded), so all rows' LDS reads are in flight together,
    // and only the store and the max are predicated.
    // q / d for q < 2^22 with a precomputed 1 / d (exact: the fraction part of
    // (q + 0.5) / d stays 0.5 / d away from an integer)
    const float inv_nzh = 1.0f / (float)nzh, inv_tj = 1.0f / (float)TJ, inv_tj2 = 1.0f / (float)(TJ - 2);
    auto fdiv = [](int q, int d, float inv, int cd) -> int {
        if (cd) return q / cd;
        (void)d;
        return (int)(((float)q + 0.5f) * inv);
    };
    auto row_at_lane = [&](bool inner, int q, int &li, int &lj) {
        if (inner) {
            li = 1 + fdiv(q, TJ - 2, inv_tj2, CTJ ? CTJ - 2 : 0);
            lj = 1 + q - (li - 1) * (TJ - 2);
        } else if (!split_rows) {
            li = fdiv(q, TJ, inv_tj, CTJ);
            lj = q - li * TJ;
        } else {
            const int w = q - 2 * TJ;
            li = q < TJ ? 0 : (q < 2 * TJ ? TI - 1 : 1 + (w >> 1));
            lj = q < TJ ? q : (q < 2 * TJ ? q - TJ : ((w & 1) ? TJ - 1 : 0));
        }
    };
    // the same update over (row, slot) elements: element f of the row set is
    // slot f % nzh of its row f / nzh
    auto sweep_flat = [&](int colour, bool inner, float m) -> float {
        float *dst = pc[colour];
        const float *oth = pc[colour ^ 1];
        const float *src = rc[colour];
        const int nel = (inner ? n_inner : n_ring) * nzh;
        for (int f0 = wave * kWave; f0 < nel; f0 += nwaves * kWave * RF) {
            float nv[RF], po[RF];
            int ee[RF];
            bool ok[RF];
#pragma unroll
            for (int u = 0; u < RF; u++) {
                int f = f0 + u * nwaves * kWave + lane;
                const bool fv = f < nel;
                f = fv ? f : 0;
                const int q = fdiv(f, nzh, inv_nzh, CNZH), h = f - q * nzh;
                int li, lj;
                row_at_lane(inner, q, li, lj);
                const int r = li * TJ + lj;
                const int i = i0 + li, j = j0 + lj;
                const int off = (i + j + colour) & 1;
                const int k = 2 * h + off;
                ok[u] = fv && i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2;
                const int e = r * nzh + h;
                const int ez = e + off - 1 > 0 ? e + off - 1 : 0;
                float xm, xp, ym, yp;
                if (inner) {
                    xm = oth[e - TJ * nzh];
                    xp = oth[e + TJ * nzh];
                    ym = oth[e - nzh];
                    yp = oth[e + nzh];
                } else {
                    xm = li > 0 ? oth[e - TJ * nzh] : h_im[lj * nzh + h];
                    xp = li < TI - 1 ? oth[e + TJ * nzh] : h_ip[lj * nzh + h];
                    ym = lj > 0 ? oth[e - nzh] : h_jm[li * nzh + h];
                    yp = lj < TJ - 1 ? oth[e + nzh] : h_jp[li * nzh + h];
                }
                const float zm = oth[ez], zp = oth[e + off];
                const float sv = sor_f32_sum(xm, xp, ym, yp, zm, zp, src[e], yx, yy, yz);
                po[u] = dst[e];
                nv[u] = sor_f32_relax(po[u], sv, inv, om, om1);
                ee[u] = e;
            }
#pragma unroll
            for (int u = 0; u < RF; u++) {
                if (!ok[u]) continue;
                dst[ee[u]] = nv[u];
                const float d = fabsf(nv[u] - po[u]);
                m = d > m ? d : m;
            }
        }
        return m;
    };

    auto sweep = [&](int colour, bool inner, float m) -> float {
        if constexpr (FLAT) return sweep_flat(colour, inner, m);
        float *dst = pc[colour];
        const float *oth = pc[colour ^ 1];
        const float *src = rc[colour];
        const int nrows = inner ? n_inner : n_ring;
        for (int q0 = wave; q0 < nrows; q0 += nwaves * RR) {
            for (int h = lane; h < nzh; h += kWave) {
                float nv[RR], po[RR];
                int ee[RR];
                bool ok[RR];
#pragma unroll
                for (int u = 0; u < RR; u++) {
                    int q = q0 + u * nwaves;
                    const bool rv = q < nrows;
                    q = rv ? q : q0;
                    int li, lj;
                    row_at(inner, q, li, lj);
                    const int r = li * TJ + lj;
                    const int i = i0 + li, j = j0 + lj;
                    const int off = (i + j + colour) & 1;
                    const int k = 2 * h + off;
                    ok[u] = rv && i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2;
                    const int e = r * nzh + h;
                    const int ez = e + off - 1 > 0 ? e + off - 1 : 0;
                    float xm, xp, ym, yp;
                    if (inner) {
                        xm = oth[e - TJ * nzh];
                        xp = oth[e + TJ * nzh];
                        ym = oth[e - nzh];
                        yp = oth[e + nzh];
                    } else {
                        xm = li > 0 ? oth[e - TJ * nzh] : h_im[lj * nzh + h];
                        xp = li < TI - 1 ? oth[e + TJ * nzh] : h_ip[lj * nzh + h];
                        ym = lj > 0 ? oth[e - nzh] : h_jm[li * nzh + h];
                        yp = lj < TJ - 1 ? oth[e + nzh] : h_jp[li * nzh + h];
                    }
                    const float zm = oth[ez], zp = oth[e + off];
                    const float s = sor_f32_sum(xm, xp, ym, yp, zm, zp, src[e], yx, yy, yz);
                    po[u] = dst[e];
                    nv[u] = sor_f32_relax(po[u], s, inv, om, om1);
                    ee[u] = e;
                }
#pragma unroll
                for (int u = 0; u < RR; u++) {
                    if (!ok[u]) continue;
                    dst[ee[u]] = nv[u];
                    const float d = fabsf(nv[u] - po[u]);
                    m = d > m ? d : m;
                }
            }
        }
        return m;
    };

    // store the tile's colour faces as tagged words (no completion wait: each
    // word carries its own tag)
    auto publish = [&](int colour, unsigned tag) {
        unsigned long long *x = (colour ? sy.xb : sy.xr) + (long long)t * xstride;
        const float *v = pc[colour];
        for (int e = tid; e < side_i; e += NT) {
            const int lj = e / nzh, h = e - lj * nzh;
            st_tagged(x + e, v[(0 * TJ + lj) * nzh + h], tag);
            st_tagged(x + side_i + e, v[((TI - 1) * TJ + lj) * nzh + h], tag);
        }
        for (int e = tid; e < side_j; e += NT) {
            const int li = e / nzh, h = e - li * nzh;
            st_tagged(x + 2 * side_i + e, v[(li * TJ + 0) * nzh + h], tag);
            st_tagged(x + 2 * side_i + side_j + e, v[(li * TJ + TJ - 1) * nzh + h], tag);
        }
    };

    // poll this thread's halo words until they carry `tag`; all loads of a
    // round are in flight together.  Returns nonzero (block-uniform) on error.
    auto fetch = [&](int colour, unsigned tag) -> int {
        const unsigned long long *xs = colour ? sy.xb : sy.xr;
        unsigned pending = 0;
#pragma unroll
        for (int q = 0; q < kFetchMax; q++)
            if (q < nfetch && src_off[q] >= 0) pending |= 1u << q;
        long long spin = 0;
        int bad = 0;
        while (pending && !bad) {
            unsigned long long w[kFetchMax];
#pragma unroll
            for (int q = 0; q < kFetchMax; q++)
                if (pending & (1u << q)) w[q] = ld_tagged(xs + src_off[q]);
#pragma unroll
            for (int q = 0; q < kFetchMax; q++) {
                if ((pending & (1u << q)) && (unsigned)(w[q] >> 32) >= tag) {
                    halo[tid + q * NT] = __uint_as_float((unsigned)w[q]);
                    pending &= ~(1u << q);
                }
            }
            if (++spin > kSpinLimit) bad = 1;
            else if ((spin & 255) == 0 && ld_dev_i(sy.err)) bad = 1;
            // a wave still waiting yields its issue slots (about 64 clocks)
            // to the waves of its SIMD that have work
            if (pending) __builtin_amdgcn_s_sleep(1);
        }
        if (bad) {
            st_dev_i(sy.err, 1);
            s_bad = 1;
        }
        __syncthreads();
        return s_bad;
    };

#ifdef HV_RES_TRACE
    // timestamps go to LDS during the loop (a global store would make the
    // next barrier wait for it) and to memory at the end
    __shared__ unsigned long long s_trace[256];
#define HV_TRACE32(slot)                                                                                   \
    do {                                                                                                   \
        if (t == 0 && tid == 0 && it >= 64 && it < 80) s_trace[(it - 64) * 16 + (slot)] = wall_clock64(); \
    } while (0)
#else
#define HV_TRACE32(slot) \
    do {                  \
    } while (0)
#endif
    // wave 0: max over all tiles of the write-once slot row `row_it`
    auto poll_slots = [&](int row_it, int parity) {
        constexpr int kSlotPoll32 = slot_poll32<NT>();
        float v[kSlotPoll32];
        int got = 0, bad = 0;
        const float *row = sy.slots + (long long)row_it * nT;
#pragma unroll
        for (int q = 0; q < kSlotPoll32; q++) {
            v[q] = 0.0f;
            if (lane + q * kWave >= nT) got |= 1 << q;
        }
        long long spin = 0;
        while (!__all(got == (1 << kSlotPoll32) - 1) && !bad) {
#pragma unroll
            for (int q = 0; q < kSlotPoll32; q++)
                if (!(got & (1 << q))) v[q] = ld_dev_f(row + lane + q * kWave);
#pragma unroll
            for (int q = 0; q < kSlotPoll32; q++)
                if (!(got & (1 << q)) && v[q] >= 0.0f) got |= 1 << q;
            if (++spin > kSpinLimit || ((spin & 255) == 0 && ld_dev_i(sy.err))) bad = 1;
            if (got != (1 << kSlotPoll32) - 1) __builtin_amdgcn_s_sleep(1);
        }
        float g = 0.0f;
#pragma unroll
        for (int q = 0; q < kSlotPoll32; q++) g = v[q] > g ? v[q] : g;
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const float w = __shfl_xor(g, o);
            g = w > g ? w : g;
        }
        bad = __any(bad);
        if (lane == 0) {
            if (bad) {
                st_dev_i(sy.err, 1);
                s_bad = 1;
            }
            s_max[parity] = g;
        }
    };

    // Iteration it:
    //   red interior -> [black halo of it-1] -> red ring -> publish red
    //   -> black interior -> [red halo of it] -> black ring -> publish black
    // The interior sweeps need no halo, so they run while the neighbours'
    // faces are in flight.  A tile publishes a colour's faces for iteration it
    // only after reading its neighbours' faces that were published after they
    // read its faces of iteration it-1, so one exchange slice per colour holds.
    int iters = max_it;
    float tested = -1.0f;
    for (int it = 0; it < max_it; it++) {
        const unsigned tag = (unsigned)it + 1u;
        HV_TRACE32(0);
        float m = sweep(0, true, 0.0f);
        HV_TRACE32(1);
        // lagged convergence test over all tiles (wave 0; acted on at the end
        // of the iteration): every tile reads the same write-once row, so all
        // stop after the same iteration.  Row it - lag (lag >= 1) is complete
        // once every tile has published black faces of it - lag, which waits
        // on nothing this tile has not published yet.
        if (it >= lag && wave == 0) poll_slots(it - lag, it & 1);
        if (it > 0) {
            if (fetch(1, tag - 1u)) break;
        } else {
            __syncthreads();
        }
        HV_TRACE32(2);
        m = sweep(0, false, m);
        __syncthreads();
        HV_TRACE32(3);
        publish(0, tag);
        m = sweep(1, true, m);
        HV_TRACE32(4);
        if (fetch(0, tag)) break;
        HV_TRACE32(5);
        m = sweep(1, false, m);
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const float w = __shfl_xor(m, o);
            m = w > m ? w : m;
        }
        if (lane == 0) red[wave] = m;
        __syncthreads();
        HV_TRACE32(6);
        publish(1, tag);
        if (tid == 0) {
            float bm = 0.0f;
            for (int w = 0; w < nwaves; w++) bm = red[w] > bm ? red[w] : bm;
            st_dev_f(sy.slots + (long long)it * nT + t, bm);
        }
        HV_TRACE32(7);
        if (it >= lag) {
            tested = s_max[it & 1];
            if (tested < tol) {
                iters = it + 1;
                break;
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < tsz; e += NT) {
        const int r = e / nzh, h = e - r * nzh;
        const int li = r / TJ, lj = r - li * TJ;
        const int i = i0 + li, j = j0 + lj;
        if (i < nx && j < ny) {
            const long long q = ((long long)i * ny + j) * nzh + h;
            delta[q] = pc[0][e];
            delta[sp.Gh + q] = pc[1][e];
        }
    }
    if (t == 0 && tid == 0) {
        sy.result[0] = (float)iters;
        sy.result[1] = tested;
    }
#ifdef HV_RES_TRACE
    if (t == 0 && sy.trace)
        for (int q = tid; q < 256; q += NT) sy.trace[q] = s_trace[q];
#endif
}

// NT threads per block: 1024 with one tile per CU, 512 with two (a tile's
// halo wait then overlaps the other tile's sweeps)
template <int CTI, int CTJ, int CNZH, int NT = kRes32Threads>
__global__ void __launch_bounds__(NT, 4)  // 4 waves per SIMD: two 512-thread blocks per CU
k_sor_resident_f32(float *__restrict__ delta, const float *__restrict__ R, Split sp, ResTiles tl, SorCoef cf,
                   int max_it, int lag, ResSync32 sy) {
    res32_solve_tile<CTI, CTJ, CNZH, NT>(delta, R, sp, tl, cf, max_it, lag, sy, (int)blockIdx.x);
}

}  // namespace hv
