// hv_device.h -- device-side building blocks of the gfx950 step loop: grid
// geometry, the CIC cell/weight rule, the on-the-fly finite-difference force,
// and exact (order-free) integer reductions.
//
// All expressions that produce physics values keep the operation order of
// the reference, so that with -ffp-contract=off the FP64 results equal the
// CPU restatement bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hv_detmath.h"

namespace hv {

constexpr int kWave = 64;         // CDNA wavefront width
constexpr int kBlock = 256;       // threads per block of the streaming kernels

// Grid geometry in the particle precision T (include/sim_structs.hpp:73-76,
// include/physics.hpp:102).
template <typename T>
struct Geom {
    int nx, ny, nz;
    long long G;      // nx*ny*nz
    T dmx, dmy, dmz;  // half-widths d_max
    T dx, dy, dz;     // spacings 2 d_max / n
    T two_dx, two_dy, two_dz;
};

__device__ __forceinline__ long long idx3(int i, int j, int k, int ny, int nz) {
    return ((long long)i * ny + j) * nz + k;
}

// Colour-split storage of the Coulomb potential (and of the SOR source): the
// red nodes ((i+j+k) even) and the black nodes live in two arrays, each row
// (i, j) holding its nodes of one colour at consecutive slots k >> 1.  A red
// sweep then reads the black array and the red array with unit stride.
struct Split {
    int nx, ny, nz, nzh;  // nzh = ceil(nz / 2)
    long long Gh;         // nx * ny * nzh: offset of the black array
};

__host__ __device__ __forceinline__ long long split_index(const Split &s, int i, int j, int k) {
    return (((i + j + k) & 1) ? s.Gh : 0) + ((long long)i * s.ny + j) * s.nzh + (k >> 1);
}

// CIC cell of a particle (include/tools.hpp:50-60 and :108-120).  Returns false
// when the particle is outside [0, n) on any axis or non-finite.
template <typename T>
struct Cell {
    int x0, y0, z0;
    T d_x, d_y, d_z, t_x, t_y, t_z;
};

template <typename T>
__device__ __forceinline__ bool cic_cell(T x, T y, T z, const Geom<T> &g, Cell<T> &c) {
    T cx = (g.nx / T(2.0)) * (x / g.dmx + T(1.0));
    T cy = (g.ny / T(2.0)) * (y / g.dmy + T(1.0));
    T cz = (g.nz / T(2.0)) * (z / g.dmz + T(1.0));
    if (!(cx >= T(0.0) && cy >= T(0.0) && cz >= T(0.0) && cx < (T)g.nx && cy < (T)g.ny && cz < (T)g.nz))
        return false;
    c.x0 = (int)cx;
    c.y0 = (int)cy;
    c.z0 = (int)cz;
    c.d_x = cx - c.x0;
    c.d_y = cy - c.y0;
    c.d_z = cz - c.z0;
    c.t_x = T(1.0) - c.d_x;
    c.t_y = T(1.0) - c.d_y;
    c.t_z = T(1.0) - c.d_z;
    return true;
}

// The eight corner weights in the reference's order (include/tools.hpp:64-94):
// 000, 100, 010, 110, 001, 101, 011, 111; each a left-to-right triple product.
template <typename T>
__device__ __forceinline__ T corner_weight(const Cell<T> &c, int corner) {
    T wx = (corner & 1) ? c.d_x : c.t_x;
    T wy = (corner & 2) ? c.d_y : c.t_y;
    T wz = (corner & 4) ? c.d_z : c.t_z;
    return wx * wy * wz;
}

// F = -grad U at one node (include/physics.hpp:172-195): central differences
// inside, one-sided first-order differences on the faces.
template <typename T, typename UT>
__device__ __forceinline__ void fdm_force(const UT *__restrict__ u, int i, int j, int k,
                                          const Geom<T> &g, T &fx, T &fy, T &fz) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long idx = idx3(i, j, k, ny, nz);
    T gx, gy, gz;
    if (i == 0) gx = ((T)u[idx3(1, j, k, ny, nz)] - (T)u[idx]) / g.dx;
    else if (i == nx - 1) gx = ((T)u[idx] - (T)u[idx3(nx - 2, j, k, ny, nz)]) / g.dx;
    else gx = ((T)u[idx3(i + 1, j, k, ny, nz)] - (T)u[idx3(i - 1, j, k, ny, nz)]) / g.two_dx;
    if (j == 0) gy = ((T)u[idx3(i, 1, k, ny, nz)] - (T)u[idx]) / g.dy;
    else if (j == ny - 1) gy = ((T)u[idx] - (T)u[idx3(i, ny - 2, k, ny, nz)]) / g.dy;
    else gy = ((T)u[idx3(i, j + 1, k, ny, nz)] - (T)u[idx3(i, j - 1, k, ny, nz)]) / g.two_dy;
    if (k == 0) gz = ((T)u[idx3(i, j, 1, ny, nz)] - (T)u[idx]) / g.dz;
    else if (k == nz - 1) gz = ((T)u[idx] - (T)u[idx3(i, j, nz - 2, ny, nz)]) / g.dz;
    else gz = ((T)u[idx3(i, j, k + 1, ny, nz)] - (T)u[idx3(i, j, k - 1, ny, nz)]) / g.two_dz;
    fx = -gx;
    fy = -gy;
    fz = -gz;
}

// CIC gather of the force computed on the fly from U (fuses
// compute_volumetric_forces_fdm + distribute_forces_to_particles_cic,
// include/physics.hpp:160-200 and include/tools.hpp:99-168): same per-node
// values and the same accumulation order as the two-pass form.
template <typename T, typename UT>
__device__ __forceinline__ void gather_force_from_u(const UT *__restrict__ u_species, const Cell<T> &c,
                                                    const Geom<T> &g, T &ax, T &ay, T &az) {
    ax = T(0.0);
    ay = T(0.0);
    az = T(0.0);
#pragma unroll
    for (int corner = 0; corner < 8; corner++) {
        int i = c.x0 + (corner & 1), j = c.y0 + ((corner >> 1) & 1), k = c.z0 + ((corner >> 2) & 1);
        if (i >= g.nx || j >= g.ny || k >= g.nz) continue;
        T w = corner_weight(c, corner);
        T fx, fy, fz;
        fdm_force<T, UT>(u_species, i, j, k, g, fx, fy, fz);
        ax += w * fx;
        ay += w * fy;
        az += w * fz;
    }
}

// ---- exact integer reductions ------------------------------------------------

__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
    for (int s = kWave / 2; s > 0; s >>= 1) v += __shfl_xor(v, s, kWave);
    return v;
}

// Block-wide sum of NV int64 values (exact; any association gives the same
// result).  Result valid in thread 0.  blockDim.x must be a multiple of 64,
// at most 1024.
template <int NV>
__device__ __forceinline__ void block_sum_i64(long long (&v)[NV]) {
    __shared__ long long red[NV][16];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int nwaves = blockDim.x / kWave;
#pragma unroll
    for (int c = 0; c < NV; c++) v[c] = wave_sum_i64(v[c]);
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < NV; c++) red[c][wave] = v[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NV; c++) {
            long long s = 0;
            for (int w = 0; w < nwaves; w++) s += red[c][w];
            v[c] = s;
        }
    }
}

// 128-bit exact accumulation into (lo, hi) words with two commutative atomics:
// the carry out of the low word is detected per add, so the final pair is the
// exact two's-complement sum whatever the order of the adds.
__device__ __forceinline__ void atomic_add_i128(unsigned long long *lo_hi, long long v) {
    unsigned long long lo_v = (unsigned long long)v;
    unsigned long long hi_v = (v < 0) ? ~0ULL : 0ULL;
    unsigned long long old = atomicAdd(&lo_hi[0], lo_v);
    unsigned long long carry = (old + lo_v < old) ? 1ULL : 0ULL;
    unsigned long long hi_add = hi_v + carry;
    if (hi_add) atomicAdd(&lo_hi[1], hi_add);
}

__device__ __forceinline__ double wave_max_f64(double v) {
#pragma unroll
    for (int s = kWave / 2; s > 0; s >>= 1) {
        double o = __shfl_xor(v, s, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// Exact 128-bit accumulation in registers: (lo, hi) += v (sign-extended).
__device__ __forceinline__ void add_i128(unsigned long long &lo, unsigned long long &hi, long long v) {
    unsigned long long nlo = lo + (unsigned long long)v;
    hi += ((v < 0) ? ~0ULL : 0ULL) + (nlo < lo ? 1ULL : 0ULL);
    lo = nlo;
}

// Same-address device-scope atomics from thousands of blocks serialise at the
// coherence point (8 XCDs share it), so block maxima are folded into one of
// kMaxSlots addresses, each on its own 128-byte line; readers take the max of
// the slots.
constexpr int kMaxSlots = 64;
constexpr int kSlotStride = 16;  // unsigned long longs (128 B)

// max over the block of non-negative doubles, folded into *dst (bit pattern of
// a non-negative double orders like the unsigned integer).
__device__ __forceinline__ void block_max_atomic(double v, unsigned long long *dst) {
    __shared__ double red[16];
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int lane = tid & (kWave - 1), wave = tid / kWave;
    const int nwaves = (blockDim.x * blockDim.y * blockDim.z + kWave - 1) / kWave;
    v = wave_max_f64(v);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (tid == 0) {
        double m = 0.0;
        for (int w = 0; w < nwaves; w++) m = red[w] > m ? red[w] : m;
        if (m > 0.0) atomicMax(dst, (unsigned long long)__double_as_longlong(m));
    }
}

}  // namespace hv
