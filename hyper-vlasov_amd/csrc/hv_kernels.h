// hv_kernels.h -- gfx950 kernels of the hyper-vlasov step loop.
//
// Precision templates: T is the particle / grid precision (double for the
// plain FP64 path, float for the mixed path); the Coulomb potential, the SOR
// update and every diagnostics reduction stay in double.
#pragma once

#include <type_traits>
#include <utility>

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

// Block partial sums of the six diagnostics, written (not accumulated) to
// partials[6 * bid + c] (bid: the block index within the pass); k_stats_finalize
// folds them into a StatSlot.
// Each block value is exact: |v| <= 256 * 2^54 < 2^63.
__device__ __forceinline__ void stat_commit(long long (&v)[6], long long *partials, int bid) {
    block_sum_i64<6>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < 6; c++) partials[6 * bid + c] = v[c];
    }
}

// One block of kStatThreads: exact sum of nblocks rows of partials into *slot
// (128-bit sums; integer addition, so the order does not matter: each thread
// sums a strided share, then a tree over the threads' sums).
constexpr int kStatThreads = 256;
__device__ __forceinline__ void stats_finalize_block(const long long *__restrict__ partials, int nblocks,
                                                     StatSlot *__restrict__ slot) {
    __shared__ unsigned long long lo[6][kStatThreads], hi[6][kStatThreads];
    const int tid = threadIdx.x;
    unsigned long long l[6] = {0, 0, 0, 0, 0, 0}, h[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 4
    for (int b = tid; b < nblocks; b += kStatThreads)
#pragma unroll
        for (int c = 0; c < 6; c++) add_i128(l[c], h[c], partials[6 * b + c]);
#pragma unroll
    for (int c = 0; c < 6; c++) {
        lo[c][tid] = l[c];
        hi[c][tid] = h[c];
    }
    __syncthreads();
    for (int s = kStatThreads / 2; s > 0; s >>= 1) {
        if (tid < s)
#pragma unroll
            for (int c = 0; c < 6; c++) {
                const unsigned long long a = lo[c][tid], nl = a + lo[c][tid + s];
                hi[c][tid] += hi[c][tid + s] + (nl < a ? 1ULL : 0ULL);
                lo[c][tid] = nl;
            }
        __syncthreads();
    }
    if (tid < 6) {
        const int c = tid;
        const unsigned long long L = lo[c][0], H = hi[c][0];
        if (c < 4) {
            slot->s[c][0] = L;
            slot->s[c][1] = H;
        } else {
            slot->inbox[c - 4] = L;
        }
    }
}

__global__ void __launch_bounds__(kStatThreads)
k_stats_finalize(const long long *__restrict__ partials, int nblocks, StatSlot *__restrict__ slot) {
    stats_finalize_block(partials, nblocks, slot);
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

// ---- count box ------------------------------------------------------------------
// The node ranges that a deposit may have touched, kept in unsigned words after
// the 2G fixed-point counts.  Per species kCboxWords = 6 words: 0xffff - the
// least i, j, k, then the greatest i, j, k + 1; all zero for an empty box.  The
// box bounds every non-zero count (it may be larger than the tight box, never
// smaller).  Grids are at most 65535 nodes along an axis.
constexpr int kCboxWords = 6;
__host__ __device__ inline int cbox_species_words(int) { return kCboxWords; }
// u64 words after the 2G counts that hold both species' boxes
__host__ __device__ inline long long cbox_acc_words(int) { return kCboxWords; }
__device__ __forceinline__ int cbox_lo(unsigned w) { return 0xffff - (int)w; }  // empty: 65535
__device__ __forceinline__ int cbox_hi(unsigned w) { return (int)w - 1; }        // empty: -1

// A deposit writes one record per block, [s0, s1, x0, x1, y0, y1, z0, z1]:
// the species s0..s1 and the node box x0..x1, y0..y1, z0..z1 its particles'
// cells touch (x1 < x0: none); k_cbox_build folds the records into the box
// words before the smearing, rewriting them.
constexpr int kBoxRec = 8;

// One block of kCboxThreads per species: the box of every record of it, four
// record loads in flight per thread.
constexpr int kCboxThreads = 1024;

__device__ __forceinline__ void cbox_build_block(const int *__restrict__ brec, int nrec, unsigned *__restrict__ cbox,
                                                 const int sp) {
    constexpr int kBig = 0x7fffffff;
    int lo[3] = {kBig, kBig, kBig}, hi[3] = {-1, -1, -1};
    for (int r0 = threadIdx.x; r0 < nrec; r0 += 4 * kCboxThreads) {
        int4 ra[4], rb[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int r = r0 + u * kCboxThreads;
            ra[u] = r < nrec ? *(const int4 *)(brec + kBoxRec * r) : int4{0, -1, 0, -1};
            rb[u] = r < nrec ? *(const int4 *)(brec + kBoxRec * r + 4) : int4{0, -1, 0, -1};
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int4 a = ra[u], c = rb[u];
            if (a.w < a.z || sp < a.x || sp > a.y) continue;
            lo[0] = min(lo[0], a.z);
            hi[0] = max(hi[0], a.w);
            lo[1] = min(lo[1], c.x);
            hi[1] = max(hi[1], c.y);
            lo[2] = min(lo[2], c.z);
            hi[2] = max(hi[2], c.w);
        }
    }
    __shared__ int red[6][kCboxThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        int l = lo[q], h = hi[q];
#pragma unroll
        for (int s = kWave / 2; s > 0; s >>= 1) {
            l = min(l, __shfl_xor(l, s, kWave));
            h = max(h, __shfl_xor(h, s, kWave));
        }
        if (lane == 0) {
            red[q][wave] = l;
            red[3 + q][wave] = h;
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int q = threadIdx.x;
        int v = red[q][0];
        for (int w = 1; w < kCboxThreads / kWave; w++) v = q < 3 ? min(v, red[q][w]) : max(v, red[q][w]);
        red[q][0] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int q = threadIdx.x;
        const bool empty = red[3][0] < red[0][0];
        cbox[sp * kCboxWords + q] = empty ? 0u : q < 3 ? 0xffffu - (unsigned)red[q][0] : (unsigned)red[q][0] + 1u;
    }
}

__global__ void __launch_bounds__(kCboxThreads)
k_cbox_build(const int *__restrict__ brec, int nrec, unsigned *__restrict__ cbox) {
    cbox_build_block(brec, nrec, cbox, blockIdx.x);
}

// stats_finalize_block's sums by a block of kCboxThreads: each thread sums a
// strided share, the lanes of a wave are summed by shuffles and the waves'
// sums in LDS (128-bit integer sums: the same values in any order).
__device__ __forceinline__ void stats_finalize_wide(const long long *__restrict__ partials, int nblocks,
                                                    StatSlot *__restrict__ slot) {
    __shared__ unsigned long long wlo[6][kCboxThreads / kWave], whi[6][kCboxThreads / kWave];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    unsigned long long l[6] = {0, 0, 0, 0, 0, 0}, h[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 4
    for (int b = tid; b < nblocks; b += kCboxThreads)
#pragma unroll
        for (int c = 0; c < 6; c++) add_i128(l[c], h[c], partials[6 * b + c]);
#pragma unroll
    for (int c = 0; c < 6; c++) {
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) {
            const unsigned long long ol = __shfl_xor(l[c], d, kWave), oh = __shfl_xor(h[c], d, kWave);
            const unsigned long long nl = l[c] + ol;
            h[c] += oh + (nl < ol ? 1ULL : 0ULL);
            l[c] = nl;
        }
        if (lane == 0) {
            wlo[c][wave] = l[c];
            whi[c][wave] = h[c];
        }
    }
    __syncthreads();
    if (tid < 6) {
        const int c = tid;
        unsigned long long L = 0, H = 0;
        for (int w = 0; w < kCboxThreads / kWave; w++) {
            const unsigned long long nl = L + wlo[c][w];
            H += whi[c][w] + (nl < L ? 1ULL : 0ULL);
            L = nl;
        }
        if (c < 4) {
            slot->s[c][0] = L;
            slot->s[c][1] = H;
        } else {
            slot->inbox[c - 4] = L;
        }
    }
}

// After a particle pass: the count box of its deposit (blocks 0, 1, as
// k_cbox_build) and its diagnostics (block 2, as k_stats_finalize), in one
// launch.
__global__ void __launch_bounds__(kCboxThreads)
k_stats_cbox(const long long *__restrict__ partials, int nblocks, StatSlot *__restrict__ slot,
             const int *__restrict__ brec, unsigned *__restrict__ cbox) {
    if (blockIdx.x < 2)
        cbox_build_block(brec, nblocks, cbox, blockIdx.x);
    else
        stats_finalize_wide(partials, nblocks, slot);
}

// ---- block-local CIC deposit ----------------------------------------------------
// When the particles of a block are sorted by cell (see sort_key_of), their
// cells fill a small box; the block then accumulates the box of nodes in LDS
// (exact int64 adds) and flushes only the non-zero nodes to HBM, one global
// atomic per node instead of eight per particle.  Blocks whose box is too big
// or that hold both species fall back to per-particle global atomics.  All
// adds are integer adds, so the result is the same bits either way.
constexpr int kBoxCap = 1536;  // nodes (12 KB of int64)

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int s = kWave / 2; s > 0; s >>= 1) v = min(v, __shfl_xor(v, s, kWave));
    return v;
}

__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int s = kWave / 2; s > 0; s >>= 1) v = max(v, __shfl_xor(v, s, kWave));
    return v;
}

template <typename T>
__device__ __forceinline__ void block_deposit(bool has, const Cell<T> &c, int species, const Geom<T> &g,
                                              unsigned long long *__restrict__ acc, int *__restrict__ brec,
                                              int bid) {
    __shared__ unsigned long long box[kBoxCap];
    __shared__ int red[8][16];
    __shared__ int bb[8];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
    int v[8];
    v[0] = has ? c.x0 : 0x7fffffff;
    v[1] = has ? c.y0 : 0x7fffffff;
    v[2] = has ? c.z0 : 0x7fffffff;
    v[3] = has ? species : 0x7fffffff;
    v[4] = has ? c.x0 : -1;
    v[5] = has ? c.y0 : -1;
    v[6] = has ? c.z0 : -1;
    v[7] = has ? species : -1;
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = wave_min_i(v[q]);
#pragma unroll
    for (int q = 4; q < 8; q++) v[q] = wave_max_i(v[q]);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 8; q++) red[q][wave] = v[q];
    __syncthreads();
    if (threadIdx.x < 8) {
        const int q = threadIdx.x;
        int r = red[q][0];
        for (int w = 1; w < nw; w++) r = (q < 4) ? min(r, red[q][w]) : max(r, red[q][w]);
        bb[q] = r;
    }
    __syncthreads();
    const int x0 = bb[0], y0 = bb[1], z0 = bb[2];
    const int ex = bb[4] - x0 + 2, ey = bb[5] - y0 + 2, ez = bb[6] - z0 + 2;
    // the block's record: its species and the nodes of its cells (none: x1 < x0)
    if (brec && threadIdx.x == 0) {
        int4 *r = (int4 *)(brec + kBoxRec * bid);
        if (bb[4] < 0) {
            r[0] = int4{0, -1, 0, -1};
            r[1] = int4{0, -1, 0, -1};
        } else {
            r[0] = int4{bb[3], bb[7], x0, min(bb[4] + 1, g.nx - 1)};
            r[1] = int4{y0, min(bb[5] + 1, g.ny - 1), z0, min(bb[6] + 1, g.nz - 1)};
        }
    }
    if (bb[4] < 0) return;  // no particle of this block is inside the grid
    const long long off = bb[3] ? g.G : 0;
    const bool lds = (bb[3] == bb[7]) && (long long)ex * ey * ez <= kBoxCap;
    if (!lds) {
        if (has) {
#pragma unroll
            for (int corner = 0; corner < 8; corner++) {
                int i = c.x0 + (corner & 1), j = c.y0 + ((corner >> 1) & 1), k = c.z0 + ((corner >> 2) & 1);
                if (i >= g.nx || j >= g.ny || k >= g.nz) continue;
                long long q = hv_fix_weight((double)corner_weight(c, corner));
                atomicAdd(&acc[(species ? g.G : 0) + idx3(i, j, k, g.ny, g.nz)], (unsigned long long)q);
            }
        }
        return;
    }
    const int vol = ex * ey * ez;
    for (int e = threadIdx.x; e < vol; e += blockDim.x) box[e] = 0ULL;
    __syncthreads();
    if (has) {
#pragma unroll
        for (int corner = 0; corner < 8; corner++) {
            int i = c.x0 + (corner & 1), j = c.y0 + ((corner >> 1) & 1), k = c.z0 + ((corner >> 2) & 1);
            if (i >= g.nx || j >= g.ny || k >= g.nz) continue;
            long long q = hv_fix_weight((double)corner_weight(c, corner));
            atomicAdd(&box[((i - x0) * ey + (j - y0)) * ez + (k - z0)], (unsigned long long)q);
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < vol; e += blockDim.x) {
        const unsigned long long val = box[e];
        if (!val) continue;
        const int i = x0 + e / (ey * ez), j = y0 + (e / ez) % ey, k = z0 + e % ez;
        if (i < g.nx && j < g.ny && k < g.nz) atomicAdd(&acc[off + idx3(i, j, k, g.ny, g.nz)], val);
    }
}

// ---- particle sort by (species, brick) ------------------------------------------
// Key = species * (NB + 1) + brick of the particle's cell (bricks of
// 4 x 4 x 4 cells, row-major), NB + 1 for particles outside the grid; protons
// keep [0, np) because the species is the leading part of the key.
constexpr int kBrickShift = 2;

// Sort key of particle i: its species (protons first), then the 4^3-node
// brick of its cell; particles outside the grid come last within their
// species, spread over kSortOutside keys by the wave of their slot (they all
// share one place in the order: one key would make every count atomic for
// them hit one word, and a key per slot would leave each lane an atomic of its
// own); keys lie in [0, 2 (bricks + kSortOutside)).
constexpr int kSortOutside = 64;

template <typename T>
__device__ __forceinline__ unsigned sort_key_of(const Particles<T> &p, int i, int np, const Geom<T> &g) {
    const int bx = (g.nx + 3) >> kBrickShift, by = (g.ny + 3) >> kBrickShift, bz = (g.nz + 3) >> kBrickShift;
    const unsigned nb = (unsigned)(bx * by * bz);
    Cell<T> c;
    unsigned b = nb + (((unsigned)i / kWave) & (kSortOutside - 1));
    if (cic_cell(p.x[i], p.y[i], p.z[i], g, c))
        b = (unsigned)(((c.x0 >> kBrickShift) * by + (c.y0 >> kBrickShift)) * bz + (c.z0 >> kBrickShift));
    return (i < np ? 0u : nb + (unsigned)kSortOutside) + b;
}

// ---- counting sort by key ------------------------------------------------------
// Three launches (counts with ranks, an exclusive scan of the key counts, the
// scatter into the other buffer set); no memset and no host decision, so a
// captured graph can hold it.  The order within a key follows the order of
// the count atomics: no result depends on the storage order (the deposits and
// diagnostics are exact integer sums, every other operation is per particle),
// and pid keeps each slot's original index for the download.

// Particle q's key into skey[q] and its rank within the key into srank[q];
// buck[k] counts key k.  The block's keys go through a table in LDS (open
// addressing, twice as many slots as threads, so it never fills): a particle
// takes its rank among the block's particles of its key from an LDS atomic,
// then one global atomic per key present adds the block's count and returns
// the block's first rank (particles are mostly in key order, so a block holds
// few keys: a handful of global atomics per block, not one per run of equal
// keys in a wave).  Called by every thread of the block (valid: q is a
// particle).
constexpr int kSortHashBits = 9;
constexpr int kSortHash = 1 << kSortHashBits;
static_assert(kSortHash >= 2 * kBlock, "the key table must hold twice the block's particles");

template <typename T>
__device__ __forceinline__ void sort_count_body(const Particles<T> &p, int np, const Geom<T> &g, int q, bool valid,
                                                unsigned *__restrict__ buck, unsigned *__restrict__ skey,
                                                int *__restrict__ srank) {
    __shared__ unsigned hkey[kSortHash], hcnt[kSortHash], hbase[kSortHash];
    constexpr unsigned kEmpty = 0xffffffffu;
    for (int e = threadIdx.x; e < kSortHash; e += blockDim.x) {
        hkey[e] = kEmpty;
        hcnt[e] = 0u;
    }
    __syncthreads();
    const unsigned k = valid ? sort_key_of(p, q, np, g) : kEmpty;
    int slot = -1;
    unsigned local = 0;
    if (valid) {
        unsigned h = (k * 2654435761u) >> (32 - kSortHashBits);  // in [0, kSortHash)
        for (;;) {
            const unsigned old = atomicCAS(&hkey[h], kEmpty, k);
            if (old == kEmpty || old == k) break;
            h = (h + 1) & (kSortHash - 1);
        }
        slot = (int)h;
        local = atomicAdd(&hcnt[h], 1u);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kSortHash; e += blockDim.x)
        if (hkey[e] != kEmpty) hbase[e] = atomicAdd(&buck[hkey[e]], hcnt[e]);
    __syncthreads();
    if (!valid) return;
    skey[q] = k;
    srank[q] = (int)(hbase[slot] + local);
}

// The m key counts at b become the first slot of each key (exclusive scan),
// by one block of kSortScanThreads.  Up to kSortScanPer counts per thread:
// each thread loads its run of consecutive counts at once, the runs' totals
// are scanned over the block (one pass, two barriers), and each thread
// writes its run's offsets.  Longer: tiles of 4 kSortScanThreads counts, each
// thread four consecutive ones, scanned over the block with the running total
// of the earlier tiles carried.
constexpr int kSortScanThreads = 1024;
constexpr int kSortScanPer = 64;

__device__ __forceinline__ void sort_scan_body(unsigned *__restrict__ b, int m) {
    __shared__ unsigned wsum[kSortScanThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int per = (m + kSortScanThreads - 1) / kSortScanThreads;
    if (per <= kSortScanPer) {
        // (the run is read twice, for its total and for its offsets: the
        // second read hits the cache, and no run is held in registers)
        const int e0 = (int)threadIdx.x * per;
        unsigned s = 0;
#pragma unroll 16
        for (int u = 0; u < per; u++) s += e0 + u < m ? b[e0 + u] : 0u;
        unsigned inc = s;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const unsigned t = __shfl_up(inc, d, kWave);
            if (lane >= d) inc += t;
        }
        if (lane == kWave - 1) wsum[wave] = inc;
        __syncthreads();
        unsigned before = 0;
        for (int w = 0; w < wave; w++) before += wsum[w];
        unsigned run = before + inc - s;
#pragma unroll 16
        for (int u = 0; u < per; u++) {
            if (e0 + u < m) {
                const unsigned c = b[e0 + u];
                b[e0 + u] = run;
                run += c;
            }
        }
        return;
    }
    // tiles of 4 counts per thread, four tiles' loads issued together (the
    // scans of a group then wait on one load latency, not four)
    constexpr int kTile = 4 * kSortScanThreads, kGroup = 4;
    unsigned carry = 0;
    for (int gbase = 0; gbase < m; gbase += kGroup * kTile) {
        unsigned vg[kGroup][4];
#pragma unroll
        for (int gi = 0; gi < kGroup; gi++)
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int e = gbase + gi * kTile + 4 * (int)threadIdx.x + u;
                vg[gi][u] = e < m ? b[e] : 0u;
            }
#pragma unroll
        for (int gi = 0; gi < kGroup; gi++) {
            const int base = gbase + gi * kTile;  // (the same in every thread)
            if (base >= m) break;
            const int e0 = base + 4 * (int)threadIdx.x;
            const unsigned *v = vg[gi];
            const unsigned s = v[0] + v[1] + v[2] + v[3];
            // inclusive scan over the wave, then over the waves' totals
            unsigned inc = s;
#pragma unroll
            for (int d = 1; d < kWave; d <<= 1) {
                const unsigned t = __shfl_up(inc, d, kWave);
                if (lane >= d) inc += t;
            }
            if (lane == kWave - 1) wsum[wave] = inc;
            __syncthreads();
            unsigned before = 0, total = 0;
            for (int w = 0; w < kSortScanThreads / kWave; w++) {
                const unsigned t = wsum[w];
                before += w < wave ? t : 0u;
                total += t;
            }
            unsigned run = carry + before + inc - s;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (e0 + u < m) b[e0 + u] = run;
                run += v[u];
            }
            carry += total;
            __syncthreads();  // wsum is rewritten by the next tile
        }
    }
}

template <typename T>
__device__ __forceinline__ void sort_scatter_body(const Particles<T> &src, const Particles<T> &dst,
                                                  const int *__restrict__ pid_src, int *__restrict__ pid_dst, int q,
                                                  const unsigned *__restrict__ buck, const unsigned *__restrict__ skey,
                                                  const int *__restrict__ srank) {
    const int d = (int)buck[skey[q]] + srank[q];
    dst.x[d] = src.x[q]; dst.y[d] = src.y[q]; dst.z[d] = src.z[q];
    dst.kx[d] = src.kx[q]; dst.ky[d] = src.ky[q]; dst.kz[d] = src.kz[q];
    dst.fx[d] = src.fx[q]; dst.fy[d] = src.fy[q]; dst.fz[d] = src.fz[q];
    pid_dst[d] = pid_src[q];
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_sort_count(Particles<T> p, int np, int total, Geom<T> g, unsigned *__restrict__ buck, unsigned *__restrict__ skey,
             int *__restrict__ srank) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    sort_count_body(p, np, g, q, q < total, buck, skey, srank);
}

__global__ void __launch_bounds__(kSortScanThreads) k_sort_scan(unsigned *__restrict__ buck, int m) {
    sort_scan_body(buck, m);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_sort_scatter(Particles<T> src, Particles<T> dst, const int *__restrict__ pid_src, int *__restrict__ pid_dst,
               int total, const unsigned *__restrict__ buck, const unsigned *__restrict__ skey,
               const int *__restrict__ srank) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < total) sort_scatter_body(src, dst, pid_src, pid_dst, q, buck, skey, srank);
}

// storage order -> original particle order
template <typename T>
__global__ void k_unpermute(const T *__restrict__ src, T *__restrict__ dst, const int *__restrict__ pid, int total) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < total) dst[pid[s]] = src[s];
}

__global__ void k_iota(int *__restrict__ a, int n) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) a[s] = s;
}

// ---- force grid: F = -grad U at the nodes a particle pass gathers from ---------
// fdm_force (the reference's compute_volumetric_forces_fdm,
// include/physics.hpp:160-200) at every node of each species' count box (fbox:
// the deposit's box of the particles' cells, copied by k_clear_counts), stored
// as (fx, fy, fz, 0) per node so that a gather takes one aligned vector load
// per corner.  Nodes outside the box are left as they are: no particle of the
// pass that follows has a corner there.
template <typename T>
__global__ void __launch_bounds__(256)
k_force_grid(const T *__restrict__ U, T *__restrict__ F4, const unsigned *__restrict__ fbox, Geom<T> g) {
    typedef T V4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
#pragma unroll
    for (int sp = 0; sp < 2; sp++) {
        const unsigned *cb = fbox + 6 * sp;
        const int i0 = cbox_lo(cb[0]), j0 = cbox_lo(cb[1]), k0 = cbox_lo(cb[2]);
        const int i1 = cbox_hi(cb[3]), j1 = cbox_hi(cb[4]), k1 = cbox_hi(cb[5]);
        if (i1 < i0) continue;
        const int nj = j1 - j0 + 1, rows = (i1 - i0 + 1) * nj;
        const T *u = U + sp * g.G;
        V4 *f = (V4 *)F4 + sp * g.G;
        for (int row = blockIdx.x * nw + wave; row < rows; row += gridDim.x * nw) {
            const int i = i0 + row / nj, j = j0 + row % nj;
            for (int k = k0 + lane; k <= k1; k += kWave) {
                T fx, fy, fz;
                fdm_force<T, T>(u, i, j, k, g, fx, fy, fz);
                f[idx3(i, j, k, g.ny, g.nz)] = V4{fx, fy, fz, T(0.0)};
            }
        }
    }
}

// CIC gather of the force from the force grid: the values and the order of
// gather_force_from_u (F4s: the species' part of the grid)
template <typename T>
__device__ __forceinline__ void gather_force_from_f(const T *__restrict__ F4s, const Cell<T> &c, const Geom<T> &g,
                                                    T &ax, T &ay, T &az) {
    typedef T V4 __attribute__((ext_vector_type(4)));
    const V4 *f = (const V4 *)F4s;
    ax = T(0.0);
    ay = T(0.0);
    az = T(0.0);
    V4 fv[8];
    // the eight loads first (32-bit node offsets: G < 2^31)
#pragma unroll
    for (int corner = 0; corner < 8; corner++) {
        const int i = c.x0 + (corner & 1), j = c.y0 + ((corner >> 1) & 1), k = c.z0 + ((corner >> 2) & 1);
        const bool in = i < g.nx && j < g.ny && k < g.nz;
        fv[corner] = in ? f[(unsigned)((i * g.ny + j) * g.nz + k)] : V4{T(0.0), T(0.0), T(0.0), T(0.0)};
    }
#pragma unroll
    for (int corner = 0; corner < 8; corner++) {
        const int i = c.x0 + (corner & 1), j = c.y0 + ((corner >> 1) & 1), k = c.z0 + ((corner >> 2) & 1);
        if (i >= g.nx || j >= g.ny || k >= g.nz) continue;
        const T w = corner_weight(c, corner);
        ax += w * fv[corner].x;
        ay += w * fv[corner].y;
        az += w * fv[corner].z;
    }
}

// Sets the status word's outside bit when a lane of the wave holds a particle
// outside the grid: one atomic per wave, none once the bit is set (particles
// that left the grid would otherwise all hit the one word).  Called by every
// lane of the wave.
__device__ __forceinline__ void flag_outside(bool out, unsigned *status) {
    if (__ballot(out) != 0ULL && (threadIdx.x & (kWave - 1)) == 0 &&
        !(__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kStatusOutside))
        atomicOr(status, kStatusOutside);
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
// FG: the force comes from the force grid F4 (computed from the same U for the
// particles' cells), else from U directly; the same bits either way.
template <typename T, typename UT, bool FG = false>
__device__ __forceinline__ void particle_step_body(const int bid, Particles<T> p, int np, int total, Geom<T> g,
                                                   const UT *__restrict__ U, const T *__restrict__ F4,
                unsigned long long *__restrict__ acc, int *__restrict__ brec, T kick_fac, T drift_fac,
                int end_kick, int boost, T boost_p, T boost_n, long long *partials, unsigned *status) {
    const int i = bid * blockDim.x + threadIdx.x;
    long long v[6] = {0, 0, 0, 0, 0, 0};
    Cell<T> c;
    bool has = false;
    if (i < total) {
        const bool proton = i < np;
        const long long off = proton ? 0 : g.G;
        T x = p.x[i], y = p.y[i], z = p.z[i];
        T kx = p.kx[i], ky = p.ky[i], kz = p.kz[i];
        T fx = T(0.0), fy = T(0.0), fz = T(0.0);
        if (cic_cell(x, y, z, g, c)) {
            if constexpr (FG) gather_force_from_f<T>(F4 + 4 * off, c, g, fx, fy, fz);
            else gather_force_from_u<T, UT>(U + off, c, g, fx, fy, fz);
        }
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
        has = cic_cell(x, y, z, g, c);
        stat_values(x, y, z, proton, g, v);
    }
    flag_outside(i < total && !has, status);
    block_deposit(has, c, i < np ? 0 : 1, g, acc, brec, bid);
    stat_commit(v, partials, bid);
}

template <typename T, typename UT, bool FG>
__global__ void __launch_bounds__(kBlock)
k_particle_step(Particles<T> p, int np, int total, Geom<T> g, const UT *__restrict__ U, const T *__restrict__ F4,
                unsigned long long *__restrict__ acc, int *__restrict__ brec, T kick_fac, T drift_fac, int end_kick,
                int boost, T boost_p, T boost_n, long long *partials, unsigned *status) {
    particle_step_body<T, UT, FG>(blockIdx.x, p, np, total, g, U, F4, acc, brec, kick_fac, drift_fac, end_kick,
                                  boost, boost_p, boost_n, partials, status);
}

// deposit (and optionally diagnostics) of the current positions, no motion
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_deposit(Particles<T> p, int np, int total, Geom<T> g, unsigned long long *__restrict__ acc,
          long long *partials, unsigned *status, int *__restrict__ brec = nullptr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    long long v[6] = {0, 0, 0, 0, 0, 0};
    Cell<T> c;
    bool has = false;
    if (i < total) {
        const bool proton = i < np;
        T x = p.x[i], y = p.y[i], z = p.z[i];
        has = cic_cell(x, y, z, g, c);
        stat_values(x, y, z, proton, g, v);
    }
    if (acc) flag_outside(i < total && !has, status);
    if (acc) block_deposit(has, c, i < np ? 0 : 1, g, acc, brec, blockIdx.x);
    if (partials) stat_commit(v, partials, blockIdx.x);
}

// gather F(x) from U into p.f, optionally followed by the half kick
template <typename T, typename UT>
__device__ __forceinline__ void gather_u_body(const int bid, Particles<T> p, int np, int total, Geom<T> g,
                                              const UT *__restrict__ U, T kick_fac, int kick) {
    const int i = bid * blockDim.x + threadIdx.x;
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

template <typename T, typename UT>
__global__ void __launch_bounds__(kBlock)
k_gather_u(Particles<T> p, int np, int total, Geom<T> g, const UT *__restrict__ U, T kick_fac, int kick) {
    gather_u_body<T, UT>(blockIdx.x, p, np, total, g, U, kick_fac, kick);
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

// counts = 0 after a smearing: only the count box of each species can hold a
// non-zero count, so only its rows are cleared (block 0 first copies the
// global boxes to fbox when given).
__global__ void __launch_bounds__(256)
k_clear_counts(unsigned long long *__restrict__ acc, int nx, int ny, int nz, unsigned *__restrict__ fbox) {
    const long long G = (long long)nx * ny * nz;
    const unsigned *cbox = (const unsigned *)(acc + 2 * G);
    const int cw = cbox_species_words(nx);
    if (fbox && blockIdx.x == 0 && threadIdx.x < 12) fbox[threadIdx.x] = cbox[(threadIdx.x / 6) * cw + threadIdx.x % 6];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
#pragma unroll
    for (int sp = 0; sp < 2; sp++) {
        const unsigned *cb = cbox + sp * cw;
        const int i0 = cbox_lo(cb[0]), j0 = cbox_lo(cb[1]), k0 = cbox_lo(cb[2]);
        const int i1 = cbox_hi(cb[3]), j1 = cbox_hi(cb[4]), k1 = cbox_hi(cb[5]);
        if (i1 < i0) continue;
        const int nj = j1 - j0 + 1, rows = (i1 - i0 + 1) * nj;
        // one row (i, j) per wave, its columns k0..k1 over the lanes
        for (int row = blockIdx.x * nw + wave; row < rows; row += gridDim.x * nw) {
            const int i = i0 + row / nj, j = j0 + row % nj;
            unsigned long long *rp = acc + sp * G + ((long long)i * ny + j) * nz;
            for (int k = k0 + lane; k <= k1; k += kWave) rp[k] = 0ULL;
        }
    }
}

// ---- Gaussian smearing (include/physics.hpp:203-234) ----------------------------
// Device copy of the tap table, in T.
template <typename T>
struct SmearTab {
    int rx, ry, rz;
    int kb[HV_SMEAR_MAXR + 1];
    int kc[HV_SMEAR_MAXR + 1][HV_SMEAR_MAXR + 1];
    T gx[HV_SMEAR_MAXR + 1], gy[HV_SMEAR_MAXR + 1], gz[HV_SMEAR_MAXR + 1];
    T term;
    // channel a's taps in summation order, t = 0..2 kb(a) for b = t - kb(a):
    // the element offset of its H row in the vector tile, (kc(a, |b|) R + t)
    // TK with R = 8 + 2 ry rows and TK = 8 VW columns, and the weight gy(|b|)
    // (consecutive, so that a run of taps loads with a few wide scalar loads)
    int toff[HV_SMEAR_MAXR + 1][2 * HV_SMEAR_MAXR + 1];
    T tg[HV_SMEAR_MAXR + 1][2 * HV_SMEAR_MAXR + 1];
    // the same offsets for the 16-row vector tile (R = 16 + 2 ry)
    int toff16[HV_SMEAR_MAXR + 1][2 * HV_SMEAR_MAXR + 1];
    // the channels wave w of a 256-thread block sums in the vector tile, -1
    // after the last (longest first to the least loaded wave: the channels'
    // tap counts differ, so a round robin leaves waves idle)
    int wch[4][HV_SMEAR_MAXR + 2];
};

// One tap of a smearing sum, acc + g h: FP64 rounds the product and the sum
// (the oracle's and the reference's arithmetic); the FP32 form of the mixed
// precision path rounds once (a fused multiply-add), which halves the work per
// tap.  Every smearing kernel takes its taps through here, so both forms and
// every tiling give the same bits in each precision.
template <typename T>
__device__ __forceinline__ T smear_madd(T acc, T g, T h) {
    if constexpr (sizeof(T) == 4)
        return __builtin_fmaf(g, h, acc);
    else
        return acc + g * h;
}

// The tap table as the smearing bodies read it: through the constant address
// space, so that wave-uniform reads go through the scalar cache even where the
// table pointer was itself loaded from memory (ensembles).
template <typename T>
using SmearTabC = const __attribute__((address_space(4))) SmearTab<T>;

template <typename T>
__device__ __forceinline__ SmearTabC<T> *smear_tab_c(const SmearTab<T> *p) {
    return (SmearTabC<T> *)p;
}

// CT is the count type: unsigned long long (fixed point) or T.
template <typename T, typename CT>
__device__ __forceinline__ T count_value(const CT *c, long long q) {
    if constexpr (sizeof(CT) == 8 && !std::is_floating_point<CT>::value)
        return (T)hv_fix_to_double((long long)c[q]);
    else
        return (T)c[q];
}

// The z factor: H_0..H_rz of the R x TK columns of a tile from its staged
// counts C (R x W, the tile's columns plus rz on each side), taps through
// smear_madd.  The chain of
// each column runs in its order; the counts and weights of four levels are
// loaded ahead of their four steps.
template <typename T>
__device__ __forceinline__ void smear_h_levels(const T *__restrict__ C, T *__restrict__ Hs,
                                               SmearTabC<T> *__restrict__ tb, int R, int W, int TK, int rz,
                                               int nthr) {
    // four chains of a thread at a time, two levels per step: the counts of
    // sixteen taps load together ahead of their sums (a chain past the
    // R x TK columns runs on a safe column and stores nothing)
    const int lvl = R * TK;
    const T gz0 = tb->gz[0];
    for (int e0 = threadIdx.x; e0 < lvl; e0 += 4 * nthr) {
        const T *cp[4];
        T h[4];
        int e[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            e[u] = e0 + u * nthr;
            const int ee = e[u] < lvl ? e[u] : 0;
            const int r = ee / TK, cc = ee - r * TK;
            cp[u] = C + r * W + cc + rz;
            h[u] = gz0 * cp[u][0];
            if (e[u] < lvl) Hs[e[u]] = h[u];
        }
        int K = 1;
        for (; K + 1 <= rz; K += 2) {
            const T g0 = tb->gz[K], g1 = tb->gz[K + 1];
            T lo0[4], hi0[4], lo1[4], hi1[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                lo0[u] = cp[u][-K];
                hi0[u] = cp[u][K];
                lo1[u] = cp[u][-(K + 1)];
                hi1[u] = cp[u][K + 1];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                h[u] = smear_madd(smear_madd(h[u], g0, lo0[u]), g0, hi0[u]);
                if (e[u] < lvl) Hs[K * lvl + e[u]] = h[u];
                h[u] = smear_madd(smear_madd(h[u], g1, lo1[u]), g1, hi1[u]);
                if (e[u] < lvl) Hs[(K + 1) * lvl + e[u]] = h[u];
            }
        }
        if (K <= rz) {
            const T g0 = tb->gz[K];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                h[u] = smear_madd(smear_madd(h[u], g0, cp[u][-K]), g0, cp[u][K]);
                if (e[u] < lvl) Hs[K * lvl + e[u]] = h[u];
            }
        }
    }
}

// The y factor of channel a for VW consecutive columns of one row, with
// 16-byte LDS reads: acc[v] = sum over b = -kb(a)..kb(a) of
// gy(|b|) H_{kc(a,|b|)}(row + b, column + v), in ascending b.  h0 is the Hs
// offset of (level 0, row + ry, column); Hs and h0 are VW-aligned.
template <typename T>
constexpr int smear_vw() { return 16 / (int)sizeof(T); }

// (the vector tile: TK = 8 VW columns and R = 8 + 2 ry rows, the shape the
// table's tap offsets are built for)
template <typename T, bool T16 = false>
__device__ __forceinline__ void smear_channel_vec(const T *__restrict__ Hs, SmearTabC<T> *__restrict__ tb, int a,
                                                  int h0, int R, int TK, T (&acc)[smear_vw<T>()]) {
    constexpr int VW = smear_vw<T>();
    typedef T V __attribute__((ext_vector_type(VW)));
    (void)R;
    const int kb = tb->kb[a], nt = 2 * kb + 1;
#pragma unroll
    for (int v = 0; v < VW; v++) acc[v] = T(0.0);
    const T *hb = Hs + h0 - kb * TK;  // tap t reads element toff[a][t] of this base
    // runs of U taps: their table entries, then their H vectors, load
    // together ahead of the sums (one load latency per run, not per tap)
    auto run = [&](auto uu, int t) {
        constexpr int U = decltype(uu)::value;
        int o[U];
        T g[U];
        V h[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            o[u] = (T16 ? tb->toff16[a] : tb->toff[a])[t + u];
            g[u] = tb->tg[a][t + u];
        }
#pragma unroll
        for (int u = 0; u < U; u++) h[u] = *(const V *)(hb + o[u]);
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int v = 0; v < VW; v++) acc[v] = smear_madd(acc[v], g[u], h[u][v]);
    };
    int t = 0;
    // runs of 8: the next run's table entries load while this run's H vectors
    // do (the sums wait once for both, not for one and then the other)
    constexpr int U8 = 8;
    if (t + U8 <= nt) {
        int o[U8];
        T g[U8];
#pragma unroll
        for (int u = 0; u < U8; u++) {
            o[u] = (T16 ? tb->toff16[a] : tb->toff[a])[u];
            g[u] = tb->tg[a][u];
        }
        for (;;) {
            V h[U8];
#pragma unroll
            for (int u = 0; u < U8; u++) h[u] = *(const V *)(hb + o[u]);
            const bool more = t + 2 * U8 <= nt;
            int on[U8];
            T gn[U8];
            if (more) {
#pragma unroll
                for (int u = 0; u < U8; u++) {
                    on[u] = (T16 ? tb->toff16[a] : tb->toff[a])[t + U8 + u];
                    gn[u] = tb->tg[a][t + U8 + u];
                }
            }
#pragma unroll
            for (int u = 0; u < U8; u++)
#pragma unroll
                for (int v = 0; v < VW; v++) acc[v] = smear_madd(acc[v], g[u], h[u][v]);
            t += U8;
            if (!more) break;
#pragma unroll
            for (int u = 0; u < U8; u++) {
                o[u] = on[u];
                g[u] = gn[u];
            }
        }
    }
    if (t + 4 <= nt) {
        run(std::integral_constant<int, 4>{}, t);
        t += 4;
    }
    for (; t < nt; t++) run(std::integral_constant<int, 1>{}, t);
}

// The vector D stage needs TJ = 8 rows of TK = 8 VW columns: one wave covers
// the tile's 64 VW nodes of one channel.
template <typename T>
__host__ __device__ constexpr bool smear_vec_tile(int TJ, int TK) {
    return (TJ == 8 || TJ == 16) && TK == 8 * smear_vw<T>();
}

// The VW sums of a lane into D at dp: one 16-byte store when the row length
// nz is a multiple of VW (dp is then aligned), else element by element up to
// the row's end (n valid elements).
template <typename T>
__device__ __forceinline__ void smear_store_vec(T *__restrict__ dp, const T (&acc)[smear_vw<T>()], int nz, int n) {
    constexpr int VW = smear_vw<T>();
    typedef T V __attribute__((ext_vector_type(VW)));
    if (nz % VW == 0 && n >= VW) {
        V o;
#pragma unroll
        for (int v = 0; v < VW; v++) o[v] = acc[v];
        *(V *)dp = o;
    } else {
#pragma unroll
        for (int v = 0; v < VW; v++)
            if (v < n) dp[v] = acc[v];
    }
}

// counts staged in LDS take R x W values rounded up to a multiple of VW, so
// that the H levels after them are 16-byte aligned
__host__ __device__ inline int smear_c_words(int R, int W, int vw) { return (R * W + vw - 1) / vw * vw; }

// The spherical-cut Gaussian factorises (hv_smear_table_build), and the sum
// is taken as three nested factor sums, in this order of operations:
//   z:  H_K(c) = (H_{K-1} + gz(K) c(k-K)) + gz(K) c(k+K),  H_0 = gz(0) c(k)
//   y:  D_a = sum over b = -kb(a)..kb(a) of gy(|b|) H_{kc(a,|b|)}(j+b)
//   x:  rho = (sum over a = -rx..rx, 0 <= i+a < nx of gx(|a|) D_|a|(i+a)) term
// each sum from 0 in ascending offset order, every tap acc + g h through
// smear_madd.  One block owns the TJ x TK
// (row, column) tile of one species over a chunk of IC planes: it walks the
// planes p that reach the chunk in ascending order, builds H and the D_a
// channels of plane p in LDS from the counts of the tile plus its (ry, rz)
// halo, and adds gx(|a|) D_|a|(p) to the running x sums of the chunk planes
// p -+ a, kept in an LDS ring of 2 rx + 1 planes; the sums of plane p - rx
// are complete after plane p and written out.  The D channels never leave
// the CU.  A plane whose tile and halo hold no count contributes exact zeros
// and is skipped.  Every x sum receives its terms in ascending a, so the
// result does not depend on the tiling or the chunking.
struct SmearTiles {
    int TJ, TK;     // tile rows, tile columns (TJ * TK a multiple of 64, at most 512)
    int njt, nkt;   // tiles per plane along j and k
    int IC, nic;    // planes per chunk, chunks along i
};

constexpr int kSmearThreads = 256;
constexpr int kSmearNpl = 8;            // nodes per lane at most (TJ * TK <= 512)
constexpr size_t kSmearLds = 81920;     // LDS per block: two blocks per CU
constexpr size_t kSmearLdsVec = 114688;  // fused form, vector tile: one block per CU
constexpr size_t kSmear2LdsVec = 98304;  // two-pass form, vector tile

// Two-pass form: staged counts in rows of smear_zy_wp = TK + 2 RZB words, the
// halo's column cc (k = k0 - rz + cc) at word RZB - rz + cc of its row, RZB
// the radius bucket of rz; tile column j's counts at j +- K then sit at word
// RZB + j +- K whatever rz is, 16-byte aligned for j a multiple of VW.
__host__ __device__ constexpr int smear_rzb(int rz) {
    return rz <= 4 ? 4 : rz <= 8 ? 8 : rz <= 12 ? 12 : rz <= 16 ? 16 : 32;
}
__host__ __device__ inline int smear_zy_wp(int TK, int rz) { return TK + 2 * smear_rzb(rz); }

// The z factor of the two-pass form: H_0..H_rz as smear_h_levels (the same
// operations in the same order per column), a thread taking VW adjacent
// columns of a row at a time with their counts j - RZB .. j + VW - 1 + RZB in
// registers (2 RZB / VW + 1 aligned vector reads) and each level stored as one
// vector.
template <typename T, int RZB>
__device__ __forceinline__ void smear_h_levels_win(const T *__restrict__ C, T *__restrict__ Hs,
                                                   SmearTabC<T> *__restrict__ tb, int R, int TK, int rz) {
    constexpr int VW = 16 / (int)sizeof(T), NW = 2 * RZB / VW + 1;
    static_assert(RZB % VW == 0, "the window starts on a vector boundary");
    typedef T V __attribute__((ext_vector_type(VW)));
    const int Wp = TK + 2 * RZB, gpr = TK / VW, groups = R * gpr, lvl = R * TK;
    // the weights of all levels up front (scalar loads issued together, not
    // one wait per level)
    T gz[RZB + 1];
#pragma unroll
    for (int K = 0; K <= RZB; K++) gz[K] = tb->gz[K];
    const T gz0 = gz[0];
    for (int gi = threadIdx.x; gi < groups; gi += kSmearThreads) {
        const int r = gi / gpr, j0 = (gi - r * gpr) * VW;
        const V *src = (const V *)(C + r * Wp + j0);
        T w[NW * VW];
#pragma unroll
        for (int q = 0; q < NW; q++) {
            const V t = src[q];
#pragma unroll
            for (int v = 0; v < VW; v++) w[q * VW + v] = t[v];
        }
        V h;
#pragma unroll
        for (int v = 0; v < VW; v++) h[v] = gz0 * w[RZB + v];
        T *hp = Hs + r * TK + j0;
        *(V *)hp = h;
#pragma unroll
        for (int K = 1; K <= RZB; K++) {
            if (K <= rz) {  // (predicated, not a loop exit: the window stays in registers)
                const T g = gz[K];
#pragma unroll
                for (int v = 0; v < VW; v++)
                    h[v] = smear_madd(smear_madd(h[v], g, w[RZB + v - K]), g, w[RZB + v + K]);
                *(V *)(hp + K * lvl) = h;
            }
        }
    }
}

// (radii beyond 16: smear_h_levels on the same layout, its rows Wp words
// apart from word RZB - rz)
template <typename T>
__device__ __forceinline__ void smear_h_levels_zy(const T *__restrict__ C, T *__restrict__ Hs,
                                                  SmearTabC<T> *__restrict__ tb, int R, int TK, int rz) {
    switch (smear_rzb(rz)) {  // uniform: one instance per block
    case 4: smear_h_levels_win<T, 4>(C, Hs, tb, R, TK, rz); break;
    case 8: smear_h_levels_win<T, 8>(C, Hs, tb, R, TK, rz); break;
    case 12: smear_h_levels_win<T, 12>(C, Hs, tb, R, TK, rz); break;
    case 16: smear_h_levels_win<T, 16>(C, Hs, tb, R, TK, rz); break;
    default:
        smear_h_levels<T>(C + (smear_rzb(rz) - rz), Hs, tb, R, smear_zy_wp(TK, rz), TK, rz, kSmearThreads);
        break;
    }
}


inline size_t smear_lds_bytes(const SmearTiles &st, int rx, int ry, int rz, size_t tsize) {
    const int R = st.TJ + 2 * ry, W = st.TK + 2 * rz;
    return tsize * ((size_t)smear_c_words(R, W, (int)(16 / tsize)) + (size_t)(rz + 1) * R * st.TK +
                    (size_t)(2 * rx + 1) * st.TJ * st.TK);
}

// the two-pass form's zy kernel: counts and H levels
inline size_t smear2_lds_bytes(const SmearTiles &st, int ry, int rz, size_t tsize) {
    const int R = st.TJ + 2 * ry, Wp = smear_zy_wp(st.TK, rz);
    return tsize * ((size_t)smear_c_words(R, Wp, (int)(16 / tsize)) + (size_t)(rz + 1) * R * st.TK);
}

inline long long smear_blocks(const SmearTiles &st) { return 2LL * st.njt * st.nkt * st.nic; }

template <typename T, typename CT>
__device__ __forceinline__ void smear_fused_body(const int bid, const CT *__restrict__ cnt, T *__restrict__ rho,
                                                 SmearTabC<T> *__restrict__ tb, int nx, int ny, int nz,
                                                 SmearTiles st) {
    extern __shared__ unsigned char smem[];
    const int rx = tb->rx, ry = tb->ry, rz = tb->rz;
    const int TJ = st.TJ, TK = st.TK;
    const int R = TJ + 2 * ry, W = TK + 2 * rz, NS = 2 * rx + 1;
    const int nnode = TJ * TK, npl = nnode / kWave;
    constexpr int VW = smear_vw<T>();
    const bool vec = smear_vec_tile<T>(TJ, TK);
    T *C = (T *)smem;                          // R x W counts of plane p
    T *Hs = C + smear_c_words(R, W, VW);       // (rz + 1) x R x TK
    T *ring = Hs + (rz + 1) * R * TK;  // NS x nnode running x sums, plane o in slot o % NS
    int t = bid;
    const int kt = t % st.nkt;
    t /= st.nkt;
    const int jt = t % st.njt;
    t /= st.njt;
    const int ic = t % st.nic;
    const int sp = t / st.nic;
    const long long G = (long long)nx * ny * nz, plane = (long long)ny * nz;
    const long long sbase = sp * G;
    const int j0 = jt * TJ, k0 = kt * TK;
    const int ia = ic * st.IC, ib = ia + st.IC < nx ? ia + st.IC : nx;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), nwaves = kSmearThreads / kWave;
    for (int e = threadIdx.x; e < NS * nnode; e += kSmearThreads) ring[e] = T(0.0);
    const T term = tb->term;
    auto flush = [&](int o) {  // write plane o of the chunk and clear its slot
        T *rs = ring + (o % NS) * nnode;
        for (int e = threadIdx.x; e < nnode; e += kSmearThreads) {
            const int r = e / TK, cc = e - r * TK;
            const int j = j0 + r, k = k0 + cc;
            const T v = rs[e];
            rs[e] = T(0.0);
            if (j < ny && k < nz) rho[sbase + o * plane + (long long)j * nz + k] = v * term;
        }
    };
    // this lane's nodes: Hs offset of (row + ry, column) and ring offset
    int hoff[kSmearNpl];
#pragma unroll
    for (int u = 0; u < kSmearNpl; u++) {
        const int node = lane + u * kWave;
        const int r = node / TK, cc = node - r * TK;
        hoff[u] = (r + ry) * TK + cc;
    }
    const int p_lo = ia - rx > 0 ? ia - rx : 0, p_hi = ib - 1 + rx < nx - 1 ? ib - 1 + rx : nx - 1;
    __syncthreads();
    for (int p = p_lo; p <= p_hi; p++) {
        const long long base = sbase + (long long)p * plane;
        int nonzero = 0;
        for (int e = threadIdx.x; e < R * W; e += kSmearThreads) {
            const int r = e / W, cc = e - r * W;
            const int j = j0 - ry + r, k = k0 - rz + cc;
            T v = T(0.0);
            if (j >= 0 && j < ny && k >= 0 && k < nz) v = count_value<T, CT>(cnt, base + (long long)j * nz + k);
            C[e] = v;
            nonzero |= (v != T(0.0));
        }
        // (read back as a wave-uniform value: the channel loops below then stay
        // scalar, with the tap tables in scalar loads)
        if (__builtin_amdgcn_readfirstlane(__syncthreads_or(nonzero))) {
            smear_h_levels<T>(C, Hs, tb, R, W, TK, rz, kSmearThreads);
            __syncthreads();
            if (vec) {
                // channel a per wave, VW consecutive columns per lane
                const int r = lane >> 3, c0 = (lane & 7) * VW;
                const int h0 = (r + ry) * TK + c0, n0 = r * TK + c0;
                for (int n = 0; n <= HV_SMEAR_MAXR; n++) {
                    const int a = tb->wch[wave][n];
                    if (a < 0) break;
                    T acc[VW];
                    smear_channel_vec<T>(Hs, tb, a, h0, R, TK, acc);
                    const T gxa = tb->gx[a];
                    const int o1 = p - a, o2 = p + a;
                    const bool w1 = o1 >= ia && o1 < ib, w2 = a > 0 && o2 >= ia && o2 < ib;
                    T *r1 = ring + (w1 ? o1 % NS : 0) * nnode + n0, *r2 = ring + (w2 ? o2 % NS : 0) * nnode + n0;
#pragma unroll
                    for (int v = 0; v < VW; v++) {
                        if (w1) r1[v] = smear_madd(r1[v], gxa, acc[v]);
                        if (w2) r2[v] = smear_madd(r2[v], gxa, acc[v]);
                    }
                }
            } else {
            // channel a per wave (uniform table reads), npl nodes per lane
            for (int a = wave; a <= rx; a += nwaves) {
                const int kb = tb->kb[a];
                T acc[kSmearNpl];
#pragma unroll
                for (int u = 0; u < kSmearNpl; u++) acc[u] = T(0.0);
                for (int b = -kb; b <= kb; b++) {
                    const int ab = b < 0 ? -b : b;
                    const T g = tb->gy[ab];
                    const T *hb = Hs + (tb->kc[a][ab] * R + b) * TK;
#pragma unroll
                    for (int u = 0; u < kSmearNpl; u++)
                        if (u < npl) acc[u] = smear_madd(acc[u], g, hb[hoff[u]]);
                }
                const T gxa = tb->gx[a];
                const int o1 = p - a, o2 = p + a;
                const bool w1 = o1 >= ia && o1 < ib, w2 = a > 0 && o2 >= ia && o2 < ib;
                T *r1 = ring + (w1 ? o1 % NS : 0) * nnode, *r2 = ring + (w2 ? o2 % NS : 0) * nnode;
#pragma unroll
                for (int u = 0; u < kSmearNpl; u++) {
                    if (u >= npl) continue;
                    const int node = lane + u * kWave;
                    if (w1) r1[node] = smear_madd(r1[node], gxa, acc[u]);
                    if (w2) r2[node] = smear_madd(r2[node], gxa, acc[u]);
                }
            }
            }
        }
        __syncthreads();
        if (p - rx >= ia && p - rx < ib) flush(p - rx);
    }
    // planes whose last contributing plane is nx - 1
    for (int o = (p_hi - rx + 1 > ia ? p_hi - rx + 1 : ia); o < ib; o++) flush(o);
}

template <typename T, typename CT>
__global__ void __launch_bounds__(kSmearThreads)
k_smear_fused(const CT *__restrict__ cnt, T *__restrict__ rho, const SmearTab<T> *__restrict__ tb, int nx, int ny,
              int nz, SmearTiles st) {
    smear_fused_body<T, CT>(blockIdx.x, cnt, rho, smear_tab_c(tb), nx, ny, nz, st);
}

// f(std::integral_constant<int, 0>) ... f(std::integral_constant<int, N - 1>),
// in order: indices known at compile time (register arrays stay in registers)
template <typename F, int... Is>
__device__ __forceinline__ void static_for_seq(F &&f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    static_for_seq(f, std::make_integer_sequence<int, N>{});
}

template <typename T, typename CT>
__device__ __forceinline__ T count_of(CT v) {
    if constexpr (sizeof(CT) == 8 && !std::is_floating_point<CT>::value)
        return (T)hv_fix_to_double((long long)v);
    else
        return (T)v;
}

// ---- two-pass smearing: z and y factors in LDS, then a sparse x pass ----------
// The same factor sums in the same order (bit-identical to k_smear_fused),
// with the D channels stored between two launches; no plane walk, so every
// (plane, tile) is a block of its own (the faster form for a single grid):
//   k_smear_zy: one block per (species, plane i, row tile of TJ rows, column
//     tile of TK columns).  A tile whose halo misses the count box of its
//     plane (written by the deposit; none: every tile reads its counts) stops
//     at once.  Otherwise it stages the counts of the tile plus its (ry rows,
//     rz columns) halo in LDS (four loads per thread in flight, none outside
//     the count box), builds every H_K of the (TJ + 2 ry) x TK columns in LDS
//     and writes the D_a channels of its TJ x TK nodes.  Each block sets or
//     clears bit i of its tile column's plane mask: set when its halo holds a
//     count.  A cleared tile writes nothing: its H and D are exactly zero.
//   k_smear_xs / k_smear_xs_vec: the x factor, skipping planes whose bit is
//     clear (their terms are exact zeros, so skipping them changes no bit);
//     the vector form takes VW nodes per thread and issues the loads of all
//     2 RXM + 1 planes at once, and writes zeros outside the count box grown
//     by the radius without reading anything.

// plane-mask words per tile column
__host__ __device__ inline int smear_mask_words(int nx) { return (nx + 31) / 32; }

// one (species sp, plane i, tile jt, kt) item of the z and y factors; pw: the
// count-box words of species sp (null: every count is read).
// Every branch is uniform over the block.
// A workgroup barrier that orders LDS accesses only: global loads issued
// before it stay in flight, so a prefetch survives it (__syncthreads fences
// every address space and waits for them).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// the OR of v over a workgroup of kSmearThreads threads, through LDS-only
// barriers (slot: kSmearThreads / 64 ints of LDS, not reused before the
// caller's next barrier)
__device__ __forceinline__ int block_or_lds(int v, int *slot) {
    const bool b = __ballot(v) != 0ULL;
    if ((threadIdx.x & (kWave - 1)) == 0) slot[threadIdx.x / kWave] = b ? 1 : 0;
    lds_barrier();
    int r = 0;
#pragma unroll
    for (int w = 0; w < kSmearThreads / kWave; w++) r |= slot[w];
    return r;
}

// staged count elements per thread with precomputed coordinates
constexpr int kZyStaged = 8;

template <typename T>
__device__ __forceinline__ T *smem_zy() {
    extern __shared__ unsigned char smem[];
    return (T *)smem;
}

// (row << 16 | column) of the halo elements thread + 256 m (-1 past the R x W)
template <int NS>
__device__ __forceinline__ void smear_zy_coords(int R, int W, int (&ecoord)[NS]) {
#pragma unroll
    for (int m = 0; m < NS; m++) {
        const int e = (int)threadIdx.x + m * kSmearThreads;
        const int r = e / W;
        ecoord[m] = e < R * W ? (r << 16 | (e - r * W)) : -1;
    }
}

// The H levels and D channels of a staged item (counts in C, `any` set when
// one of them is non-zero): sets or clears its plane bit.
template <typename T, bool V16 = false>
__device__ __forceinline__ void smear_zy_finish(bool any, int sp, int i, int jt, int kt, T *__restrict__ D,
                                                unsigned *__restrict__ fmask, SmearTabC<T> *__restrict__ tb, int nx,
                                                int ny, int nz, const SmearTiles &st) {
    // (LDS-only barriers: the next item's count loads stay in flight)
    extern __shared__ unsigned char smem[];
    const int ry = tb->ry, rz = tb->rz;
    const int TJ = st.TJ, TK = st.TK;
    const int R = TJ + 2 * ry, Wp = smear_zy_wp(TK, rz);
    constexpr int VW = smear_vw<T>();
    T *C = (T *)smem;
    T *Hs = C + smear_c_words(R, Wp, VW);
    const long long G = (long long)nx * ny * nz, n2 = 2 * G;
    const long long base = sp * G + (long long)i * ny * nz;
    const int j0 = jt * TJ, k0 = kt * TK;
    unsigned *mw = fmask + (((long long)sp * st.njt + jt) * st.nkt + kt) * smear_mask_words(nx) + (i >> 5);
    const unsigned bit = 1u << (i & 31);
    if (!any) {
        if (threadIdx.x == 0) atomicAnd(mw, ~bit);
        return;
    }
    smear_h_levels_zy<T>(C, Hs, tb, R, TK, rz);
    lds_barrier();
    if (threadIdx.x == 0) atomicOr(mw, bit);
    // channel a per wave, VW consecutive columns per lane (the vector tile)
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), nwaves = kSmearThreads / kWave;
    const int r = lane >> 3, c0 = (lane & 7) * VW;
    (void)nwaves;
    // a 16-row tile: its two 8-row halves in turn
    for (int n = 0; n <= HV_SMEAR_MAXR; n++) {
        const int a = tb->wch[wave][n];
        if (a < 0) break;
        // (V16: the caller's tiles may have 16 rows; the ensemble's have 8)
        for (int hh = 0; hh < (V16 ? TJ : 8); hh += 8) {
            const int j = j0 + hh + r;
            const long long q = base + (long long)j * nz + k0 + c0;
            T acc[VW];
            if (V16 && TJ == 16)
                smear_channel_vec<T, true>(Hs, tb, a, (hh + r + ry) * TK + c0, R, TK, acc);
            else
                smear_channel_vec<T>(Hs, tb, a, (r + ry) * TK + c0, R, TK, acc);
            if (j < ny) smear_store_vec<T>(D + (long long)a * n2 + q, acc, nz, nz - (k0 + c0));
        }
    }
}

template <typename T, typename CT>
__device__ __forceinline__ void smear_zy_item(int sp, int i, int jt, int kt, const CT *__restrict__ cnt,
                                              T *__restrict__ D, unsigned *__restrict__ fmask,
                                              const unsigned *__restrict__ pw, SmearTabC<T> *__restrict__ tb,
                                              int nx, int ny, int nz, const SmearTiles &st,
                                              const int *ecoord = nullptr) {
    extern __shared__ unsigned char smem[];
    const int rx = tb->rx, ry = tb->ry, rz = tb->rz;
    const int TJ = st.TJ, TK = st.TK;
    const int R = TJ + 2 * ry, W = TK + 2 * rz;
    // counts: R rows of Wp words, halo column cc at word pad + cc
    const int Wp = smear_zy_wp(TK, rz), pad = smear_rzb(rz) - rz;
    constexpr int VW = smear_vw<T>();
    T *C = (T *)smem;
    T *Hs = C + smear_c_words(R, Wp, VW);   // (rz + 1) x R x TK
    const long long G = (long long)nx * ny * nz, n2 = 2 * G;
    const long long base = sp * G + (long long)i * ny * nz;
    const int j0 = jt * TJ, k0 = kt * TK;
    unsigned *mw = fmask + (((long long)sp * st.njt + jt) * st.nkt + kt) * smear_mask_words(nx) + (i >> 5);
    const unsigned bit = 1u << (i & 31);
    // the count box of the species: rows pj0..pj1, columns pk0..pk1 may hold counts
    int pj0 = 0, pj1 = ny - 1, pk0 = 0, pk1 = nz - 1;
    if (pw) {
        pj0 = cbox_lo(pw[1]);
        pj1 = cbox_hi(pw[4]);
        pk0 = cbox_lo(pw[2]);
        pk1 = cbox_hi(pw[5]);
    }
    if (pj1 < pj0 || pj0 > j0 + TJ - 1 + ry || pj1 < j0 - ry || pk0 > k0 + TK - 1 + rz || pk1 < k0 - rz) {
        if (threadIdx.x == 0) atomicAnd(mw, ~bit);
        return;
    }
    int nonzero = 0;
    const int RW = R * W;
    if (ecoord && RW <= kZyStaged * kSmearThreads) {
        // element e = thread + 256 m at (row, column) = ecoord[m] of the halo
        // (rows and columns relative to (j0 - ry, k0 - rz)), all loads at once
        const CT *cb = cnt + base + (long long)(j0 - ry) * nz + (k0 - rz);
        const int rlo = pj0 - (j0 - ry), rhi = pj1 - (j0 - ry), clo = pk0 - (k0 - rz), chi = pk1 - (k0 - rz);
        CT v[kZyStaged];
#pragma unroll
        for (int m = 0; m < kZyStaged; m++) {
            const int r = ecoord[m] >> 16, cc = ecoord[m] & 0xffff;
            v[m] = (ecoord[m] >= 0 && r >= rlo && r <= rhi && cc >= clo && cc <= chi) ? cb[r * nz + cc] : CT(0);
        }
#pragma unroll
        for (int m = 0; m < kZyStaged; m++) {
            if (ecoord[m] >= 0) {
                const T x = count_of<T, CT>(v[m]);
                C[(ecoord[m] >> 16) * Wp + pad + (ecoord[m] & 0xffff)] = x;
                nonzero |= (x != T(0.0));
            }
        }
    } else {
        for (int e0 = 0; e0 < RW; e0 += 4 * kSmearThreads) {
            CT v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int e = e0 + u * kSmearThreads + (int)threadIdx.x;
                const int r = e / W, cc = e - r * W;
                const int j = j0 - ry + r, k = k0 - rz + cc;
                v[u] = (e < RW && j >= pj0 && j <= pj1 && k >= pk0 && k <= pk1) ? cnt[base + (long long)j * nz + k]
                                                                                : CT(0);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int e = e0 + u * kSmearThreads + (int)threadIdx.x;
                if (e < RW) {
                    const T x = count_of<T, CT>(v[u]);
                    const int r = e / W;
                    C[r * Wp + pad + (e - r * W)] = x;
                    nonzero |= (x != T(0.0));
                }
            }
        }
    }
    // (read back as a wave-uniform value: the channel loop below then stays
    // scalar, with the tap tables in scalar loads)
    if (!__builtin_amdgcn_readfirstlane(__syncthreads_or(nonzero))) {
        if (threadIdx.x == 0) atomicAnd(mw, ~bit);
        return;
    }
    smear_h_levels_zy<T>(C, Hs, tb, R, TK, rz);
    __syncthreads();
    if (threadIdx.x == 0) atomicOr(mw, bit);
    if (smear_vec_tile<T>(TJ, TK)) {
        // channel a per wave, VW consecutive columns per lane
        const int lane = threadIdx.x & (kWave - 1);
        const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), nwaves = kSmearThreads / kWave;
        const int r = lane >> 3, c0 = (lane & 7) * VW;
        (void)nwaves;
        // a 16-row tile: its two 8-row halves in turn
        for (int n = 0; n <= HV_SMEAR_MAXR; n++) {
            const int a = tb->wch[wave][n];
            if (a < 0) break;
            for (int hh = 0; hh < TJ; hh += 8) {
                const int j = j0 + hh + r;
                const long long q = base + (long long)j * nz + k0 + c0;
                T acc[VW];
                if (TJ == 16)
                    smear_channel_vec<T, true>(Hs, tb, a, (hh + r + ry) * TK + c0, R, TK, acc);
                else
                    smear_channel_vec<T>(Hs, tb, a, (r + ry) * TK + c0, R, TK, acc);
                if (j < ny) smear_store_vec<T>(D + (long long)a * n2 + q, acc, nz, nz - (k0 + c0));
            }
        }
        return;
    }
    for (int e = threadIdx.x; e < TJ * TK; e += kSmearThreads) {
        const int r = e / TK, cc = e - r * TK;
        const int j = j0 + r, k = k0 + cc;
        if (j >= ny || k >= nz) continue;
        const long long q = base + (long long)j * nz + k;
        for (int a = 0; a <= rx; a++) {
            const int kb = tb->kb[a];
            T acc = T(0.0);
            for (int b = -kb; b <= kb; b++) {
                const int ab = b < 0 ? -b : b;
                acc = smear_madd(acc, tb->gy[ab], Hs[(tb->kc[a][ab] * R + (r + ry + b)) * TK + cc]);
            }
            D[(long long)a * n2 + q] = acc;
        }
    }
}

// one block per (species, plane, tile) of the whole grid (ensembles)
template <typename T, typename CT>
__device__ __forceinline__ void smear_zy_body(const int bid, const CT *__restrict__ cnt, T *__restrict__ D,
                                              unsigned *__restrict__ fmask, const unsigned *__restrict__ cbox, int cw,
           SmearTabC<T> *__restrict__ tb, int nx, int ny, int nz, SmearTiles st) {
    int t = bid;
    const int kt = t % st.nkt;
    t /= st.nkt;
    const int jt = t % st.njt;
    t /= st.njt;
    const int i = t % nx;
    const int sp = t / nx;
    // the halo's coordinates once per item when the staged form fits (no
    // division per count element)
    const int R = st.TJ + 2 * tb->ry, W = st.TK + 2 * tb->rz;
    if (smear_vec_tile<T>(st.TJ, st.TK) && R * W <= kZyStaged * kSmearThreads) {
        int ecoord[kZyStaged];
        smear_zy_coords(R, W, ecoord);
        smear_zy_item<T, CT>(sp, i, jt, kt, cnt, D, fmask, cbox ? cbox + (long long)sp * cw : nullptr, tb, nx, ny,
                             nz, st, ecoord);
    } else {
        smear_zy_item<T, CT>(sp, i, jt, kt, cnt, D, fmask, cbox ? cbox + (long long)sp * cw : nullptr, tb, nx, ny,
                             nz, st);
    }
}

// The items of the count box only (planes that hold counts times the tiles
// whose halo meets the box; without a box every item), a grid of a few
// blocks per CU striding over them: no block is spent on empty space, and
// the plane masks of exactly these items are what the x pass reads.
template <typename T, typename CT>
__global__ void __launch_bounds__(kSmearThreads)
k_smear_zy(const CT *__restrict__ cnt, T *__restrict__ D, unsigned *__restrict__ fmask,
           const unsigned *__restrict__ cbox, int cw, const SmearTab<T> *__restrict__ tbp, int nx, int ny, int nz,
           SmearTiles st) {
    SmearTabC<T> *tb = smear_tab_c(tbp);
    const int ry = tb->ry, rz = tb->rz;
    int i0[2], jt0[2], kt0[2], ni[2], njt[2], nkt[2], nit[2];
#pragma unroll
    for (int sp = 0; sp < 2; sp++) {
        int ci0 = 0, ci1 = nx - 1, cj0 = 0, cj1 = ny - 1, ck0 = 0, ck1 = nz - 1;
        if (cbox) {
            const unsigned *cb = cbox + (long long)sp * cw;
            ci0 = cbox_lo(cb[0]);
            cj0 = cbox_lo(cb[1]);
            ck0 = cbox_lo(cb[2]);
            ci1 = cbox_hi(cb[3]);
            cj1 = cbox_hi(cb[4]);
            ck1 = cbox_hi(cb[5]);
        }
        const int ja = cj0 - ry > 0 ? (cj0 - ry) / st.TJ : 0, jb = min(st.njt - 1, (cj1 + ry) / st.TJ);
        const int ka = ck0 - rz > 0 ? (ck0 - rz) / st.TK : 0, kb = min(st.nkt - 1, (ck1 + rz) / st.TK);
        i0[sp] = ci0;
        jt0[sp] = ja;
        kt0[sp] = ka;
        ni[sp] = ci1 >= ci0 ? ci1 - ci0 + 1 : 0;
        njt[sp] = jb - ja + 1;
        nkt[sp] = kb - ka + 1;
        nit[sp] = ni[sp] * njt[sp] * nkt[sp];  // (at most 2 nx njt nkt items: 32-bit)
    }
    const int nall = nit[0] + nit[1];
    auto decode = [&](int it, int &sp, int &i, int &jt, int &kt) {
        sp = it >= nit[0] ? 1 : 0;
        const int u = it - (sp ? nit[0] : 0);
        const int ul = u / nkt[sp];
        kt = kt0[sp] + (u - ul * nkt[sp]);
        const int uj = ul / njt[sp];
        jt = jt0[sp] + (ul - uj * njt[sp]);
        i = i0[sp] + uj;
    };
    const int R = st.TJ + 2 * ry, W = st.TK + 2 * rz;
    if (!smear_vec_tile<T>(st.TJ, st.TK) || R * W > kZyStaged * kSmearThreads) {
        int ecoord[kZyStaged];
        smear_zy_coords(R, W, ecoord);
        for (int it = blockIdx.x; it < nall; it += gridDim.x) {
            int sp, i, jt, kt;
            decode(it, sp, i, jt, kt);
            smear_zy_item<T, CT>(sp, i, jt, kt, cnt, D, fmask, cbox ? cbox + (long long)sp * cw : nullptr, tb, nx,
                                 ny, nz, st, ecoord);
            __syncthreads();  // the next item's counts overwrite C and H
        }
        return;
    }
    // vector tiles with at most 8 staged counts per thread: the counts of the
    // next item load into registers while this item builds H and D
    int ecoord[kZyStaged];
    smear_zy_coords(R, W, ecoord);
    T *C = (T *)smem_zy<T>();
    const int Wp = smear_zy_wp(st.TK, rz), pad = smear_rzb(rz) - rz;
    __shared__ int orslot[kSmearThreads / kWave];
    CT pre[kZyStaged];
    // stage item it: its coordinates, whether its halo meets the count box of
    // its plane (hit), and the loads of its counts into pre
    auto stage = [&](int it, int &sp, int &i, int &jt, int &kt, bool &hit) {
        decode(it, sp, i, jt, kt);
        const int j0 = jt * st.TJ, k0 = kt * st.TK;
        int pj0 = 0, pj1 = ny - 1, pk0 = 0, pk1 = nz - 1;
        if (cbox) {
            const unsigned *pw = cbox + (long long)sp * cw;
            pj0 = cbox_lo(pw[1]);
            pj1 = cbox_hi(pw[4]);
            pk0 = cbox_lo(pw[2]);
            pk1 = cbox_hi(pw[5]);
        }
        hit = !(pj1 < pj0 || pj0 > j0 + st.TJ - 1 + ry || pj1 < j0 - ry || pk0 > k0 + st.TK - 1 + rz ||
                pk1 < k0 - rz);
        const long long G = (long long)nx * ny * nz;
        const CT *cb = cnt + sp * G + (long long)i * ny * nz + (long long)(j0 - ry) * nz + (k0 - rz);
        const int rlo = pj0 - (j0 - ry), rhi = pj1 - (j0 - ry), clo = pk0 - (k0 - rz), chi = pk1 - (k0 - rz);
#pragma unroll
        for (int m = 0; m < kZyStaged; m++) {
            const int r = ecoord[m] >> 16, cc = ecoord[m] & 0xffff;
            pre[m] = (hit && ecoord[m] >= 0 && r >= rlo && r <= rhi && cc >= clo && cc <= chi) ? cb[r * nz + cc]
                                                                                                : CT(0);
        }
    };
    int sp, i, jt, kt;
    bool hit = false;
    int it = blockIdx.x;
    if (it < nall) stage(it, sp, i, jt, kt, hit);
    for (; it < nall; it += gridDim.x) {
        int nonzero = 0;
        if (hit) {
#pragma unroll
            for (int m = 0; m < kZyStaged; m++) {
                if (ecoord[m] >= 0) {
                    const T x = count_of<T, CT>(pre[m]);
                    C[(ecoord[m] >> 16) * Wp + pad + (ecoord[m] & 0xffff)] = x;
                    nonzero |= (x != T(0.0));
                }
            }
        }
        const int csp = sp, ci = i, cjt = jt, ckt = kt;
        const bool chit = hit;
        if (it + (int)gridDim.x < nall) stage(it + gridDim.x, sp, i, jt, kt, hit);
        const bool any = chit && __builtin_amdgcn_readfirstlane(block_or_lds(nonzero, orslot));
        smear_zy_finish<T, true>(any, csp, ci, cjt, ckt, D, fmask, tb, nx, ny, nz, st);
        lds_barrier();  // the next item's counts overwrite C and H
    }
}

// one node per thread, any radius; with a count box (cbox, cw words per
// species) only the planes of the box are read, whose bits the zy pass wrote
template <typename T>
__device__ __forceinline__ void smear_xs_body(const int bid, const T *__restrict__ D, T *__restrict__ rho,
                                              const unsigned *__restrict__ fmask,
                                              const unsigned *__restrict__ cbox, int cw,
           SmearTabC<T> *__restrict__ tb, int nx, int ny, int nz, SmearTiles st) {
    const long long G = (long long)nx * ny * nz, n2 = 2 * G;
    const long long q = (long long)bid * blockDim.x + threadIdx.x;
    if (q >= n2) return;
    const long long plane = (long long)ny * nz;
    const int sp = q >= G ? 1 : 0;
    const long long node = q - sp * G;
    const int i = (int)(node / plane), j = (int)((node / nz) % ny), k = (int)(node % nz);
    const unsigned *mw = fmask + (((long long)sp * st.njt + j / st.TJ) * st.nkt + k / st.TK) * smear_mask_words(nx);
    int plo = 0, phi = nx - 1;
    if (cbox) {
        const unsigned *cb = cbox + (long long)sp * cw;
        plo = cbox_lo(cb[0]);
        phi = cbox_hi(cb[3]);
    }
    T acc = T(0.0);
    for (int a = -tb->rx; a <= tb->rx; a++) {
        const int aa = a < 0 ? -a : a;
        if (tb->kb[aa] < 0) continue;
        const int ia = i + a;
        if (ia < plo || ia > phi) continue;
        if (!((mw[ia >> 5] >> (ia & 31)) & 1u)) continue;
        acc = smear_madd(acc, tb->gx[aa], D[(long long)aa * n2 + q + (long long)a * plane]);
    }
    rho[q] = acc * tb->term;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_smear_xs(const T *__restrict__ D, T *__restrict__ rho, const unsigned *__restrict__ fmask,
           const unsigned *__restrict__ cbox, int cw, const SmearTab<T> *__restrict__ tb, int nx, int ny, int nz,
           SmearTiles st) {
    smear_xs_body<T>(blockIdx.x, D, rho, fmask, cbox, cw, smear_tab_c(tb), nx, ny, nz, st);
}

// VW nodes per thread (nz a multiple of VW), radius at most RXM <= 16: the
// plane bits i - RXM .. i + RXM of the tile column come from two mask words
// (planes outside the count box, or outside 0..nx-1 without one, are excluded
// by their index: the zy pass writes the bits of the box's planes only),
// the loads of all its planes are issued before the sum, and a term whose
// plane is skipped is an exact zero (acc + gx 0 = acc for acc >= +0)
template <typename T, int RXM>
__device__ __forceinline__ void smear_xs_vec_body(const int bid, const T *__restrict__ D, T *__restrict__ rho,
                                                  const unsigned *__restrict__ fmask,
                                                  const unsigned *__restrict__ cbox, int cw,
                                                  SmearTabC<T> *__restrict__ tb, int nx, int ny, int nz,
                                                  SmearTiles st, T *__restrict__ rs = nullptr, Split sps = Split{},
                                                  unsigned long long *__restrict__ acc_clear = nullptr) {
    constexpr int VW = smear_vw<T>(), NA = 2 * RXM + 1;
    static_assert(RXM <= 16, "two mask words cover the planes of a node");
    typedef T V __attribute__((ext_vector_type(VW)));
    const int nzq = nz / VW;
    const long long t = (long long)bid * blockDim.x + threadIdx.x;
    if (t >= 2LL * nx * ny * nzq) return;
    const int kq = (int)(t % nzq);
    long long rr = t / nzq;
    const int j = (int)(rr % ny);
    rr /= ny;
    const int i = (int)(rr % nx), sp = (int)(rr / nx);
    const int k0 = kq * VW;
    const long long G = (long long)nx * ny * nz, n2 = 2 * G, plane = (long long)ny * nz;
    const long long q = sp * G + (long long)i * plane + (long long)j * nz + k0;
    const int rx = tb->rx;
    V *out = (V *)(rho + q);
    // rs: the protons' densities also into the SOR's colour-split source (the
    // values k_rho_split would copy there)
    // acc_clear: the deposit's counts of these nodes (which the zy pass has
    // read) cleared for the next deposit; with a count box, only the nodes
    // inside it (every count outside it is zero already)
    bool clr = acc_clear != nullptr;
    auto put = [&](const V &v) {
        *out = v;
        if (rs && sp == 0)
#pragma unroll
            for (int u = 0; u < VW; u++) rs[split_index(sps, i, j, k0 + u)] = v[u];
        if (clr) {
            typedef unsigned long long U2 __attribute__((ext_vector_type(2)));
#pragma unroll
            for (int u = 0; u < VW; u += 2) *(U2 *)(acc_clear + q + u) = U2(0ULL);
        }
    };
    int plo = 0, phi = nx - 1;  // planes that may hold counts
    if (cbox) {
        const unsigned *cb = cbox + (long long)sp * cw;
        const int ry = tb->ry, rz = tb->rz;
        const int ci0 = cbox_lo(cb[0]), cj0 = cbox_lo(cb[1]), ck0 = cbox_lo(cb[2]);
        const int ci1 = cbox_hi(cb[3]), cj1 = cbox_hi(cb[4]), ck1 = cbox_hi(cb[5]);
        if (ci1 < ci0 || i < ci0 - rx || i > ci1 + rx || j < cj0 - ry || j > cj1 + ry || k0 + VW - 1 < ck0 - rz ||
            k0 > ck1 + rz) {
            clr = false;
            put(V(T(0.0)));
            return;
        }
        clr = clr && i >= ci0 && i <= ci1 && j >= cj0 && j <= cj1 && k0 + VW - 1 >= ck0 && k0 <= ck1;
        plo = ci0;
        phi = ci1;
    }
    const unsigned *mw = fmask + (((long long)sp * st.njt + j / st.TJ) * st.nkt + k0 / st.TK) * smear_mask_words(nx);
    // bits of planes i - RXM ..: word wlo = floor((i - RXM) / 32) and the next
    const int p0 = i - RXM, wlo = p0 >= 0 ? p0 >> 5 : -1, nmw = smear_mask_words(nx);
    const unsigned long long m = (wlo >= 0 ? (unsigned long long)mw[wlo] : 0ULL) |
                                 (wlo + 1 < nmw ? (unsigned long long)mw[wlo + 1] << 32 : 0ULL);
    const unsigned long long win = m >> (p0 - 32 * wlo);  // bit d: plane i - RXM + d
    V d[NA];
    static_for<NA>([&](auto dd) {
        constexpr int a = decltype(dd)::value - RXM, aa = a < 0 ? -a : a;
        const bool take = aa <= rx && i + a >= plo && i + a <= phi && ((win >> decltype(dd)::value) & 1ULL);
        d[decltype(dd)::value] = take ? *(const V *)(D + (long long)aa * n2 + q + (long long)a * plane) : V(T(0.0));
    });
    V acc = V(T(0.0));
    static_for<NA>([&](auto dd) {
        constexpr int a = decltype(dd)::value - RXM, aa = a < 0 ? -a : a;
        if (aa <= rx) {
            const T g = tb->gx[aa];
#pragma unroll
            for (int v = 0; v < VW; v++) acc[v] = smear_madd(acc[v], g, d[decltype(dd)::value][v]);
        }
    });
    const T term = tb->term;
#pragma unroll
    for (int v = 0; v < VW; v++) acc[v] = acc[v] * term;
    put(acc);
}

template <typename T, int RXM>
__global__ void __launch_bounds__(kBlock)
k_smear_xs_vec(const T *__restrict__ D, T *__restrict__ rho, const unsigned *__restrict__ fmask,
               const unsigned *__restrict__ cbox, int cw, const SmearTab<T> *__restrict__ tb, int nx, int ny, int nz,
               SmearTiles st, T *__restrict__ rs, Split sps, unsigned long long *__restrict__ acc_clear,
               unsigned *__restrict__ fbox) {
    // fbox: the count boxes kept for the force grid (as k_clear_counts keeps
    // them when it clears the counts)
    if (fbox && blockIdx.x == 0 && threadIdx.x < 12) fbox[threadIdx.x] = cbox[(threadIdx.x / 6) * cw + threadIdx.x % 6];
    smear_xs_vec_body<T, RXM>(blockIdx.x, D, rho, fmask, cbox, cw, smear_tab_c(tb), nx, ny, nz, st, rs, sps,
                              acc_clear);
}

// ---- Skyrme potentials (include/physics.hpp:273-281), optionally + phi ------
template <typename T>
__device__ __forceinline__ void skyrme_body(const int bid, const T *__restrict__ rho, T *__restrict__ U,
                                            const double *__restrict__ phi, double A, double B,
         double C, double gamma, long long G, Split sp) {
    long long i = (long long)bid * blockDim.x + threadIdx.x;
    if (i >= G) return;
    double rp = (double)rho[i], rn = (double)rho[i + G];
    double up, un;
    hv_skyrme_pair_f64(A, B, C, gamma, rp, rn, &up, &un);
    if (phi) {
        // merge_volumetric_potentials (include/tools.hpp:378-384), in the
        // grid precision like the separate pass
        const int k = (int)(i % sp.nz), j = (int)((i / sp.nz) % sp.ny), ii = (int)(i / ((long long)sp.ny * sp.nz));
        U[i] = (T)up + (T)phi[split_index(sp, ii, j, k)];
    } else {
        U[i] = (T)up;
    }
    U[i + G] = (T)un;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_skyrme(const T *__restrict__ rho, T *__restrict__ U, const double *__restrict__ phi, double A, double B,
         double C, double gamma, long long G, Split sp) {
    skyrme_body<T>(blockIdx.x, rho, U, phi, A, B, C, gamma, G, sp);
}

// ---- total energy (no counterpart in the reference) -----------------------------
// Grid part: per node dV * e_sky and dV * e_coul (hv_energy_density) of the
// smeared density and phi, quantised to 2^-30 and summed exactly; partials
// [2 * block + {0, 1}].
template <typename T>
__global__ void k_energy_grid(const T *__restrict__ rho, const double *__restrict__ phi, long long G, Split sp,
                              double A, double B, double C, double gamma, double dV, long long *__restrict__ partials) {
    long long v[2] = {0, 0};
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < G; i += stride) {
        const int k = (int)(i % sp.nz), j = (int)((i / sp.nz) % sp.ny), ii = (int)(i / ((long long)sp.ny * sp.nz));
        double es, ec;
        hv_energy_density(A, B, C, gamma, (double)rho[i], (double)rho[i + G], phi[split_index(sp, ii, j, k)], &es, &ec);
        v[0] += hv_stat_fix(es * dV);
        v[1] += hv_stat_fix(ec * dV);
    }
    block_sum_i64<2>(v);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = v[0];
        partials[2 * blockIdx.x + 1] = v[1];
    }
}

// Kinetic part: sum of |k|^2 over all test particles, quantised to 2^-30;
// partials[block].
template <typename T>
__global__ void k_energy_kin(const T *__restrict__ kx, const T *__restrict__ ky, const T *__restrict__ kz, int total,
                             long long *__restrict__ partials) {
    long long v[1] = {0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const double a = (double)kx[i], b = (double)ky[i], c = (double)kz[i];
        v[0] += hv_stat_fix(a * a + b * b + c * c);
    }
    block_sum_i64<1>(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

// exact 128-bit sums of nv interleaved components over nblocks partials:
// out[2 * c] = low word, out[2 * c + 1] = high word of component c
__global__ void k_sum_partials(const long long *__restrict__ partials, int nblocks, int nv,
                               unsigned long long *__restrict__ out) {
    const int c = threadIdx.x;
    if (c >= nv) return;
    unsigned long long lo = 0, hi = 0;
    for (int b = 0; b < nblocks; b++) add_i128(lo, hi, partials[(long long)nv * b + c]);
    out[2 * c] = lo;
    out[2 * c + 1] = hi;
}

// U_p += phi (include/tools.hpp:378-384)
template <typename T>
__global__ void k_merge(T *__restrict__ U, const double *__restrict__ phi, long long G, Split sp) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) return;
    const int k = (int)(i % sp.nz), j = (int)((i / sp.nz) % sp.ny), ii = (int)(i / ((long long)sp.ny * sp.nz));
    U[i] += (T)phi[split_index(sp, ii, j, k)];
}

// proton density (first G values of rho, natural order) -> colour-split copy
template <typename T>
__device__ __forceinline__ void rho_split_body(const int bid, const T *__restrict__ rho, T *__restrict__ rs,
                                               long long G, Split sp) {
    long long i = (long long)bid * blockDim.x + threadIdx.x;
    if (i >= G) return;
    const int k = (int)(i % sp.nz), j = (int)((i / sp.nz) % sp.ny), ii = (int)(i / ((long long)sp.ny * sp.nz));
    rs[split_index(sp, ii, j, k)] = rho[i];
}

template <typename T>
__global__ void k_rho_split(const T *__restrict__ rho, T *__restrict__ rs, long long G, Split sp) {
    rho_split_body<T>(blockIdx.x, rho, rs, G, sp);
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
    double ydx2, ydy2, ydz2;  // 1 / dx2 etc. (for hv_div_fixed)
    double inv_dx2;         // 1 / (2/dx^2 + 2/dy^2 + 2/dz^2)
    double src;             // 4 pi 1.44
    double omega, one_minus_omega;
    double tol;
};

struct SorCtl {
    int done;
    int iters;
};

// The FP32 correction sweep of the mixed form (L delta = R): the FP64 rule's
// stencil sum and relaxation with each product fused into its sum (one
// rounding each; the FP32 stage only brings delta near the solution, the
// FP64 finishing sweeps apply the exact rule).  Every FP32 sweep form goes
// through these two functions, so all of them give the same bits.
__device__ __forceinline__ float sor_f32_sum(float xm, float xp, float ym, float yp, float zm, float zp, float src,
                                             float yx, float yy, float yz) {
    return __builtin_fmaf(zp + zm, yz, __builtin_fmaf(yp + ym, yy, (xp + xm) * yx)) + src;
}

__device__ __forceinline__ float sor_f32_relax(float po, float s, float inv, float om, float om1) {
    return __builtin_fmaf(om1, po, om * (s * inv));
}

// Max |dphi| of iteration it: max over the kMaxSlots slots of row it.
// Called by whole waves: lane l reads slot l (kMaxSlots == wave width).
__device__ __forceinline__ double sor_row_max(const unsigned long long *__restrict__ maxd, int it) {
    const int lane = threadIdx.x & (kWave - 1);
    double v = __longlong_as_double((long long)maxd[((long long)it * kMaxSlots + lane) * kSlotStride]);
    return wave_max_f64(v);
}

// One colour of one iteration, in place, on the colour-split arrays
// (include/physics.hpp:106-151).  Each block is 64 k-slots x 4 rows and walks
// kSorPlanes planes.  Blocks are dispatched round-robin over the 8 XCDs, so
// block b runs on XCD b % 8; tiles are numbered so that each XCD always gets the
// same contiguous slab of planes, keeping its slab in that XCD's L2 from one
// sweep to the next.  Iteration `it` first checks iteration it-1's max |dphi|:
// once it is below tol the solve is over, and this and every later launch
// return at once (iteration count = the reference's stop rule).
constexpr int kSorPlanes = 2;       // planes per block, single-instance launches
constexpr int kSorPlanesBatch = 8;  // planes per block, ensemble launches
constexpr int kSorRows = 4;
constexpr int kXcds = 8;

struct SorTiles {
    int n_ic, n_jb, n_hb;  // plane chunks, row blocks, k-slot blocks
    int per_xcd;           // tiles per XCD (grid = kXcds * per_xcd blocks)
};

// P planes per block (kSorPlanes for the single-instance launches, more for
// the ensemble launches, whose grids are large enough to keep every CU busy
// with fewer, longer blocks).  All P planes' operands are loaded before the
// first update (the colour arrays do not alias), so the loads are in flight
// together.
// One colour of tile t (P planes x kSorRows rows x 64 k-slots, one node per
// thread and plane): the thread's max |dphi|.
template <typename RT, int P>
__device__ __forceinline__ double sor_tile_sweep(const int t, double *__restrict__ phi, const RT *__restrict__ rho,
                                                 const Split &sp, const SorTiles &tl, const SorCoef &cf,
                                                 int colour) {
    const int ntiles = tl.n_ic * tl.n_jb * tl.n_hb;
    double diff = 0.0;
    if (t < ntiles) {
        const int ic = t / (tl.n_jb * tl.n_hb);
        const int rem = t - ic * tl.n_jb * tl.n_hb;
        const int jb = rem / tl.n_hb, hb = rem - jb * tl.n_hb;
        const int j = 1 + jb * kSorRows + threadIdx.y;
        const int h = hb * 64 + threadIdx.x;
        const int ny = sp.ny, nz = sp.nz, nzh = sp.nzh;
        double *__restrict__ dst = phi + (colour ? sp.Gh : 0);
        const double *__restrict__ oth = phi + (colour ? 0 : sp.Gh);
        const RT *__restrict__ src = rho + (colour ? sp.Gh : 0);
        const long long sx = (long long)ny * nzh;
        double star[P], old[P];
        long long qq[P];
        bool ok[P];
#pragma unroll
        for (int di = 0; di < P; di++) {
            const int i = 1 + ic * P + di;
            const int off = (i + j + colour) & 1;
            const int k = 2 * h + off;
            ok[di] = j < ny - 1 && i < sp.nx - 1 && k >= 1 && k <= nz - 2;
            const long long q = ((long long)i * ny + j) * nzh + h;
            qq[di] = q;
            star[di] = 0.0;
            old[di] = 0.0;
            if (ok[di]) {
                const double r = (double)src[q];
                const double phi_x = hv_div_fixed(oth[q + sx] + oth[q - sx], cf.dx2, cf.ydx2);
                const double phi_y = hv_div_fixed(oth[q + nzh] + oth[q - nzh], cf.dy2, cf.ydy2);
                const double phi_z = hv_div_fixed(oth[q + off] + oth[q + off - 1], cf.dz2, cf.ydz2);
                star[di] = (phi_x + phi_y + phi_z + cf.src * r) * cf.inv_dx2;
                old[di] = dst[q];
            }
        }
#pragma unroll
        for (int di = 0; di < P; di++) {
            if (!ok[di]) continue;
            const double nv = cf.one_minus_omega * old[di] + cf.omega * star[di];
            dst[qq[di]] = nv;
            const double d = fabs(nv - old[di]);
            diff = d > diff ? d : diff;
        }
    }
    return diff;
}

template <typename RT, int P = kSorPlanes>
__device__ __forceinline__ void sor_colour_body(const int bid, double *__restrict__ phi, const RT *__restrict__ rho,
                                                Split sp, SorTiles tl, SorCoef cf, int colour, int it,
                                                unsigned long long *__restrict__ maxd, SorCtl *__restrict__ ctl) {
    if (ctl->done) return;
    if (colour == 0 && it > 0) {
        const double prev = sor_row_max(maxd, it - 1);
        if (prev < cf.tol) {
            if (threadIdx.x == 0 && threadIdx.y == 0 && bid == 0) {
                ctl->done = 1;
                ctl->iters = it;
            }
            return;
        }
    }
    const int b = bid;
    const int t = (b % kXcds) * tl.per_xcd + b / kXcds;
    const double diff = sor_tile_sweep<RT, P>(t, phi, rho, sp, tl, cf, colour);
    block_max_atomic(diff, &maxd[((long long)it * kMaxSlots + (b % kMaxSlots)) * kSlotStride]);
}

template <typename RT>
__global__ void __launch_bounds__(256)
k_sor_colour(double *__restrict__ phi, const RT *__restrict__ rho, Split sp, SorTiles tl, SorCoef cf,
             int colour, int it, unsigned long long *__restrict__ maxd, SorCtl *__restrict__ ctl) {
    sor_colour_body<RT>(blockIdx.x, phi, rho, sp, tl, cf, colour, it, maxd, ctl);
}

// ---- mixed precision SOR: defect correction --------------------------------
// With the FP64 potential phi held fixed, the red-black iteration for
// phi + delta, delta starting at 0, is
//   delta* = (sum_axes (delta_+ + delta_-)/h^2 + R) inv,
//   delta  = (1 - omega) delta + omega delta*,
// where R = sum_axes (phi_+ + phi_-)/h^2 + 4 pi e^2 rho - phi / inv is the
// FP64 residual of phi: the same iterates as the SOR on phi itself and the same
// max |update|, carried in FP32 on the small correction delta (half the bytes
// per sweep).  R is computed once per solve in FP64 and stored in FP32.

// The FP32 residual at node qq of colour `colour` (split index; its (i, j, h)
// given): R for an interior node, zero elsewhere.
template <typename RT>
__device__ __forceinline__ float sor_defect_at(const double *__restrict__ phi, const RT *__restrict__ rho, Split sp,
                                               const SorCoef &cf, int colour, long long qq, int i, int j, int h) {
    const int off = (i + j + colour) & 1;
    const int k = 2 * h + off;
    if (!(i >= 1 && i <= sp.nx - 2 && j >= 1 && j <= sp.ny - 2 && k >= 1 && k <= sp.nz - 2)) return 0.0f;
    const long long q = qq + (colour ? sp.Gh : 0);
    const double *oth = phi + (colour ? 0 : sp.Gh);
    const long long sx = (long long)sp.ny * sp.nzh;
    const double ax = (oth[qq + sx] + oth[qq - sx]) * cf.ydx2;
    const double ay = (oth[qq + sp.nzh] + oth[qq - sp.nzh]) * cf.ydy2;
    const double az = (oth[qq + off] + oth[qq + off - 1]) * cf.ydz2;
    return (float)(ax + ay + az + cf.src * (double)rho[q] - phi[q] / cf.inv_dx2);
}

// R (FP32, split) for every node; delta set to zero.
template <typename RT>
__device__ __forceinline__ void sor_defect_body(const int bid, const double *__restrict__ phi,
                                                const RT *__restrict__ rho, float *__restrict__ R,
             float *__restrict__ delta, Split sp, SorCoef cf) {
    const long long q = (long long)bid * blockDim.x + threadIdx.x;
    if (q >= 2 * sp.Gh) return;
    const int colour = q >= sp.Gh ? 1 : 0;
    const long long qq = q - (colour ? sp.Gh : 0);
    const int h = (int)(qq % sp.nzh);
    const long long row = qq / sp.nzh;
    const int j = (int)(row % sp.ny), i = (int)(row / sp.ny);
    R[q] = sor_defect_at<RT>(phi, rho, sp, cf, colour, qq, i, j, h);
    delta[q] = 0.0f;
}

template <typename RT>
__global__ void __launch_bounds__(kBlock)
k_sor_defect(const double *__restrict__ phi, const RT *__restrict__ rho, float *__restrict__ R,
             float *__restrict__ delta, Split sp, SorCoef cf) {
    sor_defect_body<RT>(blockIdx.x, phi, rho, R, delta, sp, cf);
}

// The words the LDS-resident FP32 solve expects cleared (its slot rows to
// -1, its exchange words and error word to 0), cleared by the defect launch
// that precedes it instead of by launches of their own.
struct ResClears {
    float *slots;
    long long nslots;
    unsigned long long *x;
    long long nx;
    int *err;
    // (k_sor_extrapolate only) the first maxd_rows rows of the FP64 sweeps'
    // slot maxima and their SorCtl, as k_sor_chunk_reset clears them at the
    // start of a solve
    unsigned long long *maxd;
    int maxd_rows;
    SorCtl *ctl;
};

template <typename RT>
__global__ void __launch_bounds__(kBlock)
k_sor_defect_clear(const double *__restrict__ phi, const RT *__restrict__ rho, float *__restrict__ R,
                   float *__restrict__ delta, Split sp, SorCoef cf, ResClears cl) {
    const long long q0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = q0; e < cl.nslots; e += stride) cl.slots[e] = -1.0f;
    for (long long e = q0; e < cl.nx; e += stride) cl.x[e] = 0ULL;
    if (q0 < 4) cl.err[q0] = 0;
    sor_defect_body<RT>(blockIdx.x, phi, rho, R, delta, sp, cf);
}

template <typename RT, int P = kSorPlanes>
__device__ __forceinline__ void sor_colour_f32_body(const int bid, float *__restrict__ delta,
                                                    const float *__restrict__ R, Split sp, SorTiles tl, SorCoef cf,
                                                    int colour, int it, unsigned long long *__restrict__ maxd,
                                                    SorCtl *__restrict__ ctl) {
    if (ctl->done) return;
    if (colour == 0 && it > 0) {
        const double prev = sor_row_max(maxd, it - 1);
        if (prev < cf.tol) {
            if (threadIdx.x == 0 && threadIdx.y == 0 && bid == 0) {
                ctl->done = 1;
                ctl->iters = it;
            }
            return;
        }
    }
    const int b = bid;
    const int t = (b % kXcds) * tl.per_xcd + b / kXcds;
    const int ntiles = tl.n_ic * tl.n_jb * tl.n_hb;
    float diff = 0.0f;
    if (t < ntiles) {
        const int ic = t / (tl.n_jb * tl.n_hb);
        const int rem = t - ic * tl.n_jb * tl.n_hb;
        const int jb = rem / tl.n_hb, hb = rem - jb * tl.n_hb;
        const int j = 1 + jb * kSorRows + threadIdx.y;
        const int h = hb * 64 + threadIdx.x;
        const int ny = sp.ny, nz = sp.nz, nzh = sp.nzh;
        float *__restrict__ dst = delta + (colour ? sp.Gh : 0);
        const float *__restrict__ oth = delta + (colour ? 0 : sp.Gh);
        const float *__restrict__ src = R + (colour ? sp.Gh : 0);
        const float yx = (float)cf.ydx2, yy = (float)cf.ydy2, yz = (float)cf.ydz2, inv = (float)cf.inv_dx2;
        const float om = (float)cf.omega, om1 = (float)cf.one_minus_omega;
        const long long sx = (long long)ny * nzh;
        float sum[P], old[P];
        long long qq[P];
        bool ok[P];
#pragma unroll
        for (int di = 0; di < P; di++) {
            const int i = 1 + ic * P + di;
            const int off = (i + j + colour) & 1;
            const int k = 2 * h + off;
            ok[di] = j < ny - 1 && i < sp.nx - 1 && k >= 1 && k <= nz - 2;
            const long long q = ((long long)i * ny + j) * nzh + h;
            qq[di] = q;
            sum[di] = 0.0f;
            old[di] = 0.0f;
            if (ok[di]) {
                sum[di] = sor_f32_sum(oth[q - sx], oth[q + sx], oth[q - nzh], oth[q + nzh], oth[q + off - 1],
                                      oth[q + off], src[q], yx, yy, yz);
                old[di] = dst[q];
            }
        }
#pragma unroll
        for (int di = 0; di < P; di++) {
            if (!ok[di]) continue;
            const float nv = sor_f32_relax(old[di], sum[di], inv, om, om1);
            dst[qq[di]] = nv;
            const float d = fabsf(nv - old[di]);
            diff = d > diff ? d : diff;
        }
    }
    block_max_atomic((double)diff, &maxd[((long long)it * kMaxSlots + (b % kMaxSlots)) * kSlotStride]);
}

template <typename RT>
__global__ void __launch_bounds__(256)
k_sor_colour_f32(float *__restrict__ delta, const float *__restrict__ R, Split sp, SorTiles tl, SorCoef cf,
                 int colour, int it, unsigned long long *__restrict__ maxd, SorCtl *__restrict__ ctl) {
    sor_colour_f32_body<RT>(blockIdx.x, delta, R, sp, tl, cf, colour, it, maxd, ctl);
}

// phi += delta (both colours; delta is zero off the interior)
__device__ __forceinline__ void sor_fold_body(const int bid, double *__restrict__ phi,
                                              const float *__restrict__ delta, long long n) {
    const long long q = (long long)bid * blockDim.x + threadIdx.x;
    if (q < n) phi[q] += (double)delta[q];
}

__global__ void k_sor_fold(double *__restrict__ phi, const float *__restrict__ delta, long long n) {
    sor_fold_body(blockIdx.x, phi, delta, n);
}

// Start of a mixed solve in a time loop, extrapolated in time from the
// potentials of the last solves: P_0 = phi (the last solve, phi_n) and
// P_k = phi_(n-k), k = 1..4, held in a ring of five history buffers (P_k in
// hist[(pos + k - 1) % 5]).  Predictor `id` (HV_OPT_SOR_EXTRAPOLATE) gives
// weights w_k (sum 1) over its first npts entries, the value at t_(n+1) of a
// polynomial fitted to them:
//   1  linear through 2 points          2 phi_n - phi_(n-1)
//   2  quadratic through 3 points       3 phi_n - 3 phi_(n-1) + phi_(n-2)
//   3  least-squares quadratic, 5 points (1.8, 0, -0.8, -0.6, 0.6)
//   4  least-squares quadratic, 4 points (2.25, -0.75, -1.25, 0.75)
//   5  least-squares line, 3 points      (4, 1, -2) / 3
// Each solve leaves an error of the stop rule's size that differs from step to
// step; a predictor multiplies that part by sqrt(sum w_k^2) (2.2, 4.4, 2.1,
// 2.8, 1.5), so the least-squares forms keep the next start close where the
// interpolating quadratic amplifies the previous solves' errors.
// The guess is phi_n + sum_k w_k (P_k - phi_n), exactly phi_n where the
// entries agree (the fixed Dirichlet faces).  phi_n then enters the ring at
// hist[(pos + 4) % 5] (the oldest entry, which no predictor reads).
__host__ __device__ __forceinline__ int sor_predictor(int id, double w[5]) {
    for (int k = 0; k < 5; k++) w[k] = 0.0;
    switch (id) {
    case 1: w[0] = 2.0; w[1] = -1.0; return 2;
    case 2: w[0] = 3.0; w[1] = -3.0; w[2] = 1.0; return 3;
    case 3: w[0] = 1.8; w[1] = 0.0; w[2] = -0.8; w[3] = -0.6; w[4] = 0.6; return 5;
    case 4: w[0] = 2.25; w[1] = -0.75; w[2] = -1.25; w[3] = 0.75; return 4;
    case 5: w[0] = 4.0 / 3.0; w[1] = 1.0 / 3.0; w[2] = -2.0 / 3.0; return 3;
    default: w[0] = 1.0; return 1;
    }
}

// Iteration budget of the FP64 sweeps that end a mixed solve: what is left of
// max_it after the FP32 correction iterations, and at least a few sweeps, so
// that a solve that already spent its budget in FP32 stops (not converged)
// about where the FP64 form would.
__host__ __device__ __forceinline__ int mixed_fp64_budget(int max_it, int iters32) {
    const int left = max_it - iters32;
    return left > 4 ? left : 4;
}

// The predictor a loop solve uses: the chosen one once the ring holds its
// points (hist_n counts phi and the ring's converged solves), else a line
// through the last two, else none.
__host__ __device__ __forceinline__ int sor_predictor_for(int chosen, int hist_n) {
    double w[5];
    const int need = sor_predictor(chosen, w);
    if (hist_n >= need) return chosen;
    return hist_n >= 2 ? 1 : 0;
}

struct SorHist {
    double *h[5];
};

// hist.h[k] holds P_(k+1) for the predictor's reads (k < npts - 1); hist.h[4]
// is the slot phi_n moves into (sor_hist_order).  NPTS: the predictor's point
// count.
template <int NPTS>
__device__ __forceinline__ void sor_extrapolate_body(const int bid, double *__restrict__ phi, const SorHist &hist,
                                                     int id, long long n) {
    const long long q = (long long)bid * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double w[5];
    (void)sor_predictor(id, w);
    double pk[NPTS > 1 ? NPTS - 1 : 1];
#pragma unroll
    for (int k = 1; k < NPTS; k++) pk[k - 1] = hist.h[k - 1][q];
    const double p0 = phi[q];
    double g = p0;
#pragma unroll
    for (int k = 1; k < NPTS; k++) g += w[k] * (pk[k - 1] - p0);
    hist.h[4][q] = p0;
    phi[q] = g;
}

template <int NPTS>
__global__ void __launch_bounds__(kBlock) k_sor_extrapolate(double *__restrict__ phi, SorHist hist, int id,
                                                            long long n, ResClears cl) {
    // (with the LDS-resident correction that computes its own residual, this
    // launch also clears that solve's words, as k_sor_defect_clear does)
    const long long q0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = q0; e < cl.nslots; e += stride) cl.slots[e] = -1.0f;
    for (long long e = q0; e < cl.nx; e += stride) cl.x[e] = 0ULL;
    if (cl.err && q0 < 4) cl.err[q0] = 0;
    for (long long e = q0; e < (long long)cl.maxd_rows * kMaxSlots; e += stride) cl.maxd[e * kSlotStride] = 0ULL;
    if (cl.ctl && q0 == 0) {
        cl.ctl->done = 0;
        cl.ctl->iters = 0;
    }
    sor_extrapolate_body<NPTS>(blockIdx.x, phi, hist, id, n);
}

// the ring's entries in predictor order for position pos: P_k (k = 1..4) at
// ring[(pos + k - 1) % 5], the newest entry's slot at ring[(pos + 4) % 5]
__host__ __device__ __forceinline__ SorHist sor_hist_order(const SorHist &ring, int pos) {
    SorHist o;
    for (int k = 0; k < 5; k++) {
        int slot = pos + k;
        slot -= slot >= 5 ? 5 : 0;
        o.h[k] = ring.h[slot];
    }
    return o;
}

// ---- one launch per SOR iteration: red and black fused ------------------------
// Out of place (in -> out).  A block owns FT x FT grid rows (all k): it first
// computes the new red values of its rows and of the one-row ring around them
// into LDS (the ring's are the same values the neighbouring blocks compute for
// their own rows: same operands, same operations), then the new black values
// of its rows from them, and writes both colours of its rows to `out`.
// Boundary nodes are copied, so `out` is complete.  Per iteration it reads the
// potential and the source once and writes the potential once (plus ring
// re-reads served by L2), and pays one launch instead of two.
//
// EXACT (FP64 potential): the reference's update on phi, with the division by
// dx^2 as hv_div_fixed (bit-identical).  !EXACT (FP32 correction of the mixed
// form): the delta iteration with the FP64 residual R as source.
constexpr int kFT = 8;            // tile side in rows
constexpr int kFThreads = 512;

struct FusedTiles {
    int ti, tj;      // tiles along i and j
    int per_xcd;     // tiles per XCD slice of the grid
};

template <bool EXACT, typename VT, typename ST>
__device__ __forceinline__ VT sor_update(VT xm, VT xp, VT ym, VT yp, VT zm, VT zp, VT old, ST srcv,
                                         const SorCoef &cf) {
    if constexpr (EXACT) {
        const double phi_x = hv_div_fixed(xp + xm, cf.dx2, cf.ydx2);
        const double phi_y = hv_div_fixed(yp + ym, cf.dy2, cf.ydy2);
        const double phi_z = hv_div_fixed(zp + zm, cf.dz2, cf.ydz2);
        const double phi_star = (phi_x + phi_y + phi_z + cf.src * (double)srcv) * cf.inv_dx2;
        return cf.one_minus_omega * old + cf.omega * phi_star;
    } else {
        const float s = sor_f32_sum(xm, xp, ym, yp, zm, zp, (float)srcv, (float)cf.ydx2, (float)cf.ydy2,
                                    (float)cf.ydz2);
        return sor_f32_relax(old, s, (float)cf.inv_dx2, (float)cf.omega, (float)cf.one_minus_omega);
    }
}

template <bool EXACT, typename VT, typename ST>
__global__ void __launch_bounds__(kFThreads)
k_sor_fused(const VT *__restrict__ in, VT *__restrict__ out, const ST *__restrict__ src, Split sp, FusedTiles ft,
            SorCoef cf, int it, unsigned long long *__restrict__ maxd, SorCtl *__restrict__ ctl) {
    extern __shared__ unsigned char smem[];
    VT *rn = (VT *)smem;  // new red values, (kFT + 2)^2 rows x nzh
    if (ctl->done) return;
    if (it > 0) {
        const double prev = sor_row_max(maxd, it - 1);
        if (prev < cf.tol) {
            if (threadIdx.x == 0 && blockIdx.x == 0) {
                ctl->done = 1;
                ctl->iters = it;
            }
            return;
        }
    }
    const int b = blockIdx.x;
    const int t = (b % kXcds) * ft.per_xcd + b / kXcds;
    const int nx = sp.nx, ny = sp.ny, nz = sp.nz, nzh = sp.nzh;
    const long long sx = (long long)ny * nzh;
    const VT *in_r = in, *in_b = in + sp.Gh;
    const ST *src_r = src, *src_b = src + sp.Gh;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nwaves = blockDim.x / kWave;
    const int RW = kFT + 2;
    double diff = 0.0;
    const bool active = t < ft.ti * ft.tj;
    const int i0 = active ? (t / ft.tj) * kFT : 0, j0 = active ? (t % ft.tj) * kFT : 0;
    if (active) {
        // phase A: red on own rows + ring
        for (int r = wave; r < RW * RW; r += nwaves) {
            const int li = r / RW - 1, lj = r % RW - 1;
            const int i = i0 + li, j = j0 + lj;
            if (i < 0 || i >= nx || j < 0 || j >= ny) continue;
            const int off = (i + j) & 1;
            const long long row = ((long long)i * ny + j) * nzh;
            const bool inner_row = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2;
            for (int h = lane; h < nzh; h += kWave) {
                const long long q = row + h;
                const int k = 2 * h + off;
                VT v = in_r[q];
                if (inner_row && k >= 1 && k <= nz - 2) {
                    const VT nv = sor_update<EXACT, VT, ST>(in_b[q - sx], in_b[q + sx], in_b[q - nzh], in_b[q + nzh],
                                                           in_b[q + off - 1], in_b[q + off], v, src_r[q], cf);
                    if (li >= 0 && li < kFT && lj >= 0 && lj < kFT) {
                        const double d = fabs((double)nv - (double)v);
                        diff = d > diff ? d : diff;
                    }
                    v = nv;
                }
                rn[r * nzh + h] = v;
            }
        }
    }
    __syncthreads();
    if (active) {
        // phase B: black on own rows; write both colours
        for (int r = wave; r < kFT * kFT; r += nwaves) {
            const int li = r / kFT, lj = r % kFT;
            const int i = i0 + li, j = j0 + lj;
            if (i >= nx || j >= ny) continue;
            const int off = (i + j + 1) & 1;
            const long long row = ((long long)i * ny + j) * nzh;
            const int rr = (li + 1) * RW + (lj + 1);  // LDS row of (i, j)
            const bool inner_row = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2;
            for (int h = lane; h < nzh; h += kWave) {
                const long long q = row + h;
                const int k = 2 * h + off;
                VT v = in_b[q];
                if (inner_row && k >= 1 && k <= nz - 2) {
                    const VT *c0 = rn + rr * nzh;
                    const VT nv = sor_update<EXACT, VT, ST>(rn[(rr - RW) * nzh + h], rn[(rr + RW) * nzh + h],
                                                           rn[(rr - 1) * nzh + h], rn[(rr + 1) * nzh + h],
                                                           c0[h + off - 1], c0[h + off], v, src_b[q], cf);
                    const double d = fabs((double)nv - (double)v);
                    diff = d > diff ? d : diff;
                    v = nv;
                }
                out[sp.Gh + q] = v;
                out[q] = rn[rr * nzh + h];
            }
        }
    }
    block_max_atomic(diff, &maxd[((long long)it * kMaxSlots + (b % kMaxSlots)) * kSlotStride]);
}

// Folds the slot maxima of iterations [first, last) into out[0..last-first).
__device__ __forceinline__ void sor_collect_body(const int bid, const unsigned long long *__restrict__ maxd,
                                                 double *__restrict__ out, int first,
                              int last) {
    const int it = first + bid;
    if (it >= last) return;
    const double m = sor_row_max(maxd, it);
    if (threadIdx.x == 0) out[bid] = m;
}

__global__ void k_sor_collect(const unsigned long long *__restrict__ maxd, double *__restrict__ out, int first,
                              int last) {
    sor_collect_body(blockIdx.x, maxd, out, first, last);
}

// The same, and (xres non-null) the LDS-resident FP32 solve's iteration
// count, tested max and error word after the chunk's maxima, so that one
// read-back carries both.
__global__ void k_sor_collect_x(const unsigned long long *__restrict__ maxd, double *__restrict__ out, int first,
                                int last, const float *__restrict__ xres, const int *__restrict__ xerr) {
    if (xres && blockIdx.x == 0 && threadIdx.x == 0) {
        out[last - first] = (double)xres[0];
        out[last - first + 1] = (double)xres[1];
        out[last - first + 2] = (double)xerr[0];
    }
    sor_collect_body(blockIdx.x, maxd, out, first, last);
}

// Start of a chunk of SOR iterations [r0, r1): their rows of the slot maxima
// zeroed (one word per slot), and at the start of a solve (reset_ctl) SorCtl
// cleared; grid of at least (r1 - r0) kMaxSlots threads.
__global__ void k_sor_chunk_reset(unsigned long long *__restrict__ maxd, int r0, int r1, SorCtl *__restrict__ ctl,
                                  int reset_ctl) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (long long)(r1 - r0) * kMaxSlots) maxd[((long long)r0 * kMaxSlots + e) * kSlotStride] = 0ULL;
    if (reset_ctl && e == 0) {
        ctl->done = 0;
        ctl->iters = 0;
    }
}

}  // namespace hv
