// hv_sor_owned.h -- the mixed form's FP32 correction iteration with every
// node of a tile owned by a fixed lane.
//
// res32o_solve_tile runs the solve of res32_solve_tile (the same red-black
// updates on the same operand values, the same half-sweep order, exchange
// words and lagged stop test, so a tile solved by either body gives the same
// bits), arranged so that a node update costs little index arithmetic:
//  * LDS holds each colour's values of the tile with a one-row border on all
//    four sides, where the neighbours' face values land: tile row (li, lj) is
//    LDS row (li + 1) * QJ + lj + 1 of P slots.  A node then reads its six
//    neighbours at fixed offsets, on the tile's faces as inside it.
//  * The (row, slot) elements of each row set (interior rows, ring rows) are
//    cut into 64-element chunks; chunk c is slot c / nwaves of wave
//    c % nwaves.  Per slot a lane keeps one descriptor word (LDS address, the
//    k parity and update flag of each colour) and both colours' right-hand
//    sides in registers, set up once per solve.
#pragma once

#include "hv_sor_resident.h"

namespace hv {

// LDS of the lane-owned body, in floats: the colour-0 array of (IM + 2) * QJ
// rows of P slots, 64 sink words, the colour-1 array, 64 sink words and a
// guard of QJ * P + 2 words, then the wave partials, the error flag and the
// two tested maxima (all in the dynamic LDS; the body declares none of its
// own).  Lanes without a node (the tail of a partial chunk) read around and
// store into the sinks.
template <int P, int QJ, int IM>
struct OwnedLayout {
    static constexpr int E = (IM + 2) * QJ * P;  // one colour's array
    static constexpr int ext1 = E + kWave;       // start of the colour-1 array
    static constexpr int red = ext1 + E + kWave + QJ * P + 2;
    static constexpr size_t bytes(int nwaves) { return sizeof(float) * ((size_t)red + (size_t)nwaves + 3); }
};

// does tiling t fit the lane-owned body <SI, SR, P, QJ, IM> of NT threads
// (slots per lane: SI interior, SR ring)
template <int SI, int SR, int P, int QJ, int IM, int NT>
__host__ __device__ inline bool owned_fits(const ResTiles &t) {
    if (t.TI > IM || t.TJ + 2 > QJ || t.nzh > P || t.TI < 1 || t.TJ < 1) return false;
    if (t.xstride > kFetchMax * NT) return false;
    const bool split_rows = t.TI >= 3 && t.TJ >= 3;
    const int n_inner = split_rows ? (t.TI - 2) * (t.TJ - 2) : 0;
    const int n_ring = t.TI * t.TJ - n_inner;
    const int nw = NT / kWave;
    const int ch_in = (n_inner * t.nzh + kWave - 1) / kWave, ch_rg = (n_ring * t.nzh + kWave - 1) / kWave;
    // the descriptors' LDS address field has 20 bits: the arrays stay below 2^18 words
    return ch_in <= SI * nw && ch_rg <= SR * nw && OwnedLayout<P, QJ, IM>::red < (1 << 18);
}

typedef __attribute__((address_space(3))) float lds_float;

// max over a wave of non-negative, non-NaN floats (their bit patterns order
// them), in every lane: DPP row shifts and row broadcasts into lane 63
__device__ __forceinline__ float wave_max_nonneg(float v) {
    unsigned u = __float_as_uint(v);
    u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x111, 0xf, 0xf, false));  // row_shr:1
    u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x112, 0xf, 0xf, false));  // row_shr:2
    u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x114, 0xf, 0xf, false));  // row_shr:4
    u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x118, 0xf, 0xf, false));  // row_shr:8
    u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x142, 0xa, 0xf, false));  // row_bcast:15
    u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return __uint_as_float((unsigned)__builtin_amdgcn_readlane((int)u, 63));
}

template <int SI, int SR, int P, int QJ, int IM, int NT = kRes32Threads>
__device__ __forceinline__ void res32o_solve_tile(float *__restrict__ delta, const float *__restrict__ R, Split sp,
                                                  ResTiles tl, SorCoef cf, int max_it, int lag, ResSync32 sy,
                                                  const int t) {
    using L = OwnedLayout<P, QJ, IM>;
    extern __shared__ unsigned char smem[];
    float *const x0 = (float *)smem;  // colour-0 values (bordered rows)
    float *const x1 = x0 + L::ext1;   // colour-1 values
    float *const red = x0 + L::red;   // wave partials
    constexpr int nwaves = NT / kWave;
    constexpr int XS = QJ * P;  // distance of the i-neighbours
    constexpr int S = SI + SR;
    const int TI = tl.TI, TJ = tl.TJ, nzh = tl.nzh;
    const int side_i = TJ * nzh, side_j = TI * nzh, xstride = 2 * side_i + 2 * side_j;
    const int nx = sp.nx, ny = sp.ny, nz = sp.nz;
    const int nT = tl.PI * tl.PJ;
    const int pi = t / tl.PJ, pj = t - pi * tl.PJ;
    const int i0 = pi * TI, j0 = pj * TJ;
    const int tid = threadIdx.x, lane = tid & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    const bool has_im = pi > 0, has_ip = pi < tl.PI - 1, has_jm = pj > 0, has_jp = pj < tl.PJ - 1;
    const float yx = (float)cf.ydx2, yy = (float)cf.ydy2, yz = (float)cf.ydz2, inv = (float)cf.inv_dx2;
    const float om = (float)cf.omega, om1 = (float)cf.one_minus_omega, tol = (float)cf.tol;
    int &s_bad = *(int *)(red + nwaves);
    float *const s_max = red + nwaves + 1;  // by iteration parity: written in iteration it, read at its end
    if (tid == 0) s_bad = 0;

    // the tile's values of both colours (rows outside the grid hold 0)
    for (int e = tid; e < TI * TJ * nzh; e += NT) {
        const int r = e / nzh, h = e - r * nzh;
        const int li = r / TJ, lj = r - li * TJ;
        const int i = i0 + li, j = j0 + lj;
        float v0 = 0.0f, v1 = 0.0f;
        if (i < nx && j < ny) {
            const long long q = ((long long)i * ny + j) * nzh + h;
            v0 = delta[q];
            v1 = delta[sp.Gh + q];
        }
        const int w = ((li + 1) * QJ + lj + 1) * P + h;
        x0[w] = v0;
        x1[w] = v1;
    }
    // black values around the tile for the first red sweep
    {
        const float *src = delta + sp.Gh;
        for (int e = tid; e < side_i; e += NT) {
            const int lj = e / nzh, h = e - lj * nzh, j = j0 + lj;
            if (j >= ny) continue;
            if (has_im) x1[(lj + 1) * P + h] = src[((long long)(i0 - 1) * ny + j) * nzh + h];
            if (has_ip) x1[((TI + 1) * QJ + lj + 1) * P + h] = src[((long long)(i0 + TI) * ny + j) * nzh + h];
        }
        for (int e = tid; e < side_j; e += NT) {
            const int li = e / nzh, h = e - li * nzh, i = i0 + li;
            if (i >= nx) continue;
            if (has_jm) x1[(li + 1) * QJ * P + h] = src[((long long)i * ny + (j0 - 1)) * nzh + h];
            if (has_jp) x1[((li + 1) * QJ + TJ + 1) * P + h] = src[((long long)i * ny + (j0 + TJ)) * nzh + h];
        }
    }

    // Row sets: the ring of rows on the tile's faces (the only rows that read
    // border values) and the interior rows.  Tiles thinner than 3 rows in i
    // or j have no interior; their ring is every row.
    const bool split_rows = TI >= 3 && TJ >= 3;
    const int n_inner = split_rows ? (TI - 2) * (TJ - 2) : 0;
    const int nel_in = n_inner * nzh, nel_rg = (TI * TJ - n_inner) * nzh;
    const int nch_in = (nel_in + kWave - 1) / kWave, nch_rg = (nel_rg + kWave - 1) / kWave;
    // slots this wave holds in each set (wave-uniform)
    const int ns_in = nch_in > wave ? (nch_in - wave + nwaves - 1) / nwaves : 0;
    const int ns_rg = nch_rg > wave ? (nch_rg - wave + nwaves - 1) / nwaves : 0;

    // Descriptor of a slot: bits 0-19 the LDS address of the node's
    // i-minus neighbour in the colour-0 array, bit 20 + c the k parity of colour c (its
    // k neighbours are slots h + parity - 1 and h + parity of the other
    // colour), bit 22 + c set when colour c's node is updated (not on the
    // grid's faces, not outside the grid, not a sink).
    const unsigned lds0 = (unsigned)(size_t)(lds_float *)x0;
    unsigned D[S];
    float R0[S], R1[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const bool inner = s < SI;
        const int ss = inner ? s : s - SI;
        const int f = (ss * nwaves + wave) * kWave + lane;
        unsigned d = lds0 + 4u * (unsigned)(L::E + lane - XS);  // sink
        float r0 = 0.0f, r1 = 0.0f;
        if (f < (inner ? nel_in : nel_rg)) {
            const int q = f / nzh, h = f - q * nzh;
            int li, lj;
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
            const int i = i0 + li, j = j0 + lj;
            const unsigned o0 = (unsigned)(i + j) & 1u, o1 = o0 ^ 1u;
            const int k0 = 2 * h + (int)o0, k1 = 2 * h + (int)o1;
            const bool okij = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2;
            const unsigned ok0 = okij && k0 >= 1 && k0 <= nz - 2, ok1 = okij && k1 >= 1 && k1 <= nz - 2;
            const int w = ((li + 1) * QJ + lj + 1) * P + h;
            d = (lds0 + 4u * (unsigned)(w - XS)) | o0 << 20 | o1 << 21 | ok0 << 22 | ok1 << 23;
            if (i < nx && j < ny) {
                const long long g = ((long long)i * ny + j) * nzh + h;
                r0 = R[g];
                r1 = R[sp.Gh + g];
            }
        }
        D[s] = d;
        R0[s] = r0;
        R1[s] = r1;
    }

    // this thread's exchange words (at most kFetchMax; the host checks that
    // xstride fits): word x = tid + q * NT of the tile's slice, the byte
    // offset of its counterpart in the exchange array (bit q of fmask: there
    // is a neighbour), the border slot it fills and the face slot it
    // publishes.  32-bit offsets from the array bases keep them one register
    // each.
    unsigned src_b[kFetchMax], fmask = 0;
    int fdst[kFetchMax], psrc[kFetchMax];
    int nfetch = 0;
    for (int x = tid; x < xstride && nfetch < kFetchMax; x += NT) {
        int off = -1, fd, ps;
        if (x < side_i) {
            const int lj = x / nzh, h = x - lj * nzh;
            if (has_im) off = (t - tl.PJ) * xstride + side_i + x;  // i-high face of (pi-1, pj)
            fd = (lj + 1) * P + h;
            ps = (QJ + lj + 1) * P + h;
        } else if (x < 2 * side_i) {
            const int y = x - side_i, lj = y / nzh, h = y - lj * nzh;
            if (has_ip) off = (t + tl.PJ) * xstride + y;  // i-low face of (pi+1, pj)
            fd = ((TI + 1) * QJ + lj + 1) * P + h;
            ps = (TI * QJ + lj + 1) * P + h;
        } else if (x < 2 * side_i + side_j) {
            const int y = x - 2 * side_i, li = y / nzh, h = y - li * nzh;
            if (has_jm) off = (t - 1) * xstride + 2 * side_i + side_j + y;
            fd = (li + 1) * QJ * P + h;
            ps = ((li + 1) * QJ + 1) * P + h;
        } else {
            const int y = x - 2 * side_i - side_j, li = y / nzh, h = y - li * nzh;
            if (has_jp) off = (t + 1) * xstride + 2 * side_i + y;
            fd = ((li + 1) * QJ + TJ + 1) * P + h;
            ps = ((li + 1) * QJ + TJ) * P + h;
        }
        src_b[nfetch] = off >= 0 ? 8u * (unsigned)off : 0u;
        if (off >= 0) fmask |= 1u << nfetch;
        fdst[nfetch] = fd;
        psrc[nfetch] = ps;
        nfetch++;
    }
    __syncthreads();

    // the update of k_sor_colour_f32 over this lane's slots of one row set:
    // colour C reads the other colour's array and writes its own
    auto sweep = [&](auto cc, auto inn, float m) -> float {
        constexpr int C = decltype(cc)::value;
        constexpr bool INNER = decltype(inn)::value;
        constexpr int B = INNER ? 0 : SI, NS = INNER ? SI : SR;
        const int ns = INNER ? ns_in : ns_rg;
        constexpr unsigned OB = C ? 0 : sizeof(float) * L::ext1, DB = C ? sizeof(float) * L::ext1 : 0;
        // addresses and flags are derived from the descriptors in every sweep
        // rather than held across sweeps (registers for the slots)
#pragma unroll
        for (int u = 0; u < NS; u++) asm volatile("" : "+v"(D[B + u]));
#pragma unroll
        for (int u = 0; u < NS; u++) {
            if (u >= ns) break;
            const int s = B + u;
            const unsigned d = D[s];
            const unsigned a = d & 0xfffffu;
            const lds_float *o = (const lds_float *)(size_t)(a + OB);
            const float xm = o[0], ym = o[XS - P], yp = o[XS + P], xp = o[2 * XS];
            const lds_float *oz = o + ((d >> (20 + C)) & 1u);
            const float zm = oz[XS - 1], zp = oz[XS];
            const float sv = sor_f32_sum(xm, xp, ym, yp, zm, zp, C ? R1[s] : R0[s], yx, yy, yz);
            lds_float *dv = (lds_float *)(size_t)(a + DB + sizeof(float) * XS);
            const float po = *dv;
            float nv = sor_f32_relax(po, sv, inv, om, om1);
            nv = ((d >> (22 + C)) & 1u) ? nv : po;
            *dv = nv;
            m = fmaxf(m, fabsf(nv - po));
        }
        return m;
    };
    using C0 = std::integral_constant<int, 0>;
    using C1 = std::integral_constant<int, 1>;
    using In = std::integral_constant<bool, true>;
    using Rg = std::integral_constant<bool, false>;

    // store the tile's colour faces as tagged words (no completion wait: each
    // word carries its own tag)
    // (the opaque copies keep the compiler from holding 64-bit addresses of
    // every exchange word across the iteration loop)
    auto publish = [&](const float *xc, unsigned long long *xw, unsigned tag) {
        unsigned b = 8u * (unsigned)(t * xstride + tid);
        asm volatile("" : "+v"(b));
#pragma unroll
        for (int q = 0; q < kFetchMax; q++)
            if (q < nfetch) st_tagged((unsigned long long *)((char *)xw + (b + 8u * q * NT)), xc[psrc[q]], tag);
    };

    // poll this thread's neighbour words until they carry `tag` and put them
    // in the border of array xc; all loads of a round are in flight together.
    // Returns nonzero (block-uniform) on error.
    auto fetch = [&](float *xc, const unsigned long long *xs, unsigned tag) -> int {
        unsigned pending = fmask;
        unsigned sb[kFetchMax];
#pragma unroll
        for (int q = 0; q < kFetchMax; q++) {
            sb[q] = src_b[q];
            asm volatile("" : "+v"(sb[q]));
        }
        long long spin = 0;
        int bad = 0;
        while (pending && !bad) {
            unsigned long long w[kFetchMax];
#pragma unroll
            for (int q = 0; q < kFetchMax; q++)
                if (pending & (1u << q)) w[q] = ld_tagged((const unsigned long long *)((const char *)xs + sb[q]));
#pragma unroll
            for (int q = 0; q < kFetchMax; q++) {
                if ((pending & (1u << q)) && (unsigned)(w[q] >> 32) >= tag) {
                    xc[fdst[q]] = __uint_as_float((unsigned)w[q]);
                    pending &= ~(1u << q);
                }
            }
            if (++spin > kSpinLimit) bad = 1;
            else if ((spin & 255) == 0 && ld_dev_i(sy.err)) bad = 1;
            // a wave still waiting yields its issue slots to the waves of
            // its SIMD that have work
            if (pending) __builtin_amdgcn_s_sleep(1);
        }
        if (bad) {
            st_dev_i(sy.err, 1);
            s_bad = 1;
        }
        __syncthreads();
        return s_bad;
    };

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

    // Iteration it (as res32_solve_tile):
    //   red interior -> [black border of it-1] -> red ring -> publish red
    //   -> black interior -> [red border of it] -> black ring -> publish black
    int iters = max_it;
    float tested = -1.0f;
    for (int it = 0; it < max_it; it++) {
        const unsigned tag = (unsigned)it + 1u;
        float m = sweep(C0{}, In{}, 0.0f);
        if (it >= lag && wave == 0) poll_slots(it - lag, it & 1);
        if (it > 0) {
            if (fetch(x1, sy.xb, tag - 1u)) break;
        } else {
            __syncthreads();
        }
        m = sweep(C0{}, Rg{}, m);
        __syncthreads();
        publish(x0, sy.xr, tag);
        m = sweep(C1{}, In{}, m);
        if (fetch(x0, sy.xr, tag)) break;
        m = sweep(C1{}, Rg{}, m);
        m = wave_max_nonneg(m);  // the lanes' maxima are never NaN (fmaxf)
        if (lane == 0) red[wave] = m;
        __syncthreads();
        publish(x1, sy.xb, tag);
        if (tid == 0) {
            float bm = 0.0f;
            for (int w = 0; w < nwaves; w++) bm = red[w] > bm ? red[w] : bm;
            st_dev_f(sy.slots + (long long)it * nT + t, bm);
        }
        if (it >= lag) {
            tested = s_max[it & 1];
            if (tested < tol) {
                iters = it + 1;
                break;
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < TI * TJ * nzh; e += NT) {
        const int r = e / nzh, h = e - r * nzh;
        const int li = r / TJ, lj = r - li * TJ;
        const int i = i0 + li, j = j0 + lj;
        if (i < nx && j < ny) {
            const long long q = ((long long)i * ny + j) * nzh + h;
            const int w = ((li + 1) * QJ + lj + 1) * P + h;
            delta[q] = x0[w];
            delta[sp.Gh + q] = x1[w];
        }
    }
    if (t == 0 && tid == 0) {
        sy.result[0] = (float)iters;
        sy.result[1] = tested;
    }
}

// ---- one exchange per iteration: the deep-halo lane-owned body ----------------
// res32d_solve_tile runs the same solve as res32o_solve_tile (same updates on
// the same operands, the same lagged stop test, so the same bits) with one
// neighbour exchange per iteration instead of two.  After its black sweep a
// tile publishes its black values of the two rows along each face and of its
// four corner rows (three values to two 64-bit words, each word with the low
// 16 bits of the iteration tag); a neighbour then has every black value within two rows
// of its tile (and its diagonal corners) and recomputes, redundantly, the red
// values of the one-row ring around its tile from them (the update its owner
// makes, on the same operands in the same order, so the same bits).  Its own
// black sweep then reads red values it holds, without a wait:
//   red interior -> [black border of it-1] -> red ring and border ring
//   -> black everything -> publish black faces (buffer it & 1)
// The exchange words are double-buffered by iteration parity: a tile
// overwrites a buffer two iterations after filling it, by which time each
// neighbour has read it (the neighbour's faces that this tile waited for in
// between were published after that read).  A word a reader polls therefore
// holds the tag it waits for or the one two iterations older (or the clear
// value 0 before its first write), so 16 bits of tag tell them apart.
//
// LDS: each colour's array holds (IM + 4) x QJ rows of P slots, tile row
// (li, lj) at (li + 2) * QJ + lj + 2 (two border rows on each side), then
// the sinks, guard and partials as in OwnedLayout.
template <int P, int QJ, int IM>
struct DeepLayout {
    static constexpr int E = (IM + 4) * QJ * P;
    static constexpr int ext1 = E + kWave;
    static constexpr int red = ext1 + E + kWave + QJ * P + 2;
    static constexpr size_t bytes(int nwaves) { return sizeof(float) * ((size_t)red + (size_t)nwaves + 3); }
};

// exchange words per thread of the deep-halo body: 4 with 1024-thread
// blocks (one tile per CU), 6 with 512-thread blocks (two per CU)
template <int NT>
constexpr int fetch_max_d() { return NT == 512 ? 6 : 4; }

// Exchange words of a section of n values: three values ride in two 64-bit
// words.  Word 2g holds value 3g (bits 0-31) and the low half of value 3g + 1
// (bits 32-47), word 2g + 1 value 3g + 2 and the high half of 3g + 1; bits
// 48-63 of each carry the low 16 bits of the iteration tag.
__host__ __device__ inline int deep_words(int n) { return 2 * ((n + 2) / 3); }

// exchange words per tile and buffer: two rows along each face, four corners
__host__ __device__ inline int deep_xstride(const ResTiles &t) {
    return 4 * deep_words(t.TJ * t.nzh) + 4 * deep_words(t.TI * t.nzh) + 4 * deep_words(t.nzh);
}

// does tiling t fit the deep-halo body <SI, SR, SH, P, QJ, IM> of NT threads
// (slots per lane: SI interior, SR ring, SH border ring)
template <int SI, int SR, int SH, int P, int QJ, int IM, int NT>
__host__ __device__ inline bool deep_fits(const ResTiles &t) {
    if (t.TI > IM || t.TJ + 4 > QJ || t.nzh > P || t.TI < 3 || t.TJ < 3) return false;
    if (deep_xstride(t) > fetch_max_d<NT>() * NT) return false;
    const int n_inner = (t.TI - 2) * (t.TJ - 2), n_ring = t.TI * t.TJ - n_inner, n_halo = 2 * (t.TI + t.TJ);
    const int nw = NT / kWave;
    auto ch = [&](int rows) { return (rows * t.nzh + kWave - 1) / kWave; };
    return ch(n_inner) <= SI * nw && ch(n_ring) <= SR * nw && ch(n_halo) <= SH * nw &&
           DeepLayout<P, QJ, IM>::red < (1 << 18);
}

// MP: the convergence poll rides in the rounds of the neighbour fetch (one
// round trip fewer per iteration); without it wave 0 polls the slot row on
// its own before the fetch
// FUSE: the solve also takes the defect launch's and the fold's part for its
// tile: delta starts at zero in LDS, each slot's R is the FP32 residual of phi
// and rho computed here (sor_defect_at, the defect launch's arithmetic), and
// the result is added to phi (phi += (double) delta, the fold's arithmetic)
// instead of being stored; the same bits as defect, solve and fold in three
// launches.  No tile folds before every tile of the solve has computed its
// residual: a tile stops only after reading every tile's maximum of an
// iteration, which each tile writes after its set-up.
template <int SI, int SR, int SH, int P, int QJ, int IM, bool MP = true, int NT = kRes32Threads, bool FUSE = false>
__device__ __forceinline__ void res32d_solve_tile(float *__restrict__ delta, const float *__restrict__ R, Split sp,
                                                  ResTiles tl, SorCoef cf, int max_it, int lag, ResSync32 sy,
                                                  const int t, double *__restrict__ phi = nullptr,
                                                  const float *__restrict__ rho = nullptr) {
    using L = DeepLayout<P, QJ, IM>;
    static_assert(L::E < (1 << 16), "the exchange words' LDS slots are 16-bit fields");
    constexpr int kFetchMaxD = fetch_max_d<NT>();
    extern __shared__ unsigned char smem[];
    float *const x0 = (float *)smem;  // red values (bordered rows)
    float *const x1 = x0 + L::ext1;   // black values
    float *const red = x0 + L::red;   // wave partials
    constexpr int nwaves = NT / kWave;
    constexpr int XS = QJ * P;
    constexpr int S = SI + SR + SH;
    const int TI = tl.TI, TJ = tl.TJ, nzh = tl.nzh;
    const int side_i = TJ * nzh, side_j = TI * nzh, xstride = tl.xstride;
    const int nx = sp.nx, ny = sp.ny, nz = sp.nz;
    const int nT = tl.PI * tl.PJ, PJ = tl.PJ;
    const int pi = t / PJ, pj = t - pi * PJ;
    const int i0 = pi * TI, j0 = pj * TJ;
    const int tid = threadIdx.x, lane = tid & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    const bool has_im = pi > 0, has_ip = pi < tl.PI - 1, has_jm = pj > 0, has_jp = pj < PJ - 1;
    const float yx = (float)cf.ydx2, yy = (float)cf.ydy2, yz = (float)cf.ydz2, inv = (float)cf.inv_dx2;
    const float om = (float)cf.omega, om1 = (float)cf.one_minus_omega, tol = (float)cf.tol;
    int &s_bad = *(int *)(red + nwaves);
    float *const s_max = red + nwaves + 1;
    if (tid == 0) s_bad = 0;
    auto lrow = [&](int li, int lj) { return ((li + 2) * QJ + lj + 2) * P; };

    // the tile's values of both colours, the red values of the ring around it
    // and the black values of the two rows around it and of its corners
    // (values outside the grid are never read by an update)
    for (int e = tid; e < (TI + 4) * (TJ + 4) * nzh; e += NT) {
        const int r = e / nzh, h = e - r * nzh;
        const int li = r / (TJ + 4) - 2, lj = r - (li + 2) * (TJ + 4) - 2;
        const bool in_i = li >= 0 && li < TI, in_j = lj >= 0 && lj < TJ;
        const bool own = in_i && in_j;
        const bool ring1 = (in_j && (li == -1 || li == TI)) || (in_i && (lj == -1 || lj == TJ));
        const bool black = own || (in_j && !in_i) || (in_i && !in_j) ||
                           ((li == -1 || li == TI) && (lj == -1 || lj == TJ));
        const int i = i0 + li, j = j0 + lj;
        if (!(i >= 0 && i < nx && j >= 0 && j < ny)) continue;
        const long long q = ((long long)i * ny + j) * nzh + h;
        const int w = lrow(li, lj) + h;
        if constexpr (FUSE) {
            if (own || ring1) x0[w] = 0.0f;
            if (black) x1[w] = 0.0f;
        } else {
            if (own || ring1) x0[w] = delta[q];
            if (black) x1[w] = delta[sp.Gh + q];
        }
    }

    // Row sets: interior rows, the ring of the tile's face rows, and the ring
    // of rows just outside the tile (no corners), whose red values the black
    // sweep reads.
    const int n_inner = (TI - 2) * (TJ - 2);
    const int nel_in = n_inner * nzh, nel_rg = (TI * TJ - n_inner) * nzh, nel_hr = 2 * (TI + TJ) * nzh;
    auto nslots = [&](int nel) {
        const int nch = (nel + kWave - 1) / kWave;
        return nch > wave ? (nch - wave + nwaves - 1) / nwaves : 0;
    };
    const int ns_in = nslots(nel_in), ns_rg = nslots(nel_rg), ns_hr = nslots(nel_hr);

    // slot descriptors as in res32o_solve_tile; R1 for tile rows only (the
    // border ring is updated in red only)
    const unsigned lds0 = (unsigned)(size_t)(lds_float *)x0;
    unsigned D[S];
    float R0[S], R1[SI + SR];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int set = s < SI ? 0 : (s < SI + SR ? 1 : 2);
        const int ss = set == 0 ? s : (set == 1 ? s - SI : s - SI - SR);
        const int f = (ss * nwaves + wave) * kWave + lane;
        unsigned d = lds0 + 4u * (unsigned)(L::E + lane - XS);  // sink
        float r0 = 0.0f, r1 = 0.0f;
        if (f < (set == 0 ? nel_in : (set == 1 ? nel_rg : nel_hr))) {
            const int q = f / nzh, h = f - q * nzh;
            int li, lj;
            if (set == 0) {
                li = 1 + q / (TJ - 2);
                lj = 1 + q - (li - 1) * (TJ - 2);
            } else if (set == 1) {
                if (q < TJ) {
                    li = 0;
                    lj = q;
                } else if (q < 2 * TJ) {
                    li = TI - 1;
                    lj = q - TJ;
                } else {
                    li = 1 + ((q - 2 * TJ) >> 1);
                    lj = ((q - 2 * TJ) & 1) ? TJ - 1 : 0;
                }
            } else {
                if (q < TJ) {
                    li = -1;
                    lj = q;
                } else if (q < 2 * TJ) {
                    li = TI;
                    lj = q - TJ;
                } else {
                    li = (q - 2 * TJ) >> 1;
                    lj = ((q - 2 * TJ) & 1) ? TJ : -1;
                }
            }
            const int i = i0 + li, j = j0 + lj;
            const unsigned o0 = (unsigned)(i + j) & 1u, o1 = o0 ^ 1u;
            const int k0 = 2 * h + (int)o0, k1 = 2 * h + (int)o1;
            const bool okij = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2;
            const unsigned ok0 = okij && k0 >= 1 && k0 <= nz - 2;
            const unsigned ok1 = set < 2 && okij && k1 >= 1 && k1 <= nz - 2;
            const int w = lrow(li, lj) + h;
            d = (lds0 + 4u * (unsigned)(w - XS)) | o0 << 20 | o1 << 21 | ok0 << 22 | ok1 << 23;
            if (i >= 0 && i < nx && j >= 0 && j < ny) {
                const long long g = ((long long)i * ny + j) * nzh + h;
                if constexpr (FUSE) {
                    r0 = sor_defect_at<float>(phi, rho, sp, cf, 0, g, i, j, h);
                    if (set < 2) r1 = sor_defect_at<float>(phi, rho, sp, cf, 1, g, i, j, h);
                } else {
                    r0 = R[g];
                    if (set < 2) r1 = R[sp.Gh + g];
                }
            }
        }
        D[s] = d;
        R0[s] = r0;
        if (s < SI + SR) R1[s] = r1;
    }

    // exchange words of this thread: word x = tid + q * NT of a tile's slice,
    // laid out as [im rows d = 0, 1][ip rows][jm rows][jp rows][corners
    // (0,0), (0,TJ-1), (TI-1,0), (TI-1,TJ-1)], each section packed three
    // values to two words (deep_words; sections start at even words, so
    // word x carries the half x & 1 = tid & 1 of its middle value).  Value v
    // of a section is element v of its rows, row-major.  Per word: fa and fm
    // hold the LDS slots of its whole value and of its middle value in the
    // rows it publishes (low 16 bits) and in the border rows it receives
    // from the neighbour's matching word (high 16 bits; values past a
    // section's end map to the sink word E); src_b the byte offset of that
    // word; bit q of fmask: the neighbour exists; bit q of pmask: a neighbour
    // reads the word this thread publishes (tiles on the grid's faces skip
    // the stores nobody reads).
    const int wi = deep_words(side_i), wj = deep_words(side_j), wc = deep_words(nzh);
    const int oIM = 0, oIP = 2 * wi, oJM = 4 * wi, oJP = 4 * wi + 2 * wj, oC = oJP + 2 * wj;
    unsigned src_b[kFetchMaxD], fa[kFetchMaxD], fm[kFetchMaxD], fmask = 0, pmask = 0;
#pragma unroll
    for (int q = 0; q < kFetchMaxD; q++) src_b[q] = fa[q] = fm[q] = 0u;
    int nfetch = 0;
    for (int x = tid; x < xstride && nfetch < kFetchMaxD; x += NT) {
        int off = -1, kind, sec, y;  // kind 0: i face, 1: j face, 2: corner
        if (x < 4 * wi) {
            kind = 0;
            sec = x / wi;
            y = x - sec * wi;
            if (sec < 2) {  // my rows li = d; I receive li = -1 - d from (pi-1)'s ip rows
                if (has_im) off = (t - PJ) * xstride + oIP + (sec & 1) * wi + y;
            } else {  // my rows li = TI-1-d; I receive li = TI + d from (pi+1)'s im rows
                if (has_ip) off = (t + PJ) * xstride + oIM + (sec & 1) * wi + y;
            }
        } else if (x < oC) {
            kind = 1;
            sec = (x - oJM) / wj;
            y = x - oJM - sec * wj;
            if (sec < 2) {  // my rows lj = d; I receive lj = -1 - d from (pj-1)'s jp rows
                if (has_jm) off = (t - 1) * xstride + oJP + (sec & 1) * wj + y;
            } else {  // my rows lj = TJ-1-d; I receive lj = TJ + d from (pj+1)'s jm rows
                if (has_jp) off = (t + 1) * xstride + oJM + (sec & 1) * wj + y;
            }
        } else {
            // corner c of mine; corner c of the diagonal tile that it touches
            // lands on my opposite corner
            kind = 2;
            sec = (x - oC) / wc;
            y = x - oC - sec * wc;
            const int di = (sec & 2) ? -1 : 1, dj = (sec & 1) ? -1 : 1;  // the tile whose corner c faces me
            const bool ex = (di > 0 ? has_ip : has_im) && (dj > 0 ? has_jp : has_jm);
            if (ex) off = (t + di * PJ + dj) * xstride + oC + sec * wc + y;
            // my corner c is read by the diagonal tile (pi - di, pj - dj)
            if ((di > 0 ? has_im : has_ip) && (dj > 0 ? has_jm : has_jp)) pmask |= 1u << nfetch;
        }
        // a face word is read by the neighbour it is fetched from
        if (kind < 2 && off >= 0) pmask |= 1u << nfetch;
        // LDS slot of value v of this word's section: in the rows published
        // (recv 0) or in the border rows received (recv 1)
        auto slot = [&](int v, int recv) -> unsigned {
            const int d = sec & 1;
            int r;
            if (kind == 0) {
                if (v >= side_i) return (unsigned)L::E;
                const int lj = v / nzh;
                r = (sec < 2 ? (recv ? lrow(-1 - d, lj) : lrow(d, lj))
                             : (recv ? lrow(TI + d, lj) : lrow(TI - 1 - d, lj))) + v - lj * nzh;
            } else if (kind == 1) {
                if (v >= side_j) return (unsigned)L::E;
                const int li = v / nzh;
                r = (sec < 2 ? (recv ? lrow(li, -1 - d) : lrow(li, d))
                             : (recv ? lrow(li, TJ + d) : lrow(li, TJ - 1 - d))) + v - li * nzh;
            } else {
                if (v >= nzh) return (unsigned)L::E;
                const int cli = (sec & 2) ? TI - 1 : 0, clj = (sec & 1) ? TJ - 1 : 0;
                r = (recv ? lrow((sec & 2) ? -1 : TI, (sec & 1) ? -1 : TJ) : lrow(cli, clj)) + v;
            }
            return (unsigned)r;
        };
        const int g = y >> 1, va = 3 * g + 2 * (y & 1), vm = 3 * g + 1;
        src_b[nfetch] = off >= 0 ? 8u * (unsigned)off : 0u;
        if (off >= 0) fmask |= 1u << nfetch;
        fa[nfetch] = slot(va, 0) | slot(va, 1) << 16;
        fm[nfetch] = slot(vm, 0) | slot(vm, 1) << 16;
        nfetch++;
    }
    __syncthreads();

    // the update over this lane's slots of one row set; colour C reads the
    // other colour's array and writes its own
    auto sweep = [&](auto cc, auto set_, float m) -> float {
        constexpr int C = decltype(cc)::value;
        constexpr int SET = decltype(set_)::value;
        constexpr int B = SET == 0 ? 0 : (SET == 1 ? SI : SI + SR), NS = SET == 0 ? SI : (SET == 1 ? SR : SH);
        const int ns = SET == 0 ? ns_in : (SET == 1 ? ns_rg : ns_hr);
        constexpr unsigned OB = C ? 0 : sizeof(float) * L::ext1, DB = C ? sizeof(float) * L::ext1 : 0;
#pragma unroll
        for (int u = 0; u < NS; u++) asm volatile("" : "+v"(D[B + u]));
#pragma unroll
        for (int u = 0; u < NS; u++) {
            if (u >= ns) break;
            const int s = B + u;
            const unsigned d = D[s];
            const unsigned a = d & 0xfffffu;
            const lds_float *o = (const lds_float *)(size_t)(a + OB);
            const float xm = o[0], ym = o[XS - P], yp = o[XS + P], xp = o[2 * XS];
            const lds_float *oz = o + ((d >> (20 + C)) & 1u);
            const float zm = oz[XS - 1], zp = oz[XS];
            float rs;
            if constexpr (C == 0)
                rs = R0[s];
            else
                rs = R1[s < SI + SR ? s : 0];
            const float sv = sor_f32_sum(xm, xp, ym, yp, zm, zp, rs, yx, yy, yz);
            lds_float *dv = (lds_float *)(size_t)(a + DB + sizeof(float) * XS);
            const float po = *dv;
            float nv = sor_f32_relax(po, sv, inv, om, om1);
            nv = ((d >> (22 + C)) & 1u) ? nv : po;
            *dv = nv;
            m = fmaxf(m, fabsf(nv - po));
        }
        return m;
    };
    using C0 = std::integral_constant<int, 0>;
    using C1 = std::integral_constant<int, 1>;
    using In = std::integral_constant<int, 0>;
    using Rg = std::integral_constant<int, 1>;
    using Hr = std::integral_constant<int, 2>;

    const unsigned hsel = (unsigned)tid & 1u;  // which half of its middle value a word carries
    auto publish = [&](unsigned long long *xw, unsigned tag) {
        unsigned b = 8u * (unsigned)(t * xstride + tid);
        asm volatile("" : "+v"(b));
        const unsigned long long tg = (unsigned long long)(tag & 0xffffu) << 48;
#pragma unroll
        for (int q = 0; q < kFetchMaxD; q++)
            if (q < nfetch && ((pmask >> q) & 1u)) {
                const unsigned a = __float_as_uint(x1[fa[q] & 0xffffu]);
                const unsigned mh = (__float_as_uint(x1[fm[q] & 0xffffu]) >> (16u * hsel)) & 0xffffu;
                __hip_atomic_store((unsigned long long *)((char *)xw + (b + 8u * q * NT)),
                                   tg | (unsigned long long)mh << 32 | a, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
    };
    // a received word into the border rows: its whole value, and its half of
    // the middle value (16-bit LDS store into that half)
    auto land = [&](unsigned long long w, int q) {
        x1[fa[q] >> 16] = __uint_as_float((unsigned)w);
        ((unsigned short *)(x1 + (fm[q] >> 16)))[hsel] = (unsigned short)(w >> 32);
    };

    // Poll this thread's neighbour words until they carry `tag`, into the
    // black border, and (poll_row >= 0) in wave 0 the write-once slot row
    // poll_row of every tile's max: both in the same rounds of loads, so the
    // convergence test costs no round trip of its own.  Wave 0 then leaves
    // the max over all tiles in s_max[parity].  Returns nonzero
    // (block-uniform) on error.
    auto fetch = [&](const unsigned long long *xs, unsigned tag, int poll_row, int parity) -> int {
        constexpr int kSlotPoll32 = slot_poll32<NT>();
        unsigned pending = fmask;
        unsigned sb[kFetchMaxD];
#pragma unroll
        for (int q = 0; q < kFetchMaxD; q++) {
            sb[q] = src_b[q];
            asm volatile("" : "+v"(sb[q]));
        }
        const bool slots = MP && poll_row >= 0 && wave == 0;
        const float *row = sy.slots + (long long)(poll_row > 0 ? poll_row : 0) * nT;
        float v[kSlotPoll32];
        unsigned spend = 0;
#pragma unroll
        for (int q = 0; q < kSlotPoll32; q++) {
            v[q] = 0.0f;
            if (slots && lane + q * kWave < nT) spend |= 1u << q;
        }
        long long spin = 0;
        int bad = 0;
        if constexpr (MP) {
            while ((pending || spend) && !bad) {
                unsigned long long w[kFetchMaxD];
#pragma unroll
                for (int q = 0; q < kFetchMaxD; q++)
                    if (pending & (1u << q)) w[q] = ld_tagged((const unsigned long long *)((const char *)xs + sb[q]));
                float sv[kSlotPoll32];
#pragma unroll
                for (int q = 0; q < kSlotPoll32; q++)
                    if (spend & (1u << q)) sv[q] = ld_dev_f(row + lane + q * kWave);
#pragma unroll
                for (int q = 0; q < kFetchMaxD; q++) {
                    if ((pending & (1u << q)) && (unsigned)(w[q] >> 48) == (tag & 0xffffu)) {
                        land(w[q], q);
                        pending &= ~(1u << q);
                    }
                }
#pragma unroll
                for (int q = 0; q < kSlotPoll32; q++) {
                    if ((spend & (1u << q)) && sv[q] >= 0.0f) {
                        v[q] = sv[q];
                        spend &= ~(1u << q);
                    }
                }
                if (++spin > kSpinLimit) bad = 1;
                else if ((spin & 255) == 0 && ld_dev_i(sy.err)) bad = 1;
                if (pending || spend) __builtin_amdgcn_s_sleep(1);
            }
        } else {
            while ((pending || spend) && !bad) {
                unsigned long long w[kFetchMaxD];
#pragma unroll
                for (int q = 0; q < kFetchMaxD; q++)
                    if (pending & (1u << q)) w[q] = ld_tagged((const unsigned long long *)((const char *)xs + sb[q]));
#pragma unroll
                for (int q = 0; q < kFetchMaxD; q++) {
                    if ((pending & (1u << q)) && (unsigned)(w[q] >> 48) == (tag & 0xffffu)) {
                        land(w[q], q);
                        pending &= ~(1u << q);
                    }
                }
                if (++spin > kSpinLimit) bad = 1;
                else if ((spin & 255) == 0 && ld_dev_i(sy.err)) bad = 1;
                if (pending) __builtin_amdgcn_s_sleep(1);
            }
        }
        if (MP && slots) {
            float g = 0.0f;
#pragma unroll
            for (int q = 0; q < kSlotPoll32; q++) g = v[q] > g ? v[q] : g;
            for (int o = kWave / 2; o > 0; o >>= 1) {
                const float u = __shfl_xor(g, o);
                g = u > g ? u : g;
            }
            if (lane == 0) s_max[parity] = g;
        }
        if (bad) {
            st_dev_i(sy.err, 1);
            s_bad = 1;
        }
        __syncthreads();
        return s_bad;
    };

#ifdef HV_RES_TRACE
    // phase clocks of tile 0, thread 0, iterations 64-79 (trace builds only;
    // kept in LDS during the loop so that no global store joins a barrier)
    __shared__ unsigned long long s_trace[256];
#define HV_TRACE_D(slot)                                                                                   \
    do {                                                                                                   \
        if (t == 0 && tid == 0 && it >= 64 && it < 80) s_trace[(it - 64) * 16 + (slot)] = wall_clock64(); \
    } while (0)
#else
#define HV_TRACE_D(slot) \
    do {                  \
    } while (0)
#endif
    // wave 0, without MP: max over all tiles of the write-once slot row
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

    int iters = max_it;
    float tested = -1.0f;
    for (int it = 0; it < max_it; it++) {
        const unsigned tag = (unsigned)it + 1u;
        HV_TRACE_D(0);
        float m = sweep(C0{}, In{}, 0.0f);
        HV_TRACE_D(1);
        if (!MP && it >= lag && wave == 0) poll_slots(it - lag, it & 1);
        HV_TRACE_D(2);
        // (lag >= 1: with MP, every iteration that tests has a fetch)
        if (it > 0) {
            if (fetch((it - 1) & 1 ? sy.xb : sy.xr, tag - 1u, MP && it >= lag ? it - lag : -1, it & 1)) break;
        } else {
            __syncthreads();
        }
        HV_TRACE_D(3);
        m = sweep(C0{}, Rg{}, m);
        (void)sweep(C0{}, Hr{}, 0.0f);  // the border ring: not this tile's nodes
        HV_TRACE_D(4);
        __syncthreads();
        HV_TRACE_D(5);
        m = sweep(C1{}, In{}, m);
        m = sweep(C1{}, Rg{}, m);
        HV_TRACE_D(6);
        m = wave_max_nonneg(m);  // the lanes' maxima are never NaN (fmaxf)
        if (lane == 0) red[wave] = m;
        __syncthreads();
        HV_TRACE_D(7);
        publish(it & 1 ? sy.xb : sy.xr, tag);
        HV_TRACE_D(8);
        if (tid == 0) {
            float bm = 0.0f;
            for (int w = 0; w < nwaves; w++) bm = red[w] > bm ? red[w] : bm;
            st_dev_f(sy.slots + (long long)it * nT + t, bm);
        }
        if (it >= lag) {
            tested = s_max[it & 1];
            if (tested < tol) {
                iters = it + 1;
                break;
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < TI * TJ * nzh; e += NT) {
        const int r = e / nzh, h = e - r * nzh;
        const int li = r / TJ, lj = r - li * TJ;
        const int i = i0 + li, j = j0 + lj;
        if (i < nx && j < ny) {
            const long long q = ((long long)i * ny + j) * nzh + h;
            const int w = lrow(li, lj) + h;
            if constexpr (FUSE) {
                phi[q] += (double)x0[w];
                phi[sp.Gh + q] += (double)x1[w];
            } else {
                delta[q] = x0[w];
                delta[sp.Gh + q] = x1[w];
            }
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
#undef HV_TRACE_D
}

// lane-owned shapes: the ensemble kernel's (tiles up to 16 x 12 rows of 48
// slots: 96^3 grids and smaller) and the single-grid kernel's (8 x 8 rows of
// 64 slots: 128^3 on 256 CUs)
#define HV_OWNED_BATCH 7, 3, 48, 14, 16
#define HV_OWNED_SOLO 3, 2, 64, 10, 8
constexpr int kOwnedBatchP = 48, kOwnedBatchQJ = 14, kOwnedBatchIM = 16;
constexpr int kOwnedSoloP = 64, kOwnedSoloQJ = 10, kOwnedSoloIM = 8;
// deep-halo shapes: <SI, SR, SH, P, QJ, IM>
#define HV_DEEP_BATCH 7, 3, 3, 48, 16, 16
#define HV_DEEP_SOLO 3, 2, 2, 64, 12, 8
#define HV_DEEP_SOLO2 2, 3, 3, 64, 8, 8  // two tiles per CU (512 threads): up to 8 x 4 rows
constexpr int kDeepBatchP = 48, kDeepBatchQJ = 16, kDeepBatchIM = 16;
constexpr int kDeepSoloP = 64, kDeepSoloQJ = 12, kDeepSoloIM = 8;
constexpr int kDeepSolo2P = 64, kDeepSolo2QJ = 8, kDeepSolo2IM = 8;

template <int SI, int SR, int P, int QJ, int IM>
__global__ void __launch_bounds__(kRes32Threads)
k_sor_resident_f32o(float *__restrict__ delta, const float *__restrict__ R, Split sp, ResTiles tl, SorCoef cf,
                    int max_it, int lag, ResSync32 sy) {
    res32o_solve_tile<SI, SR, P, QJ, IM>(delta, R, sp, tl, cf, max_it, lag, sy, (int)blockIdx.x);
}

// NT threads per block: 1024 (one tile per CU) or 512 (two per CU: one
// tile's wait for its neighbours' faces overlaps the other's sweeps)
template <int SI, int SR, int SH, int P, int QJ, int IM, int NT = kRes32Threads>
__global__ void __launch_bounds__(NT, kRes32Threads / NT)
k_sor_resident_f32d(float *__restrict__ delta, const float *__restrict__ R, Split sp, ResTiles tl, SorCoef cf,
                    int max_it, int lag, ResSync32 sy) {
    res32d_solve_tile<SI, SR, SH, P, QJ, IM, true, NT>(delta, R, sp, tl, cf, max_it, lag, sy, (int)blockIdx.x);
}

// The same with the defect and the fold in the solve (FUSE above): phi is
// read for the residual and receives the correction, rho is the split source.
template <int SI, int SR, int SH, int P, int QJ, int IM, int NT = kRes32Threads>
__global__ void __launch_bounds__(NT, kRes32Threads / NT)
k_sor_resident_f32d_fold(double *__restrict__ phi, const float *__restrict__ rho, Split sp, ResTiles tl, SorCoef cf,
                         int max_it, int lag, ResSync32 sy) {
    res32d_solve_tile<SI, SR, SH, P, QJ, IM, true, NT, true>(nullptr, nullptr, sp, tl, cf, max_it, lag, sy,
                                                              (int)blockIdx.x, phi, rho);
}

}  // namespace hv
