// hv_batch_kernels.h -- one launch per loop stage for a whole ensemble of
// nuclei (each with its own grid, particle count, Skyrme set and boost).
//
// Every batched kernel runs the same block body as the single-instance kernel
// (hv_kernels.h): a launch covers the blocks of all instances back to back,
// a block finds its instance in a prefix table of per-instance block counts
// (first[i] = blocks of instances before i) and runs the body with its local
// block index and the instance's arguments from a device array.  The node and
// particle arithmetic is therefore the single-instance arithmetic, and an
// instance of a batch produces the bits of a solo run of it.
#pragma once

#include "hv_kernels.h"

namespace hv {

// All per-instance arguments of the batched stages.
template <typename T>
struct BInst {
    // particle pass
    Particles<T> p;
    int np, total;
    // batched sort: the other buffer set, the storage-slot ids of both, each
    // particle's key and rank within its key, the instance's first key bucket
    Particles<T> palt;
    int *pid, *pid_alt;
    unsigned *skey;
    int *srank;
    long long sbase;
    Geom<T> g;
    T *U;
    unsigned long long *acc;
    T kick_fac, drift_fac, boost_p, boost_n;
    int boost;  // 1 during the step that applies the boost
    long long *partials;
    unsigned *status;
    StatSlot *slots;
    int part_blocks;
    // smearing
    const SmearTab<T> *tab;
    T *rho;
    SmearTiles st;        // fused form
    SmearTiles st2;       // two-pass form
    T *D;                 // two-pass D channels
    unsigned *fmask;      // two-pass plane masks of the tile columns
    // grid, SOR, potentials
    long long G;
    Split sp;
    double *phi;
    SorHist hist;         // mixed form: ring of the potentials of the five solves before the last
    T *rho_s;
    float *R, *delta;
    SorTiles tl;
    SorCoef cf;
    SorCoef cf32;  // the FP32 correction's (its relaxation factor, mixed_omega32)
    unsigned long long *maxd;
    SorCtl *ctl;
    double A, B, C, gamma;
};

// instance of block b: first[i] <= b < first[i + 1] (wave-uniform)
__device__ __forceinline__ int batch_instance(const int *__restrict__ first, int n, int b) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

// (the instance index is read back as a wave-uniform value, so that the
// instance's arguments and the tables they point to load through the scalar
// cache, not per lane)
#define HV_BATCH_LOCATE()                                                           \
    const int inst = __builtin_amdgcn_readfirstlane(batch_instance(first, n, blockIdx.x)); \
    const int bid = (int)blockIdx.x - first[inst];                                  \
    const BInst<T> &A = a[inst]

// launches over a subset of the instances (those still iterating): ids[k] is
// the instance of the k-th entry of the prefix table
#define HV_BATCH_LOCATE_IDS()                                                        \
    const int slot_ = __builtin_amdgcn_readfirstlane(batch_instance(first, n, blockIdx.x)); \
    const int bid = (int)blockIdx.x - first[slot_];                                  \
    const BInst<T> &A = a[__builtin_amdgcn_readfirstlane(ids[slot_])]

template <typename T>
__global__ void __launch_bounds__(kBlock)
kb_particle_step(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n, int end_kick) {
    HV_BATCH_LOCATE();
    particle_step_body<T, T, false>(bid, A.p, A.np, A.total, A.g, A.U, nullptr, A.acc, nullptr, A.kick_fac, A.drift_fac, end_kick, A.boost,
                             A.boost_p, A.boost_n, A.partials, A.status);
}

// one block per instance; the diagnostics slot is the step counter + 1 (the
// counter lives on the device, so a captured step graph needs no update)
template <typename T>
__global__ void __launch_bounds__(kStatThreads) kb_stats_finalize(const BInst<T> *__restrict__ a, const int *__restrict__ step) {
    const BInst<T> &A = a[blockIdx.x];
    stats_finalize_block(A.partials, A.part_blocks, A.slots + *step + 1);
}

__global__ void kb_step_advance(int *step) {
    if (threadIdx.x == 0) *step += 1;
}

template <typename T>
__global__ void __launch_bounds__(kSmearThreads)
kb_smear_zy(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n) {
    HV_BATCH_LOCATE();
    smear_zy_body<T, unsigned long long>(bid, A.acc, A.D, A.fmask, nullptr, 0, smear_tab_c(A.tab), A.g.nx,
                                         A.g.ny, A.g.nz, A.st2);
}

// the same over a subset of the instances (ids: one LDS class of them, so
// that the launch's LDS size fits its instances and not the largest radius)
template <typename T>
__global__ void __launch_bounds__(kSmearThreads)
kb_smear_zy_ids(const BInst<T> *__restrict__ a, const int *__restrict__ first, const int *__restrict__ ids, int n) {
    HV_BATCH_LOCATE_IDS();
    smear_zy_body<T, unsigned long long>(bid, A.acc, A.D, A.fmask, nullptr, 0, smear_tab_c(A.tab), A.g.nx,
                                         A.g.ny, A.g.nz, A.st2);
}

// The zy pass of one LDS class of instances as a loop over its items
// (instance, species, plane, tile), a few blocks per CU striding over them:
// the next item's instance is located and its counts load into registers
// while this item builds its H levels and D channels (one block per item
// waits for its counts with nothing of its own to overlap).  Every instance
// of the class has the vector tile and at most kZyStagedB x kSmearThreads
// staged counts; an item's results are those of smear_zy_item (the same
// staging, smear_zy_finish).
constexpr int kZyStagedB = 12;

template <typename T>
__global__ void __launch_bounds__(kSmearThreads)
kb_smear_zy_loop(const BInst<T> *__restrict__ a, const int *__restrict__ first, const int *__restrict__ ids, int n) {
    const int nall = first[n];
    T *C = smem_zy<T>();
    __shared__ int orslot[kSmearThreads / kWave];
    int ecoord[kZyStagedB];
    unsigned long long pre[kZyStagedB];
    int cR = -1, cW = -1;  // the halo shape ecoord was computed for
    int s_inst = 0, s_sp = 0, s_i = 0, s_jt = 0, s_kt = 0;
    // locate item it, and load its counts (zero outside the grid) into pre
    auto stage = [&](int it) {
        const int slot = __builtin_amdgcn_readfirstlane(batch_instance(first, n, it));
        int t = it - first[slot];
        const int inst = __builtin_amdgcn_readfirstlane(ids[slot]);
        const BInst<T> &A = a[inst];
        SmearTabC<T> *tb = smear_tab_c(A.tab);
        const int ry = tb->ry, rz = tb->rz;
        const SmearTiles st = A.st2;
        const int nx = A.g.nx, ny = A.g.ny, nz = A.g.nz;
        const int kt = t % st.nkt;
        t /= st.nkt;
        const int jt = t % st.njt;
        t /= st.njt;
        const int i = t % nx, sp = t / nx;
        const int R = st.TJ + 2 * ry, W = st.TK + 2 * rz;
        if (R != cR || W != cW) {
            smear_zy_coords(R, W, ecoord);
            cR = R;
            cW = W;
        }
        const int j0 = jt * st.TJ, k0 = kt * st.TK;
        const long long G = (long long)nx * ny * nz;
        const unsigned long long *cb =
            A.acc + sp * G + (long long)i * ny * nz + (long long)(j0 - ry) * nz + (k0 - rz);
        const int rlo = ry - j0, rhi = ny - 1 - (j0 - ry), clo = rz - k0, chi = nz - 1 - (k0 - rz);
#pragma unroll
        for (int m = 0; m < kZyStagedB; m++) {
            const int r = ecoord[m] >> 16, cc = ecoord[m] & 0xffff;
            pre[m] = (ecoord[m] >= 0 && r >= rlo && r <= rhi && cc >= clo && cc <= chi) ? cb[r * nz + cc] : 0ULL;
        }
        s_inst = inst;
        s_sp = sp;
        s_i = i;
        s_jt = jt;
        s_kt = kt;
    };
    int it = blockIdx.x;
    if (it < nall) stage(it);
    for (; it < nall; it += gridDim.x) {
        const int inst = s_inst, sp = s_sp, i = s_i, jt = s_jt, kt = s_kt;
        const BInst<T> &A = a[inst];
        SmearTabC<T> *tb = smear_tab_c(A.tab);
        const int rz = tb->rz;
        const int Wp = smear_zy_wp(A.st2.TK, rz), pad = smear_rzb(rz) - rz;
        int nonzero = 0;
#pragma unroll
        for (int m = 0; m < kZyStagedB; m++) {
            if (ecoord[m] >= 0) {
                const T x = count_of<T, unsigned long long>(pre[m]);
                C[(ecoord[m] >> 16) * Wp + pad + (ecoord[m] & 0xffff)] = x;
                nonzero |= (x != T(0.0));
            }
        }
        if (it + (int)gridDim.x < nall) stage(it + (int)gridDim.x);
        const bool any = __builtin_amdgcn_readfirstlane(block_or_lds(nonzero, orslot));
        smear_zy_finish<T, true>(any, sp, i, jt, kt, A.D, A.fmask, tb, A.g.nx, A.g.ny, A.g.nz, A.st2);
        lds_barrier();  // the next item's counts overwrite C and H
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
kb_smear_xs(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n) {
    HV_BATCH_LOCATE();
    smear_xs_body<T>(bid, A.D, A.rho, A.fmask, nullptr, 0, smear_tab_c(A.tab), A.g.nx, A.g.ny, A.g.nz, A.st2);
}

// The vector x pass (smear_xs_vec_body: VW nodes per thread, the loads of
// all planes issued together) with each instance's radius bucket; blocks of
// the 2G / VW table.  The host uses it when every instance has nz % VW == 0
// and a radius of at most 14.
template <typename T>
__global__ void __launch_bounds__(kBlock)
kb_smear_xs_vec(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n) {
    // (the protons' densities also into the colour-split source, and the
    // deposit's counts and box words cleared: no kb_rho_split or kb_zero_acc
    // launch after it)
    HV_BATCH_LOCATE();
    if (bid == 0 && threadIdx.x < cbox_acc_words(A.g.nx)) A.acc[2 * A.G + threadIdx.x] = 0ULL;
    SmearTabC<T> *tb = smear_tab_c(A.tab);
    const int rx = tb->rx;  // wave-uniform: one instance per block
    if (rx <= 4)
        smear_xs_vec_body<T, 4>(bid, A.D, A.rho, A.fmask, nullptr, 0, tb, A.g.nx, A.g.ny, A.g.nz, A.st2,
                                 (T *)A.rho_s, A.sp, A.acc);
    else if (rx <= 9)
        smear_xs_vec_body<T, 9>(bid, A.D, A.rho, A.fmask, nullptr, 0, tb, A.g.nx, A.g.ny, A.g.nz, A.st2,
                                 (T *)A.rho_s, A.sp, A.acc);
    else
        smear_xs_vec_body<T, 14>(bid, A.D, A.rho, A.fmask, nullptr, 0, tb, A.g.nx, A.g.ny, A.g.nz, A.st2,
                                 (T *)A.rho_s, A.sp, A.acc);
}

template <typename T>
__global__ void __launch_bounds__(kSmearThreads)
kb_smear_fused(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n) {
    HV_BATCH_LOCATE();
    smear_fused_body<T, unsigned long long>(bid, A.acc, A.rho, smear_tab_c(A.tab), A.g.nx, A.g.ny, A.g.nz, A.st);
}

// acc = 0 (2G words, then the count box), blocks of the 2G table
template <typename T>
__global__ void __launch_bounds__(kBlock)
kb_zero_acc(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n) {
    HV_BATCH_LOCATE();
    const long long q = (long long)bid * blockDim.x + threadIdx.x;
    if (q < 2 * A.G) A.acc[q] = 0ULL;
    if (q < cbox_acc_words(A.g.nx)) A.acc[2 * A.G + q] = 0ULL;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
kb_rho_split(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n) {
    HV_BATCH_LOCATE();
    rho_split_body<T>(bid, A.rho, A.rho_s, A.G, A.sp);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
kb_skyrme(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n) {
    HV_BATCH_LOCATE();
    skyrme_body<T>(bid, A.rho, A.U, A.phi, A.A, A.B, A.C, A.gamma, A.G, A.sp);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
kb_sor_defect(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n) {
    HV_BATCH_LOCATE();
    sor_defect_body<T>(bid, A.phi, A.rho_s, A.R, A.delta, A.sp, A.cf);
}

// extrapolated start of every instance's mixed solve (sor_extrapolate_body):
// xs[2 i] the predictor of instance i, xs[2 i + 1] its ring position
template <typename T>
__global__ void __launch_bounds__(kBlock)
kb_sor_extrapolate(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n, const int *__restrict__ xs) {
    HV_BATCH_LOCATE();
    const int id = __builtin_amdgcn_readfirstlane(xs[2 * inst]);
    const SorHist hs = sor_hist_order(A.hist, __builtin_amdgcn_readfirstlane(xs[2 * inst + 1]));
    double w[5];
    switch (sor_predictor(id, w)) {  // wave-uniform: one instance per block
    case 1: sor_extrapolate_body<1>(bid, A.phi, hs, id, 2 * A.sp.Gh); break;
    case 2: sor_extrapolate_body<2>(bid, A.phi, hs, id, 2 * A.sp.Gh); break;
    case 3: sor_extrapolate_body<3>(bid, A.phi, hs, id, 2 * A.sp.Gh); break;
    case 4: sor_extrapolate_body<4>(bid, A.phi, hs, id, 2 * A.sp.Gh); break;
    default: sor_extrapolate_body<5>(bid, A.phi, hs, id, 2 * A.sp.Gh); break;
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
kb_sor_fold(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n) {
    HV_BATCH_LOCATE();
    sor_fold_body(bid, A.phi, A.delta, 2 * A.sp.Gh);
}

// Start of a solve: SorCtl cleared and row 0 of the slot maxima zeroed (later
// rows are zeroed one iteration ahead by the black sweep); one block per
// instance.
template <typename T>
__global__ void kb_sor_reset(const BInst<T> *__restrict__ a) {
    const BInst<T> &A = a[blockIdx.x];
    if (threadIdx.x < kMaxSlots) A.maxd[(long long)threadIdx.x * kSlotStride] = 0ULL;
    if (threadIdx.x == 0) {
        A.ctl->done = 0;
        A.ctl->iters = 0;
    }
}

// zero row it + 1 of the slot maxima (no block of this launch touches it; the
// next iteration's launches accumulate into it)
__device__ __forceinline__ void zero_next_row(unsigned long long *maxd, int colour, int bid, int it, int max_it) {
    if (colour == 1 && bid == 0 && threadIdx.y == 0 && it + 1 < max_it)
        maxd[((long long)(it + 1) * kMaxSlots + threadIdx.x) * kSlotStride] = 0ULL;
}

// one colour of FP64 SOR (source: the colour-split proton density), over the
// n instances listed in ids
template <typename T>
__global__ void __launch_bounds__(256)
kb_sor_colour(const BInst<T> *__restrict__ a, const int *__restrict__ first, const int *__restrict__ ids, int n,
              int colour, int it, int max_it) {
    HV_BATCH_LOCATE_IDS();
    zero_next_row(A.maxd, colour, bid, it, max_it);
    sor_colour_body<T, kSorPlanesBatch>(bid, A.phi, A.rho_s, A.sp, A.tl, A.cf, colour, it, A.maxd, A.ctl);
}

// one colour of the mixed form's FP32 correction iteration, over the n
// instances listed in ids
template <typename T>
__global__ void __launch_bounds__(256)
kb_sor_colour_f32(const BInst<T> *__restrict__ a, const int *__restrict__ first, const int *__restrict__ ids, int n,
                  int colour, int it, int max_it) {
    HV_BATCH_LOCATE_IDS();
    zero_next_row(A.maxd, colour, bid, it, max_it);
    sor_colour_f32_body<T, kSorPlanesBatch>(bid, A.delta, A.R, A.sp, A.tl, A.cf32, colour, it, A.maxd, A.ctl);
}

// per-iteration maxima of iterations [it0, it1) of every instance into
// out[inst * width + (it - it0)]; grid = n * (it1 - it0)
template <typename T>
__global__ void kb_sor_collect(const BInst<T> *__restrict__ a, double *__restrict__ out, int width, int it0,
                               int it1) {
    const int m = it1 - it0;
    const int inst = blockIdx.x / m, r = blockIdx.x - inst * m;
    sor_collect_body(r, a[inst].maxd, out + (long long)inst * width, it0, it1);
}

// ---- batched particle sort ---------------------------------------------------
// The counting sort of hv_kernels.h (sort_count_body, sort_scan_body,
// sort_scatter_body) over every instance, each with its own run of key
// buckets from sbase, and a pointer swap on the device.
template <typename T>
__global__ void __launch_bounds__(kBlock)
kb_sort_count(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n, unsigned *__restrict__ buck) {
    HV_BATCH_LOCATE();
    const int q = bid * (int)blockDim.x + (int)threadIdx.x;
    sort_count_body(A.p, A.np, A.g, q, q < A.total, buck + A.sbase, A.skey, A.srank);
}

// one block per instance: its key counts (sfirst[i] .. sfirst[i + 1])
__global__ void __launch_bounds__(kSortScanThreads)
kb_sort_scan(const long long *__restrict__ sfirst, unsigned *__restrict__ buck) {
    sort_scan_body(buck + sfirst[blockIdx.x], (int)(sfirst[blockIdx.x + 1] - sfirst[blockIdx.x]));
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
kb_sort_scatter(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n,
                const unsigned *__restrict__ buck) {
    HV_BATCH_LOCATE();
    const int q = bid * (int)blockDim.x + (int)threadIdx.x;
    if (q < A.total) sort_scatter_body(A.p, A.palt, A.pid, A.pid_alt, q, buck + A.sbase, A.skey, A.srank);
}

// the sorted set becomes the current one (one thread per instance)
template <typename T>
__global__ void kb_sort_swap(BInst<T> *__restrict__ a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Particles<T> t = a[i].p;
    a[i].p = a[i].palt;
    a[i].palt = t;
    int *u = a[i].pid;
    a[i].pid = a[i].pid_alt;
    a[i].pid_alt = u;
}

// closing half kick of a run of steps (gather of the final U)
template <typename T>
__global__ void __launch_bounds__(kBlock)
kb_gather_kick(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n) {
    HV_BATCH_LOCATE();
    gather_u_body<T, T>(bid, A.p, A.np, A.total, A.g, A.U, A.kick_fac, 1);
}

#undef HV_BATCH_LOCATE
#undef HV_BATCH_LOCATE_IDS

}  // namespace hv
