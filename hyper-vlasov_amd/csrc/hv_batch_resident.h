// hv_batch_resident.h -- the mixed form's FP32 correction iteration for a
// whole ensemble in one persistent, cooperatively launched kernel.
//
// Each instance is cut into tiles of TI x TJ grid rows (all k) that fit the
// LDS of one CU; a block runs the LDS-resident solve of one tile
// (res32_solve_tile) and the tiles of an instance, run at the same time by
// different blocks, exchange faces through that instance's own tagged words
// and stop after the same iteration (the lagged test over that instance's own
// slot rows).  The host plans, per solve, a queue of (instance, tile) jobs for
// every block: instances are taken in order of decreasing predicted work and
// given the blocks that become free first, so that all blocks finish at about
// the same time.  All queues follow that one instance order, so the first
// unfinished instance always has every one of its tiles running: no block
// waits on a block that waits on it.
//
// The iterates and the iteration count of an instance do not depend on its
// tiling (every node update reads the same operand values), so instance i
// gets the bits of a single-instance LDS-resident solve with the same lag.
#pragma once

#include "hv_sor_owned.h"
#include "hv_sor_resident.h"

namespace hv {

struct BRes32 {
    float *delta;
    const float *R;
    double *phi;       // (the deep-halo kernel's FUSE form: the residual from phi and rho,
    const float *rho;  // the correction added to phi)
    Split sp;
    ResTiles tl;
    SorCoef cf;
    int max_it;
    ResSync32 sy;  // the instance's exchange words, slot rows and result; err shared by all
#ifdef HV_BRES_TRACE
    unsigned long long *jtrace;  // start and end clock of every job (trace builds only)
#endif
};

// CTI, CTJ, CNZH: one tile shape with a compile-time body (0: none); tiles of
// other shapes run the run-time-shape body.  (One compile-time shape per
// kernel: bodies of several shapes in one kernel exceed the register budget
// of 1024-thread blocks.)
template <int CTI, int CTJ, int CNZH>
__global__ void __launch_bounds__(kRes32Threads)
kb_sor_resident_f32(const BRes32 *__restrict__ inst, const int *__restrict__ job_first,
                    const int2 *__restrict__ jobs, int lag) {
    const int j1 = job_first[blockIdx.x + 1];
    for (int j = job_first[blockIdx.x]; j < j1; j++) {
        const int2 jb = jobs[j];
        const BRes32 &I = inst[jb.x];
#ifdef HV_BRES_TRACE
        if (threadIdx.x == 0) I.jtrace[2 * j] = wall_clock64();
#endif
        if (CTI && I.tl.TI == CTI && I.tl.TJ == CTJ && I.tl.nzh == CNZH)
            res32_solve_tile<CTI, CTJ, CNZH>(I.delta, I.R, I.sp, I.tl, I.cf, I.max_it, lag, I.sy, jb.y);
        else
            res32_solve_tile<0, 0, 0>(I.delta, I.R, I.sp, I.tl, I.cf, I.max_it, lag, I.sy, jb.y);
        __syncthreads();  // the tile's write-back has read the LDS the next job overwrites
#ifdef HV_BRES_TRACE
        if (threadIdx.x == 0) I.jtrace[2 * j + 1] = wall_clock64();
#endif
    }
}

// The same job loop with the lane-owned body <SI, SR, P, QJ, IM>; the host
// launches it only when every instance's tiling fits that body (a tile that
// does not is skipped with the error word set: its instance's other tiles
// skip too, so nothing waits on it).
template <int SI, int SR, int P, int QJ, int IM>
__global__ void __launch_bounds__(kRes32Threads)
kb_sor_resident_f32o(const BRes32 *__restrict__ inst, const int *__restrict__ job_first,
                     const int2 *__restrict__ jobs, int lag) {
    const int j1 = job_first[blockIdx.x + 1];
    for (int j = job_first[blockIdx.x]; j < j1; j++) {
        const int2 jb = jobs[j];
        const BRes32 &I = inst[jb.x];
#ifdef HV_BRES_TRACE
        if (threadIdx.x == 0) I.jtrace[2 * j] = wall_clock64();
#endif
        if (owned_fits<SI, SR, P, QJ, IM, kRes32Threads>(I.tl))
            res32o_solve_tile<SI, SR, P, QJ, IM>(I.delta, I.R, I.sp, I.tl, I.cf, I.max_it, lag, I.sy, jb.y);
        else if (threadIdx.x == 0)
            st_dev_i(I.sy.err, 1);
        __syncthreads();  // the tile's write-back has read the LDS the next job overwrites
#ifdef HV_BRES_TRACE
        if (threadIdx.x == 0) I.jtrace[2 * j + 1] = wall_clock64();
#endif
    }
}

// The same job loop with the deep-halo body <SI, SR, SH, P, QJ, IM> (tilings
// with the deep exchange stride; the host launches it only when every
// instance fits).  FUSE: each tile computes its residual from the instance's
// phi and rho and adds its correction to phi (the defect and fold launches'
// parts, the same bits).
template <int SI, int SR, int SH, int P, int QJ, int IM, bool FUSE = false>
__global__ void __launch_bounds__(kRes32Threads)
kb_sor_resident_f32d(const BRes32 *__restrict__ inst, const int *__restrict__ job_first,
                     const int2 *__restrict__ jobs, int lag) {
    const int j1 = job_first[blockIdx.x + 1];
    for (int j = job_first[blockIdx.x]; j < j1; j++) {
        const int2 jb = jobs[j];
        const BRes32 &I = inst[jb.x];
#ifdef HV_BRES_TRACE
        if (threadIdx.x == 0) I.jtrace[2 * j] = wall_clock64();
#endif
        if (deep_fits<SI, SR, SH, P, QJ, IM, kRes32Threads>(I.tl) && I.tl.xstride == deep_xstride(I.tl))
            res32d_solve_tile<SI, SR, SH, P, QJ, IM, true, kRes32Threads, FUSE>(I.delta, I.R, I.sp, I.tl, I.cf,
                                                                               I.max_it, lag, I.sy, jb.y, I.phi,
                                                                               I.rho);
        else if (threadIdx.x == 0)
            st_dev_i(I.sy.err, 1);
        __syncthreads();
#ifdef HV_BRES_TRACE
        if (threadIdx.x == 0) I.jtrace[2 * j + 1] = wall_clock64();
#endif
    }
}

}  // namespace hv
