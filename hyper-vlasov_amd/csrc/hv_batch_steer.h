// hv_batch_steer.h -- the end of an ensemble's mixed-precision Coulomb solve
// without the host: the FP64 finishing sweeps of every instance in one
// cooperative launch that applies the reference stop rule on the device, and
// a per-step record that keeps the extrapolation history on the device.
//
// The reference iterates red-black SOR until the max |dphi| of an iteration
// is below tol (include/physics.hpp:100-157).  In the mixed form that rule is
// applied by FP64 sweeps after the FP32 correction; run by the host it takes
// a read-back of every iteration's maxima.  Here each instance owns a group of
// blocks of one cooperative launch; the group sweeps the instance's red and
// black nodes in global memory (the same tile sweep, sor_tile_sweep, as the
// per-sweep launches, so the same bits), meets at an instance barrier after
// each colour, reduces the iteration's max into the instance's slot row and
// stops after the iteration whose max is below tol, or at the FP64 budget
// left by its FP32 iterations (mixed_fp64_budget).  Instances converge
// independently; no block waits on a block of another instance, and all
// blocks of a cooperative launch are resident, so every barrier completes.
#pragma once

#include "hv_batch_kernels.h"
#include "hv_sor_resident.h"

namespace hv {

// per instance, per finishing launch
struct BFin {
    int bar;    // arrivals at the instance's barriers (zeroed before the launch)
    int it64;   // FP64 iterations done
    double md;  // max |dphi| of the last of them
};

// one step of one instance (read back by the host at its sync points)
struct BLog {
    int it32;   // FP32 correction iterations
    int it64;   // FP64 finishing iterations
    int pred;   // predictor of the step's extrapolated start (0: none)
    int pad;
    double md;  // max |dphi| of the solve's last FP64 iteration
};

constexpr int kFinThreads = 64 * kSorRows;  // blocks of (64, kSorRows) threads, as the sweep launches

// A barrier of the nb blocks of one instance: every thread's global stores
// are released at agent scope, one arrival per block is counted, and after
// the count reaches target every thread acquires.  False when the wait
// exceeded its spin limit or another block raised the error word (the caller
// then stops; every wait of the launch ends the same way).
__device__ __forceinline__ bool inst_barrier(int *ctr, int target, int *err) {
    __shared__ int bad;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        long long spin = 0;
        int b = 0;
        while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spin > kSpinLimit || ((spin & 255) == 0 && ld_dev_i(err))) {
                b = 1;
                break;
            }
        }
        if (b) st_dev_i(err, 1);
        bad = b;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return bad == 0;
}

// Zeroes the barrier words and the slot rows (stride maxima) of every
// instance before a finishing launch; one block per instance.
__global__ void kb_fin_reset(BFin *__restrict__ fin, unsigned long long *__restrict__ slots, int stride) {
    const int i = blockIdx.x;
    for (int q = threadIdx.x; q < stride; q += blockDim.x) slots[(long long)i * stride + q] = 0ULL;
    if (threadIdx.x == 0) {
        fin[i].bar = 0;
        fin[i].it64 = 0;
        fin[i].md = 0.0;
    }
}

// The FP64 finishing sweeps of every instance (phi already holds the folded
// FP32 correction).  first: prefix table of the instances' block groups;
// rres[2 i]: instance i's FP32 iterations (the resident FP32 launch's
// result); slots: rows of `stride` maxima per instance (u64 images of
// non-negative doubles, maxed atomically; stride >= the largest budget).
template <typename T>
__global__ void __launch_bounds__(kFinThreads)
kb_sor_finish(const BInst<T> *__restrict__ a, const int *__restrict__ first, int n, const float *__restrict__ rres,
              BFin *__restrict__ fin, unsigned long long *__restrict__ slots, int stride, int max_it,
              int *__restrict__ err) {
    const int inst = __builtin_amdgcn_readfirstlane(batch_instance(first, n, blockIdx.x));
    const int lb = (int)blockIdx.x - first[inst];
    const int nb = first[inst + 1] - first[inst];
    const BInst<T> &A = a[inst];
    const int budget = mixed_fp64_budget(max_it, (int)rres[2 * inst]);
    const int ntiles = A.tl.n_ic * A.tl.n_jb * A.tl.n_hb;
    unsigned long long *row = slots + (long long)inst * stride;
    int *bar = &fin[inst].bar;
    int nbar = 0, done = 0;
    double md = 0.0;
    while (done < budget) {
        double diff = 0.0;
        for (int t = lb; t < ntiles; t += nb) {
            const double d = sor_tile_sweep<T, kSorPlanesBatch>(t, A.phi, A.rho_s, A.sp, A.tl, A.cf, 0);
            diff = d > diff ? d : diff;
        }
        if (!inst_barrier(bar, nb * ++nbar, err)) break;
        for (int t = lb; t < ntiles; t += nb) {
            const double d = sor_tile_sweep<T, kSorPlanesBatch>(t, A.phi, A.rho_s, A.sp, A.tl, A.cf, 1);
            diff = d > diff ? d : diff;
        }
        block_max_atomic(diff, row + done);
        if (!inst_barrier(bar, nb * ++nbar, err)) break;
        md = __longlong_as_double(
            (long long)__hip_atomic_load(row + done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        done++;
        if (md < A.cf.tol) break;
    }
    if (lb == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
        fin[inst].it64 = done;
        fin[inst].md = md;
    }
}

// End of a step: instance i's iteration counts and last FP64 max into row
// (step counter - 1) of the log; the count of consecutive converged loop
// solves (hist[i]); with extrapolation (extrap > 0) the predictor and ring
// position of the next step's start (xorder[2 i], xorder[2 i + 1]), as the
// host advances them for a solo run; the resident launch's error word folded
// into the sticky error word serr.
template <typename T>
__global__ void kb_step_record(const BInst<T> *__restrict__ a, int n, const float *__restrict__ rres,
                               const BFin *__restrict__ fin, BLog *__restrict__ log, const int *__restrict__ step,
                               int extrap, int *__restrict__ hist, int *__restrict__ xorder,
                               const int *__restrict__ rerr, int *__restrict__ serr) {
    const long long row = (long long)(*step - 1) * n;
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int)(gridDim.x * blockDim.x)) {
        BLog L;
        L.it32 = (int)rres[2 * i];
        L.it64 = fin[i].it64;
        L.pred = extrap ? xorder[2 * i] : 0;
        L.pad = 0;
        L.md = fin[i].md;
        log[row + i] = L;
        int hn = hist[i];
        hn = L.md > a[i].cf.tol ? 0 : (hn < 6 ? hn + 1 : 6);
        hist[i] = hn;
        if (extrap) {
            xorder[2 * i] = sor_predictor_for(extrap, hn);
            xorder[2 * i + 1] = (xorder[2 * i + 1] + 4) % 5;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && *rerr) *serr = 1;
}

}  // namespace hv
