// hv_batch_impl.h -- ensembles of nuclei stepped together (included at the end
// of hv_abi.hip: it uses that file's context and helpers).
//
// A batch owns one context per instance (its own grid, particle counts,
// buffers), all on the batch's stream.  hv_batch_run_steps runs the step loop
// of every instance in lock step with one launch per loop stage for the whole
// ensemble (hv_batch_kernels.h): particle pass, diagnostics, smearing, the
// SOR sweeps (instances converge independently; a converged instance's blocks
// return at once), the Skyrme + Coulomb potential.  Sorting, the pre-loop
// solve and the per-instance read-backs stay per instance.
//
// Per-instance results equal a solo hv_run_steps of the same instance with the
// same options bit for bit: the same block bodies run on the same operands and
// the SOR stop rule is evaluated per instance.  The mixed form's FP32
// correction runs LDS-resident for all instances in one persistent launch
// (hv_batch_resident.h; the bits of a solo LDS-resident solve), or as
// per-sweep launches when that is switched off or an instance does not tile;
// the FP64 sweeps use per-sweep launches.

#include "hv_batch_kernels.h"
#include "hv_batch_resident.h"
#include "hv_batch_steer.h"

enum { kBPart, kBSmear, kBZy, kBN2, kBG, kBGh2, kBSor, kBN2V, kBKinds };

struct hv_batch {
    int n = 0;
    int precision = HV_PREC_F64;
    std::vector<hv_ctx *> inst;
    hipStream_t stream = nullptr;
    void *args = nullptr;                 // BInst<T>[n]
    std::vector<unsigned char> hargs;     // host image of args
    int *first = nullptr;                 // kBKinds prefix tables of n + 1 ints
    unsigned blocks[kBKinds] = {};
    size_t smear_lds = 0, smear2_lds = 0;
    int smear_fused = 1;                  // option: fused smearing (else two-pass)
    bool xs_vec = false;                  // two-pass: the vector x pass fits every instance
    // two-pass zy launches by LDS class: the instances whose tiles allow the
    // same number of blocks per CU run in one launch sized for them (one
    // launch sized for the largest radius would hold every instance to its
    // occupancy)
    struct ZyClass {
        size_t lds;    // dynamic LDS of the launch
        unsigned blocks;
        size_t off;    // int offset of its prefix table (m + 1) then ids (m) in zytab
        int m;         // instances
        bool loop;     // every instance fits kb_smear_zy_loop (then launched with `grid` blocks)
        unsigned grid;
    };
    std::vector<ZyClass> zycls;
    int *zytab = nullptr;
    size_t zytab_cap = 0;
    double *maxc = nullptr;               // collected SOR maxima, n x maxc_width
    int maxc_width = 0;
    int *act = nullptr;                   // SOR launch subset: prefix table (n + 1) then ids (n)
    int *hact = nullptr;                  // pinned staging of act
    std::vector<int> sor_blocks;          // per-instance SOR sweep blocks
    double *pinned = nullptr;
    size_t pinned_cap = 0;
    std::vector<int> pred32, pred64;      // per-instance SOR iteration predictions
    std::vector<int> boost;               // per-instance boost flag of the current step
    bool ready = false;
    int extrap = 0;                       // option HV_OPT_SOR_EXTRAPOLATE (mixed form)
    int *xorder = nullptr;                // per instance: predictor and ring position of the current step
    int *hxorder = nullptr;               // pinned staging of xorder
    int sort_every = 8;
    int since_sort = 1 << 30;             // steps since the last batched sort (as maybe_sort)
    unsigned *sbuck = nullptr;            // key counts / first slots of every instance's keys
    long long *sfirst = nullptr;          // per instance: its first key bucket (n + 1)
    long long sbuck_n = 0;
    std::vector<long long> hsfirst;       // host copy of sfirst
    int group_kib = 0;  // SOR working-set budget (KiB) of one instance group (0: all instances together)
    // LDS-resident FP32 correction of the whole batch (mixed precision)
    int res32 = 1;                        // option: use it when every instance tiles
    int res32_owned = 2;                  // option: 2 deep-halo, 1 lane-owned, 0 per-row (when every instance fits)
    int res32_is_owned = 0;               // the body the chosen kernel runs (same values)
    bool res32_fold = false;              // that kernel computes its residual and folds its correction
    int res32_lag = 2;                    // option: lag of its convergence test
    int res_state = 0;                    // 0: not set up, 1: ready, -1: unavailable
    int res_max_it = 0;
    int res_nblk = 0;                     // blocks of the persistent launch (one per CU)
    size_t res_lds = 0;
    std::vector<ResTiles> rtl;            // per-instance tiling
    const void *rkernel = nullptr;        // ensemble kernel instance (its compile-time shape)
    std::vector<size_t> rx_off, rslot_off;
    size_t rx_words = 0, rslot_words = 0, rjobs_n = 0;
    void *rinst = nullptr;                // BRes32[n]
    unsigned long long *rx = nullptr;     // every instance's tagged exchange words
    float *rslots = nullptr;              // every instance's slot rows
    float *rres = nullptr;                // per instance: iterations, tested max
    int *rerr = nullptr;
    int *rjob = nullptr;                  // job_first (nblk + 1) then jobs (int2 per tile)
    int *hrjob = nullptr;                 // pinned staging of rjob
    float *hrres = nullptr;               // pinned read-back of rres and rerr
    void *hrinst = nullptr;               // pinned staging of rinst
    // device-steered mixed step (the FP64 finishing sweeps and the solve's
    // bookkeeping on the device; the host reads back every kReplanSteps steps)
    int steer = 1;                        // option HV_OPT_BATCH_DEVICE_STEER
    int fin_state = 0;                    // 0: not set up, 1: ready, -1: unavailable
    int fin_nblk = 0;                     // blocks of the cooperative finishing launch
    int fin_stride = 0;                   // slot row length (>= sor_max_it and >= 4)
    int cur_max_it = 0;                   // sor_max_it of the running call
    int *ffirst = nullptr;                // prefix table of the instances' block groups (n + 1)
    BFin *fin = nullptr;                  // per instance
    unsigned long long *fslots = nullptr; // n x fin_stride iteration maxima
    BLog *log = nullptr;                  // per step and instance
    size_t log_cap = 0;                   // records
    int *dhist = nullptr;                 // per instance: consecutive converged loop solves
    int *serr = nullptr;                  // sticky error word of the call
    // step graphs
    int use_graph = 1;                    // option: replay a captured graph of each step's device part
    int *dstep = nullptr;                 // device step counter (diagnostics slot of the step - 1)
    hipGraphExec_t gexec = nullptr;       // the step without a sort
    hipGraphExec_t gexec_sort = nullptr;  // the step that sorts first
    long long graph_steps = 0, captures = 0;
#ifdef HV_BRES_TRACE
    unsigned long long *jtrace = nullptr;
#endif
    // stage timing (hipEvent pairs per stage, when enabled)
    bool prof = false;
    double prof_ms[HV_KCLASS_COUNT] = {};
    long long prof_launches[HV_KCLASS_COUNT] = {};
    std::vector<hipEvent_t> ev;
    struct Sample {
        int cls;
        hipEvent_t a, e;
    };
    std::vector<Sample> samples;  // sampled sweep launches, folded at the next read-back
    long long sweeps[2] = {0, 0};  // sweep launches since hv_batch_profile_enable: FP32 correction, FP64
    char err[512] = {0};
};

namespace {

int batch_err(hv_batch *b, int code, const char *fmt, ...) {
    if (b) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(b->err, sizeof b->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

// a failing instance call: copy its message
int batch_fwd(hv_batch *b, hv_ctx *c, int rc) {
    if (rc < 0) batch_err(b, rc, "%s", c->err);
    return rc;
}

#define HV_BHIP(b, expr)                                                                       \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return batch_err((b), HV_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                             __FILE__, __LINE__);                                              \
    } while (0)

int batch_pinned(hv_batch *b, size_t doubles) {
    if (b->pinned_cap >= doubles) return HV_OK;
    if (b->pinned) (void)hipHostFree(b->pinned);
    b->pinned = nullptr;
    b->pinned_cap = 0;
    HV_BHIP(b, hipHostMalloc((void **)&b->pinned, doubles * sizeof(double), hipHostMallocDefault));
    b->pinned_cap = doubles;
    return HV_OK;
}

double step_dt(const hv_params &p) { return p.t_f / p.steps; }

// Per-instance argument image (host) for the current buffers and parameters.
template <typename T>
void batch_fill_args(hv_batch *b, const hv_skyrme *sk, const hv_params *prm) {
    b->hargs.assign(sizeof(BInst<T>) * (size_t)b->n, 0);
    BInst<T> *h = (BInst<T> *)b->hargs.data();
    for (int i = 0; i < b->n; i++) {
        hv_ctx *c = b->inst[i];
        const hv_params &p = prm[i];
        BInst<T> &A = h[i];
        A.p = particles<T>(c);
        A.np = c->np;
        A.total = c->total;
        A.palt = particles_alt<T>(c);
        A.pid = c->pid;
        A.pid_alt = c->pid_alt;
        A.skey = c->keys[0];
        A.srank = c->ord[0];
        A.sbase = b->hsfirst.empty() ? 0 : b->hsfirst[i];
        A.g = make_geom<T>(c);
        A.U = (T *)c->U;
        A.acc = c->acc;
        // the factors of run_steps, bit for bit
        const double dt = step_dt(p);
        A.kick_fac = (T)(dt / (2.0 * HV_HBARC));
        A.drift_fac = (T)((dt * HV_HBARC) / HV_MC2);
        double bp, bn;
        if (p.boost_mode == HV_BOOST_SYMMETRIC) {
            bp = p.eta_exc;
            bn = p.eta_exc;
        } else {
            double zz = p.z, nnuc = p.n;
            bp = p.eta_exc * nnuc / (zz + nnuc);
            bn = p.eta_exc * zz / (zz + nnuc);
        }
        A.boost_p = (T)bp;
        A.boost_n = (T)bn;
        A.boost = b->boost[i];
        A.partials = c->partials;
        A.status = c->status;
        A.slots = c->slots;
        A.part_blocks = (int)blocks_for(c->total);
        A.tab = (const SmearTab<T> *)c->smear_tab;
        A.rho = (T *)c->density;
        A.st = c->stf;
        A.st2 = c->stl;
        A.D = (T *)c->D;
        A.fmask = c->sflags;
        A.G = c->G;
        A.sp = c->sp;
        A.phi = c->phi;
        for (int k = 0; k < 5; k++) A.hist.h[k] = c->phi_h[k];
        A.rho_s = (T *)c->rho_s;
        A.R = c->sor_R;
        A.delta = c->sor_delta;
        A.tl = make_sor_tiles(c, kSorPlanesBatch);
        A.cf = make_sor_coef(c, p.sor_omega, p.sor_tol);
        A.cf32 = A.cf;
        A.cf32.tol = ldexp(A.cf.tol, -c->f32_tol_shift);
        A.cf32.omega = mixed_omega32(c, p.sor_omega);
        A.cf32.one_minus_omega = 1.0 - A.cf32.omega;
        A.maxd = c->maxd;
        A.ctl = c->ctl;
        A.A = sk[i].A;
        A.B = sk[i].B;
        A.C = sk[i].C;
        A.gamma = sk[i].gamma;
    }
}

template <typename T>
int batch_upload_args(hv_batch *b) {
    HV_BHIP(b, hipMemcpyAsync(b->args, b->hargs.data(), b->hargs.size(), hipMemcpyHostToDevice, b->stream));
    HV_BHIP(b, hipStreamSynchronize(b->stream));
    return HV_OK;
}

// Buffers, block tables and the argument array for a run with these params.
template <typename T>
int batch_build(hv_batch *b, const hv_skyrme *sk, const hv_params *prm, int nslots) {
    std::vector<int> hf((size_t)kBKinds * (b->n + 1), 0);
    size_t lds = 0, lds2 = 0;
    bool xs_vec = true;  // every instance fits the vector x pass
    for (int i = 0; i < b->n; i++) {
        hv_ctx *c = b->inst[i];
        int rc = ensure_smear<T>(c, prm[i].sigma_r, prm[i].part_per_nucleon);
        if (rc) return batch_fwd(b, c, rc);
        // one chunk of planes per tile: the ensemble has blocks enough
        c->stf.IC = c->world.n[0];
        c->stf.nic = 1;
        rc = ensure_maxd(c, prm[i].sor_max_it);
        if (rc) return batch_fwd(b, c, rc);
        rc = ensure_slots(c, nslots);
        if (rc) return batch_fwd(b, c, rc);
        const SorTiles tl = make_sor_tiles(c, kSorPlanesBatch);
        const long long cnt[kBKinds] = {
            (long long)blocks_for(c->total),
            smear_blocks(c->stf),
            2LL * c->world.n[0] * c->stl.njt * c->stl.nkt,
            (long long)blocks_for(2 * c->G),
            (long long)blocks_for(c->G),
            (long long)blocks_for(2 * c->sp.Gh),
            (long long)tl.per_xcd * kXcds,  // a multiple of 8: local block b runs on XCD b % 8
            (long long)blocks_for(2 * c->G / smear_vw<T>()),
        };
        xs_vec = xs_vec && c->world.n[2] % smear_vw<T>() == 0 && c->xs_rxm > 0;
        for (int k = 0; k < kBKinds; k++) {
            const long long nx = hf[(size_t)k * (b->n + 1) + i] + cnt[k];
            if (nx > 0x7fffffffLL) return batch_err(b, HV_ERR_ARG, "batch too large for one launch");
            hf[(size_t)k * (b->n + 1) + i + 1] = (int)nx;
        }
        lds = c->stf_lds > lds ? c->stf_lds : lds;
        lds2 = c->stl_lds > lds2 ? c->stl_lds : lds2;
    }
    for (int k = 0; k < kBKinds; k++) b->blocks[k] = (unsigned)hf[(size_t)k * (b->n + 1) + b->n];
    b->xs_vec = xs_vec;
    b->sor_blocks.resize(b->n);
    for (int i = 0; i < b->n; i++)
        b->sor_blocks[i] = hf[(size_t)kBSor * (b->n + 1) + i + 1] - hf[(size_t)kBSor * (b->n + 1) + i];
    if (!b->act) HV_BHIP(b, hipMalloc((void **)&b->act, sizeof(int) * (2 * (size_t)b->n + 1)));
    if (!b->hact)
        HV_BHIP(b, hipHostMalloc((void **)&b->hact, sizeof(int) * (2 * (size_t)b->n + 1), hipHostMallocDefault));
    b->smear_lds = lds;
    b->smear2_lds = lds2;
    // zy launch classes: blocks per CU that the instance's LDS allows (at most 8)
    {
        int dev = 0, ncu = 256;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            ncu = 256;
        std::vector<int> cls(b->n);
        std::vector<int> kinds;
        for (int i = 0; i < b->n; i++) {
            const size_t l = b->inst[i]->stl_lds > 0 ? b->inst[i]->stl_lds : 1;
            int k = (int)((size_t)(160 * 1024) / l);
            k = k < 1 ? 1 : (k > 8 ? 8 : k);
            cls[i] = k;
            if (std::find(kinds.begin(), kinds.end(), k) == kinds.end()) kinds.push_back(k);
        }
        std::sort(kinds.begin(), kinds.end());
        std::vector<int> tab;
        b->zycls.clear();
        for (int k : kinds) {
            hv_batch::ZyClass zc{};
            std::vector<int> ids;
            for (int i = 0; i < b->n; i++)
                if (cls[i] == k) ids.push_back(i);
            zc.m = (int)ids.size();
            zc.off = tab.size();
            long long acc = 0;
            tab.push_back(0);
            zc.loop = true;
            for (int i : ids) {
                const hv_ctx *c = b->inst[i];
                acc += 2LL * c->world.n[0] * c->stl.njt * c->stl.nkt;
                if (acc > 0x7fffffffLL) return batch_err(b, HV_ERR_ARG, "batch too large for one launch");
                tab.push_back((int)acc);
                zc.lds = c->stl_lds > zc.lds ? c->stl_lds : zc.lds;
                zc.loop = zc.loop && smear_vec_tile<T>(c->stl.TJ, c->stl.TK) &&
                          c->zy_staged <= kZyStagedB * kSmearThreads;
            }
            for (int i : ids) tab.push_back(i);
            zc.blocks = (unsigned)acc;
            // the loop form: k blocks per CU (its class) on every CU, each
            // striding over the items
            zc.grid = (unsigned)std::min<long long>(acc, (long long)k * ncu);
            b->zycls.push_back(zc);
        }
        if (b->zytab_cap < tab.size()) {
            if (b->zytab) (void)hipFree(b->zytab);
            b->zytab = nullptr;
            b->zytab_cap = 0;
            HV_BHIP(b, hipMalloc((void **)&b->zytab, sizeof(int) * tab.size()));
            b->zytab_cap = tab.size();
        }
        HV_BHIP(b, hipMemcpy(b->zytab, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice));
    }
    HV_BHIP(b, hipFuncSetAttribute((const void *)kb_smear_fused<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kSmearLdsVec));
    HV_BHIP(b, hipFuncSetAttribute((const void *)kb_smear_zy<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kSmear2LdsVec));
    HV_BHIP(b, hipFuncSetAttribute((const void *)kb_smear_zy_ids<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kSmear2LdsVec));
    HV_BHIP(b, hipFuncSetAttribute((const void *)kb_smear_zy_loop<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kSmear2LdsVec));
    if (!b->first) HV_BHIP(b, hipMalloc((void **)&b->first, sizeof(int) * hf.size()));
    HV_BHIP(b, hipMemcpyAsync(b->first, hf.data(), sizeof(int) * hf.size(), hipMemcpyHostToDevice, b->stream));
    // key buckets of the batched sort: sort_buckets per instance
    {
        std::vector<long long> sf(b->n + 1, 0);
        for (int i = 0; i < b->n; i++) sf[i + 1] = sf[i] + sort_buckets(b->inst[i]);
        if (b->hsfirst != sf) {
            if (b->sbuck) (void)hipFree(b->sbuck);
            if (b->sfirst) (void)hipFree(b->sfirst);
            b->sbuck = nullptr;
            b->sfirst = nullptr;
            b->hsfirst.clear();
            HV_BHIP(b, hipMalloc((void **)&b->sbuck, sizeof(unsigned) * (size_t)sf[b->n]));
            HV_BHIP(b, hipMalloc((void **)&b->sfirst, sizeof(long long) * sf.size()));
            HV_BHIP(b, hipMemcpy(b->sfirst, sf.data(), sizeof(long long) * sf.size(), hipMemcpyHostToDevice));
            b->hsfirst = sf;
            b->sbuck_n = sf[b->n];
        }
    }
    if (!b->args) HV_BHIP(b, hipMalloc(&b->args, sizeof(BInst<T>) * (size_t)b->n));
    batch_fill_args<T>(b, sk, prm);
    return batch_upload_args<T>(b);
}

const int *btab(const hv_batch *b, int kind) { return b->first + (size_t)kind * (b->n + 1); }

hipEvent_t batch_event(hv_batch *b) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    b->ev.push_back(e);
    return e;
}

struct BStage {
    hv_batch *b;
    int cls;
    hipEvent_t a = nullptr, e = nullptr;
    BStage(hv_batch *bb, int k) : b(bb), cls(k) {
        if (!b->prof) return;
        a = batch_event(b);
        if (a) (void)hipEventRecord(a, b->stream);
    }
    ~BStage() {
        if (!a) return;
        e = batch_event(b);
        if (!e) return;
        (void)hipEventRecord(e, b->stream);
        float ms = 0.0f;
        if (hipEventSynchronize(e) == hipSuccess && hipEventElapsedTime(&ms, a, e) == hipSuccess) {
            b->prof_ms[cls] += ms;
            b->prof_launches[cls] += 1;
        }
        for (hipEvent_t x : b->ev) (void)hipEventDestroy(x);
        b->ev.clear();
    }
};

// SOR iterations of every instance in chunks of batched sweeps; iters[i]
// receives instance i's count under the reference stop rule, capped at
// budget[i].  Instances are iterated in groups whose sweep working set (the
// iterate and its source) fits the budget group_kib KiB, one group after the
// other (meant to keep a group's sweeps in the memory-side cache; on the
// 256-nucleus ensemble one group of all instances measured fastest, hence the
// default 0); each instance's iterates are the same whatever the grouping.
// Each chunk's launches cover only the instances of the group still
// iterating at its start (a converged instance's blocks would return at once,
// but would still be dispatched).
template <typename T>
int batch_sor_iterate(hv_batch *b, const hv_params *prm, bool f32, const std::vector<int> &pred,
                      const std::vector<int> &budget, std::vector<int> &iters, std::vector<double> &lastmax) {
    const int n = b->n;
    const BInst<T> *a = (const BInst<T> *)b->args;
    const int max_it = prm[0].sor_max_it;
    kb_sor_reset<T><<<n, 64, 0, b->stream>>>(a);
    HV_BHIP(b, hipGetLastError());
    iters.assign(n, -1);
    lastmax.assign(n, 0.0);
    // groups of consecutive instances under the working-set budget
    std::vector<int> grp(n, 0);
    int ngroups = 1;
    {
        const double cap = b->group_kib > 0 ? b->group_kib * 1024.0 : 1e300;
        double acc = 0.0;
        for (int i = 0; i < n; i++) {
            const double bytes = (f32 ? 8.0 : (double)(sizeof(double) + sizeof(T))) * 2.0 * (double)b->inst[i]->sp.Gh;
            if (acc > 0.0 && acc + bytes > cap) {
                ngroups++;
                acc = 0.0;
            }
            acc += bytes;
            grp[i] = ngroups - 1;
        }
    }
    const dim3 sblk(64, kSorRows, 1);
    for (int g = 0; g < ngroups; g++) {
        int remaining = 0, it = 0, chunk = 1;
        for (int i = 0; i < n; i++) {
            if (grp[i] != g) continue;
            remaining++;
            chunk = pred[i] - 1 > chunk ? pred[i] - 1 : chunk;
        }
        while (remaining > 0) {
            // launch subset: instances not converged and under budget
            int m = 0, blocks = 0, room = max_it;
            b->hact[0] = 0;
            for (int i = 0; i < n; i++) {
                if (iters[i] >= 0 || grp[i] != g) continue;
                b->hact[n + 1 + m] = i;
                blocks += b->sor_blocks[i];
                b->hact[++m] = blocks;
                room = budget[i] - it < room ? budget[i] - it : room;
            }
            // ids right after the m + 1 prefix entries
            for (int k = 0; k < m; k++) b->hact[m + 1 + k] = b->hact[n + 1 + k];
            HV_BHIP(b, hipMemcpyAsync(b->act, b->hact, sizeof(int) * (2 * (size_t)m + 1), hipMemcpyHostToDevice,
                                      b->stream));
            const int start = it;
            const int len = chunk < room ? chunk : room;
            const int end = it + len;
            const dim3 sgrd((unsigned)blocks, 1, 1);
            const int *afirst = b->act, *aids = b->act + m + 1;
            for (; it < end; it++)
                for (int colour = 0; colour < 2; colour++) {
                    // one launch in kSorSample is timed by start/stop events of its own dispatch
                    hipEvent_t ea = nullptr, ee = nullptr;
                    if (b->prof && ((it * 2 + colour) % kSorSample) == 3 &&
                        hipEventCreate(&ea) == hipSuccess && hipEventCreate(&ee) == hipSuccess) {
                        b->samples.push_back({f32 ? HV_KCLASS_SOR_F32 : HV_KCLASS_SOR, ea, ee});
                    } else if (ea) {
                        (void)hipEventDestroy(ea);
                        ea = nullptr;
                    }
                    if (f32)
                        hipExtLaunchKernelGGL(kb_sor_colour_f32<T>, sgrd, sblk, 0u, b->stream, ea, ee, 0u,
                                              (const BInst<T> *)a, afirst, aids, m, colour, it, max_it);
                    else
                        hipExtLaunchKernelGGL(kb_sor_colour<T>, sgrd, sblk, 0u, b->stream, ea, ee, 0u,
                                              (const BInst<T> *)a, afirst, aids, m, colour, it, max_it);
                }
            HV_BHIP(b, hipGetLastError());
            b->sweeps[f32 ? 0 : 1] += 2LL * len;
            if (b->maxc_width < len) {
                if (b->maxc) (void)hipFree(b->maxc);
                b->maxc = nullptr;
                b->maxc_width = 0;
                HV_BHIP(b, hipMalloc((void **)&b->maxc, sizeof(double) * (size_t)n * len));
                b->maxc_width = len;
            }
            int rc = batch_pinned(b, (size_t)n * len);
            if (rc) return rc;
            kb_sor_collect<T><<<(unsigned)(n * len), 64, 0, b->stream>>>(a, b->maxc, len, start, end);
            HV_BHIP(b, hipGetLastError());
            HV_BHIP(b, hipMemcpyAsync(b->pinned, b->maxc, sizeof(double) * (size_t)n * len, hipMemcpyDeviceToHost,
                                      b->stream));
            HV_BHIP(b, hipStreamSynchronize(b->stream));
            for (auto &smp : b->samples) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, smp.a, smp.e) == hipSuccess) {
                    b->prof_ms[smp.cls] += ms;
                    b->prof_launches[smp.cls] += 1;
                }
                (void)hipEventDestroy(smp.a);
                (void)hipEventDestroy(smp.e);
            }
            b->samples.clear();
            for (int i = 0; i < n; i++) {
                if (iters[i] >= 0 || grp[i] != g) continue;
                const double *row = b->pinned + (size_t)i * len;
                // (the FP32 correction stops at its own tolerance, as in sor_solve)
                const double tl = f32 ? ldexp(prm[i].sor_tol, -b->inst[i]->f32_tol_shift) : prm[i].sor_tol;
                for (int q = 0; q < len; q++) {
                    if (row[q] < tl) {
                        iters[i] = start + q + 1;
                        lastmax[i] = row[q];
                        break;
                    }
                }
                if (iters[i] < 0) {
                    lastmax[i] = row[len - 1];
                    if (end >= budget[i]) iters[i] = budget[i];  // out of budget: not converged
                }
                if (iters[i] >= 0) remaining--;
            }
            chunk = 4;
        }
    }
    return HV_OK;
}

// ---- LDS-resident FP32 correction of the whole batch ---------------------------
// tile shapes with a compile-time body: (TI, TJ, nzh)
constexpr int kRes32Shapes[][3] = {{16, 12, 48}, {16, 16, 32}, {32, 16, 16}};
constexpr int kRes32NShapes = 3;

const void *res32_batch_kernel(int shape) {
    switch (shape) {
        case 0: return (const void *)kb_sor_resident_f32<16, 12, 48>;
        case 1: return (const void *)kb_sor_resident_f32<16, 16, 32>;
        case 2: return (const void *)kb_sor_resident_f32<32, 16, 16>;
        default: return (const void *)kb_sor_resident_f32<0, 0, 0>;
    }
}

int res32_shape_of(const ResTiles &t) {
    for (int k = 0; k < kRes32NShapes; k++)
        if (t.TI == kRes32Shapes[k][0] && t.TJ == kRes32Shapes[k][1] && t.nzh == kRes32Shapes[k][2]) return k;
    return -1;
}

// Tiling of one instance for the ensemble kernel: the fewest tiles (the
// largest LDS-resident tiles) whose LDS and halo words fit, a shape with a
// compile-time body counting as 30% fewer tiles.
int batch_res_tiles(const hv_ctx *c, size_t maxlds, int maxblk, ResTiles *out) {
    const int nx = c->world.n[0], ny = c->world.n[1];
    long long best_cost = -1;
    for (int PI = 1; PI <= nx; PI++) {
        for (int PJ = 1; PJ <= ny; PJ++) {
            ResTiles t;
            t.TI = (nx + PI - 1) / PI;
            t.TJ = (ny + PJ - 1) / PJ;
            t.PI = (nx + t.TI - 1) / t.TI;
            t.PJ = (ny + t.TJ - 1) / t.TJ;
            if (t.PI != PI || t.PJ != PJ) continue;
            const int nT = t.PI * t.PJ;
            if (nT > maxblk || nT > kSlotPoll * kWave) continue;
            t.nzh = c->sp.nzh;
            t.side_i = t.TJ * t.nzh;
            t.side_j = t.TI * t.nzh;
            t.xstride = 2 * t.side_i + 2 * t.side_j;
            if (res32_lds_bytes(t) > maxlds || t.xstride > kFetchMax * kRes32Threads) continue;
            const bool fixed = res32_shape_of(t) >= 0;
            const long long cost = (long long)nT * (fixed ? 100 : 130) * 4096 + t.TI + t.TJ;
            if (best_cost < 0 || cost < best_cost) {
                best_cost = cost;
                *out = t;
            }
        }
    }
    return best_cost < 0 ? 1 : HV_OK;
}

// Tiling of one instance for the lane-owned ensemble body (deep: the
// deep-halo body, with its exchange stride): the fewest tiles that fit it.
int batch_owned_tiles(const hv_ctx *c, int maxblk, ResTiles *out, bool deep = false) {
    const int nx = c->world.n[0], ny = c->world.n[1];
    long long best_cost = -1;
    for (int PI = 1; PI <= nx; PI++) {
        for (int PJ = 1; PJ <= ny; PJ++) {
            ResTiles t;
            t.TI = (nx + PI - 1) / PI;
            t.TJ = (ny + PJ - 1) / PJ;
            t.PI = (nx + t.TI - 1) / t.TI;
            t.PJ = (ny + t.TJ - 1) / t.TJ;
            if (t.PI != PI || t.PJ != PJ) continue;
            const int nT = t.PI * t.PJ;
            if (nT > maxblk || nT > slot_poll32<kRes32Threads>() * kWave) continue;
            t.nzh = c->sp.nzh;
            t.side_i = t.TJ * t.nzh;
            t.side_j = t.TI * t.nzh;
            t.xstride = 2 * t.side_i + 2 * t.side_j;
            if (deep) {
                t.xstride = deep_xstride(t);
                if (!deep_fits<HV_DEEP_BATCH, kRes32Threads>(t)) continue;
            } else if (!owned_fits<HV_OWNED_BATCH, kRes32Threads>(t)) {
                continue;
            }
            const long long cost = (long long)nT * 4096 + t.TI + t.TJ;
            if (best_cost < 0 || cost < best_cost) {
                best_cost = cost;
                *out = t;
            }
        }
    }
    return best_cost < 0 ? 1 : HV_OK;
}

// job table layout: job_first[nblk + 1], then the (instance, tile) pairs from
// an 8-byte aligned offset
inline size_t res32_jobs_at(int nblk) { return ((size_t)nblk + 2) & ~(size_t)1; }

// Tilings and buffers for solves of up to max_it iterations; nonzero when the
// device or an instance does not allow the resident form (the batch then
// sweeps).
int batch_res32_setup(hv_batch *b, int max_it) {
    if (b->res_state < 0) return 1;
    if (b->res_state > 0 && b->res_max_it >= max_it) return HV_OK;
    b->res_state = -1;
    int dev = 0, ncu = 0, maxlds = 0, coop = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop) return 1;
    const int n = b->n;
    b->rtl.assign(n, ResTiles{});
    b->rx_off.assign(n + 1, 0);
    b->rslot_off.assign(n + 1, 0);
    size_t lds = 0;
    const void *kern = nullptr;
    // the deep-halo body, else the lane-owned body, when every instance fits it
    bool deep = b->res32_owned >= 2;
    for (int i = 0; i < n && deep; i++) deep = batch_owned_tiles(b->inst[i], ncu, &b->rtl[i], true) == HV_OK;
    bool owned = !deep && b->res32_owned >= 1;
    for (int i = 0; i < n && owned; i++) owned = batch_owned_tiles(b->inst[i], ncu, &b->rtl[i]) == HV_OK;
    b->res32_is_owned = deep ? 2 : (owned ? 1 : 0);
    b->res32_fold = deep;
    if (deep) {
        kern = (const void *)kb_sor_resident_f32d<HV_DEEP_BATCH, true>;
        lds = DeepLayout<kDeepBatchP, kDeepBatchQJ, kDeepBatchIM>::bytes(kRes32Threads / kWave);
    } else if (owned) {
        kern = (const void *)kb_sor_resident_f32o<HV_OWNED_BATCH>;
        lds = OwnedLayout<kOwnedBatchP, kOwnedBatchQJ, kOwnedBatchIM>::bytes(kRes32Threads / kWave);
    } else {
        double share[kRes32NShapes + 1] = {};  // nodes per compile-time shape
        for (int i = 0; i < n; i++) {
            if (batch_res_tiles(b->inst[i], (size_t)maxlds - 1024, ncu, &b->rtl[i]) != HV_OK) return 1;
            const int sh = res32_shape_of(b->rtl[i]);
            if (sh >= 0) share[sh] += (double)b->inst[i]->G;
            const size_t l = res32_lds_bytes(b->rtl[i]);
            lds = l > lds ? l : lds;
        }
        int best = -1;
        for (int k = 0; k < kRes32NShapes; k++)
            if (share[k] > 0.0 && (best < 0 || share[k] > share[best])) best = k;
        kern = res32_batch_kernel(best);
    }
    for (int i = 0; i < n; i++) {
        const size_t nT = (size_t)b->rtl[i].PI * b->rtl[i].PJ;
        b->rx_off[i + 1] = b->rx_off[i] + 2 * nT * (size_t)b->rtl[i].xstride;
        b->rslot_off[i + 1] = b->rslot_off[i] + (size_t)max_it * nT;
    }
    if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) return 1;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kRes32Threads, lds) != hipSuccess || per_cu < 1)
        return 1;
    b->res_nblk = ncu * per_cu;
    size_t tiles = 0;
    for (int i = 0; i < n; i++) tiles += (size_t)b->rtl[i].PI * b->rtl[i].PJ;
    if (b->rx) (void)hipFree(b->rx);
    if (b->rslots) (void)hipFree(b->rslots);
    if (b->rjob) (void)hipFree(b->rjob);
    if (b->hrjob) (void)hipHostFree(b->hrjob);
    b->rx = nullptr;
    b->rslots = nullptr;
    b->rjob = nullptr;
    b->hrjob = nullptr;
    const size_t jw = res32_jobs_at(b->res_nblk) + 2 * tiles;
    if (hipMalloc((void **)&b->rx, sizeof(unsigned long long) * b->rx_off[n]) != hipSuccess) return 1;
    if (hipMalloc((void **)&b->rslots, sizeof(float) * b->rslot_off[n]) != hipSuccess) return 1;
    if (hipMalloc((void **)&b->rjob, sizeof(int) * jw) != hipSuccess) return 1;
    if (hipHostMalloc((void **)&b->hrjob, sizeof(int) * jw, hipHostMallocDefault) != hipSuccess) return 1;
    if (!b->rinst && hipMalloc(&b->rinst, sizeof(BRes32) * (size_t)n) != hipSuccess) return 1;
    if (!b->rres && hipMalloc((void **)&b->rres, sizeof(float) * (2 * (size_t)n + 4)) != hipSuccess) return 1;
    if (!b->hrres && hipHostMalloc((void **)&b->hrres, sizeof(float) * (2 * (size_t)n + 4), hipHostMallocDefault) !=
                         hipSuccess)
        return 1;
    if (!b->hrinst && hipHostMalloc(&b->hrinst, sizeof(BRes32) * (size_t)n, hipHostMallocDefault) != hipSuccess)
        return 1;
    b->rerr = (int *)(b->rres + 2 * (size_t)n);
#ifdef HV_BRES_TRACE
    if (b->jtrace) (void)hipFree(b->jtrace);
    if (hipMalloc((void **)&b->jtrace, sizeof(unsigned long long) * 2 * tiles) != hipSuccess) return 1;
#endif
    b->rx_words = b->rx_off[n];
    b->rslot_words = b->rslot_off[n];
    b->rjobs_n = tiles;
    b->res_lds = lds;
    b->rkernel = kern;
    b->res_max_it = max_it;
    b->res_state = 1;
    return HV_OK;
}

// Job queues of one solve.  Instance i's tiles all run for about pred[i]
// iterations of a cost that grows with the tile size; instances are placed in
// order of decreasing work on the blocks that are free first (list
// scheduling), which keeps one global instance order on every queue.
void batch_res32_plan(hv_batch *b, const std::vector<int> &pred) {
    const int n = b->n, nb = b->res_nblk;
    std::vector<double> w(n);
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) {
        const ResTiles &t = b->rtl[i];
        w[i] = (double)(pred[i] > 1 ? pred[i] : 1) * ((double)t.TI * t.TJ * t.nzh + 512.0);
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return w[x] > w[y]; });
    std::vector<double> freet(nb, 0.0);
    std::vector<std::vector<std::pair<int, int>>> q(nb);
    std::vector<int> blk(nb);
    for (int i : order) {
        const int nT = b->rtl[i].PI * b->rtl[i].PJ;
        for (int k = 0; k < nb; k++) blk[k] = k;
        std::partial_sort(blk.begin(), blk.begin() + nT, blk.end(), [&](int x, int y) {
            return freet[x] < freet[y] || (freet[x] == freet[y] && x < y);
        });
        double start = 0.0;
        for (int k = 0; k < nT; k++) start = freet[blk[k]] > start ? freet[blk[k]] : start;
        for (int k = 0; k < nT; k++) {
            q[blk[k]].push_back({i, k});
            freet[blk[k]] = start + w[i];
        }
    }
    int *jf = b->hrjob, *jobs = b->hrjob + res32_jobs_at(nb);
    int pos = 0;
    for (int k = 0; k < nb; k++) {
        jf[k] = pos;
        for (const auto &e : q[k]) {
            jobs[2 * pos] = e.first;
            jobs[2 * pos + 1] = e.second;
            pos++;
        }
    }
    jf[nb] = pos;
}

// The FP32 correction iteration of every instance (delta, R of its context)
// in one cooperative launch, in three parts: the host image of the instance
// arguments and the job plan (pinned staging), the device work (copies from
// the staging, clears, the launch, the results to pinned memory; it takes no
// host decision and may be captured into a graph), and the read-back.
int batch_res32_prepare(hv_batch *b, const hv_params *prm) {
    const int n = b->n;
    const int max_it = prm[0].sor_max_it;
    BRes32 *h = (BRes32 *)b->hrinst;
    for (int i = 0; i < n; i++) {
        hv_ctx *c = b->inst[i];
        BRes32 &I = h[i];
        I.delta = c->sor_delta;
        I.R = c->sor_R;
        I.phi = c->phi;
        I.rho = (const float *)c->rho_s;
        I.sp = c->sp;
        I.tl = b->rtl[i];
        I.cf = make_sor_coef(c, prm[i].sor_omega, prm[i].sor_tol);
        I.cf.tol = ldexp(I.cf.tol, -c->f32_tol_shift);
        I.cf.omega = mixed_omega32(c, prm[i].sor_omega);
        I.cf.one_minus_omega = 1.0 - I.cf.omega;
        I.max_it = max_it;
        const size_t nT = (size_t)I.tl.PI * I.tl.PJ;
        I.sy.xr = b->rx + b->rx_off[i];
        I.sy.xb = I.sy.xr + nT * (size_t)I.tl.xstride;
        I.sy.slots = b->rslots + b->rslot_off[i];
        I.sy.err = b->rerr;
        I.sy.result = b->rres + 2 * (size_t)i;
#ifdef HV_RES_TRACE
        I.sy.trace = nullptr;
#endif
#ifdef HV_BRES_TRACE
        I.jtrace = b->jtrace;
#endif
    }
    batch_res32_plan(b, b->pred32);
    return HV_OK;
}

// the instance arguments and job plan from their pinned staging (the host
// rewrites the staging only after the stream has passed these copies)
int batch_res32_upload(hv_batch *b) {
    const size_t jw = res32_jobs_at(b->res_nblk) + 2 * b->rjobs_n;
    HV_BHIP(b, hipMemcpyAsync(b->rinst, b->hrinst, sizeof(BRes32) * (size_t)b->n, hipMemcpyHostToDevice, b->stream));
    HV_BHIP(b, hipMemcpyAsync(b->rjob, b->hrjob, sizeof(int) * jw, hipMemcpyHostToDevice, b->stream));
    return HV_OK;
}

// host_io: upload the plan first and copy the results to pinned memory after
// (the host-steered step); without it the launch uses the plan already on the
// device and leaves its results there (the device-steered step)
int batch_res32_enqueue(hv_batch *b, bool host_io = true) {
    const int n = b->n, nb = b->res_nblk;
    if (host_io) {
        int rc = batch_res32_upload(b);
        if (rc) return rc;
    }
    k_fill<float><<<4096, kThreads, 0, b->stream>>>(b->rslots, -1.0f, (long long)b->rslot_words);
    k_fill<unsigned long long><<<4096, kThreads, 0, b->stream>>>(b->rx, 0ULL, (long long)b->rx_words);
    k_fill<int><<<1, 64, 0, b->stream>>>(b->rerr, 0, 4);
    HV_BHIP(b, hipGetLastError());
    const BRes32 *ri = (const BRes32 *)b->rinst;
    const int *jf = b->rjob;
    const int2 *jobs = (const int2 *)(b->rjob + res32_jobs_at(nb));
    int lag = b->res32_lag;
    void *args[] = {(void *)&ri, (void *)&jf, (void *)&jobs, &lag};
    // timed by events of its own (a stage timer inside the SOR stage would
    // release the stage's events); never while a graph is captured
    hipEvent_t ea = nullptr, ee = nullptr;
    if (b->prof && (hipEventCreate(&ea) != hipSuccess || hipEventCreate(&ee) != hipSuccess)) {
        if (ea) (void)hipEventDestroy(ea);
        ea = ee = nullptr;
    }
    if (ea) (void)hipEventRecord(ea, b->stream);
    HV_BHIP(b, hipLaunchCooperativeKernel(b->rkernel, dim3((unsigned)nb), dim3(kRes32Threads), args,
                                          (unsigned)b->res_lds, b->stream));
    if (ee) (void)hipEventRecord(ee, b->stream);
    if (ea) b->samples.push_back({HV_KCLASS_SOR_RES_F32, ea, ee});
    if (host_io)
        HV_BHIP(b, hipMemcpyAsync(b->hrres, b->rres, sizeof(float) * (2 * (size_t)n + 4), hipMemcpyDeviceToHost,
                                  b->stream));
    return HV_OK;
}

// the sampled launch times into the stage totals (the stream has passed them)
void batch_fold_samples(hv_batch *b) {
    for (auto &smp : b->samples) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, smp.a, smp.e) == hipSuccess) {
            b->prof_ms[smp.cls] += ms;
            b->prof_launches[smp.cls] += 1;
        }
        (void)hipEventDestroy(smp.a);
        (void)hipEventDestroy(smp.e);
    }
    b->samples.clear();
}

int batch_res32_collect(hv_batch *b, std::vector<int> &iters) {
    const int n = b->n;
    HV_BHIP(b, hipStreamSynchronize(b->stream));
    batch_fold_samples(b);
    if (*(const int *)(b->hrres + 2 * (size_t)n))
        return batch_err(b, HV_ERR_HIP, "batched LDS-resident FP32 SOR: a tile wait exceeded its spin limit");
    iters.assign(n, 0);
    for (int i = 0; i < n; i++) iters[i] = (int)b->hrres[2 * (size_t)i];
    return HV_OK;
}

// ---- device-steered end of the mixed solve (hv_batch_steer.h) ---------------
constexpr int kReplanSteps = 16;  // steps between the host's read-backs of a steered call

// The cooperative finishing launch: every instance gets blocks in proportion
// to its nodes (at least one, at most its tile count) out of the resident
// blocks at up to four per CU; nonzero (the batch steers from the host) when
// the device has no cooperative launch or the instances outnumber the blocks.
template <typename T>
int batch_fin_setup(hv_batch *b, int max_it) {
    const int stride = max_it > 4 ? max_it : 4;
    if (b->fin_state < 0) return 1;
    if (b->fin_state > 0 && b->fin_stride >= stride) return HV_OK;
    b->fin_state = -1;
    int dev = 0, ncu = 0, coop = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop) return 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kb_sor_finish<T>, kFinThreads, 0) !=
            hipSuccess ||
        per_cu < 1)
        return 1;
    const int n = b->n;
    const long long cap = (long long)ncu * (per_cu < 4 ? per_cu : 4);
    if (n > cap) return 1;
    std::vector<double> w(n);
    std::vector<long long> nbk(n), most(n);
    double tot = 0.0;
    for (int i = 0; i < n; i++) {
        const SorTiles tl = make_sor_tiles(b->inst[i], kSorPlanesBatch);
        most[i] = (long long)tl.n_ic * tl.n_jb * tl.n_hb;
        w[i] = (double)b->inst[i]->sp.Gh;
        tot += w[i];
    }
    long long sum = 0;
    for (int i = 0; i < n; i++) {
        long long k = (long long)((double)cap * w[i] / tot);
        k = k < 1 ? 1 : (k > most[i] ? most[i] : k);
        nbk[i] = k;
        sum += k;
    }
    while (sum > cap) {  // (each instance keeps at least one block: n <= cap)
        int j = 0;
        for (int i = 1; i < n; i++)
            if (nbk[i] > nbk[j]) j = i;
        nbk[j]--;
        sum--;
    }
    std::vector<int> hf(n + 1, 0);
    for (int i = 0; i < n; i++) hf[i + 1] = hf[i] + (int)nbk[i];
    if (b->ffirst) (void)hipFree(b->ffirst);
    if (b->fin) (void)hipFree(b->fin);
    if (b->fslots) (void)hipFree(b->fslots);
    if (b->dhist) (void)hipFree(b->dhist);
    if (b->serr) (void)hipFree(b->serr);
    b->ffirst = nullptr;
    b->fin = nullptr;
    b->fslots = nullptr;
    b->dhist = nullptr;
    b->serr = nullptr;
    if (hipMalloc((void **)&b->ffirst, sizeof(int) * (size_t)(n + 1)) != hipSuccess ||
        hipMalloc((void **)&b->fin, sizeof(BFin) * (size_t)n) != hipSuccess ||
        hipMalloc((void **)&b->fslots, sizeof(unsigned long long) * (size_t)n * stride) != hipSuccess ||
        hipMalloc((void **)&b->dhist, sizeof(int) * (size_t)n) != hipSuccess ||
        hipMalloc((void **)&b->serr, sizeof(int)) != hipSuccess)
        return 1;
    if (hipMemcpy(b->ffirst, hf.data(), sizeof(int) * hf.size(), hipMemcpyHostToDevice) != hipSuccess) return 1;
    b->fin_nblk = hf[n];
    b->fin_stride = stride;
    b->fin_state = 1;
    return HV_OK;
}

// the FP64 finishing sweeps of every instance (device stop rule), after the fold
template <typename T>
int batch_fin_enqueue(hv_batch *b) {
    const int n = b->n;
    kb_fin_reset<<<n, 256, 0, b->stream>>>(b->fin, b->fslots, b->fin_stride);
    HV_BHIP(b, hipGetLastError());
    const BInst<T> *a = (const BInst<T> *)b->args;
    const int *ff = b->ffirst;
    const float *rr = b->rres;
    BFin *fi = b->fin;
    unsigned long long *sl = b->fslots;
    int st = b->fin_stride, mi = b->cur_max_it;
    int *er = b->serr;
    int nn = n;
    void *args[] = {(void *)&a, (void *)&ff, (void *)&nn, (void *)&rr, (void *)&fi, (void *)&sl, &st, &mi, (void *)&er};
    HV_BHIP(b, hipLaunchCooperativeKernel((const void *)kb_sor_finish<T>, dim3((unsigned)b->fin_nblk),
                                          dim3(64, kSorRows), args, 0u, b->stream));
    return HV_OK;
}

// The device part of one step (no host decision, no host wait: the part a
// graph captures): particle pass with the start-of-step kick (end_kick) and
// the diagnostics of its positions, smearing, the colour-split source and,
// with the resident FP32 correction (res), the defect, that solve and the
// fold of its correction; steered (steer), also the FP64 finishing sweeps,
// the Skyrme + Coulomb potentials and the step's record, i.e. the whole step.
template <typename T>
int batch_step_device(hv_batch *b, int end_kick, bool res, bool steer = false, bool sort = false) {
    const int n = b->n;
    const BInst<T> *a = (const BInst<T> *)b->args;
    {
        BStage st(b, HV_KCLASS_PARTICLE);
        if (sort) {
            // every instance's particles by key into the other buffer set,
            // which then becomes the current one
            k_fill<unsigned><<<1024, kThreads, 0, b->stream>>>(b->sbuck, 0u, b->sbuck_n);
            kb_sort_count<T><<<b->blocks[kBPart], kThreads, 0, b->stream>>>(a, btab(b, kBPart), n, b->sbuck);
            kb_sort_scan<<<n, kSortScanThreads, 0, b->stream>>>(b->sfirst, b->sbuck);
            kb_sort_scatter<T><<<b->blocks[kBPart], kThreads, 0, b->stream>>>(a, btab(b, kBPart), n, b->sbuck);
            kb_sort_swap<T><<<(unsigned)((n + 255) / 256), 256, 0, b->stream>>>((BInst<T> *)b->args, n);
            HV_BHIP(b, hipGetLastError());
        }
        kb_particle_step<T><<<b->blocks[kBPart], kThreads, 0, b->stream>>>(a, btab(b, kBPart), n, end_kick);
        kb_stats_finalize<T><<<n, kStatThreads, 0, b->stream>>>(a, b->dstep);
        kb_step_advance<<<1, 64, 0, b->stream>>>(b->dstep);
        HV_BHIP(b, hipGetLastError());
    }
    {
        BStage st(b, HV_KCLASS_SMEAR);
        if (b->smear_fused) {
            kb_smear_fused<T><<<b->blocks[kBSmear], kSmearThreads, b->smear_lds, b->stream>>>(a, btab(b, kBSmear),
                                                                                              n);
        } else {
            // one launch per LDS class of the instances
            for (const auto &zc : b->zycls) {
                const int *zf = b->zytab + zc.off;
                if (zc.loop)
                    kb_smear_zy_loop<T><<<zc.grid, kSmearThreads, zc.lds, b->stream>>>(a, zf, zf + zc.m + 1, zc.m);
                else
                    kb_smear_zy_ids<T><<<zc.blocks, kSmearThreads, zc.lds, b->stream>>>(a, zf, zf + zc.m + 1, zc.m);
            }
            if (b->xs_vec)
                kb_smear_xs_vec<T><<<b->blocks[kBN2V], kThreads, 0, b->stream>>>(a, btab(b, kBN2V), n);
            else
                kb_smear_xs<T><<<b->blocks[kBN2], kThreads, 0, b->stream>>>(a, btab(b, kBN2), n);
        }
        if (b->smear_fused || !b->xs_vec)  // (the vector x pass clears the counts itself)
            kb_zero_acc<T><<<b->blocks[kBN2], kThreads, 0, b->stream>>>(a, btab(b, kBN2), n);
        HV_BHIP(b, hipGetLastError());
    }
    if (b->smear_fused || !b->xs_vec) {  // (the vector x pass writes the split source itself)
        kb_rho_split<T><<<b->blocks[kBG], kThreads, 0, b->stream>>>(a, btab(b, kBG), n);
        HV_BHIP(b, hipGetLastError());
    }
    if (res) {
        BStage st(b, HV_KCLASS_SOR_STAGE);
        if (b->extrap) {
            kb_sor_extrapolate<T><<<b->blocks[kBGh2], kThreads, 0, b->stream>>>(a, btab(b, kBGh2), n, b->xorder);
            HV_BHIP(b, hipGetLastError());
        }
        // (the deep-halo resident kernel computes the residual and folds the
        // correction itself)
        if (!b->res32_fold) {
            kb_sor_defect<T><<<b->blocks[kBGh2], kThreads, 0, b->stream>>>(a, btab(b, kBGh2), n);
            HV_BHIP(b, hipGetLastError());
        }
        int rc = batch_res32_enqueue(b, !steer);
        if (rc) return rc;
        if (!b->res32_fold) {
            kb_sor_fold<T><<<b->blocks[kBGh2], kThreads, 0, b->stream>>>(a, btab(b, kBGh2), n);
            HV_BHIP(b, hipGetLastError());
        }
        if (steer) {
            rc = batch_fin_enqueue<T>(b);
            if (rc) return rc;
        }
    }
    if (steer) {
        {
            BStage st(b, HV_KCLASS_POTENTIAL);
            kb_skyrme<T><<<b->blocks[kBG], kThreads, 0, b->stream>>>(a, btab(b, kBG), n);
            HV_BHIP(b, hipGetLastError());
        }
        kb_step_record<T><<<(unsigned)((n + 255) / 256), 256, 0, b->stream>>>(
            a, n, b->rres, b->fin, b->log, b->dstep, b->extrap, b->dhist, b->xorder, b->rerr, b->serr);
        HV_BHIP(b, hipGetLastError());
    }
    return HV_OK;
}

// The rest of the step, after batch_step_device: the SOR iterations the host
// steers (the FP32 correction by sweeps when it is not resident, the FP64
// finishing sweeps of the mixed form, the FP64 solve), then the Skyrme +
// Coulomb potentials.
template <typename T>
int batch_step_finish(hv_batch *b, const hv_params *prm, bool res, int *iters_out, int *warn) {
    const int n = b->n;
    const BInst<T> *a = (const BInst<T> *)b->args;
    std::vector<int> it1, it2;
    std::vector<double> md1, md2;
    {
        BStage st(b, HV_KCLASS_SOR_STAGE);
        if (b->precision == HV_PREC_MIXED) {
            int rc;
            if (res) {
                rc = batch_res32_collect(b, it1);
            } else {
                if (b->extrap) {
                    kb_sor_extrapolate<T><<<b->blocks[kBGh2], kThreads, 0, b->stream>>>(a, btab(b, kBGh2), n,
                                                                                        b->xorder);
                    HV_BHIP(b, hipGetLastError());
                }
                kb_sor_defect<T><<<b->blocks[kBGh2], kThreads, 0, b->stream>>>(a, btab(b, kBGh2), n);
                HV_BHIP(b, hipGetLastError());
                const std::vector<int> cap(n, prm[0].sor_max_it);
                rc = batch_sor_iterate<T>(b, prm, true, b->pred32, cap, it1, md1);
                if (!rc) {
                    kb_sor_fold<T><<<b->blocks[kBGh2], kThreads, 0, b->stream>>>(a, btab(b, kBGh2), n);
                    if (hipGetLastError() != hipSuccess) rc = batch_err(b, HV_ERR_HIP, "fold launch failed");
                }
            }
            if (rc) return rc;
            std::vector<int> two(n, 2), budget(n);
            for (int i = 0; i < n; i++) budget[i] = mixed_fp64_budget(prm[0].sor_max_it, it1[i]);
            rc = batch_sor_iterate<T>(b, prm, false, two, budget, it2, md2);
            if (rc) return rc;
            for (int i = 0; i < n; i++) {
                b->pred32[i] = it1[i];
                b->inst[i]->cnt_it32 += it1[i];
                b->inst[i]->cnt_it64 += it2[i];
                b->inst[i]->mixed_fp64_sweeps = it2[i];
                iters_out[i] = it1[i] + it2[i];
                if (md2[i] > prm[i].sor_tol) *warn |= HV_WARN_SOR_NOT_CONVERGED;
                // as sor_solve: consecutive converged loop solves
                hv_ctx *c = b->inst[i];
                c->last_extrap = b->extrap ? b->hxorder[2 * i] : 0;
                c->hist_n = md2[i] > prm[i].sor_tol ? 0 : (c->hist_n < 6 ? c->hist_n + 1 : 6);
            }
        } else {
            const std::vector<int> cap(n, prm[0].sor_max_it);
            int rc = batch_sor_iterate<T>(b, prm, false, b->pred64, cap, it1, md1);
            if (rc) return rc;
            for (int i = 0; i < n; i++) {
                b->pred64[i] = it1[i];
                b->inst[i]->cnt_it64 += it1[i];
                iters_out[i] = it1[i];
                if (md1[i] > prm[i].sor_tol) *warn |= HV_WARN_SOR_NOT_CONVERGED;
            }
        }
        for (int i = 0; i < n; i++) b->inst[i]->cnt_solves++;
    }
    {
        BStage st(b, HV_KCLASS_POTENTIAL);
        kb_skyrme<T><<<b->blocks[kBG], kThreads, 0, b->stream>>>(a, btab(b, kBG), n);
        HV_BHIP(b, hipGetLastError());
    }
    return HV_OK;
}

// Captures batch_step_device into b->gexec; nonzero (and no graph) when the
// runtime refuses any part of it, the stream then back in normal mode.
template <typename T>
int batch_capture_step(hv_batch *b, bool res, bool steer, bool sort) {
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    const int rc = batch_step_device<T>(b, 1, res, steer, sort);
    const hipError_t ec = hipStreamEndCapture(b->stream, &g);
    (void)hipGetLastError();
    if (rc || ec != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        b->err[0] = 0;
        return 1;
    }
    hipGraphExec_t ge = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess || !ge) {
        (void)hipGetLastError();
        return 1;
    }
    (sort ? b->gexec_sort : b->gexec) = ge;
    b->captures++;
    return HV_OK;
}

void batch_drop_graph(hv_batch *b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->gexec_sort) (void)hipGraphExecDestroy(b->gexec_sort);
    b->gexec = nullptr;
    b->gexec_sort = nullptr;
}

int batch_check(hv_batch *b, const hv_skyrme *sk, const hv_params *prm) {
    if (!sk || !prm) return batch_err(b, HV_ERR_ARG, "null skyrme or params");
    for (int i = 0; i < b->n; i++) {
        int rc = check_params(b->inst[i], &sk[i], &prm[i]);
        if (rc) return batch_fwd(b, b->inst[i], rc);
        if (prm[i].sor_max_it != prm[0].sor_max_it)
            return batch_err(b, HV_ERR_ARG, "the instances of a batch share sor_max_it");
    }
    return HV_OK;
}

template <typename T>
int batch_run_init(hv_batch *b, const hv_skyrme *sk, const hv_params *prm, int *sor_iters) {
    int rc = batch_check(b, sk, prm);
    if (rc) return rc;
    int warn = 0;
    for (int i = 0; i < b->n; i++) {
        int it = 0;
        rc = run_init<T>(b->inst[i], &sk[i], &prm[i], &it);
        if (rc < 0) return batch_fwd(b, b->inst[i], rc);
        warn |= rc;
        if (sor_iters) sor_iters[i] = it;
        b->pred32[i] = 16;
        b->pred64[i] = it;
        b->boost[i] = 0;
    }
    b->ready = true;
    return warn;
}

// Reads back the log rows [r0, r1) of a steered call (the stream is synchronised).
int batch_read_log(hv_batch *b, int r0, int r1, std::vector<BLog> &out) {
    const size_t cnt = (size_t)(r1 - r0) * b->n;
    out.resize(cnt);
    if (!cnt) return HV_OK;
    HV_BHIP(b, hipMemcpyAsync(out.data(), b->log + (size_t)r0 * b->n, sizeof(BLog) * cnt, hipMemcpyDeviceToHost,
                              b->stream));
    HV_BHIP(b, hipStreamSynchronize(b->stream));
    batch_fold_samples(b);
    int e = 0;
    HV_BHIP(b, hipMemcpy(&e, b->serr, sizeof(int), hipMemcpyDeviceToHost));
    if (e) return batch_err(b, HV_ERR_HIP, "steered ensemble SOR: a wait exceeded its spin limit");
    return HV_OK;
}

template <typename T>
int batch_run_steps(hv_batch *b, const hv_skyrme *sk, const hv_params *prm, int first, int nsteps, double *stats,
                    int *sor_iters) {
    if (!b->ready) return batch_err(b, HV_ERR_STATE, "hv_batch_run_init first");
    int rc = batch_check(b, sk, prm);
    if (rc) return rc;
    if (nsteps <= 0) return HV_OK;
    const int n = b->n;
    for (int i = 0; i < n; i++) {
        hv_ctx *c = b->inst[i];
        rc = ensure_slots(c, nsteps + 1);
        if (rc) return batch_fwd(b, c, rc);
        rc = stats_into<T>(c, 0);  // diagnostics of the positions at the start of the first step
        if (rc) return batch_fwd(b, c, rc);
        b->boost[i] = 0;
    }
    rc = batch_build<T>(b, sk, prm, nsteps + 1);
    if (rc) return rc;
    const BInst<T> *a = (const BInst<T> *)b->args;
    const bool res = b->precision == HV_PREC_MIXED && b->res32 && batch_res32_setup(b, prm[0].sor_max_it) == HV_OK;
    // the whole mixed step on the device: the host reads back every
    // kReplanSteps steps (to re-plan the resident launch) and at the end
    const bool steer = res && b->steer && batch_fin_setup<T>(b, prm[0].sor_max_it) == HV_OK;
    k_fill<int><<<1, 64, 0, b->stream>>>(b->dstep, 0, 1);  // diagnostics slot of the first step: 1
    HV_BHIP(b, hipGetLastError());
    if (steer) {
        b->cur_max_it = prm[0].sor_max_it;
        if (b->log_cap < (size_t)n * nsteps) {
            if (b->log) (void)hipFree(b->log);
            b->log = nullptr;
            b->log_cap = 0;
            HV_BHIP(b, hipMalloc((void **)&b->log, sizeof(BLog) * (size_t)n * nsteps));
            b->log_cap = (size_t)n * nsteps;
        }
        // the history as the host holds it, and the first step's predictor and
        // ring position (the device advances them after every step)
        std::vector<int> hh(n);
        for (int i = 0; i < n; i++) {
            hv_ctx *c = b->inst[i];
            hh[i] = c->hist_n;
            b->hxorder[2 * i] = b->extrap ? sor_predictor_for(b->extrap, c->hist_n) : 0;
            b->hxorder[2 * i + 1] = c->hist_pos;
        }
        HV_BHIP(b, hipMemcpyAsync(b->dhist, hh.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, b->stream));
        HV_BHIP(b, hipMemcpyAsync(b->xorder, b->hxorder, sizeof(int) * 2 * (size_t)n, hipMemcpyHostToDevice,
                                  b->stream));
        k_fill<int><<<1, 64, 0, b->stream>>>(b->serr, 0, 1);
        HV_BHIP(b, hipGetLastError());
        HV_BHIP(b, hipStreamSynchronize(b->stream));
    }
    // steps after the first of this call (start-of-step kick) replay one
    // captured graph of the device part, unless stages are being timed
    bool graphs = b->use_graph && !b->prof;
    batch_drop_graph(b);
    std::vector<int> its(n);
    int warn = 0;
    std::vector<BLog> lg;
    for (int s = 0; s < nsteps; s++) {
        const int step = first + s;
        bool dirty = false;
        // the batched sort's cadence (as maybe_sort: the first step, then every
        // sort_every steps); the sort runs at the head of the step's particle stage
        bool sort_now = false;
        if (b->sort_every > 0) {
            if (b->since_sort < b->sort_every) {
                b->since_sort++;
            } else {
                sort_now = true;
                b->since_sort = 1;
            }
        }
        for (int i = 0; i < n; i++) {
            hv_ctx *c = b->inst[i];
            int bo = 0;
            if (step * step_dt(prm[i]) >= prm[i].t_exc && !c->excited) {
                c->excited = 1;
                bo = 1;
            }
            if (bo != b->boost[i]) dirty = true;
            b->boost[i] = bo;
        }
        if (dirty) {
            batch_fill_args<T>(b, sk, prm);
            rc = batch_upload_args<T>(b);
            if (rc) return rc;
        }
        if (steer) {
            if (s % kReplanSteps == 0) {
                // the job plan of the resident launch from the last step's counts
                if (s > 0) {
                    rc = batch_read_log(b, s - 1, s, lg);
                    if (rc) return rc;
                    for (int i = 0; i < n; i++) b->pred32[i] = lg[i].it32;
                } else {
                    HV_BHIP(b, hipStreamSynchronize(b->stream));  // the staging is free
                }
                rc = batch_res32_prepare(b, prm);
                if (rc) return rc;
                rc = batch_res32_upload(b);
                if (rc) return rc;
            }
        } else {
            if (res) {
                rc = batch_res32_prepare(b, prm);
                if (rc) return rc;
            }
            if (b->extrap) {
                // predictor and ring position of each instance's solve (as sor_solve)
                for (int i = 0; i < n; i++) {
                    hv_ctx *c = b->inst[i];
                    b->hxorder[2 * i] = sor_predictor_for(b->extrap, c->hist_n);
                    b->hxorder[2 * i + 1] = c->hist_pos;
                    c->hist_pos = (c->hist_pos + 4) % 5;
                }
                HV_BHIP(b, hipMemcpyAsync(b->xorder, b->hxorder, sizeof(int) * 2 * (size_t)n,
                                          hipMemcpyHostToDevice, b->stream));
            }
        }
        if (s > 0 && graphs && !(sort_now ? b->gexec_sort : b->gexec) &&
            batch_capture_step<T>(b, res, steer, sort_now) != HV_OK)
            graphs = false;
        if (s > 0 && graphs) {
            HV_BHIP(b, hipGraphLaunch(sort_now ? b->gexec_sort : b->gexec, b->stream));
            b->graph_steps++;
        } else {
            rc = batch_step_device<T>(b, s > 0 ? 1 : 0, res, steer, sort_now);
            if (rc) return rc;
        }
        if (sort_now) {
            // the host's view of the buffer sets follows the device's swap
            for (hv_ctx *c : b->inst) {
                for (int q = 0; q < 9; q++) std::swap(c->part[q], c->part_alt[q]);
                std::swap(c->pid, c->pid_alt);
                c->permuted = true;
            }
            batch_fill_args<T>(b, sk, prm);
        }
        if (!steer) {
            rc = batch_step_finish<T>(b, prm, res, its.data(), &warn);
            if (rc) return rc;
            if (sor_iters)
                for (int i = 0; i < n; i++) sor_iters[(size_t)i * nsteps + s] = its[i];
        }
    }
    batch_drop_graph(b);
    if (steer) {
        // the steps' records: iteration counts, warnings, the solve counters and
        // the history the host keeps for the next call
        rc = batch_read_log(b, 0, nsteps, lg);
        if (rc) return rc;
        std::vector<int> hh(n);
        HV_BHIP(b, hipMemcpy(hh.data(), b->dhist, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost));
        HV_BHIP(b, hipMemcpy(b->hxorder, b->xorder, sizeof(int) * 2 * (size_t)n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; i++) {
            hv_ctx *c = b->inst[i];
            for (int s = 0; s < nsteps; s++) {
                const BLog &L = lg[(size_t)s * n + i];
                if (sor_iters) sor_iters[(size_t)i * nsteps + s] = L.it32 + L.it64;
                if (L.md > prm[i].sor_tol) warn |= HV_WARN_SOR_NOT_CONVERGED;
                c->cnt_it32 += L.it32;
                c->cnt_it64 += L.it64;
                c->cnt_solves++;
            }
            const BLog &last = lg[(size_t)(nsteps - 1) * n + i];
            b->pred32[i] = last.it32;
            c->mixed_fp64_sweeps = last.it64;
            c->last_extrap = last.pred;
            c->hist_n = hh[i];
            if (b->extrap) c->hist_pos = b->hxorder[2 * i + 1];
        }
    }
    // closing half kick of the last step, with the force stored (src/hyper_vlasov.cpp:101-103)
    kb_gather_kick<T><<<b->blocks[kBPart], kThreads, 0, b->stream>>>(a, btab(b, kBPart), n);
    HV_BHIP(b, hipGetLastError());
    bool clear = false;
    for (int i = 0; i < n; i++) {
        clear = clear || b->boost[i];
        b->boost[i] = 0;
    }
    if (clear) {
        batch_fill_args<T>(b, sk, prm);
        rc = batch_upload_args<T>(b);
        if (rc) return rc;
    }
    for (int i = 0; i < n; i++) {
        hv_ctx *c = b->inst[i];
        if (stats) {
            std::vector<StatSlot> h((size_t)nsteps + 1);
            HV_BHIP(b, hipMemcpyAsync(h.data(), c->slots, sizeof(StatSlot) * h.size(), hipMemcpyDeviceToHost,
                                      b->stream));
            HV_BHIP(b, hipStreamSynchronize(b->stream));
            for (int s = 0; s < nsteps; s++) slot_stats(h[s], c->np, c->nn, stats + ((size_t)i * nsteps + s) * 4);
        }
        int st = 0;
        rc = take_status(c, &st);
        if (rc) return batch_fwd(b, c, rc);
        warn |= st & HV_WARN_PARTICLE_OUTSIDE;
    }
    return warn;
}

}  // namespace

extern "C" {

int hv_batch_create(int n, const hv_world *worlds, const int *n_protons, const int *n_neutrons, int precision,
                    hv_batch **out) {
    if (!out) return HV_ERR_ARG;
    *out = nullptr;
    if (n <= 0 || !worlds || !n_protons || !n_neutrons) return HV_ERR_ARG;
    if (precision != HV_PREC_F64 && precision != HV_PREC_MIXED) return HV_ERR_ARG;
    hv_batch *b = new (std::nothrow) hv_batch();
    if (!b) return HV_ERR_ALLOC;
    b->n = n;
    b->precision = precision;
    b->pred32.assign(n, 16);
    b->pred64.assign(n, 16);
    b->boost.assign(n, 0);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&b->dstep, sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&b->xorder, sizeof(int) * 2 * (size_t)n) != hipSuccess ||
        hipHostMalloc((void **)&b->hxorder, sizeof(int) * 2 * (size_t)n, hipHostMallocDefault) != hipSuccess) {
        hv_batch_destroy(b);
        return HV_ERR_HIP;
    }
    for (int i = 0; i < n; i++) {
        hv_ctx *c = nullptr;
        int rc = hv_ctx_create(&worlds[i], n_protons[i], n_neutrons[i], precision, &c);
        if (rc) {
            hv_batch_destroy(b);
            return rc;
        }
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
        c->stream = b->stream;
        c->own_stream = false;
        // solo solves on an instance (the pre-loop solve) use the per-sweep form
        c->sor_res32 = 0;
        c->sor_fused = 0;
        c->sor_resident = 0;
        c->sort_every = b->sort_every;
        b->inst.push_back(c);
    }
    b->extrap = precision == HV_PREC_MIXED ? b->inst[0]->sor_extrap : 0;
    *out = b;
    return HV_OK;
}

int hv_batch_destroy(hv_batch *b) {
    if (!b) return HV_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (hv_ctx *c : b->inst) hv_ctx_destroy(c);
    if (b->args) (void)hipFree(b->args);
    if (b->first) (void)hipFree(b->first);
    if (b->maxc) (void)hipFree(b->maxc);
    if (b->rinst) (void)hipFree(b->rinst);
    if (b->rx) (void)hipFree(b->rx);
    if (b->rslots) (void)hipFree(b->rslots);
    if (b->rres) (void)hipFree(b->rres);
    if (b->rjob) (void)hipFree(b->rjob);
    if (b->hrjob) (void)hipHostFree(b->hrjob);
    if (b->hrres) (void)hipHostFree(b->hrres);
    if (b->hrinst) (void)hipHostFree(b->hrinst);
    if (b->dstep) (void)hipFree(b->dstep);
    if (b->xorder) (void)hipFree(b->xorder);
    if (b->hxorder) (void)hipHostFree(b->hxorder);
    if (b->sbuck) (void)hipFree(b->sbuck);
    if (b->sfirst) (void)hipFree(b->sfirst);
    if (b->zytab) (void)hipFree(b->zytab);
    if (b->ffirst) (void)hipFree(b->ffirst);
    if (b->fin) (void)hipFree(b->fin);
    if (b->fslots) (void)hipFree(b->fslots);
    if (b->log) (void)hipFree(b->log);
    if (b->dhist) (void)hipFree(b->dhist);
    if (b->serr) (void)hipFree(b->serr);
    batch_fold_samples(b);
    batch_drop_graph(b);
    if (b->act) (void)hipFree(b->act);
    if (b->hact) (void)hipHostFree(b->hact);
    if (b->pinned) (void)hipHostFree(b->pinned);
    for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return HV_OK;
}

const char *hv_batch_last_error(const hv_batch *b) { return b ? b->err : "null batch"; }

int hv_batch_size(const hv_batch *b) { return b ? b->n : 0; }

hv_ctx *hv_batch_instance(hv_batch *b, int i) {
    if (!b || i < 0 || i >= b->n) return nullptr;
    return b->inst[i];
}

int hv_batch_get_option(hv_batch *b, int option, int *value) {
    if (!b || !value) return HV_ERR_ARG;
    switch (option) {
    case HV_OPT_SORT_EVERY: *value = b->sort_every; return HV_OK;
    case HV_OPT_BATCH_GROUP_KIB: *value = b->group_kib; return HV_OK;
    case HV_OPT_SOR_RESIDENT_F32: *value = b->res32 && b->res_state >= 0; return HV_OK;
    case HV_OPT_SOR_F32_LAG: *value = b->res32_lag; return HV_OK;
    case HV_OPT_SOR_F32_OWNED:  // once tiled: whether the lane-owned body runs
        *value = b->res_state > 0 ? b->res32_is_owned : b->res32_owned;
        return HV_OK;
    case HV_OPT_SMEAR_FUSED: *value = b->smear_fused; return HV_OK;
    case HV_OPT_BATCH_GRAPH: *value = b->use_graph; return HV_OK;
    case HV_OPT_SOR_EXTRAPOLATE: *value = b->extrap; return HV_OK;
    case HV_OPT_SOR_F32_OMEGA:
    case HV_OPT_SOR_F32_TOL_SHIFT:
        return b->n > 0 ? hv_get_option(b->inst[0], option, value) : HV_OK;
    case HV_OPT_BATCH_DEVICE_STEER:  // once the batch has stepped: whether the steered step runs
        *value = b->steer && b->fin_state >= 0 && b->precision == HV_PREC_MIXED && b->res32 && b->res_state >= 0;
        return HV_OK;
    default: return batch_err(b, HV_ERR_ARG, "option %d is not a batch option", option);
    }
}

int hv_batch_set_option(hv_batch *b, int option, int value) {
    if (!b) return HV_ERR_ARG;
    if (option == HV_OPT_BATCH_GROUP_KIB) {
        if (value < 0) return batch_err(b, HV_ERR_ARG, "group budget must be >= 0");
        b->group_kib = value;
        return HV_OK;
    }
    if (option == HV_OPT_SOR_RESIDENT_F32) {
        b->res32 = value ? 1 : 0;
        return HV_OK;
    }
    if (option == HV_OPT_SMEAR_FUSED) {
        b->smear_fused = value ? 1 : 0;
        return HV_OK;
    }
    if (option == HV_OPT_SOR_F32_OWNED) {
        if (value < 0 || value > 2) return batch_err(b, HV_ERR_ARG, "the FP32 SOR body must be 0, 1 or 2");
        if (value != b->res32_owned) {
            // tilings and kernel chosen again at the next solve
            if (b->stream) (void)hipStreamSynchronize(b->stream);
            batch_drop_graph(b);
            b->res_state = 0;
            b->res32_owned = value;
        }
        return HV_OK;
    }
    if (option == HV_OPT_BATCH_GRAPH) {
        b->use_graph = value ? 1 : 0;
        return HV_OK;
    }
    if (option == HV_OPT_BATCH_DEVICE_STEER) {
        b->steer = value ? 1 : 0;
        return HV_OK;
    }
    if (option == HV_OPT_SOR_EXTRAPOLATE) {
        for (hv_ctx *c : b->inst) {
            int rc = hv_set_option(c, option, value);
            if (rc) return batch_fwd(b, c, rc);
        }
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        batch_drop_graph(b);  // the step's device part gains or loses a kernel
        b->extrap = value;
        return HV_OK;
    }
    if (option == HV_OPT_SOR_F32_OMEGA || option == HV_OPT_SOR_F32_TOL_SHIFT) {
        for (hv_ctx *c : b->inst) {
            int rc = hv_set_option(c, option, value);
            if (rc) return batch_fwd(b, c, rc);
        }
        return HV_OK;  // (the coefficients are rebuilt at the start of every call)
    }
    if (option == HV_OPT_SOR_F32_LAG) {
        if (value < 1 || value > 64) return batch_err(b, HV_ERR_ARG, "lag must be in 1..64");
        b->res32_lag = value;
        return HV_OK;
    }
    if (option != HV_OPT_SORT_EVERY && option != HV_OPT_SMEAR_TILE)
        return batch_err(b, HV_ERR_ARG, "option %d is not a batch option", option);
    for (hv_ctx *c : b->inst) {
        int rc = hv_set_option(c, option, value);
        if (rc) return batch_fwd(b, c, rc);
    }
    if (option == HV_OPT_SORT_EVERY) {
        b->sort_every = value;
        b->since_sort = 1 << 30;
    }
    return HV_OK;
}

int hv_batch_run_init(hv_batch *b, const hv_skyrme *skyrme, const hv_params *params, int *sor_iters) {
    if (!b) return HV_ERR_ARG;
    return b->precision == HV_PREC_F64 ? batch_run_init<double>(b, skyrme, params, sor_iters)
                                       : batch_run_init<float>(b, skyrme, params, sor_iters);
}

int hv_batch_run_steps(hv_batch *b, const hv_skyrme *skyrme, const hv_params *params, int first_step, int nsteps,
                       double *stats, int *sor_iters) {
    if (!b) return HV_ERR_ARG;
    if (first_step < 0 || nsteps < 0) return batch_err(b, HV_ERR_ARG, "negative step range");
    return b->precision == HV_PREC_F64
               ? batch_run_steps<double>(b, skyrme, params, first_step, nsteps, stats, sor_iters)
               : batch_run_steps<float>(b, skyrme, params, first_step, nsteps, stats, sor_iters);
}

int hv_batch_profile_enable(hv_batch *b, int on) {
    if (!b) return HV_ERR_ARG;
    HV_BHIP(b, hipStreamSynchronize(b->stream));
    b->prof = on != 0;
    for (int k = 0; k < HV_KCLASS_COUNT; k++) {
        b->prof_ms[k] = 0.0;
        b->prof_launches[k] = 0;
    }
    for (hv_ctx *c : b->inst) c->cnt_solves = c->cnt_it32 = c->cnt_it64 = 0;
    b->sweeps[0] = b->sweeps[1] = 0;
    return HV_OK;
}

#ifdef HV_BRES_TRACE
// start/end clocks (100 MHz) of the last ensemble-resident launch's jobs, and
// the job table (job_first[nblk + 1] then (instance, tile) pairs)
int hv_batch_debug_trace(hv_batch *b, unsigned long long *times, int *jobs, int *nblk, int *njobs) {
    if (!b || !b->jtrace) return HV_ERR_STATE;
    HV_BHIP(b, hipStreamSynchronize(b->stream));
    HV_BHIP(b, hipMemcpy(times, b->jtrace, sizeof(unsigned long long) * 2 * b->rjobs_n, hipMemcpyDeviceToHost));
    memcpy(jobs, b->hrjob, sizeof(int) * (res32_jobs_at(b->res_nblk) + 2 * b->rjobs_n));
    *nblk = b->res_nblk;
    *njobs = (int)b->rjobs_n;
    return HV_OK;
}
#endif

int hv_batch_graph_info(hv_batch *b, long long *graph_steps, long long *captures) {
    if (!b) return HV_ERR_ARG;
    if (graph_steps) *graph_steps = b->graph_steps;
    if (captures) *captures = b->captures;
    return HV_OK;
}

int hv_batch_sor_counts(hv_batch *b, long long *sweeps_f32, long long *sweeps_f64, long long *node_iters_f32,
                        long long *node_iters_f64) {
    if (!b) return HV_ERR_ARG;
    long long n32 = 0, n64 = 0;
    for (hv_ctx *c : b->inst) {
        const long long interior = (long long)(c->world.n[0] - 2) * (c->world.n[1] - 2) * (c->world.n[2] - 2);
        n32 += c->cnt_it32 * interior;
        n64 += c->cnt_it64 * interior;
    }
    if (sweeps_f32) *sweeps_f32 = b->sweeps[0];
    if (sweeps_f64) *sweeps_f64 = b->sweeps[1];
    if (node_iters_f32) *node_iters_f32 = n32;
    if (node_iters_f64) *node_iters_f64 = n64;
    return HV_OK;
}

int hv_batch_profile_read(hv_batch *b, int kclass, double *ms, long long *launches) {
    if (!b || kclass < 0 || kclass >= HV_KCLASS_COUNT) return HV_ERR_ARG;
    HV_BHIP(b, hipStreamSynchronize(b->stream));
    if (ms) *ms = b->prof_ms[kclass];
    if (launches) *launches = b->prof_launches[kclass];
    return HV_OK;
}

}  // extern "C"
