// hv_abi.hip -- context, buffers and the extern "C" entry points of
// libhvamd.so (declared in include/hv_abi.h).
//
// The host side owns one HIP stream per context and every device buffer of
// the loop.  It launches the kernels of hv_kernels.h in the order of the
// reference's cpu_simulate (src/hyper_vlasov.cpp:53-104) and drives the SOR
// iteration count exactly as the reference's stop rule does
// (include/physics.hpp:104, 152-153).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <functional>
#include <cstring>
#include <cstdarg>
#include <new>
#include <vector>


#include "hv_abi.h"
#include "hv_kernels.h"
#include "hv_sor_resident.h"
#include "hv_sor_owned.h"

#define HV_XSTR(x) HV_STR(x)
#define HV_STR(x) #x
#define HIP_VERSION_STR HV_XSTR(HIP_VERSION_MAJOR) "." HV_XSTR(HIP_VERSION_MINOR) "." HV_XSTR(HIP_VERSION_PATCH)

using namespace hv;

struct hv_ctx {
    int precision = HV_PREC_F64;
    hv_world world{};
    int np = 0, nn = 0, total = 0;
    long long G = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // The work after a loop solve (potential, force grid), enqueued ahead of
    // the FP64 sweeps' first read-back so that the device runs it while the
    // host waits; spec_valid: that first chunk met the stop rule, so the
    // enqueued work saw the final phi (else it is enqueued again after the
    // solve and overwrites it).  ev_copy: the read-back's completion.
    std::function<int()> spec;
    bool spec_valid = false;
    bool rho_s_ready = false;           // the x pass wrote rho_s from the density just smeared
    bool cbox_ready = false;            // the count box of the last deposit is built (k_stats_cbox)
    int sor_rows_cleared = 0;           // rows of maxd (and SorCtl) the extrapolation cleared for the next sor_iterate
    hipEvent_t ev_copy = nullptr;
    char err[512] = {0};

    // particle SoA (T = double or float): x y z kx ky kz fx fy fz
    void *part[9] = {};
    unsigned long long *acc = nullptr;  // 2G fixed-point CIC counts
    void *density = nullptr;            // 2G T
    void *U = nullptr;                  // 2G T
    double *phi = nullptr;              // 2 Gh, colour-split (hv::Split)
    void *rho_s = nullptr;              // 2 Gh T: colour-split copy of the proton density
    float *sor_R = nullptr;             // mixed SOR: FP32 residual and correction (2 Gh each)
    float *sor_delta = nullptr;
    int mixed_fp64_sweeps = 0;          // FP64 iterations of the last mixed solve
    long long cnt_solves = 0, cnt_it32 = 0, cnt_it64 = 0;  // SOR work since hv_profile_enable
    bool phi_warm = false;              // phi holds a previous solve (mixed SOR corrects it in FP32)
    // mixed form in the time loop: a ring of the potentials of the five solves
    // before the last (split layout; newest at phi_h[hist_pos]) and how many
    // consecutive converged loop solves phi and the ring hold (0..6): the start
    // of the next solve is extrapolated from them (sor_extrapolate_body)
    double *phi_h[5] = {};
    int hist_n = 0;
    int hist_pos = 0;
    int sor_extrap = 3;                 // option HV_OPT_SOR_EXTRAPOLATE (mixed form only)
    int last_extrap = 0;                // predictor used by the last solve
    int f32_tol_shift = 0;              // option HV_OPT_SOR_F32_TOL_SHIFT: FP32 stage tolerance tol / 2^shift
    int f32_omega = 0;                  // option HV_OPT_SOR_F32_OMEGA: 0 grid rule, 1 the run's omega, else /1000
    int sor_fused = 0;                  // option: one launch per SOR iteration (red + black fused; slower at 128^3)
    double *phi_alt = nullptr;          // second potential buffer of the fused SOR
    float *sor_delta_alt = nullptr;
    Split sp{};
    void *D = nullptr;                  // smearing: D channels of the two-pass form, (rx + 1) 2G T
    size_t D_cap = 0;                   // bytes
    void *F[3] = {};                    // 2G T each (per-kernel API only)
    void *smear_tab = nullptr;          // SmearTab<T> on device
    double tab_sigma = -1.0;
    int tab_ppn = -1;
    int tab_warn = 0;
    SmearTiles stl{};                   // two-pass smearing tiles
    size_t stl_lds = 0;
    unsigned *sflags = nullptr;         // plane masks of the tile columns of the two-pass smearing
    size_t sflags_cap = 0;
    SmearTiles stf{};                   // fused (plane-walking) smearing tiles
    size_t stf_lds = 0;
    int smear_fused = 0;                // option: single-grid smearing in the fused form
    int smear_xs_vec = 1;               // two-pass x pass: VW nodes per thread (when nz allows)
    int xs_rxm = 0;                     // its radius bucket (0: radius too large for it)
    long long zy_blocks = 1024;         // workgroups of the zy pass (blocks per CU that its LDS allows x CUs)
    int zy_staged = 0;                  // counts a zy item stages: (TJ + 2 ry)(TK + 2 rz) of stl
    unsigned clear_blocks = 2048;       // workgroups of k_clear_counts and k_force_grid (one box row per wave)
    unsigned *fbox = nullptr;           // the count boxes of the last smearing (2 x 6 words)
    int *brec = nullptr;                // per-block box records of the last deposit (kBoxRec ints per block)
    void *fgrid = nullptr;              // force grid (fx, fy, fz, 0) per node, T, within fbox (run_steps)
    size_t fgrid_cap = 0;
    unsigned long long *maxd = nullptr; // per-iteration slot maxima of |dphi| (double bits)
    double *maxc = nullptr;             // per-iteration max |dphi|, collected for read-back
    int maxd_cap = 0;
    long long *partials = nullptr;      // per-block diagnostics partial sums (6 per block)
    long long *en_partials = nullptr;   // total energy: per-block partial sums
    unsigned long long *en_out = nullptr;  // total energy: exact 128-bit sums (3 components)
    SorCtl *ctl = nullptr;
    StatSlot *slots = nullptr;
    int slot_cap = 0;
    unsigned *status = nullptr;
    double *pinned = nullptr;           // host staging for small read-backs
    size_t pinned_cap = 0;              // doubles
    int sor_pred = 16;
    int excited = 0;
    int last_outside = 0;

    // particle sort by (species, brick): storage order differs from the
    // caller's order once sorted; pid maps storage slot -> original index
    int *pid = nullptr, *pid_alt = nullptr;
    void *part_alt[9] = {};
    unsigned *keys[2] = {};             // [0]: sort keys; [1]: key counts / first slots (sort_buckets)
    int *ord[2] = {};                   // [0]: ranks within a key
    int sort_every = 8;  // steps between sorts; 0 = never
    int smear_tk_max = 64;  // widest smearing tile (columns) tried
    bool smear_tj16 = true; // two-pass vector tile of 16 rows when its LDS fits (else 8)
    bool own_stream = true; // false for the instances of a batch (the batch's stream)
    // LDS-resident SOR (hv_sor_resident.h)
    int sor_resident = 0;     // option: use it when the grid fits (off by default: slower than per-sweep launches at 64^3..128^3)
    int res_state = 0;        // 0 unknown, 1 available, -1 not available
    ResTiles rtl{};
    size_t res_lds = 0;
    double *res_x = nullptr;  // exchange buffers (2 colours)
    int *res_flag = nullptr;
    double *res_slots = nullptr;
    int res_slots_cap = 0;
    int *res_err = nullptr;
    double *res_result = nullptr;
    unsigned long long *res_trace = nullptr;  // phase timestamps (HV_RES_TRACE builds only)
    // LDS-resident FP32 correction iteration of the mixed form
    int sor_res32 = 1;        // option: use it when the grid fits
    int res32_lag = 2;        // option: iterations between a max and its convergence test
    int res32_per_cu = 1;     // option: LDS-resident FP32 tiles per CU (1 or 2)
    int res32_owned = 2;      // option: 2 deep-halo, 1 lane-owned, 0 per-row body (the best that fits)
    int res32_is_owned = 0;   // the body the chosen kernel runs (same values)
    int res32_threads = 1024;
    int res32_state = 0;
    ResTiles rtl32{};
    size_t res32_lds = 0;
    unsigned long long *res32_x = nullptr;  // tagged exchange words (2 colours)
    const void *res32_kernel = nullptr;     // instance for the chosen tile shape
    const void *res32_kernel_fold = nullptr;  // the deep-halo body with the defect and fold in it (or none)
    float *res32_slots = nullptr;
    int res32_slots_cap = 0;
    int *res32_err = nullptr;
    float *res32_result = nullptr;
    float *res32_hres = nullptr;        // pinned: iterations, tested max, error word
    double xread[3] = {0.0, 0.0, 0.0};  // the resident FP32 solve's result read back with the FP64 sweeps' first chunk
    int since_sort = 1 << 30;
    bool permuted = false;

    // kernel-class profiling
    bool prof = false;
    double prof_ms[HV_KCLASS_COUNT] = {};
    long long prof_launches[HV_KCLASS_COUNT] = {};
    long long prof_issued[HV_KCLASS_COUNT] = {};  // launches of a sampled class issued, timed or not
    struct Pending {
        int cls;
        hipEvent_t a, b;
        long long launches;
    };
    std::vector<hipEvent_t> ev_free;
    std::vector<Pending> pending;
};

namespace {

constexpr int kThreads = 256;

// u64 words of the counts buffer: 2G fixed-point counts, then the count box
long long acc_words(const hv_ctx *c) { return 2 * c->G + cbox_acc_words(c->world.n[0]); }

int set_err(hv_ctx *c, int code, const char *fmt, ...) {
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof c->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HV_HIP(c, expr)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_err((c), HV_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                                \
    } while (0)

#define HV_CHECK_LAUNCH(c) HV_HIP(c, hipGetLastError())

inline unsigned blocks_for(long long n, int t = kThreads) {
    return (unsigned)((n + t - 1) / t);
}

template <typename T>
Geom<T> make_geom(const hv_ctx *c) {
    Geom<T> g;
    g.nx = c->world.n[0];
    g.ny = c->world.n[1];
    g.nz = c->world.n[2];
    g.G = c->G;
    g.dmx = (T)c->world.d_max[0];
    g.dmy = (T)c->world.d_max[1];
    g.dmz = (T)c->world.d_max[2];
    g.dx = T(2.0) * g.dmx / g.nx;
    g.dy = T(2.0) * g.dmy / g.ny;
    g.dz = T(2.0) * g.dmz / g.nz;
    g.two_dx = T(2.0) * g.dx;
    g.two_dy = T(2.0) * g.dy;
    g.two_dz = T(2.0) * g.dz;
    return g;
}

Split make_split(const hv_world &w) {
    Split s;
    s.nx = w.n[0];
    s.ny = w.n[1];
    s.nz = w.n[2];
    s.nzh = (w.n[2] + 1) / 2;
    s.Gh = (long long)s.nx * s.ny * s.nzh;
    return s;
}

void natural_to_split(const Split &s, const double *nat, double *spl) {
    for (int i = 0; i < s.nx; i++)
        for (int j = 0; j < s.ny; j++)
            for (int k = 0; k < s.nz; k++) spl[split_index(s, i, j, k)] = nat[((long long)i * s.ny + j) * s.nz + k];
}

void split_to_natural(const Split &s, const double *spl, double *nat) {
    for (int i = 0; i < s.nx; i++)
        for (int j = 0; j < s.ny; j++)
            for (int k = 0; k < s.nz; k++) nat[((long long)i * s.ny + j) * s.nz + k] = spl[split_index(s, i, j, k)];
}

template <typename T>
Particles<T> particles(const hv_ctx *c) {
    Particles<T> p;
    p.x = (T *)c->part[0]; p.y = (T *)c->part[1]; p.z = (T *)c->part[2];
    p.kx = (T *)c->part[3]; p.ky = (T *)c->part[4]; p.kz = (T *)c->part[5];
    p.fx = (T *)c->part[6]; p.fy = (T *)c->part[7]; p.fz = (T *)c->part[8];
    return p;
}

size_t tsize(const hv_ctx *c) { return c->precision == HV_PREC_F64 ? sizeof(double) : sizeof(float); }

template <typename V>
int fill(hv_ctx *c, V *p, V v, long long n) {
    if (n <= 0) return HV_OK;
    unsigned nb = blocks_for(n);
    if (nb > 4096) nb = 4096;
    k_fill<V><<<nb, kThreads, 0, c->stream>>>(p, v, n);
    HV_CHECK_LAUNCH(c);
    return HV_OK;
}

int ensure_pinned(hv_ctx *c, size_t doubles) {
    if (c->pinned_cap >= doubles) return HV_OK;
    if (c->pinned) (void)hipHostFree(c->pinned);
    c->pinned = nullptr;
    c->pinned_cap = 0;
    HV_HIP(c, hipHostMalloc((void **)&c->pinned, doubles * sizeof(double), hipHostMallocDefault));
    c->pinned_cap = doubles;
    return HV_OK;
}

int ensure_buf(hv_ctx *c, void **p, size_t *cap, size_t bytes) {
    if (*cap >= bytes) return HV_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return set_err(c, HV_ERR_ALLOC, "device allocation of %zu bytes failed", bytes);
    }
    *cap = bytes;
    return HV_OK;
}

// ---- kernel-class profiling ------------------------------------------------------
hipEvent_t prof_event(hv_ctx *c) {
    hipEvent_t e = nullptr;
    if (!c->ev_free.empty()) {
        e = c->ev_free.back();
        c->ev_free.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        e = nullptr;
    }
    return e;
}

// record the start event of a bracket (nullptr when profiling is off)
hipEvent_t prof_begin(hv_ctx *c) {
    if (!c->prof) return nullptr;
    hipEvent_t a = prof_event(c);
    if (a) (void)hipEventRecord(a, c->stream);
    return a;
}

void prof_end(hv_ctx *c, int cls, hipEvent_t a, long long launches) {
    if (!c->prof || !a) return;
    hipEvent_t b = prof_event(c);
    if (!b) return;
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back({cls, a, b, launches});
}

// Per-kernel timing of the SOR sweeps: one launch in kSorSample is issued
// through hipExtLaunchKernelGGL with a start/stop event pair, which the runtime
// stamps at the dispatch packet's own begin and end (the interval a kernel
// trace reports), so launch gaps stay out of the figure.
constexpr int kSorSample = 8;

template <typename K, typename... A>
void sor_launch(hv_ctx *c, int it, int colour, int kclass, K kernel, dim3 grid, dim3 block, size_t shm,
                A... args) {
    if (c->prof) c->prof_issued[kclass]++;
    if (c->prof && ((it * 2 + colour) % kSorSample) == 3) {
        hipEvent_t a = prof_event(c), b = a ? prof_event(c) : nullptr;
        if (a && b) {
            hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t)shm, c->stream, a, b, 0u, args...);
            c->pending.push_back({kclass, a, b, 1});
            return;
        }
        if (a) c->ev_free.push_back(a);
    }
    hipLaunchKernelGGL(kernel, grid, block, shm, c->stream, args...);
}

// fold completed brackets into the counters (call at a host sync point)
void prof_flush(hv_ctx *c) {
    for (auto &p : c->pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            c->prof_ms[p.cls] += ms;
            c->prof_launches[p.cls] += p.launches;
        }
        c->ev_free.push_back(p.a);
        c->ev_free.push_back(p.b);
    }
    c->pending.clear();
}

// ---- smearing ----------------------------------------------------------------
template <typename T>
int ensure_smear(hv_ctx *c, double sigma_r, int ppn) {
    if (!(sigma_r > 0.0) || ppn <= 0) return set_err(c, HV_ERR_ARG, "smearing needs sigma_r > 0 and ppn > 0");
    if (c->smear_tab && c->tab_sigma == sigma_r && c->tab_ppn == ppn && (c->smear_tj16 || c->stl.TJ != 16))
        return HV_OK;
    hv_smear_table tb;
    if (hv_smear_table_build(&tb, c->world.n, c->world.d_max, sigma_r, ppn) != 0)
        return set_err(c, HV_ERR_ARG, "smearing radius exceeds %d nodes", HV_SMEAR_MAXR);
    SmearTab<T> h;
    memset(&h, 0, sizeof h);
    h.rx = tb.rx; h.ry = tb.ry; h.rz = tb.rz;
    for (int a = 0; a <= HV_SMEAR_MAXR; a++) {
        h.kb[a] = tb.kb[a];
        for (int b = 0; b <= HV_SMEAR_MAXR; b++) h.kc[a][b] = tb.kc[a][b];
        h.gx[a] = (T)tb.gx[a];
        h.gy[a] = (T)tb.gy[a];
        h.gz[a] = (T)tb.gz[a];
    }
    h.term = (T)tb.term;
    // tap offsets for the vector tiles (8 + 2 ry or 16 + 2 ry rows of 8 VW
    // columns)
    const int vR = 8 + 2 * tb.ry, vR16 = 16 + 2 * tb.ry, vTK = 8 * smear_vw<T>();
    for (int a = 0; a <= tb.rx; a++)
        for (int t = 0; t <= 2 * tb.kb[a]; t++) {
            const int ab = t > tb.kb[a] ? t - tb.kb[a] : tb.kb[a] - t;
            h.toff[a][t] = (tb.kc[a][ab] * vR + t) * vTK;
            h.toff16[a][t] = (tb.kc[a][ab] * vR16 + t) * vTK;
            h.tg[a][t] = (T)tb.gy[ab];
        }
    // channels to the 4 waves of a smearing block: longest first (2 kb + 1
    // taps and a store), each to the least loaded wave, ties to the lower
    // channel and wave
    {
        static_assert(kSmearThreads / kWave == 4, "wave table sized for 4 waves");
        int order[HV_SMEAR_MAXR + 1], nch[4] = {0, 0, 0, 0}, load[4] = {0, 0, 0, 0};
        for (int a = 0; a <= tb.rx; a++) order[a] = a;
        std::stable_sort(order, order + tb.rx + 1, [&](int x, int y) { return tb.kb[x] > tb.kb[y]; });
        for (int w = 0; w < 4; w++)
            for (int n = 0; n < HV_SMEAR_MAXR + 2; n++) h.wch[w][n] = -1;
        for (int m = 0; m <= tb.rx; m++) {
            const int a = order[m];
            int w = 0;
            for (int v = 1; v < 4; v++)
                if (load[v] < load[w]) w = v;
            h.wch[w][nch[w]++] = a;
            load[w] += 2 * tb.kb[a] + 3;
        }
    }
    if (!c->smear_tab) HV_HIP(c, hipMalloc(&c->smear_tab, sizeof(SmearTab<T>)));
    HV_HIP(c, hipMemcpyAsync(c->smear_tab, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
    HV_HIP(c, hipStreamSynchronize(c->stream));
    size_t n2 = (size_t)(2 * c->G);
    // two-pass form: the vector tile (16 or 8 rows x 8 VW columns) when it
    // fits kSmear2LdsVec and smear_tk_max allows it, else TJ = 8 rows and TK
    // the widest of smear_tk_max, /2, ... columns whose counts and H levels
    // fit 64 KB.  16 rows stage (16 + 2 ry) rows of counts and H levels for
    // 16 output rows where 8 rows stage 8 + 2 ry for 8.
    constexpr int kVecTK = 8 * smear_vw<T>();
    SmearTiles stl{};
    stl.TJ = 16;
    size_t lds = 0;
    stl.TK = kVecTK;
    if (c->smear_tj16 && c->smear_tk_max >= kVecTK &&
        (lds = smear2_lds_bytes(stl, tb.ry, tb.rz, sizeof(T))) <= kSmear2LdsVec &&
        (size_t)(stl.TJ + 2 * tb.ry) * (stl.TK + 2 * tb.rz) <= (size_t)kZyStaged * kSmearThreads) {
    } else if ((stl.TJ = 8, c->smear_tk_max >= kVecTK) &&
               (lds = smear2_lds_bytes(stl, tb.ry, tb.rz, sizeof(T))) <= kSmear2LdsVec) {
    } else {
        for (int tk = c->smear_tk_max; tk >= 8; tk /= 2) {
            stl.TK = tk;
            lds = smear2_lds_bytes(stl, tb.ry, tb.rz, sizeof(T));
            if (lds <= 65536) break;
        }
        if (lds > 65536) return set_err(c, HV_ERR_ARG, "smearing radius too large for the LDS tiles");
    }
    if (hipFuncSetAttribute((const void *)k_smear_zy<T, unsigned long long>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSmear2LdsVec) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_smear_zy<T, T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kSmear2LdsVec) != hipSuccess)
        return set_err(c, HV_ERR_HIP, "smearing: LDS size attribute refused");
    stl.njt = (c->world.n[1] + stl.TJ - 1) / stl.TJ;
    stl.nkt = (c->world.n[2] + stl.TK - 1) / stl.TK;
    stl.IC = 1;
    stl.nic = c->world.n[0];
    // fused form: the same with the x-sum ring (vector tile within kSmearLdsVec,
    // else the widest that fits kSmearLds)
    SmearTiles stf{};
    stf.TJ = 8;
    size_t ldf = 0;
    stf.TK = kVecTK;
    if (c->smear_tk_max >= kVecTK &&
        (ldf = smear_lds_bytes(stf, tb.rx, tb.ry, tb.rz, sizeof(T))) <= kSmearLdsVec) {
    } else {
        for (int tk = c->smear_tk_max; tk >= 8; tk /= 2) {
            stf.TK = tk;
            ldf = smear_lds_bytes(stf, tb.rx, tb.ry, tb.rz, sizeof(T));
            if (ldf <= kSmearLds) break;
        }
        if (ldf > kSmearLds) return set_err(c, HV_ERR_ARG, "smearing radius too large for the LDS tiles");
    }
    stf.njt = (c->world.n[1] + stf.TJ - 1) / stf.TJ;
    stf.nkt = (c->world.n[2] + stf.TK - 1) / stf.TK;
    // chunks along i: enough blocks for two per CU, each chunk at least 2 rx
    // planes long (the planes a chunk smears beyond its own cost at most as
    // much again)
    {
        int dev = 0, ncu = 256;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            ncu = 256;
        const long long per_chunk = 2LL * stf.njt * stf.nkt;
        const int nx = c->world.n[0];
        int nic = (int)((2LL * ncu + per_chunk - 1) / per_chunk);
        const int nic_max = nx / (2 * tb.rx > 1 ? 2 * tb.rx : 1);
        if (nic > nic_max) nic = nic_max;
        if (nic < 1) nic = 1;
        stf.IC = (nx + nic - 1) / nic;
        stf.nic = (nx + stf.IC - 1) / stf.IC;
    }
    if (hipFuncSetAttribute((const void *)k_smear_fused<T, unsigned long long>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSmearLdsVec) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_smear_fused<T, T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kSmearLdsVec) != hipSuccess)
        return set_err(c, HV_ERR_HIP, "smearing: LDS size attribute refused");
    c->stl = stl;
    c->stl_lds = lds;
    c->zy_staged = (stl.TJ + 2 * tb.ry) * (stl.TK + 2 * tb.rz);
    {
        int dev = 0, ncu = 256;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            ncu = 256;
        const long long per_cu = std::max<long long>(1, std::min<long long>(8, (long long)(160 * 1024) / (long long)std::max<size_t>(lds, 1)));
        c->zy_blocks = per_cu * ncu;
    }
    c->stf = stf;
    c->stf_lds = ldf;
    // the vector x pass: radius bucket RXM >= rx
    c->xs_rxm = tb.rx <= 4 ? 4 : tb.rx <= 9 ? 9 : tb.rx <= 14 ? 14 : 0;
    const size_t nmask = (size_t)2 * stl.njt * stl.nkt * smear_mask_words(c->world.n[0]);
    int rc = ensure_buf(c, (void **)&c->sflags, &c->sflags_cap, nmask * sizeof(unsigned));
    if (rc) return rc;
    HV_HIP(c, hipMemsetAsync(c->sflags, 0, nmask * sizeof(unsigned), c->stream));
    rc = ensure_buf(c, &c->D, &c->D_cap, (size_t)(tb.rx + 1) * n2 * sizeof(T));
    if (rc) return rc;
    c->tab_sigma = sigma_r;
    c->tab_ppn = ppn;
    c->tab_warn = (tb.min_margin < 1e-9) ? HV_WARN_SMEAR_MARGIN : 0;
    return HV_OK;
}

// counts (fixed point in c->acc when from_acc, else T counts in c->density) ->
// smeared density in c->density
template <typename T>
int smear(hv_ctx *c, bool from_acc, double sigma_r, int ppn, bool *cleared = nullptr) {
    int rc = ensure_smear<T>(c, sigma_r, ppn);
    if (rc) return rc;
    const long long n2 = 2 * c->G;
    const SmearTab<T> *tb = (const SmearTab<T> *)c->smear_tab;
    T *D = (T *)c->D, *rho = (T *)c->density;
    hipEvent_t pa = prof_begin(c);
    const int nx = c->world.n[0], ny = c->world.n[1], nz = c->world.n[2];
    if (c->smear_fused) {
        const unsigned nblk = (unsigned)smear_blocks(c->stf);
        if (from_acc) {
            k_smear_fused<T, unsigned long long><<<nblk, kSmearThreads, c->stf_lds, c->stream>>>(c->acc, rho, tb,
                                                                                               nx, ny, nz, c->stf);
        } else {  // not in place: the counts go to D first
            HV_HIP(c, hipMemcpyAsync(D, rho, sizeof(T) * (size_t)n2, hipMemcpyDeviceToDevice, c->stream));
            k_smear_fused<T, T><<<nblk, kSmearThreads, c->stf_lds, c->stream>>>(D, rho, tb, nx, ny, nz, c->stf);
        }
        HV_CHECK_LAUNCH(c);
    } else {
        // a few blocks per CU stride over the zy items
        const unsigned ntiles = (unsigned)std::min<long long>(2LL * nx * c->stl.njt * c->stl.nkt, c->zy_blocks);
        // the count box of the deposit (none for counts uploaded as a field)
        const unsigned *cbox = from_acc ? (const unsigned *)(c->acc + 2 * c->G) : nullptr;
        const int cw = cbox_species_words(nx);
        if (from_acc)
            k_smear_zy<T, unsigned long long><<<ntiles, kSmearThreads, c->stl_lds, c->stream>>>(
                c->acc, D, c->sflags, cbox, cw, tb, nx, ny, nz, c->stl);
        else
            k_smear_zy<T, T><<<ntiles, kSmearThreads, c->stl_lds, c->stream>>>(rho, D, c->sflags, nullptr, 0, tb, nx, ny,
                                                                               nz, c->stl);
        HV_CHECK_LAUNCH(c);
        constexpr int VW = smear_vw<T>();
        const long long nq = n2 / VW;
        if (c->smear_xs_vec && nz % VW == 0 && c->xs_rxm > 0) {
            const unsigned nb = (unsigned)((nq + kThreads - 1) / kThreads);
            // a loop step's x pass also writes the SOR's colour-split source
            // (the solve that follows then skips k_rho_split)
            T *rs = from_acc ? (T *)c->rho_s : nullptr;
            if (rs) c->rho_s_ready = true;
            // asked to (cleared given): the x pass also clears the counts of
            // the box and keeps the boxes in fbox, in place of k_clear_counts
            const bool clr = from_acc && cleared;
            unsigned long long *ac = clr ? c->acc : nullptr;
            unsigned *fb = clr ? c->fbox : nullptr;
            if (clr) *cleared = true;
            switch (c->xs_rxm) {
            case 4: k_smear_xs_vec<T, 4><<<nb, kThreads, 0, c->stream>>>(D, rho, c->sflags, cbox, cw, tb, nx, ny, nz, c->stl, rs, c->sp, ac, fb); break;
            case 9: k_smear_xs_vec<T, 9><<<nb, kThreads, 0, c->stream>>>(D, rho, c->sflags, cbox, cw, tb, nx, ny, nz, c->stl, rs, c->sp, ac, fb); break;
            default: k_smear_xs_vec<T, 14><<<nb, kThreads, 0, c->stream>>>(D, rho, c->sflags, cbox, cw, tb, nx, ny, nz, c->stl, rs, c->sp, ac, fb); break;
            }
        } else {
            k_smear_xs<T><<<blocks_for(n2), kThreads, 0, c->stream>>>(D, rho, c->sflags, cbox, cw, tb, nx, ny, nz, c->stl);
        }
        HV_CHECK_LAUNCH(c);
    }
    prof_end(c, HV_KCLASS_SMEAR, pa, 1);
    return c->tab_warn;
}

// ---- SOR -------------------------------------------------------------------
inline size_t maxd_words(int iters) { return (size_t)iters * kMaxSlots * kSlotStride; }

int ensure_maxd(hv_ctx *c, int max_it) {
    if (c->maxd_cap >= max_it) return HV_OK;
    if (c->maxd) (void)hipFree(c->maxd);
    if (c->maxc) (void)hipFree(c->maxc);
    c->maxd = nullptr;
    c->maxc = nullptr;
    c->maxd_cap = 0;
    HV_HIP(c, hipMalloc((void **)&c->maxd, sizeof(unsigned long long) * maxd_words(max_it)));
    HV_HIP(c, hipMalloc((void **)&c->maxc, sizeof(double) * ((size_t)max_it + 4)));  // + the resident solve's result
    c->maxd_cap = max_it;
    return HV_OK;
}

// ---- LDS-resident SOR --------------------------------------------------------
template <typename RT>
size_t res_lds_bytes(const ResTiles &t) {
    const size_t tsz = (size_t)t.TI * t.TJ * t.nzh;
    return sizeof(double) * (2 * tsz + 2 * (size_t)t.side_i + 2 * (size_t)t.side_j + 32) + sizeof(RT) * 2 * tsz;
}

size_t res32_lds_bytes(const ResTiles &t) {
    const size_t tsz = (size_t)t.TI * t.TJ * t.nzh;
    return sizeof(float) * (4 * tsz + (size_t)t.xstride + 32);
}

// Picks the tiling (one block per tile, at most one tile per CU and 256
// tiles) with the least exchange per node that fits the LDS, and checks that
// every tile block can be resident at once.
template <typename LdsFn>
int pick_res_tiles(hv_ctx *c, const void *kernel, int threads, LdsFn lds_bytes, ResTiles *out, size_t *lds_out,
                   int max_xstride = 1 << 30, int per_cu = 1, int max_tiles = kSlotPoll * kWave) {
    int dev = 0, ncu = 0, maxlds = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 1;
    int coop = 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop) return 1;
    const int nx = c->world.n[0], ny = c->world.n[1];
    const long long rows = (long long)nx * ny;
    long long target = rows / (32 / per_cu);
    if (target < 1) target = 1;
    if (target > (long long)per_cu * ncu) target = (long long)per_cu * ncu;
    if (target > max_tiles) target = max_tiles;
    ResTiles best{};
    long long best_cost = -1;
    for (int PI = 1; PI <= nx && PI <= target; PI++) {
        int PJ = (int)(target / PI);
        if (PJ < 1) break;
        if (PJ > ny) PJ = ny;
        ResTiles t;
        t.TI = (nx + PI - 1) / PI;
        t.TJ = (ny + PJ - 1) / PJ;
        t.PI = (nx + t.TI - 1) / t.TI;
        t.PJ = (ny + t.TJ - 1) / t.TJ;
        t.nzh = c->sp.nzh;
        t.side_i = t.TJ * t.nzh;
        t.side_j = t.TI * t.nzh;
        t.xstride = 2 * t.side_i + 2 * t.side_j;
        if (lds_bytes(t) > (size_t)maxlds || t.xstride > max_xstride) continue;
        const long long cost = (long long)t.TI * t.TJ * 4 + t.TI + t.TJ;
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = t;
        }
    }
    if (best_cost < 0) return 1;
    const size_t lds = lds_bytes(best);
    if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) return 1;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, lds) != hipSuccess || occ < 1 ||
        (long long)best.PI * best.PJ > (long long)occ * ncu)
        return 1;
    *out = best;
    *lds_out = lds;
    return HV_OK;
}

template <typename RT>
int res_setup(hv_ctx *c) {
    if (c->res_state) return c->res_state > 0 ? HV_OK : 1;
    c->res_state = -1;
    ResTiles best{};
    size_t lds = 0;
    if (pick_res_tiles(c, (const void *)k_sor_resident<RT>, kResThreads, res_lds_bytes<RT>, &best, &lds) != HV_OK)
        return 1;
    const size_t nT = (size_t)best.PI * best.PJ;
    if (hipMalloc((void **)&c->res_x, sizeof(double) * 2 * nT * best.xstride) != hipSuccess) return 1;
    if (hipMalloc((void **)&c->res_flag, sizeof(int) * nT) != hipSuccess) return 1;
    if (hipMalloc((void **)&c->res_err, sizeof(int) * 4) != hipSuccess) return 1;
    if (hipMalloc((void **)&c->res_result, sizeof(double) * 4) != hipSuccess) return 1;
    c->rtl = best;
    c->res_lds = lds;
    c->res_state = 1;
    return HV_OK;
}

int res32_setup(hv_ctx *c) {
    if (c->res32_state) return c->res32_state > 0 ? HV_OK : 1;
    c->res32_state = -1;
    ResTiles best{};
    size_t lds = 0;
    // one tile per CU (1024-thread blocks) or two (512-thread blocks)
    const int per = c->res32_per_cu;
    const int nt = per == 2 ? 512 : kRes32Threads;
    const void *generic = per == 2 ? (const void *)k_sor_resident_f32<0, 0, 0, 512>
                                   : (const void *)k_sor_resident_f32<0, 0, 0>;
    if (pick_res_tiles(c, generic, nt, res32_lds_bytes, &best, &lds, kFetchMax * nt, per,
                       (per == 2 ? slot_poll32<512>() : slot_poll32<kRes32Threads>()) * kWave) !=
        HV_OK)
        return 1;
    // a fixed-shape instance for the common tilings (same LDS size and block)
    c->res32_kernel = generic;
    c->res32_kernel_fold = nullptr;
    if (per == 2) {
        if (best.TI == 8 && best.TJ == 4 && best.nzh == 64)
            c->res32_kernel = (const void *)k_sor_resident_f32<8, 4, 64, 512>;
    } else {
        if (best.TI == 8 && best.TJ == 8 && best.nzh == 64)
            c->res32_kernel = (const void *)k_sor_resident_f32<8, 8, 64>;
        if (best.TI == 8 && best.TJ == 4 && best.nzh == 32)
            c->res32_kernel = (const void *)k_sor_resident_f32<8, 4, 32>;
        if (best.TI == 4 && best.TJ == 8 && best.nzh == 32)
            c->res32_kernel = (const void *)k_sor_resident_f32<4, 8, 32>;
    }
    if (c->res32_kernel != generic) {
        int occ = 0, dev = 0, ncu = 0;
        const bool ok = hipFuncSetAttribute(c->res32_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) ==
                            hipSuccess &&
                        hipGetDevice(&dev) == hipSuccess &&
                        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
                        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, c->res32_kernel, nt, lds) == hipSuccess &&
                        (long long)best.PI * best.PJ <= (long long)occ * ncu;
        if (!ok) c->res32_kernel = generic;
    }
    // the deep-halo lane-owned body, else the lane-owned body, where the
    // tiling fits it (one tile per CU)
    auto try_kernel = [&](const void *k, size_t l) {
        int occ = 0, dev = 0, ncu = 0;
        return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l) == hipSuccess &&
               hipGetDevice(&dev) == hipSuccess &&
               hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
               hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, nt, l) == hipSuccess &&
               (long long)best.PI * best.PJ <= (long long)occ * ncu;
    };
    bool deep = false;
    if (c->res32_owned >= 2) {
        ResTiles d = best;
        d.xstride = deep_xstride(d);
        const void *k = per == 2 ? (const void *)k_sor_resident_f32d<HV_DEEP_SOLO2, 512>
                                 : (const void *)k_sor_resident_f32d<HV_DEEP_SOLO>;
        const size_t l = per == 2 ? DeepLayout<kDeepSolo2P, kDeepSolo2QJ, kDeepSolo2IM>::bytes(512 / kWave)
                                  : DeepLayout<kDeepSoloP, kDeepSoloQJ, kDeepSoloIM>::bytes(kRes32Threads / kWave);
        const bool fits =
            per == 2 ? deep_fits<HV_DEEP_SOLO2, 512>(d) : deep_fits<HV_DEEP_SOLO, kRes32Threads>(d);
        if (fits && try_kernel(k, l)) {
            best = d;
            c->res32_kernel = k;
            lds = l;
            deep = true;
            const void *kf = per == 2 ? (const void *)k_sor_resident_f32d_fold<HV_DEEP_SOLO2, 512>
                                      : (const void *)k_sor_resident_f32d_fold<HV_DEEP_SOLO>;
            c->res32_kernel_fold = try_kernel(kf, l) ? kf : nullptr;
        }
    }
    if (!deep && per == 1 && c->res32_owned >= 1 && owned_fits<HV_OWNED_SOLO, kRes32Threads>(best)) {
        const void *k = (const void *)k_sor_resident_f32o<HV_OWNED_SOLO>;
        const size_t l = OwnedLayout<kOwnedSoloP, kOwnedSoloQJ, kOwnedSoloIM>::bytes(kRes32Threads / kWave);
        int occ = 0, dev = 0, ncu = 0;
        const bool ok = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l) == hipSuccess &&
                        hipGetDevice(&dev) == hipSuccess &&
                        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
                        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, nt, l) == hipSuccess &&
                        (long long)best.PI * best.PJ <= (long long)occ * ncu;
        if (ok) {
            c->res32_kernel = k;
            lds = l;
        }
    }
    c->res32_is_owned = deep ? 2 : (c->res32_kernel == (const void *)k_sor_resident_f32o<HV_OWNED_SOLO> ? 1 : 0);
    c->res32_threads = nt;
    const size_t nT = (size_t)best.PI * best.PJ;
    if (hipMalloc((void **)&c->res32_x, sizeof(unsigned long long) * 2 * nT * best.xstride) != hipSuccess) return 1;
    if (hipMalloc((void **)&c->res32_err, sizeof(int) * 4) != hipSuccess) return 1;
    if (hipMalloc((void **)&c->res32_result, sizeof(float) * 4) != hipSuccess) return 1;
    if (hipHostMalloc((void **)&c->res32_hres, sizeof(float) * 4, hipHostMallocDefault) != hipSuccess) return 1;
    c->rtl32 = best;
    c->res32_lds = lds;
    c->res32_state = 1;
    return HV_OK;
}

// The mixed form's FP32 correction iteration on c->sor_delta with source
// c->sor_R, in one cooperative launch, enqueued with the copy of its
// iteration count and error word to pinned memory; sor_resident_f32_collect
// reads them after a wait on the stream.
// Slot rows for solves of up to max_it iterations, and the words the solve
// expects cleared (the defect launch before it clears them).
int sor_resident_f32_clears(hv_ctx *c, int max_it, ResClears *cl) {
    const int nT = c->rtl32.PI * c->rtl32.PJ;
    if (c->res32_slots_cap < max_it) {
        if (c->res32_slots) (void)hipFree(c->res32_slots);
        c->res32_slots = nullptr;
        c->res32_slots_cap = 0;
        HV_HIP(c, hipMalloc((void **)&c->res32_slots, sizeof(float) * (size_t)max_it * nT));
        c->res32_slots_cap = max_it;
    }
    cl->slots = c->res32_slots;
    cl->nslots = (long long)max_it * nT;
    cl->x = c->res32_x;
    cl->nx = 2LL * nT * c->rtl32.xstride;
    cl->err = c->res32_err;
    return HV_OK;
}

// (after the clears of sor_resident_f32_clears; its iteration count, tested
// max and error word are read back with the FP64 sweeps' first chunk)
int sor_resident_f32_enqueue(hv_ctx *c, const SorCoef &cf, int max_it, const float *rho_fold = nullptr) {
    const int nT = c->rtl32.PI * c->rtl32.PJ;
    ResSync32 sy;
#ifdef HV_RES_TRACE
    static unsigned long long *trace32 = nullptr;
    if (!trace32) (void)hipMalloc((void **)&trace32, 16 * 16 * sizeof(unsigned long long));
    sy.trace = trace32;
    c->res_trace = trace32;
#endif
    sy.xr = c->res32_x;
    sy.xb = c->res32_x + (size_t)nT * c->rtl32.xstride;
    sy.slots = c->res32_slots;
    sy.err = c->res32_err;
    sy.result = c->res32_result;
    float *delta = c->sor_delta;
    const float *R = c->sor_R;
    Split sp = c->sp;
    ResTiles tl = c->rtl32;
    SorCoef cfv = cf;
    int mi = max_it, lag = c->res32_lag;
    double *phi = c->phi;
    const float *rho = rho_fold;
    void *args[] = {&delta, (void *)&R, &sp, &tl, &cfv, &mi, &lag, &sy};
    // (rho_fold: the variant that computes its residual from phi and adds its
    // correction to phi, k_sor_resident_f32d_fold)
    void *args_fold[] = {&phi, (void *)&rho, &sp, &tl, &cfv, &mi, &lag, &sy};
    hipEvent_t pa = prof_begin(c);
    HV_HIP(c, hipLaunchCooperativeKernel(rho_fold ? c->res32_kernel_fold : c->res32_kernel, dim3((unsigned)nT),
                                         dim3((unsigned)c->res32_threads), rho_fold ? args_fold : args,
                                         (unsigned)c->res32_lds, c->stream));
    hipEvent_t pb = nullptr;
    if (pa) {
        pb = prof_event(c);
        if (pb) (void)hipEventRecord(pb, c->stream);
    }
    if (pa && pb) c->pending.push_back({HV_KCLASS_SOR_RES_F32, pa, pb, 1});
    return HV_OK;
}

// the resident solve's result, as read back by the FP64 sweeps' first chunk
int sor_resident_f32_collect(hv_ctx *c, int *iters_out) {
    if (c->xread[2] != 0.0) return set_err(c, HV_ERR_HIP, "LDS-resident FP32 SOR: a tile wait exceeded its spin limit");
    *iters_out = (int)c->xread[0];
    prof_flush(c);
    return HV_OK;
}

template <typename RT>
int sor_resident(hv_ctx *c, const RT *rho, const SorCoef &cf, int max_it, int *iters_out, double *md_out) {
    const int nT = c->rtl.PI * c->rtl.PJ;
    if (c->res_slots_cap < max_it) {
        if (c->res_slots) (void)hipFree(c->res_slots);
        c->res_slots = nullptr;
        c->res_slots_cap = 0;
        HV_HIP(c, hipMalloc((void **)&c->res_slots, sizeof(double) * (size_t)max_it * nT));
        c->res_slots_cap = max_it;
    }
    int rc = fill<double>(c, c->res_slots, -1.0, (long long)max_it * nT);
    if (rc) return rc;
    rc = fill<int>(c, c->res_flag, 0, nT);
    if (rc) return rc;
    rc = fill<int>(c, c->res_err, 0, 4);
    if (rc) return rc;
    ResSync sy;
#ifdef HV_RES_TRACE
    static unsigned long long *trace = nullptr;
    if (!trace) (void)hipMalloc((void **)&trace, 16 * 16 * sizeof(unsigned long long));
    sy.trace = trace;
    c->res_trace = trace;
#endif
    sy.xr = c->res_x;
    sy.xb = c->res_x + (size_t)nT * c->rtl.xstride;
    sy.flag_red = c->res_flag;
    sy.slots = c->res_slots;
    sy.err = c->res_err;
    sy.result = c->res_result;
    double *phi = c->phi;
    Split sp = c->sp;
    ResTiles tl = c->rtl;
    SorCoef cfv = cf;
    int mi = max_it;
    void *args[] = {&phi, (void *)&rho, &sp, &tl, &cfv, &mi, &sy};
    hipEvent_t pa = prof_begin(c);
    HV_HIP(c, hipLaunchCooperativeKernel((const void *)k_sor_resident<RT>, dim3((unsigned)nT), dim3(kResThreads), args,
                                         (unsigned)c->res_lds, c->stream));
    hipEvent_t pb = nullptr;
    if (pa) {
        pb = prof_event(c);
        if (pb) (void)hipEventRecord(pb, c->stream);
    }
    double res[2];
    int err = 0;
    HV_HIP(c, hipMemcpyAsync(res, c->res_result, sizeof res, hipMemcpyDeviceToHost, c->stream));
    HV_HIP(c, hipMemcpyAsync(&err, c->res_err, sizeof err, hipMemcpyDeviceToHost, c->stream));
    HV_HIP(c, hipStreamSynchronize(c->stream));
    if (err) return set_err(c, HV_ERR_HIP, "LDS-resident SOR: a tile wait exceeded its spin limit");
    const int iters = (int)res[0];
    if (pa && pb) c->pending.push_back({HV_KCLASS_SOR, pa, pb, 2LL * iters});
    prof_flush(c);
    *iters_out = iters;
    *md_out = res[1];
    return HV_OK;
}

// Runs red/black iterations it = 0, 1, ... until the first iteration whose
// max |update| is below tol (or max_it), launching them in chunks and reading
// back each chunk's per-iteration maxima.  launch(it) enqueues iteration it;
// launches past the converged iteration return at once.
// after_first_wait(max_it) runs once, after the first chunk's read-back, and
// may lower max_it (the budget of the mixed form's FP64 sweeps, known only
// when the FP32 stage's count is read back in the same wait); the first chunk
// must then not exceed any budget it can set.
struct NoHook {
    int operator()(int &) const { return HV_OK; }
};

// xres / xerr (optional): the resident FP32 solve's result and error word,
// read back with the first chunk into c->xread before after_first_wait.
template <typename Launch, typename Hook = NoHook>
int sor_iterate(hv_ctx *c, const SorCoef &cf, int max_it, int *iters_out, double *md_out, Launch launch,
                Hook after_first_wait = Hook(), const float *xres = nullptr, const int *xerr = nullptr,
                bool final_solve = false) {
    const double tol = cf.tol;
    int rc = HV_OK;
    int it = 0, iters = -1, start = 0;
    int chunk = c->sor_pred > 2 ? c->sor_pred - 1 : 1;
    bool first = true;
    // rows a launch earlier in the queue already cleared (with SorCtl)
    const int pre_cleared = c->sor_rows_cleared;
    c->sor_rows_cleared = 0;
    while (it < max_it) {
        start = it;
        int end = it + chunk < max_it ? it + chunk : max_it;
        // this chunk's rows of the slot maxima (and SorCtl at the start)
        if (!(first && end <= pre_cleared))
            k_sor_chunk_reset<<<(unsigned)(((end - start) * kMaxSlots + 255) / 256), 256, 0, c->stream>>>(
                c->maxd, start, end, c->ctl, first ? 1 : 0);
        for (; it < end; it++) launch(it);
        HV_CHECK_LAUNCH(c);
        const float *xr = first ? xres : nullptr;
        k_sor_collect_x<<<(unsigned)(end - start), 64, 0, c->stream>>>(c->maxd, c->maxc, start, end, xr, xerr);
        HV_CHECK_LAUNCH(c);
        HV_HIP(c, hipMemcpyAsync(c->pinned, c->maxc, sizeof(double) * (size_t)(end - start + (xr ? 3 : 0)),
                                 hipMemcpyDeviceToHost, c->stream));
        // the solve's final chunk of a loop step: the step's later work goes in
        // the queue first, and only the read-back is waited for
        const bool spec_now = first && final_solve && c->spec;
        if (spec_now) {
            HV_HIP(c, hipEventRecord(c->ev_copy, c->stream));
            rc = c->spec();
            if (rc) return rc;
            HV_HIP(c, hipEventSynchronize(c->ev_copy));
        } else {
            HV_HIP(c, hipStreamSynchronize(c->stream));
        }
        if (first) {
            first = false;
            if (xr)
                for (int q = 0; q < 3; q++) c->xread[q] = c->pinned[end - start + q];
            rc = after_first_wait(max_it);
            if (rc) return rc;
        }
        for (int m = 0; m < end - start; m++) {
            if (c->pinned[m] < tol) {
                iters = start + m + 1;
                break;
            }
        }
        if (spec_now) c->spec_valid = iters > 0;
        prof_flush(c);
        if (iters > 0) break;
        chunk = 4;
    }
    if (iters < 0) iters = max_it;
    *iters_out = iters;
    *md_out = c->pinned[iters - 1 - start];
    return HV_OK;
}


// The reference's SOR on c->phi (FP64), one launch per iteration (fused) or
// per sweep.  The fused form alternates phi and phi_alt; the result is
// left in c->phi.
template <typename RT, typename Hook = NoHook>
int sor_exact(hv_ctx *c, const RT *rho, const SorCoef &cf, int max_it, bool fused, dim3 sgrd, dim3 sblk,
              const SorTiles &tl, dim3 fgrd, dim3 fblk, const FusedTiles &ft, size_t rows_lds, int *iters,
              double *md, Hook after_first_wait = Hook(), const float *xres = nullptr, const int *xerr = nullptr) {
    if (fused) {
        double *pbuf[2] = {c->phi, c->phi_alt};
        int rc = sor_iterate(
            c, cf, max_it, iters, md,
            [&](int it) {
                sor_launch(c, it, 1, HV_KCLASS_SOR, k_sor_fused<true, double, RT>, fgrd, fblk,
                           rows_lds * sizeof(double), pbuf[it & 1], pbuf[(it + 1) & 1], (const RT *)rho, c->sp, ft, cf,
                           it, c->maxd, c->ctl);
            },
            after_first_wait, xres, xerr);
        if (rc) return rc;
        if (*iters & 1) std::swap(c->phi, c->phi_alt);
        return HV_OK;
    }
    // (in place on c->phi: a loop step's later work may be enqueued ahead of
    // the first read-back)
    return sor_iterate(
        c, cf, max_it, iters, md,
        [&](int it) {
            for (int colour = 0; colour < 2; colour++)
                sor_launch(c, it, colour, HV_KCLASS_SOR, k_sor_colour<RT>, sgrd, sblk, 0, c->phi, (const RT *)rho,
                           c->sp, tl, cf, colour, it, c->maxd, c->ctl);
        },
        after_first_wait, xres, xerr, true);
}

// SOR coefficients of the context's grid (include/physics.hpp:101-103)
SorCoef make_sor_coef(const hv_ctx *c, double omega, double tol) {
    const int nx = c->world.n[0], ny = c->world.n[1], nz = c->world.n[2];
    SorCoef cf;
    double dx = 2.0 * c->world.d_max[0] / nx, dy = 2.0 * c->world.d_max[1] / ny, dz = 2.0 * c->world.d_max[2] / nz;
    cf.dx2 = dx * dx;
    cf.dy2 = dy * dy;
    cf.dz2 = dz * dz;
    cf.ydx2 = 1.0 / cf.dx2;
    cf.ydy2 = 1.0 / cf.dy2;
    cf.ydz2 = 1.0 / cf.dz2;
    cf.inv_dx2 = 1.0 / (2.0 / (dx * dx) + 2.0 / (dy * dy) + 2.0 / (dz * dz));
    cf.src = 4.0 * HV_PI * HV_E2;
    cf.omega = omega;
    cf.one_minus_omega = 1.0 - omega;
    cf.tol = tol;
    return cf;
}

// The relaxation factor of the mixed form's FP32 correction (option
// HV_OPT_SOR_F32_OMEGA).  The grid rule: the Jacobi iteration of the 7-point
// Laplacian with Dirichlet faces has spectral radius mu = sum_d cos(pi /
// (n_d - 1)) / h_d^2 over sum_d 1 / h_d^2, and SOR's optimal factor is
// 2 / (1 + s), s = sqrt(1 - mu^2); the correction's error after an
// extrapolated start is not the slowest mode alone, and 2 / (1 + 2.2 s) took
// the fewest iterations on 64^3..128^3 grids.  Never below the run's own
// omega: at a smaller factor the slowest mode decays more slowly, so the
// correction stops (at the same tolerance on its updates) further from the
// fixed point and the FP64 sweeps take more iterations.
double mixed_omega32(const hv_ctx *c, double omega) {
    if (c->f32_omega == 1) return omega;
    if (c->f32_omega > 1) return c->f32_omega / 1000.0;
    double num = 0.0, den = 0.0;
    for (int d = 0; d < 3; d++) {
        const double h = 2.0 * c->world.d_max[d] / c->world.n[d];
        num += cos(HV_PI / (c->world.n[d] - 1)) / (h * h);
        den += 1.0 / (h * h);
    }
    const double mu = num / den;
    const double w = 2.0 / (1.0 + 2.2 * sqrt(1.0 - mu * mu));
    return w > omega ? w : omega;
}

// tiles of the per-sweep SOR kernels (`planes` planes x kSorRows rows x 64
// k-slots per block, XCD slabs)
SorTiles make_sor_tiles(const hv_ctx *c, int planes = kSorPlanes) {
    SorTiles tl;
    tl.n_ic = (c->world.n[0] - 2 + planes - 1) / planes;
    tl.n_jb = (c->world.n[1] - 2 + kSorRows - 1) / kSorRows;
    tl.n_hb = (c->sp.nzh + 63) / 64;
    const int ntiles = tl.n_ic * tl.n_jb * tl.n_hb;
    tl.per_xcd = (ntiles + kXcds - 1) / kXcds;
    return tl;
}

// Red-black SOR on c->phi with the proton density (first G values of `rho`).
// The host launches iterations in chunks and reads back each chunk's
// per-iteration max |dphi|; kernels launched past the converged iteration
// return immediately (SorCtl), so the result is exactly the reference's
// iterate: stop after the first iteration whose max |dphi| < tol.
//
// loop: a solve of the fused time loop.  In the mixed form its start is then
// extrapolated from the last converged loop solves (HV_OPT_SOR_EXTRAPOLATE);
// the FP32 correction and the FP64 finishing sweeps under the reference's
// stop rule run from that start.  Any other solve (the per-kernel entry
// point) starts from phi as it is and clears the history.
template <typename RT>
int sor_solve(hv_ctx *c, const RT *rho_nat, double omega, double tol, int max_it, int *iters_out,
              double *maxdiff_out, bool loop = false) {
    if (max_it <= 0) return set_err(c, HV_ERR_ARG, "sor max_it must be positive");
    c->last_extrap = 0;
    c->sor_rows_cleared = 0;
    const int nx = c->world.n[0], ny = c->world.n[1], nz = c->world.n[2];
    RT *rho = (RT *)c->rho_s;
    if (!(loop && c->rho_s_ready && rho_nat == (const RT *)c->density)) {
        k_rho_split<RT><<<blocks_for(c->G), kThreads, 0, c->stream>>>(rho_nat, rho, c->G, c->sp);
        HV_CHECK_LAUNCH(c);
    }
    c->rho_s_ready = false;
    int rc = ensure_maxd(c, max_it);
    if (rc) return rc;
    rc = ensure_pinned(c, (size_t)max_it + 8);
    if (rc) return rc;
    const SorCoef cf = make_sor_coef(c, omega, tol);
    if (c->sor_resident && res_setup<RT>(c) == HV_OK) {
        int iters = 0;
        double md = 0.0;
        rc = sor_resident<RT>(c, rho, cf, max_it, &iters, &md);
        if (rc) return rc;
        c->sor_pred = iters;
        c->cnt_it64 += iters;
        c->cnt_solves++;
        if (iters_out) *iters_out = iters;
        if (maxdiff_out) *maxdiff_out = md;
        return (md > tol) ? HV_WARN_SOR_NOT_CONVERGED : HV_OK;
    }
    const SorTiles tl = make_sor_tiles(c);
    const dim3 sblk(64, kSorRows, 1), sgrd((unsigned)(tl.per_xcd * kXcds), 1, 1);
    FusedTiles ft;
    ft.ti = (nx + kFT - 1) / kFT;
    ft.tj = (ny + kFT - 1) / kFT;
    ft.per_xcd = (ft.ti * ft.tj + kXcds - 1) / kXcds;
    const dim3 fgrd((unsigned)(ft.per_xcd * kXcds), 1, 1), fblk(kFThreads, 1, 1);
    const size_t rows_lds = (size_t)(kFT + 2) * (kFT + 2) * c->sp.nzh;
    const bool fused = c->sor_fused && rows_lds * sizeof(double) <= 65536;
    (void)nz;
    int iters = 0;
    double md = 0.0;
    if (c->precision == HV_PREC_MIXED && c->phi_warm) {
        const bool res32 = c->sor_res32 && res32_setup(c) == HV_OK;
        // the resident correction computes its own residual and folds its
        // result into phi when a launch ahead of it (the extrapolation) can
        // clear its words: no defect and no fold launch (the same bits)
        const bool infold = res32 && loop && c->sor_extrap > 0 && c->res32_kernel_fold && std::is_same<RT, float>::value;
        ResClears cl{};
        if (res32) {
            rc = sor_resident_f32_clears(c, max_it, &cl);
            if (rc) return rc;
        }
        if (loop && c->sor_extrap > 0) {
            const int id = sor_predictor_for(c->sor_extrap, c->hist_n);
            SorHist ring;
            for (int k = 0; k < 5; k++) ring.h[k] = c->phi_h[k];
            const SorHist hs = sor_hist_order(ring, c->hist_pos);
            double w[5];
            const unsigned nb = blocks_for(2 * c->sp.Gh);
            const long long nn = 2 * c->sp.Gh;
            ResClears xc = infold ? cl : ResClears{};
            // (in a fused solve the FP64 sweeps' first chunk, one iteration,
            // finds its slot row and SorCtl cleared by this launch)
            if (infold) {
                xc.maxd = c->maxd;
                xc.maxd_rows = 1;
                xc.ctl = c->ctl;
            }
            switch (sor_predictor(id, w)) {
            case 1: k_sor_extrapolate<1><<<nb, kThreads, 0, c->stream>>>(c->phi, hs, id, nn, xc); break;
            case 2: k_sor_extrapolate<2><<<nb, kThreads, 0, c->stream>>>(c->phi, hs, id, nn, xc); break;
            case 3: k_sor_extrapolate<3><<<nb, kThreads, 0, c->stream>>>(c->phi, hs, id, nn, xc); break;
            case 4: k_sor_extrapolate<4><<<nb, kThreads, 0, c->stream>>>(c->phi, hs, id, nn, xc); break;
            default: k_sor_extrapolate<5><<<nb, kThreads, 0, c->stream>>>(c->phi, hs, id, nn, xc); break;
            }
            HV_CHECK_LAUNCH(c);
            if (infold) c->sor_rows_cleared = 1;
            c->hist_pos = (c->hist_pos + 4) % 5;
            c->last_extrap = id;
        }
        // FP32 defect-correction iterations on the small correction of a warm
        // potential, then FP64 sweeps until the stop rule holds in FP64
        float *dbuf[2] = {c->sor_delta, c->sor_delta_alt};
        if (infold) {
        } else if (res32) {
            // the defect launch also clears the resident solve's words
            k_sor_defect_clear<RT><<<blocks_for(2 * c->sp.Gh), kThreads, 0, c->stream>>>(
                c->phi, rho, c->sor_R, c->sor_delta, c->sp, cf, cl);
        } else {
            k_sor_defect<RT><<<blocks_for(2 * c->sp.Gh), kThreads, 0, c->stream>>>(c->phi, rho, c->sor_R,
                                                                                  c->sor_delta, c->sp, cf);
        }
        HV_CHECK_LAUNCH(c);
        // the FP32 stage may stop below the reference tolerance (its last
        // potential then carries less of the stop rule's own error into the
        // next start); the FP64 sweeps apply the reference rule either way
        SorCoef cf32 = cf;
        cf32.tol = ldexp(cf.tol, -c->f32_tol_shift);
        cf32.omega = mixed_omega32(c, omega);
        cf32.one_minus_omega = 1.0 - cf32.omega;
        if (res32) {
            // its iteration count is read back with the FP64 sweeps' first chunk
            rc = sor_resident_f32_enqueue(c, cf32, max_it, infold ? (const float *)rho : nullptr);
        } else if (fused) {
            rc = sor_iterate(c, cf32, max_it, &iters, &md, [&](int it) {
                sor_launch(c, it, 1, HV_KCLASS_SOR_F32, k_sor_fused<false, float, float>, fgrd, fblk,
                           rows_lds * sizeof(float), dbuf[it & 1], dbuf[(it + 1) & 1], (const float *)c->sor_R, c->sp,
                           ft, cf32, it, c->maxd, c->ctl);
            });
        } else {
            rc = sor_iterate(c, cf32, max_it, &iters, &md, [&](int it) {
                for (int colour = 0; colour < 2; colour++)
                    sor_launch(c, it, colour, HV_KCLASS_SOR_F32, k_sor_colour_f32<RT>, sgrd, sblk, 0, c->sor_delta,
                               (const float *)c->sor_R, c->sp, tl, cf32, colour, it, c->maxd, c->ctl);
            });
        }
        if (rc) return rc;
        if (!infold) {
            const float *dres = (fused && !res32) ? dbuf[iters & 1] : c->sor_delta;
            k_sor_fold<<<blocks_for(2 * c->sp.Gh), kThreads, 0, c->stream>>>(c->phi, dres, 2 * c->sp.Gh);
            HV_CHECK_LAUNCH(c);
        }
        int it64 = 0;
        // FP64 sweeps: the first chunk is one iteration (prediction 2), within
        // every budget (at least 4), so it is queued before the resident
        // FP32 stage's count is known; the budget is set in the first wait
        c->sor_pred = 2;
        auto hook = [&](int &budget) -> int {
            if (!res32) return HV_OK;
            int rc2 = sor_resident_f32_collect(c, &iters);
            if (rc2) return rc2;
            budget = mixed_fp64_budget(max_it, iters);
            return HV_OK;
        };
        rc = sor_exact(c, rho, cf, res32 ? max_it : mixed_fp64_budget(max_it, iters), fused, sgrd, sblk, tl, fgrd,
                       fblk, ft, rows_lds, &it64, &md, hook, res32 ? (const float *)c->res32_result : nullptr,
                       res32 ? (const int *)c->res32_err : nullptr);
        if (rc) return rc;
        c->mixed_fp64_sweeps = it64;
        c->sor_pred = iters;
        c->cnt_it32 += iters;
        c->cnt_it64 += it64;
        iters += it64;
    } else {
        rc = sor_exact(c, rho, cf, max_it, fused, sgrd, sblk, tl, fgrd, fblk, ft, rows_lds, &iters, &md);
        if (rc) return rc;
        c->sor_pred = iters;
        c->cnt_it64 += iters;
    }
    c->cnt_solves++;
    c->phi_warm = true;
    // the history counts consecutive converged loop solves only
    c->hist_n = (loop && !(md > tol)) ? (c->hist_n < 6 ? c->hist_n + 1 : 6) : 0;
    if (iters_out) *iters_out = iters;
    if (maxdiff_out) *maxdiff_out = md;
    return (md > tol) ? HV_WARN_SOR_NOT_CONVERGED : HV_OK;
}

// ---- diagnostics -------------------------------------------------------------
int ensure_slots(hv_ctx *c, int n) {
    if (c->slot_cap >= n) return HV_OK;
    if (c->slots) (void)hipFree(c->slots);
    c->slots = nullptr;
    c->slot_cap = 0;
    HV_HIP(c, hipMalloc((void **)&c->slots, sizeof(StatSlot) * (size_t)n));
    c->slot_cap = n;
    return HV_OK;
}

double slot_mean(const StatSlot &s, int c, long long count) {
    __int128 v = (__int128)(((unsigned __int128)s.s[c][1] << 64) | (unsigned __int128)s.s[c][0]);
    return hv_stat_mean(v, count);
}

void slot_stats(const StatSlot &s, int np, int nn, double out[4]) {
    out[0] = slot_mean(s, 0, np);
    out[1] = slot_mean(s, 1, nn);
    out[2] = slot_mean(s, 2, (long long)s.inbox[0]);
    out[3] = slot_mean(s, 3, (long long)s.inbox[1]);
}

// diagnostics of the current positions into slots[slot]
template <typename T>
int stats_into(hv_ctx *c, int slot) {
    const unsigned nb = blocks_for(c->total);
    k_deposit<T><<<nb, kThreads, 0, c->stream>>>(particles<T>(c), c->np, c->total, make_geom<T>(c), nullptr,
                                                 c->partials, c->status);
    HV_CHECK_LAUNCH(c);
    k_stats_finalize<<<1, kStatThreads, 0, c->stream>>>(c->partials, (int)nb, c->slots + slot);
    HV_CHECK_LAUNCH(c);
    return HV_OK;
}

template <typename T>
int stats_now(hv_ctx *c, double out[4]) {
    int rc = ensure_slots(c, 1);
    if (rc) return rc;
    rc = stats_into<T>(c, 0);
    if (rc) return rc;
    StatSlot s;
    HV_HIP(c, hipMemcpyAsync(&s, c->slots, sizeof s, hipMemcpyDeviceToHost, c->stream));
    HV_HIP(c, hipStreamSynchronize(c->stream));
    slot_stats(s, c->np, c->nn, out);
    return HV_OK;
}

// ---- particle sort -------------------------------------------------------------
template <typename T>
Particles<T> particles_alt(const hv_ctx *c) {
    Particles<T> p;
    p.x = (T *)c->part_alt[0]; p.y = (T *)c->part_alt[1]; p.z = (T *)c->part_alt[2];
    p.kx = (T *)c->part_alt[3]; p.ky = (T *)c->part_alt[4]; p.kz = (T *)c->part_alt[5];
    p.fx = (T *)c->part_alt[6]; p.fy = (T *)c->part_alt[7]; p.fz = (T *)c->part_alt[8];
    return p;
}

// key buckets of the counting sort: 2 (bricks + kSortOutside)
long long sort_buckets(const hv_ctx *c) {
    const long long nb = (long long)((c->world.n[0] + 3) >> kBrickShift) * ((c->world.n[1] + 3) >> kBrickShift) *
                         ((c->world.n[2] + 3) >> kBrickShift);
    return 2 * (nb + kSortOutside);
}

// Reorders the particle arrays by (species, brick of the cell): the counting
// sort of hv_kernels.h (counts with ranks, exclusive scan, scatter into the
// other buffer set), then the buffer sets swap.
template <typename T>
int sort_particles(hv_ctx *c) {
    const int n = c->total;
    const long long m = sort_buckets(c);
    k_fill<unsigned><<<(unsigned)((m + kThreads - 1) / kThreads), kThreads, 0, c->stream>>>(c->keys[1], 0u, m);
    k_sort_count<T><<<blocks_for(n), kThreads, 0, c->stream>>>(particles<T>(c), c->np, n, make_geom<T>(c),
                                                               c->keys[1], c->keys[0], c->ord[0]);
    k_sort_scan<<<1, kSortScanThreads, 0, c->stream>>>(c->keys[1], (int)m);
    k_sort_scatter<T><<<blocks_for(n), kThreads, 0, c->stream>>>(particles<T>(c), particles_alt<T>(c), c->pid,
                                                                 c->pid_alt, n, c->keys[1], c->keys[0], c->ord[0]);
    HV_CHECK_LAUNCH(c);
    for (int q = 0; q < 9; q++) std::swap(c->part[q], c->part_alt[q]);
    std::swap(c->pid, c->pid_alt);
    c->permuted = true;
    c->since_sort = 0;
    return HV_OK;
}

// Puts the particle arrays back into the caller's order (pid = identity).
int restore_order(hv_ctx *c) {
    if (!c->permuted) return HV_OK;
    const int n = c->total;
    for (int q = 0; q < 9; q++) {
        if (c->precision == HV_PREC_F64)
            k_unpermute<double><<<blocks_for(n), kThreads, 0, c->stream>>>((const double *)c->part[q],
                                                                          (double *)c->part_alt[q], c->pid, n);
        else
            k_unpermute<float><<<blocks_for(n), kThreads, 0, c->stream>>>((const float *)c->part[q],
                                                                         (float *)c->part_alt[q], c->pid, n);
        HV_CHECK_LAUNCH(c);
        std::swap(c->part[q], c->part_alt[q]);
    }
    k_iota<<<blocks_for(n), kThreads, 0, c->stream>>>(c->pid, n);
    HV_CHECK_LAUNCH(c);
    c->permuted = false;
    c->since_sort = 1 << 30;
    return HV_OK;
}

template <typename T>
int maybe_sort(hv_ctx *c) {
    if (c->sort_every <= 0) return HV_OK;
    if (c->since_sort < c->sort_every) {
        c->since_sort++;
        return HV_OK;
    }
    int rc = sort_particles<T>(c);
    c->since_sort = 1;
    return rc;
}

// ---- status word -------------------------------------------------------------
int take_status(hv_ctx *c, int *bits) {
    unsigned s = 0;
    HV_HIP(c, hipMemcpyAsync(&s, c->status, sizeof s, hipMemcpyDeviceToHost, c->stream));
    HV_HIP(c, hipStreamSynchronize(c->stream));
    unsigned zero = 0;
    HV_HIP(c, hipMemcpyAsync(c->status, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
    HV_HIP(c, hipStreamSynchronize(c->stream));
    *bits = (int)s;
    return HV_OK;
}

// ---- the fused loop ------------------------------------------------------------
template <typename T>
int potential_from_density(hv_ctx *c, const hv_skyrme *s, bool with_phi) {
    hipEvent_t pa = prof_begin(c);
    k_skyrme<T><<<blocks_for(c->G), kThreads, 0, c->stream>>>((const T *)c->density, (T *)c->U,
                                                             with_phi ? c->phi : nullptr, s->A, s->B, s->C,
                                                             s->gamma, c->G, c->sp);
    HV_CHECK_LAUNCH(c);
    prof_end(c, HV_KCLASS_POTENTIAL, pa, 1);
    return HV_OK;
}

// deposit already in c->acc -> rho -> SOR -> U (src/hyper_vlasov.cpp:94-98)
template <typename T>
int field_solve(hv_ctx *c, const hv_skyrme *s, const hv_params *p, int *iters, int *warn) {
    // the count box of the deposit, from its per-block records (unless the
    // particle pass's diagnostics launch built it)
    const bool cbox_built = c->cbox_ready;
    c->cbox_ready = false;
    if (!cbox_built) {
        k_cbox_build<<<2, kCboxThreads, 0, c->stream>>>(c->brec, (int)blocks_for(c->total),
                                                         (unsigned *)(c->acc + 2 * c->G));
        HV_CHECK_LAUNCH(c);
    }
    bool cleared = false;
    int rc = smear<T>(c, true, p->sigma_r, p->part_per_nucleon, &cleared);
    if (rc < 0) return rc;
    *warn |= rc;
    // counts back to zero: the box's rows only (the global boxes kept in fbox
    // for the force grid), unless the smearing's x pass did both
    if (!cleared) {
        k_clear_counts<<<c->clear_blocks, 256, 0, c->stream>>>(c->acc, c->world.n[0], c->world.n[1], c->world.n[2],
                                                               c->fbox);
        HV_CHECK_LAUNCH(c);
    }
    double md;
    rc = sor_solve<T>(c, (const T *)c->density, p->sor_omega, p->sor_tol, p->sor_max_it, iters, &md, true);
    if (rc < 0) return rc;
    *warn |= rc;
    if (c->spec_valid) return HV_OK;  // the potential is already enqueued, from the final phi
    return potential_from_density<T>(c, s, true);
}

int check_params(hv_ctx *c, const hv_skyrme *s, const hv_params *p) {
    if (!s || !p) return set_err(c, HV_ERR_ARG, "null skyrme or params");
    if (p->steps <= 0 || !(p->t_f > 0.0)) return set_err(c, HV_ERR_ARG, "need steps > 0 and t_f > 0");
    if (p->sor_max_it <= 0 || !(p->sor_tol > 0.0)) return set_err(c, HV_ERR_ARG, "need sor_max_it > 0 and sor_tol > 0");
    if (!(p->sigma_r > 0.0) || p->part_per_nucleon <= 0) return set_err(c, HV_ERR_ARG, "need sigma_r > 0 and ppn > 0");
    return HV_OK;
}

template <typename T>
int run_init(hv_ctx *c, const hv_skyrme *s, const hv_params *p, int *sor_iters) {
    int rc = check_params(c, s, p);
    if (rc) return rc;
    // phi: zero interior, Dirichlet faces (src/hyper_vlasov.cpp:48,56)
    std::vector<double> h((size_t)c->G, 0.0), hs((size_t)(2 * c->sp.Gh), 0.0);
    hv_coulomb_faces(h.data(), c->world.n, c->world.d_max, p->z);
    natural_to_split(c->sp, h.data(), hs.data());
    HV_HIP(c, hipMemcpyAsync(c->phi, hs.data(), sizeof(double) * hs.size(), hipMemcpyHostToDevice, c->stream));
    HV_HIP(c, hipStreamSynchronize(c->stream));
    c->phi_warm = false;
    c->hist_n = 0;
    rc = fill<unsigned long long>(c, c->acc, 0ULL, acc_words(c));
    if (rc) return rc;
    if (c->sort_every > 0) {
        rc = sort_particles<T>(c);
        if (rc) return rc;
    }
    k_deposit<T><<<blocks_for(c->total), kThreads, 0, c->stream>>>(particles<T>(c), c->np, c->total, make_geom<T>(c),
                                                                  c->acc, nullptr, c->status, c->brec);
    HV_CHECK_LAUNCH(c);
    int warn = 0, it = 0;
    rc = field_solve<T>(c, s, p, &it, &warn);
    if (rc < 0) return rc;
    if (sor_iters) *sor_iters = it;
    k_gather_u<T, T><<<blocks_for(c->total), kThreads, 0, c->stream>>>(particles<T>(c), c->np, c->total,
                                                                      make_geom<T>(c), (const T *)c->U, T(0), 0);
    HV_CHECK_LAUNCH(c);
    c->excited = 0;
    int st = 0;
    rc = take_status(c, &st);
    if (rc) return rc;
    return warn | (st & HV_WARN_PARTICLE_OUTSIDE);
}

template <typename T>
int run_steps(hv_ctx *c, const hv_skyrme *s, const hv_params *p, int first, int nsteps, double *stats,
              int *sor_iters) {
    int rc = check_params(c, s, p);
    if (rc) return rc;
    if (nsteps <= 0) return HV_OK;
    rc = ensure_slots(c, nsteps + 1);
    if (rc) return rc;

    const Geom<T> g = make_geom<T>(c);
    const double dt = p->t_f / p->steps;
    const T kick_fac = (T)(dt / (2.0 * HV_HBARC));
    const T drift_fac = (T)((dt * HV_HBARC) / HV_MC2);
    double bp, bn;
    if (p->boost_mode == HV_BOOST_SYMMETRIC) {
        bp = p->eta_exc;
        bn = p->eta_exc;
    } else {
        double zz = p->z, nnuc = p->n;
        bp = p->eta_exc * nnuc / (zz + nnuc);
        bn = p->eta_exc * zz / (zz + nnuc);
    }
    // diagnostics of the positions at the start of the first step
    rc = stats_into<T>(c, 0);
    if (rc) return rc;
    const unsigned nbp = blocks_for(c->total);
    int warn = 0;
    // the force grid is valid only from a field solve of this call to the next
    // particle pass (nothing else can have moved a particle or changed U then)
    bool fgrid_ok = false;
    if (nsteps > 1) {
        rc = ensure_buf(c, &c->fgrid, &c->fgrid_cap, 4 * sizeof(T) * (size_t)(2 * c->G));
        if (rc) return rc;
    }
    for (int n = 0; n < nsteps; n++) {
        const int step = first + n;
        int boost = 0;
        if (step * dt >= p->t_exc && !c->excited) {
            c->excited = 1;
            boost = 1;
        }
        // a re-sort counts in the particle pass's time (its only purpose)
        hipEvent_t pa = prof_begin(c);
        rc = maybe_sort<T>(c);
        if (rc) return rc;
        // the force grid of the previous step's field solve when this call made
        // it, else the force from U directly
        if (fgrid_ok)
            k_particle_step<T, T, true><<<nbp, kThreads, 0, c->stream>>>(
                particles<T>(c), c->np, c->total, g, (const T *)c->U, (const T *)c->fgrid, c->acc, c->brec,
                kick_fac, drift_fac, n > 0 ? 1 : 0, boost, (T)bp, (T)bn, c->partials, c->status);
        else
            k_particle_step<T, T, false><<<nbp, kThreads, 0, c->stream>>>(
                particles<T>(c), c->np, c->total, g, (const T *)c->U, nullptr, c->acc, c->brec, kick_fac,
                drift_fac, n > 0 ? 1 : 0, boost, (T)bp, (T)bn, c->partials, c->status);
        HV_CHECK_LAUNCH(c);
        fgrid_ok = false;
        // the pass's diagnostics and the count box of its deposit in one launch
        k_stats_cbox<<<3, kCboxThreads, 0, c->stream>>>(c->partials, (int)nbp, c->slots + (n + 1), c->brec,
                                                        (unsigned *)(c->acc + 2 * c->G));
        HV_CHECK_LAUNCH(c);
        c->cbox_ready = true;
        prof_end(c, HV_KCLASS_PARTICLE, pa, 1);
        int it = 0;
        // F = -grad U at the cells the deposit above recorded: the nodes the
        // next particle pass gathers from
        auto force_grid = [&]() -> int {
            hipEvent_t pf = prof_begin(c);
            k_force_grid<T><<<c->clear_blocks, 256, 0, c->stream>>>((const T *)c->U, (T *)c->fgrid, c->fbox, g);
            HV_CHECK_LAUNCH(c);
            prof_end(c, HV_KCLASS_POTENTIAL, pf, 1);
            return HV_OK;
        };
        // the potential and the force grid go in the queue ahead of the solve's
        // last read-back (mixed form: its FP64 sweeps nearly always meet the
        // stop rule in their first chunk); when they did not, both are
        // enqueued again after the solve
        c->spec_valid = false;
        if (c->precision == HV_PREC_MIXED)
            c->spec = [&]() -> int {
                int r = potential_from_density<T>(c, s, true);
                if (r) return r;
                return n + 1 < nsteps ? force_grid() : HV_OK;
            };
        rc = field_solve<T>(c, s, p, &it, &warn);
        c->spec = nullptr;
        const bool tail_done = c->spec_valid;
        c->spec_valid = false;
        if (rc < 0) return rc;
        if (sor_iters) sor_iters[n] = it;
        if (n + 1 < nsteps) {
            if (!tail_done) {
                rc = force_grid();
                if (rc) return rc;
            }
            fgrid_ok = true;
        }
    }
    // end-of-step half kick of the last step, with the force stored (:101-103)
    k_gather_u<T, T><<<blocks_for(c->total), kThreads, 0, c->stream>>>(particles<T>(c), c->np, c->total, g,
                                                                      (const T *)c->U, kick_fac, 1);
    HV_CHECK_LAUNCH(c);
    if (stats) {
        std::vector<StatSlot> h((size_t)nsteps + 1);
        HV_HIP(c, hipMemcpyAsync(h.data(), c->slots, sizeof(StatSlot) * h.size(), hipMemcpyDeviceToHost, c->stream));
        HV_HIP(c, hipStreamSynchronize(c->stream));
        for (int n = 0; n < nsteps; n++) slot_stats(h[n], c->np, c->nn, stats + 4 * n);
    }
    int st = 0;
    rc = take_status(c, &st);
    if (rc) return rc;
    prof_flush(c);
    return warn | (st & HV_WARN_PARTICLE_OUTSIDE);
}

// ---- total energy (no counterpart in the reference) -------------------------------
constexpr int kEnergyBlocks = 1024;

template <typename T>
int total_energy(hv_ctx *c, const hv_skyrme *s, int ppn, double out[4]) {
    if (!s || !out || ppn <= 0) return set_err(c, HV_ERR_ARG, "total energy needs skyrme, out and ppn > 0");
    if (!c->en_partials) {
        HV_HIP(c, hipMalloc((void **)&c->en_partials, sizeof(long long) * 3 * kEnergyBlocks));
        HV_HIP(c, hipMalloc((void **)&c->en_out, sizeof(unsigned long long) * 6));
    }
    const Particles<T> pp = particles<T>(c);
    k_energy_kin<T><<<kEnergyBlocks, kThreads, 0, c->stream>>>(pp.kx, pp.ky, pp.kz, c->total, c->en_partials);
    HV_CHECK_LAUNCH(c);
    k_sum_partials<<<1, 64, 0, c->stream>>>(c->en_partials, kEnergyBlocks, 1, c->en_out);
    HV_CHECK_LAUNCH(c);
    const double dV = hv_spacing(c->world.d_max[0], c->world.n[0]) * hv_spacing(c->world.d_max[1], c->world.n[1]) *
                      hv_spacing(c->world.d_max[2], c->world.n[2]);
    k_energy_grid<T><<<kEnergyBlocks, kThreads, 0, c->stream>>>((const T *)c->density, c->phi, c->G, c->sp, s->A,
                                                                s->B, s->C, s->gamma, dV,
                                                                c->en_partials + kEnergyBlocks);
    HV_CHECK_LAUNCH(c);
    k_sum_partials<<<1, 64, 0, c->stream>>>(c->en_partials + kEnergyBlocks, kEnergyBlocks, 2, c->en_out + 2);
    HV_CHECK_LAUNCH(c);
    unsigned long long h[6];
    HV_HIP(c, hipMemcpyAsync(h, c->en_out, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HV_HIP(c, hipStreamSynchronize(c->stream));
    double v[3];
    for (int q = 0; q < 3; q++) {
        const __int128 S = (__int128)(((unsigned __int128)h[2 * q + 1] << 64) | (unsigned __int128)h[2 * q]);
        v[q] = (double)S * HV_STAT_INV;
    }
    out[1] = HV_KIN_FAC * v[0] / (double)ppn;
    out[2] = v[1];
    out[3] = v[2];
    out[0] = out[1] + out[2] + out[3];
    return HV_OK;
}

// dispatch on precision
#define HV_DISPATCH(c, call_f64, call_mixed) ((c)->precision == HV_PREC_F64 ? (call_f64) : (call_mixed))

int field_ptr(hv_ctx *c, int which, void **p, long long *n, bool *is_f64) {
    *is_f64 = (c->precision == HV_PREC_F64);
    switch (which) {
    case HV_FIELD_DENSITY: *p = c->density; *n = 2 * c->G; return HV_OK;
    case HV_FIELD_POTENTIALS: *p = c->U; *n = 2 * c->G; return HV_OK;
    case HV_FIELD_COULOMB: *p = c->phi; *n = c->G; *is_f64 = true; return HV_OK;
    case HV_FIELD_FORCES_X:
    case HV_FIELD_FORCES_Y:
    case HV_FIELD_FORCES_Z: {
        int a = which - HV_FIELD_FORCES_X;
        if (!c->F[a]) {
            for (int q = 0; q < 3; q++) {
                if (c->F[q]) continue;
                if (hipMalloc(&c->F[q], tsize(c) * (size_t)(2 * c->G)) != hipSuccess)
                    return set_err(c, HV_ERR_ALLOC, "force field allocation failed");
                HV_HIP(c, hipMemsetAsync(c->F[q], 0, tsize(c) * (size_t)(2 * c->G), c->stream));
            }
        }
        *p = c->F[a];
        *n = 2 * c->G;
        return HV_OK;
    }
    default: return set_err(c, HV_ERR_ARG, "unknown field %d", which);
    }
}

int copy_in(hv_ctx *c, void *dst, const double *src, long long n, bool f64) {
    if (f64) {
        HV_HIP(c, hipMemcpyAsync(dst, src, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream));
        HV_HIP(c, hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> tmp((size_t)n);
        for (long long i = 0; i < n; i++) tmp[(size_t)i] = (float)src[i];
        HV_HIP(c, hipMemcpyAsync(dst, tmp.data(), sizeof(float) * (size_t)n, hipMemcpyHostToDevice, c->stream));
        HV_HIP(c, hipStreamSynchronize(c->stream));
    }
    return HV_OK;
}

int copy_out(hv_ctx *c, double *dst, const void *src, long long n, bool f64) {
    if (f64) {
        HV_HIP(c, hipMemcpyAsync(dst, src, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
        HV_HIP(c, hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> tmp((size_t)n);
        HV_HIP(c, hipMemcpyAsync(tmp.data(), src, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
        HV_HIP(c, hipStreamSynchronize(c->stream));
        for (long long i = 0; i < n; i++) dst[i] = (double)tmp[(size_t)i];
    }
    return HV_OK;
}

}  // namespace

// A plain device copy of n 16-byte words (read n, write n): the bandwidth
// probe and the known-size kernel that scales HBM byte counters.
__global__ void __launch_bounds__(256) k_copy_calibration(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                          long long n) {
    const long long stride = (long long)gridDim.x * 256;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// =============================================================================
extern "C" {

int hv_ctx_create(const hv_world *world, int n_protons, int n_neutrons, int precision, hv_ctx **out) {
    if (!out) return HV_ERR_ARG;
    *out = nullptr;
    if (!world) return HV_ERR_ARG;
    if (precision != HV_PREC_F64 && precision != HV_PREC_MIXED) return HV_ERR_ARG;
    for (int a = 0; a < 3; a++)
        if (world->n[a] < 4 || !(world->d_max[a] > 0.0)) return HV_ERR_ARG;
    if (n_protons < 0 || n_neutrons < 0 || n_protons > HV_MAX_SPECIES_PARTICLES ||
        n_neutrons > HV_MAX_SPECIES_PARTICLES || n_protons + n_neutrons == 0)
        return HV_ERR_ARG;
    hv_ctx *c = new (std::nothrow) hv_ctx();
    if (!c) return HV_ERR_ALLOC;
    c->precision = precision;
    c->world = *world;
    c->np = n_protons;
    c->nn = n_neutrons;
    c->total = n_protons + n_neutrons;
    c->G = (long long)world->n[0] * world->n[1] * world->n[2];
    c->sp = make_split(*world);
    const size_t ts = tsize(c);
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess &&
         hipEventCreateWithFlags(&c->ev_copy, hipEventDisableTiming) == hipSuccess;
    for (int q = 0; q < 9 && ok; q++) ok = hipMalloc(&c->part[q], ts * (size_t)c->total) == hipSuccess;
    for (int q = 0; q < 9 && ok; q++) ok = hipMalloc(&c->part_alt[q], ts * (size_t)c->total) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->pid, sizeof(int) * (size_t)c->total) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->pid_alt, sizeof(int) * (size_t)c->total) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->keys[0], sizeof(unsigned) * (size_t)c->total) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->ord[0], sizeof(int) * (size_t)c->total) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->keys[1], sizeof(unsigned) * (size_t)sort_buckets(c)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->acc, sizeof(unsigned long long) * (size_t)acc_words(c)) == hipSuccess;
    ok = ok && hipMalloc(&c->density, ts * (size_t)(2 * c->G)) == hipSuccess;
    ok = ok && hipMalloc(&c->U, ts * (size_t)(2 * c->G)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->phi, sizeof(double) * (size_t)(2 * c->sp.Gh)) == hipSuccess;
    ok = ok && hipMalloc(&c->rho_s, ts * (size_t)(2 * c->sp.Gh)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->phi_alt, sizeof(double) * (size_t)(2 * c->sp.Gh)) == hipSuccess;
    if (precision == HV_PREC_MIXED) {
        ok = ok && hipMalloc((void **)&c->sor_delta_alt, sizeof(float) * (size_t)(2 * c->sp.Gh)) == hipSuccess;
        ok = ok && hipMalloc((void **)&c->sor_R, sizeof(float) * (size_t)(2 * c->sp.Gh)) == hipSuccess;
        ok = ok && hipMalloc((void **)&c->sor_delta, sizeof(float) * (size_t)(2 * c->sp.Gh)) == hipSuccess;
        for (int q = 0; q < 5; q++)
            ok = ok && hipMalloc((void **)&c->phi_h[q], sizeof(double) * (size_t)(2 * c->sp.Gh)) == hipSuccess;
    }
    if (precision != HV_PREC_MIXED) c->sor_extrap = 0;
    ok = ok && hipMalloc((void **)&c->ctl, sizeof(SorCtl)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->status, sizeof(unsigned)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->partials, sizeof(long long) * 6 * (size_t)blocks_for(c->total)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->fbox, 12 * sizeof(unsigned)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->brec, kBoxRec * sizeof(int) * (size_t)blocks_for(c->total)) == hipSuccess;
    if (!ok) {
        hv_ctx_destroy(c);
        return HV_ERR_ALLOC;
    }
    // zero state like the reference's allocators (include/init.hpp:97-99, 111, 124-128)
    for (int q = 0; q < 9 && ok; q++) ok = hipMemsetAsync(c->part[q], 0, ts * (size_t)c->total, c->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(c->acc, 0, sizeof(unsigned long long) * (size_t)acc_words(c), c->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(c->density, 0, ts * (size_t)(2 * c->G), c->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(c->U, 0, ts * (size_t)(2 * c->G), c->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(c->phi, 0, sizeof(double) * (size_t)(2 * c->sp.Gh), c->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(c->rho_s, 0, ts * (size_t)(2 * c->sp.Gh), c->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(c->ctl, 0, sizeof(SorCtl), c->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(c->status, 0, sizeof(unsigned), c->stream) == hipSuccess;
    if (ok) {
        k_iota<<<blocks_for(c->total), kThreads, 0, c->stream>>>(c->pid, c->total);
        ok = hipGetLastError() == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(c->stream) == hipSuccess;
    if (!ok) {
        hv_ctx_destroy(c);
        return HV_ERR_HIP;
    }
    *out = c;
    return HV_OK;
}

int hv_ctx_destroy(hv_ctx *c) {
    if (!c) return HV_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int q = 0; q < 9; q++) {
        if (c->part[q]) (void)hipFree(c->part[q]);
        if (c->part_alt[q]) (void)hipFree(c->part_alt[q]);
    }
    void *sbufs[] = {c->pid, c->pid_alt, c->keys[0], c->keys[1], c->ord[0],
                     c->res_x, c->res_flag, c->res_slots, c->res_err, c->res_result,
                     c->res32_x, c->res32_slots, c->res32_err, c->res32_result};
    for (void *b : sbufs)
        if (b) (void)hipFree(b);
    void *bufs[] = {c->acc, c->density, c->U, c->phi, c->phi_alt, c->rho_s, c->sor_R, c->sor_delta,
                    c->sor_delta_alt, c->phi_h[0], c->phi_h[1], c->phi_h[2], c->phi_h[3], c->phi_h[4], c->D,
                    c->F[0], c->F[1], c->F[2],
                    c->smear_tab, c->maxd, c->maxc, c->ctl, c->slots, c->status, c->partials, c->sflags,
                    c->en_partials, c->en_out, c->fbox, c->fgrid, c->brec};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->res32_hres) (void)hipHostFree(c->res32_hres);
    prof_flush(c);
    for (hipEvent_t e : c->ev_free) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev_copy) (void)hipEventDestroy(c->ev_copy);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return HV_OK;
}

const char *hv_last_error(const hv_ctx *c) { return c ? c->err : "null context"; }

int hv_sync(hv_ctx *c) {
    if (!c) return HV_ERR_ARG;
    HV_HIP(c, hipStreamSynchronize(c->stream));
    return HV_OK;
}

void *hv_stream(hv_ctx *c) { return c ? (void *)c->stream : nullptr; }

int hv_last_outside(hv_ctx *c, int *count) {
    if (!c || !count) return HV_ERR_ARG;
    *count = c->last_outside;
    return HV_OK;
}

int hv_upload_particles(hv_ctx *c, const double *x, const double *y, const double *z, const double *kx,
                        const double *ky, const double *kz, const double *fx, const double *fy,
                        const double *fz) {
    if (!c) return HV_ERR_ARG;
    int rc0 = restore_order(c);
    if (rc0) return rc0;
    c->hist_n = 0;  // a new state: the potentials of earlier solves do not extrapolate to it
    const double *src[9] = {x, y, z, kx, ky, kz, fx, fy, fz};
    for (int q = 0; q < 9; q++) {
        if (!src[q]) continue;
        int rc = copy_in(c, c->part[q], src[q], c->total, c->precision == HV_PREC_F64);
        if (rc) return rc;
    }
    return HV_OK;
}

int hv_download_particles(hv_ctx *c, double *x, double *y, double *z, double *kx, double *ky, double *kz,
                          double *fx, double *fy, double *fz) {
    if (!c) return HV_ERR_ARG;
    double *dst[9] = {x, y, z, kx, ky, kz, fx, fy, fz};
    for (int q = 0; q < 9; q++) {
        if (!dst[q]) continue;
        const void *src = c->part[q];
        if (c->permuted) {
            if (c->precision == HV_PREC_F64)
                k_unpermute<double><<<blocks_for(c->total), kThreads, 0, c->stream>>>(
                    (const double *)c->part[q], (double *)c->part_alt[q], c->pid, c->total);
            else
                k_unpermute<float><<<blocks_for(c->total), kThreads, 0, c->stream>>>(
                    (const float *)c->part[q], (float *)c->part_alt[q], c->pid, c->total);
            HV_CHECK_LAUNCH(c);
            src = c->part_alt[q];
        }
        int rc = copy_out(c, dst[q], src, c->total, c->precision == HV_PREC_F64);
        if (rc) return rc;
    }
    return HV_OK;
}

int hv_upload_field(hv_ctx *c, int which, const double *host) {
    if (!c || !host) return HV_ERR_ARG;
    if (which == HV_FIELD_COULOMB) {
        std::vector<double> hs((size_t)(2 * c->sp.Gh), 0.0);
        natural_to_split(c->sp, host, hs.data());
        HV_HIP(c, hipMemcpyAsync(c->phi, hs.data(), sizeof(double) * hs.size(), hipMemcpyHostToDevice, c->stream));
        HV_HIP(c, hipStreamSynchronize(c->stream));
        c->phi_warm = false;
        c->hist_n = 0;
        return HV_OK;
    }
    void *p;
    long long n;
    bool f64;
    int rc = field_ptr(c, which, &p, &n, &f64);
    if (rc) return rc;
    return copy_in(c, p, host, n, f64);
}

int hv_download_field(hv_ctx *c, int which, double *host) {
    if (!c || !host) return HV_ERR_ARG;
    if (which == HV_FIELD_COULOMB) {
        std::vector<double> hs((size_t)(2 * c->sp.Gh));
        HV_HIP(c, hipMemcpyAsync(hs.data(), c->phi, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, c->stream));
        HV_HIP(c, hipStreamSynchronize(c->stream));
        split_to_natural(c->sp, hs.data(), host);
        return HV_OK;
    }
    void *p;
    long long n;
    bool f64;
    int rc = field_ptr(c, which, &p, &n, &f64);
    if (rc) return rc;
    return copy_out(c, host, p, n, f64);
}

// ---- per-kernel entry points --------------------------------------------------
int hv_cic_deposit(hv_ctx *c) {
    if (!c) return HV_ERR_ARG;
    int rc = fill<unsigned long long>(c, c->acc, 0ULL, acc_words(c));
    if (rc) return rc;
    if (c->precision == HV_PREC_F64) {
        k_deposit<double><<<blocks_for(c->total), kThreads, 0, c->stream>>>(
            particles<double>(c), c->np, c->total, make_geom<double>(c), c->acc, nullptr, c->status);
        HV_CHECK_LAUNCH(c);
        k_counts_to_field<double><<<blocks_for(2 * c->G), kThreads, 0, c->stream>>>(c->acc, (double *)c->density,
                                                                                   2 * c->G);
    } else {
        k_deposit<float><<<blocks_for(c->total), kThreads, 0, c->stream>>>(
            particles<float>(c), c->np, c->total, make_geom<float>(c), c->acc, nullptr, c->status);
        HV_CHECK_LAUNCH(c);
        k_counts_to_field<float><<<blocks_for(2 * c->G), kThreads, 0, c->stream>>>(c->acc, (float *)c->density,
                                                                                  2 * c->G);
    }
    HV_CHECK_LAUNCH(c);
    int st = 0;
    rc = take_status(c, &st);
    if (rc) return rc;
    return st & HV_WARN_PARTICLE_OUTSIDE;
}

int hv_smear_density(hv_ctx *c, double sigma_r, int ppn) {
    if (!c) return HV_ERR_ARG;
    return HV_DISPATCH(c, smear<double>(c, false, sigma_r, ppn), smear<float>(c, false, sigma_r, ppn));
}

int hv_skyrme_potentials(hv_ctx *c, const hv_skyrme *s) {
    if (!c || !s) return HV_ERR_ARG;
    return HV_DISPATCH(c, potential_from_density<double>(c, s, false), potential_from_density<float>(c, s, false));
}

int hv_set_coulomb_bc(hv_ctx *c, int z) {
    if (!c) return HV_ERR_ARG;
    std::vector<double> h((size_t)c->G);
    int rc = hv_download_field(c, HV_FIELD_COULOMB, h.data());
    if (rc) return rc;
    hv_coulomb_faces(h.data(), c->world.n, c->world.d_max, z);
    return hv_upload_field(c, HV_FIELD_COULOMB, h.data());
}

int hv_coulomb_sor(hv_ctx *c, double omega, double tol, int max_it, int *iters, double *max_diff) {
    if (!c) return HV_ERR_ARG;
    // tol 0: exactly max_it iterations (no max |dphi| is below 0)
    if (!(tol >= 0.0)) return set_err(c, HV_ERR_ARG, "tol must be non-negative");
    return HV_DISPATCH(c, sor_solve<double>(c, (const double *)c->density, omega, tol, max_it, iters, max_diff),
                       sor_solve<float>(c, (const float *)c->density, omega, tol, max_it, iters, max_diff));
}

int hv_merge_potentials(hv_ctx *c) {
    if (!c) return HV_ERR_ARG;
    if (c->precision == HV_PREC_F64)
        k_merge<double><<<blocks_for(c->G), kThreads, 0, c->stream>>>((double *)c->U, c->phi, c->G, c->sp);
    else
        k_merge<float><<<blocks_for(c->G), kThreads, 0, c->stream>>>((float *)c->U, c->phi, c->G, c->sp);
    HV_CHECK_LAUNCH(c);
    return HV_OK;
}

int hv_forces_fdm(hv_ctx *c) {
    if (!c) return HV_ERR_ARG;
    void *p;
    long long n;
    bool f64;
    int rc = field_ptr(c, HV_FIELD_FORCES_X, &p, &n, &f64);
    if (rc) return rc;
    if (c->precision == HV_PREC_F64)
        k_fdm<double><<<blocks_for(2 * c->G), kThreads, 0, c->stream>>>(
            (const double *)c->U, (double *)c->F[0], (double *)c->F[1], (double *)c->F[2], make_geom<double>(c));
    else
        k_fdm<float><<<blocks_for(2 * c->G), kThreads, 0, c->stream>>>(
            (const float *)c->U, (float *)c->F[0], (float *)c->F[1], (float *)c->F[2], make_geom<float>(c));
    HV_CHECK_LAUNCH(c);
    return HV_OK;
}

int hv_write_vtk(hv_ctx *c, const char *path) {
    if (!c || !path) return HV_ERR_ARG;
    int rc = hv_forces_fdm(c);
    if (rc) return rc;
    const size_t n2 = 2 * (size_t)c->G;
    std::vector<double> d(n2), u(n2), fx(n2), fy(n2), fz(n2);
    const int which[5] = {HV_FIELD_DENSITY, HV_FIELD_POTENTIALS, HV_FIELD_FORCES_X, HV_FIELD_FORCES_Y,
                          HV_FIELD_FORCES_Z};
    double *dst[5] = {d.data(), u.data(), fx.data(), fy.data(), fz.data()};
    for (int q = 0; q < 5; q++) {
        rc = hv_download_field(c, which[q], dst[q]);
        if (rc) return rc;
    }
    rc = hv_vtk_write_fields(path, &c->world, d.data(), u.data(), fx.data(), fy.data(), fz.data(),
                             c->precision == HV_PREC_MIXED);
    if (rc) return set_err(c, rc, "cannot write the snapshot %s", path);
    return HV_OK;
}

int hv_cic_gather(hv_ctx *c) {
    if (!c) return HV_ERR_ARG;
    void *p;
    long long n;
    bool f64;
    int rc = field_ptr(c, HV_FIELD_FORCES_X, &p, &n, &f64);
    if (rc) return rc;
    if (c->precision == HV_PREC_F64)
        k_gather_f<double><<<blocks_for(c->total), kThreads, 0, c->stream>>>(
            particles<double>(c), c->np, c->total, make_geom<double>(c), (const double *)c->F[0],
            (const double *)c->F[1], (const double *)c->F[2]);
    else
        k_gather_f<float><<<blocks_for(c->total), kThreads, 0, c->stream>>>(
            particles<float>(c), c->np, c->total, make_geom<float>(c), (const float *)c->F[0],
            (const float *)c->F[1], (const float *)c->F[2]);
    HV_CHECK_LAUNCH(c);
    return HV_OK;
}

int hv_kick_half(hv_ctx *c, double dt) {
    if (!c) return HV_ERR_ARG;
    double fac = dt / (2.0 * HV_HBARC);
    if (c->precision == HV_PREC_F64)
        k_kick<double><<<blocks_for(c->total), kThreads, 0, c->stream>>>(particles<double>(c), c->total, fac);
    else
        k_kick<float><<<blocks_for(c->total), kThreads, 0, c->stream>>>(particles<float>(c), c->total, (float)fac);
    HV_CHECK_LAUNCH(c);
    return HV_OK;
}

int hv_drift(hv_ctx *c, double dt) {
    if (!c) return HV_ERR_ARG;
    double fac = (dt * HV_HBARC) / HV_MC2;
    if (c->precision == HV_PREC_F64)
        k_drift<double><<<blocks_for(c->total), kThreads, 0, c->stream>>>(particles<double>(c), c->total, fac);
    else
        k_drift<float><<<blocks_for(c->total), kThreads, 0, c->stream>>>(particles<float>(c), c->total, (float)fac);
    HV_CHECK_LAUNCH(c);
    return HV_OK;
}

int hv_excite(hv_ctx *c, double eta, int z, int n, int mode) {
    if (!c) return HV_ERR_ARG;
    double dp, dn;
    if (mode == HV_BOOST_SYMMETRIC) {
        dp = eta;
        dn = eta;
    } else if (mode == HV_BOOST_REF_WEIGHTED) {
        double zz = z, nnuc = n;
        dp = eta * nnuc / (zz + nnuc);
        dn = eta * zz / (zz + nnuc);
    } else {
        return set_err(c, HV_ERR_ARG, "unknown boost mode %d", mode);
    }
    if (c->precision == HV_PREC_F64)
        k_excite<double><<<blocks_for(c->total), kThreads, 0, c->stream>>>((double *)c->part[5], c->np, c->total, dp,
                                                                          dn);
    else
        k_excite<float><<<blocks_for(c->total), kThreads, 0, c->stream>>>((float *)c->part[5], c->np, c->total,
                                                                         (float)dp, (float)dn);
    HV_CHECK_LAUNCH(c);
    return HV_OK;
}

int hv_center_of_mass(hv_ctx *c, int type, double *out) {
    if (!c || !out || (type != HV_PROTONS && type != HV_NEUTRONS)) return HV_ERR_ARG;
    double st[4];
    int rc = HV_DISPATCH(c, stats_now<double>(c, st), stats_now<float>(c, st));
    if (rc) return rc;
    *out = st[type];
    return HV_OK;
}

int hv_mean_squared_radius(hv_ctx *c, int type, double *out) {
    if (!c || !out || (type != HV_PROTONS && type != HV_NEUTRONS)) return HV_ERR_ARG;
    double st[4];
    int rc = HV_DISPATCH(c, stats_now<double>(c, st), stats_now<float>(c, st));
    if (rc) return rc;
    *out = st[2 + type];
    return HV_OK;
}

int hv_total_energy(hv_ctx *c, const hv_skyrme *s, int part_per_nucleon, double out[4]) {
    if (!c) return HV_ERR_ARG;
    return HV_DISPATCH(c, total_energy<double>(c, s, part_per_nucleon, out),
                       total_energy<float>(c, s, part_per_nucleon, out));
}

int hv_run_init(hv_ctx *c, const hv_skyrme *s, const hv_params *p, int *sor_iters) {
    if (!c) return HV_ERR_ARG;
    return HV_DISPATCH(c, run_init<double>(c, s, p, sor_iters), run_init<float>(c, s, p, sor_iters));
}

int hv_run_steps(hv_ctx *c, const hv_skyrme *s, const hv_params *p, int first_step, int nsteps, double *stats,
                 int *sor_iters) {
    if (!c) return HV_ERR_ARG;
    if (first_step < 0 || nsteps < 0) return set_err(c, HV_ERR_ARG, "negative step range");
    return HV_DISPATCH(c, run_steps<double>(c, s, p, first_step, nsteps, stats, sor_iters),
                       run_steps<float>(c, s, p, first_step, nsteps, stats, sor_iters));
}

int hv_simulate(hv_ctx *c, const hv_skyrme *s, const hv_params *p, double *stats, int *sor_iters) {
    if (!c) return HV_ERR_ARG;
    int it0 = 0;
    int rc = hv_run_init(c, s, p, &it0);
    if (rc < 0) return rc;
    int warn = rc;
    rc = hv_run_steps(c, s, p, 0, p->steps, stats, sor_iters);
    if (rc < 0) return rc;
    return warn | rc;
}

int hv_timer_start(hv_ctx *c) {
    if (!c) return HV_ERR_ARG;
    HV_HIP(c, hipEventRecord(c->ev0, c->stream));
    return HV_OK;
}

int hv_timer_stop(hv_ctx *c, double *ms) {
    if (!c || !ms) return HV_ERR_ARG;
    HV_HIP(c, hipEventRecord(c->ev1, c->stream));
    HV_HIP(c, hipEventSynchronize(c->ev1));
    float f = 0.0f;
    HV_HIP(c, hipEventElapsedTime(&f, c->ev0, c->ev1));
    *ms = f;
    return HV_OK;
}

// drops the FP32 resident form's tiling and buffers: chosen again at the
// next solve
int res32_reset(hv_ctx *c) {
    HV_HIP(c, hipStreamSynchronize(c->stream));
    void *bufs[] = {c->res32_x, c->res32_err, c->res32_result, c->res32_slots};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->res32_hres) (void)hipHostFree(c->res32_hres);
    c->res32_x = nullptr;
    c->res32_err = nullptr;
    c->res32_result = nullptr;
    c->res32_slots = nullptr;
    c->res32_slots_cap = 0;
    c->res32_hres = nullptr;
    c->res32_state = 0;
    return HV_OK;
}

int hv_set_option(hv_ctx *c, int option, int value) {
    if (!c) return HV_ERR_ARG;
    switch (option) {
    case HV_OPT_SORT_EVERY:
        if (value < 0) return set_err(c, HV_ERR_ARG, "sort_every must be >= 0");
        c->sort_every = value;
        c->since_sort = 1 << 30;
        return HV_OK;
    case HV_OPT_SOR_RESIDENT:
        c->sor_resident = value ? 1 : 0;
        return HV_OK;
    case HV_OPT_SOR_FUSED:
        c->sor_fused = value ? 1 : 0;
        return HV_OK;
    case HV_OPT_SOR_RESIDENT_F32:
        c->sor_res32 = value ? 1 : 0;
        return HV_OK;
    case HV_OPT_SMEAR_TILE:
        if (value != 8 && value != 16 && value != 32 && value != 64)
            return set_err(c, HV_ERR_ARG, "smearing tile width must be 8, 16, 32 or 64");
        c->smear_tk_max = value;
        c->tab_sigma = -1.0;  // rebuild the tiles at the next smearing
        return HV_OK;
    case HV_OPT_SMEAR_FUSED:
        c->smear_fused = value ? 1 : 0;
        return HV_OK;
    case HV_OPT_SOR_F32_TILES_PER_CU:
        if (value != 1 && value != 2) return set_err(c, HV_ERR_ARG, "FP32 resident tiles per CU must be 1 or 2");
        if (value != c->res32_per_cu) {
            int rc = res32_reset(c);
            if (rc) return rc;
            c->res32_per_cu = value;
        }
        return HV_OK;
    case HV_OPT_SOR_F32_OWNED:
        if (value < 0 || value > 2) return set_err(c, HV_ERR_ARG, "the FP32 SOR body must be 0, 1 or 2");
        if (value != c->res32_owned) {
            int rc = res32_reset(c);
            if (rc) return rc;
            c->res32_owned = value;
        }
        return HV_OK;
    case HV_OPT_SOR_F32_LAG:
        // the test of row it - lag runs before this iteration's black sweep,
        // so row it itself is never complete then: lag >= 1
        if (value < 1 || value > 64) return set_err(c, HV_ERR_ARG, "the FP32 SOR lag must be in [1, 64]");
        c->res32_lag = value;
        return HV_OK;
    case HV_OPT_SOR_F32_TOL_SHIFT:
        if (value < 0 || value > 20) return set_err(c, HV_ERR_ARG, "the FP32 tolerance shift must be in 0..20");
        c->f32_tol_shift = value;
        return HV_OK;
    case HV_OPT_SOR_F32_OMEGA:
        if (value != 0 && value != 1 && (value < 1001 || value > 1999))
            return set_err(c, HV_ERR_ARG, "the FP32 SOR factor must be 0, 1 or 1001..1999 (thousandths)");
        c->f32_omega = value;
        return HV_OK;
    case HV_OPT_SOR_EXTRAPOLATE:
        if (value < 0 || value > 5) return set_err(c, HV_ERR_ARG, "the SOR start predictor must be in 0..5");
        if (value != 0 && c->precision != HV_PREC_MIXED)
            return set_err(c, HV_ERR_STATE, "the FP64 form starts every solve from the last potential");
        c->sor_extrap = value;
        return HV_OK;
    default: return set_err(c, HV_ERR_ARG, "unknown option %d", option);
    }
}

int hv_get_option(hv_ctx *c, int option, int *value) {
    if (!c || !value) return HV_ERR_ARG;
    switch (option) {
    case HV_OPT_SORT_EVERY: *value = c->sort_every; return HV_OK;
    case HV_OPT_SOR_RESIDENT: *value = c->sor_resident && c->res_state >= 0; return HV_OK;
    case HV_OPT_SOR_FUSED: *value = c->sor_fused; return HV_OK;
    case HV_OPT_SOR_RESIDENT_F32: *value = c->sor_res32 && c->res32_state >= 0; return HV_OK;
    case HV_OPT_SOR_F32_LAG: *value = c->res32_lag; return HV_OK;
    case HV_OPT_SOR_F32_TILES_PER_CU: *value = c->res32_per_cu; return HV_OK;
    case HV_OPT_SOR_F32_OWNED:  // once tiled: whether the lane-owned body runs
        *value = c->res32_state > 0 ? c->res32_is_owned : c->res32_owned;
        return HV_OK;
    case HV_OPT_SMEAR_TILE: *value = c->smear_tk_max; return HV_OK;
    case HV_OPT_SMEAR_FUSED: *value = c->smear_fused; return HV_OK;
    case HV_OPT_SOR_EXTRAPOLATE: *value = c->sor_extrap; return HV_OK;
    case HV_OPT_SOR_F32_TOL_SHIFT: *value = c->f32_tol_shift; return HV_OK;
    case HV_OPT_SOR_F32_OMEGA: *value = c->f32_omega; return HV_OK;
    default: return set_err(c, HV_ERR_ARG, "unknown option %d", option);
    }
}

int hv_profile_enable(hv_ctx *c, int on) {
    if (!c) return HV_ERR_ARG;
    HV_HIP(c, hipStreamSynchronize(c->stream));
    prof_flush(c);
    c->prof = on != 0;
    for (int k = 0; k < HV_KCLASS_COUNT; k++) {
        c->prof_ms[k] = 0.0;
        c->prof_launches[k] = 0;
        c->prof_issued[k] = 0;
    }
    c->cnt_solves = c->cnt_it32 = c->cnt_it64 = 0;
    return HV_OK;
}

int hv_sor_counts(hv_ctx *c, long long *solves, long long *iters_f32, long long *iters_f64) {
    if (!c) return HV_ERR_ARG;
    if (solves) *solves = c->cnt_solves;
    if (iters_f32) *iters_f32 = c->cnt_it32;
    if (iters_f64) *iters_f64 = c->cnt_it64;
    return HV_OK;
}

int hv_profile_read(hv_ctx *c, int kclass, double *ms, long long *launches) {
    if (!c || kclass < 0 || kclass >= HV_KCLASS_COUNT) return HV_ERR_ARG;
    HV_HIP(c, hipStreamSynchronize(c->stream));
    prof_flush(c);
    if (ms) *ms = c->prof_ms[kclass];
    if (launches) *launches = c->prof_launches[kclass];
    return HV_OK;
}

int hv_profile_issued(hv_ctx *c, int kclass, long long *issued) {
    if (!c || kclass < 0 || kclass >= HV_KCLASS_COUNT || !issued) return HV_ERR_ARG;
    *issued = c->prof_issued[kclass];
    return HV_OK;
}

#ifdef HV_RES_TRACE
// trace builds only: the phase timestamps of the last resident SOR solve
int hv_debug_trace(hv_ctx *c, unsigned long long *host256) {
    if (!c || !c->res_trace) return HV_ERR_STATE;
    HV_HIP(c, hipMemcpy(host256, c->res_trace, 256 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    return HV_OK;
}
#endif

int hv_copy_bandwidth(long long bytes, int reps, double *best_gbs, double *best_us) {
    if (bytes < 16 || (bytes & 15) || reps < 1) return HV_ERR_ARG;
    uint4 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = HV_OK;
    float best = 0.0f;
    const long long n = bytes / 16;
    const long long want = (n + 255) / 256;
    const unsigned grid = (unsigned)(want < 256LL * 16 ? want : 256LL * 16);
    if (hipMalloc((void **)&a, (size_t)bytes) != hipSuccess || hipMalloc((void **)&b, (size_t)bytes) != hipSuccess) {
        rc = HV_ERR_ALLOC;
        goto out;
    }
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(a, 1, (size_t)bytes, st) != hipSuccess) {
        rc = HV_ERR_HIP;
        goto out;
    }
    // the first copy is not timed (first touch of b)
    for (int r = 0; r <= reps; r++) {
        (void)hipEventRecord(e0, st);
        k_copy_calibration<<<grid, 256, 0, st>>>(a, b, n);
        (void)hipEventRecord(e1, st);
        float ms = 0.0f;
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = HV_ERR_HIP;
            goto out;
        }
        if (r > 0 && (best == 0.0f || ms < best)) best = ms;
    }
    if (best_gbs) *best_gbs = 2.0 * (double)bytes / (best * 1e-3) / 1e9;
    if (best_us) *best_us = 1e3 * best;
out:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

const char *hv_build_info(void) {
    return "libhvamd gfx950 -O3 -ffp-contract=off (HIP " HIP_VERSION_STR ")";
}

}  // extern "C"

// ---- ensembles (hv_batch_*) -------------------------------------------------------
#include "hv_batch_impl.h"

// ---- adjoint of the loop (hv_dipole_loss_grad) -------------------------------------
#include "hv_adjoint_impl.h"

// ---- diagnostics without allocation (hv_stats_noalloc) ----------------------------
#include "hv_noalloc_impl.h"

// ---- the reference's Woods-Saxon / Fermi-level initial state (hv_init_woods_saxon) ----
#include "hv_init_ws_impl.h"
