// hv_noalloc_impl.h -- hv_stats_noalloc: the diagnostics of hv_center_of_mass /
// hv_mean_squared_radius for positions already in device memory, with no
// context and no allocation.  Its scratch (the per-block partial sums and the
// 128-bit accumulator) is a file-scope device array of fixed size; each GPU
// holds its own copy for the life of the process.
// Included once, at the end of hv_abi.hip (one translation unit).
#include <mutex>

namespace {

// scratch layout: [partials: 6 x max blocks int64][StatSlot]
constexpr long long kNoallocBlocks = (HV_NOALLOC_MAX_PARTICLES + kBlock - 1) / kBlock;
constexpr size_t kPartialBytes = sizeof(long long) * 6 * (size_t)kNoallocBlocks;
constexpr size_t kSlotOffset = (kPartialBytes + 255) / 256 * 256;
constexpr size_t kScratchBytes = kSlotOffset + 256;
static_assert(sizeof(StatSlot) <= 256, "slot area");

}  // namespace

// Device scratch shared by every hv_stats_noalloc call on a GPU.  In host code
// this name is a stand-in, not a device address: kernels receive the address
// hipGetSymbolAddress returns for the current device (scratch_ptr below).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device, looked up once per
// device and kept
int scratch_ptr(void **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return HV_ERR_HIP;
    if (dev < 0 || dev >= kMaxDevices) return HV_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return HV_ERR_HIP;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return HV_OK;
}

// the exact sums of the partial rows (as k_stats_finalize), then the four
// means of hv_stat_mean: (S 2^-30) / count
__global__ void __launch_bounds__(kStatThreads)
k_stats_means(const long long *__restrict__ partials, int nblocks, StatSlot *__restrict__ slot, int np, int nn,
              double *__restrict__ out) {
    stats_finalize_block(partials, nblocks, slot);
    __syncthreads();
    if (threadIdx.x < 4) {
        const int c = threadIdx.x;
        const volatile StatSlot *s = slot;
        const __int128 S = (__int128)(((unsigned __int128)s->s[c][1] << 64) | (unsigned __int128)s->s[c][0]);
        const long long count = c == 0 ? np : c == 1 ? nn : (long long)s->inbox[c - 2];
        out[c] = ((double)S * HV_STAT_INV) / (double)count;
    }
}

template <typename T>
int stats_noalloc(const hv_world *w, const T *x, const T *y, const T *z, int np, int nn, double *out,
                  hipStream_t stream) {
    void *base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc) return rc;
    Geom<T> g;
    g.nx = w->n[0];
    g.ny = w->n[1];
    g.nz = w->n[2];
    g.G = (long long)g.nx * g.ny * g.nz;
    g.dmx = (T)w->d_max[0];
    g.dmy = (T)w->d_max[1];
    g.dmz = (T)w->d_max[2];
    g.dx = T(2.0) * g.dmx / g.nx;
    g.dy = T(2.0) * g.dmy / g.ny;
    g.dz = T(2.0) * g.dmz / g.nz;
    g.two_dx = T(2.0) * g.dx;
    g.two_dy = T(2.0) * g.dy;
    g.two_dz = T(2.0) * g.dz;
    Particles<T> p{};
    p.x = const_cast<T *>(x);
    p.y = const_cast<T *>(y);
    p.z = const_cast<T *>(z);
    const int total = np + nn;
    const int nb = (total + kBlock - 1) / kBlock;
    long long *partials = (long long *)base;
    StatSlot *slot = (StatSlot *)((unsigned char *)base + kSlotOffset);
    // every partial row the finalize reads is written by this call's pass
    k_deposit<T><<<nb, kBlock, 0, stream>>>(p, np, total, g, nullptr, partials, nullptr);
    if (hipGetLastError() != hipSuccess) return HV_ERR_HIP;
    k_stats_means<<<1, kStatThreads, 0, stream>>>(partials, nb, slot, np, nn, out);
    if (hipGetLastError() != hipSuccess) return HV_ERR_HIP;
    return HV_OK;
}

}  // namespace

extern "C" int hv_stats_noalloc(const hv_world *world, int precision, const void *x, const void *y, const void *z,
                                int n_protons, int n_neutrons, double *out, void *stream) {
    if (!world || !x || !y || !z || !out) return HV_ERR_ARG;
    if (precision != HV_PREC_F64 && precision != HV_PREC_MIXED) return HV_ERR_ARG;
    if (n_protons < 1 || n_neutrons < 1 || (long long)n_protons + n_neutrons > HV_NOALLOC_MAX_PARTICLES)
        return HV_ERR_ARG;
    for (int a = 0; a < 3; a++)
        if (world->n[a] < 2 || !(world->d_max[a] > 0.0)) return HV_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    if (precision == HV_PREC_F64)
        return stats_noalloc<double>(world, (const double *)x, (const double *)y, (const double *)z, n_protons,
                                     n_neutrons, out, st);
    return stats_noalloc<float>(world, (const float *)x, (const float *)y, (const float *)z, n_protons, n_neutrons,
                                out, st);
}
