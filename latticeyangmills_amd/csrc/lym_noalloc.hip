// lym_observables_noalloc (include/lym.h): the observables of a link field
// the caller holds on the device, with no device allocation.  The scratch of
// the reduction, one compensated partial per 64-site workgroup plus the four
// totals, is the file-scope device array `scratch`; its device address is
// looked up once per device and kept, and the kernels receive that address.
#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>

#include "lym.h"
#include "lym_kernels.hpp"

namespace lym {
const char *noalloc_last_error();
}

using lym::ksum;

constexpr long kMaxBlocks = long(LYM_NOALLOC_MAX_SITES) / lym::kObsSites;
constexpr long kScratchSlots = 4 * kMaxBlocks + 4;  // partials, then the raw totals as ksum-sized slots
constexpr int kMaxDevices = 64;

// 16 MiB per device: 4 partials of kMaxBlocks workgroups, then the totals.
static __device__ __align__(256) ksum scratch[kScratchSlots];

namespace {

// Device addresses of `scratch`, one per device (each GPU holds its own copy).
std::mutex g_scratch_lock;
ksum *g_scratch_ptr[kMaxDevices] = {};

thread_local std::string t_err;

int fail(int code, const std::string &m)
{
    t_err = m;
    return code;
}

// The address of this device's scratch array; looked up on the first call on
// the device, with the device current.
hipError_t scratch_for(int dev, ksum **out)
{
    std::lock_guard<std::mutex> g(g_scratch_lock);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<ksum *>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

// raw sums -> the normalisation of lym_observables (src/Observables/superobs.cpp:125-126),
// the same operations in the same order
__global__ void k_obs_normalise(const double *__restrict__ raw, long V, double *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    const double qn = 16 * 16 * M_PI * M_PI;
    out[0] = raw[0] / 18.0 / double(V);
    out[1] = raw[1];
    out[2] = raw[2] / qn;
    out[3] = raw[3] / qn;
}

template <typename S> void enqueue(const void *links, const lym::Geom &g, ksum *sp, double *out, hipStream_t st)
{
    const unsigned nblk = unsigned((g.V + lym::kObsSites - 1) / lym::kObsSites);
    double *raw = reinterpret_cast<double *>(sp + 4 * kMaxBlocks);
    lym::k_observables<S><<<nblk, 192, 0, st>>>(static_cast<const lym::F2<S> *>(links), g, lym::Sched{nullptr, 0}, sp);
    lym::k_reduce_partials<<<1, lym::kReduceThreads, 0, st>>>(sp, int(nblk), raw);
    k_obs_normalise<<<1, 64, 0, st>>>(raw, g.V, out);
}

}  // namespace

extern "C" {

int lym_observables_noalloc(const void *links, const int dims[4], int precision, double *out, void *stream)
{
    if (!links || !dims || !out) return fail(LYM_ERR_ARG, "lym_observables_noalloc: null links, dims or out");
    if (precision != LYM_PREC_FP64 && precision != LYM_PREC_MIXED)
        return fail(LYM_ERR_ARG, "lym_observables_noalloc: unknown precision");
    lym::Geom g{};
    long V = 1;
    for (int d = 0; d < 4; d++) {
        if (dims[d] < 2) return fail(LYM_ERR_ARG, "lym_observables_noalloc: extents must be >= 2");
        g.L[d] = dims[d];
        g.stride[d] = int(V);
        V *= dims[d];
        if (V > LYM_NOALLOC_MAX_SITES) break;
    }
    if (V > LYM_NOALLOC_MAX_SITES)
        return fail(LYM_ERR_TOO_LARGE, "lym_observables_noalloc: more than " + std::to_string(LYM_NOALLOC_MAX_SITES) +
                                           " sites, the maximum its static scratch holds");
    // 48 V bytes per buffer descriptor, as lym_create (always true below the
    // static scratch's maximum; kept for the day that maximum grows)
    if (48 * V >= (1L << 31))
        return fail(LYM_ERR_ARG, "lym_observables_noalloc: more than 44.7 M sites");
    g.V = V;
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(LYM_ERR_NODEV, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(LYM_ERR_ARG, "lym_observables_noalloc: device index out of range");
    ksum *sp = nullptr;
    if ((e = scratch_for(dev, &sp)) != hipSuccess)
        return fail(LYM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if (precision == LYM_PREC_FP64)
        enqueue<double>(links, g, sp, out, st);
    else
        enqueue<float>(links, g, sp, out, st);
    if ((e = hipGetLastError()) != hipSuccess)
        return fail(LYM_ERR_HIP, std::string("lym_observables_noalloc launch: ") + hipGetErrorString(e));
    return LYM_OK;
}

}  // extern "C"

// For lym_last_error(NULL): the last failure of a context-free call on this thread.
const char *lym::noalloc_last_error() { return t_err.empty() ? nullptr : t_err.c_str(); }
