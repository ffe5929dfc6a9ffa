// Copies between device memory and pageable host memory (a caller's array)
// through two pinned bounce buffers (host side).
//
// A copy from device memory into pageable host memory runs at ~8-10 GB/s:
// the runtime stages it and the page faults of fresh memory fall on one
// thread.  Here the DMA engine writes chunk k into one pinned buffer while
// host threads copy chunk k-1 out of the other into the caller's array, so
// the page faults are taken by several threads at once; uploads run the same
// pipeline the other way.  The buffers are allocated with hipHostMalloc on
// first use, sized to the transfer (at most kBounceBytes each), and kept
// until release().  Nothing here throws: these run inside the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

namespace lym {

constexpr size_t kBounceBytes = size_t(64) << 20;  // per buffer, at most
constexpr unsigned kCopyThreads = 8;

struct Bounce {
    char *buf[2] = {nullptr, nullptr};
    size_t cap = 0;  // bytes per buffer
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};  // an upload's DMA may still read buf[k] (ev[k] marks its end)

    // Buffers of at least min(kBounceBytes, need) bytes: a small lattice pins
    // little; a larger transfer later replaces them, after their DMAs end.
    hipError_t init(size_t need)
    {
        const size_t want = std::min(kBounceBytes, std::max<size_t>(need, 1));
        for (int k = 0; k < 2; k++)
            if (!ev[k])
                if (hipError_t e = hipEventCreateWithFlags(&ev[k], hipEventDisableTiming)) return e;
        if (cap >= want) return hipSuccess;
        for (int k = 0; k < 2; k++) {
            if (pending[k]) {
                if (hipError_t e = hipEventSynchronize(ev[k])) return e;
                pending[k] = false;
            }
            if (buf[k]) (void)hipHostFree(buf[k]);
            buf[k] = nullptr;
        }
        cap = 0;
        for (int k = 0; k < 2; k++) {
            void *p = nullptr;
            if (hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault)) return e;
            buf[k] = static_cast<char *>(p);
        }
        cap = want;
        return hipSuccess;
    }
    void release()
    {
        for (int k = 0; k < 2; k++) {
            if (pending[k] && ev[k]) (void)hipEventSynchronize(ev[k]);
            if (buf[k]) (void)hipHostFree(buf[k]);
            if (ev[k]) (void)hipEventDestroy(ev[k]);
            buf[k] = nullptr;
            ev[k] = nullptr;
            pending[k] = false;
        }
        cap = 0;
    }
};

// memcpy split over up to kCopyThreads threads (one for small copies).  If a
// thread cannot be created, the calling thread copies the parts no thread took.
inline void parallel_memcpy(char *dst, const char *src, size_t n)
{
    const unsigned nt = std::min(kCopyThreads, std::max(1u, std::thread::hardware_concurrency()));
    if (nt == 1 || n < (size_t(4) << 20)) {
        std::memcpy(dst, src, n);
        return;
    }
    std::vector<std::thread> th;
    unsigned started = 0;
    try {
        th.reserve(nt);
        for (; started < nt; started++) {
            const size_t a = n * started / nt, b = n * (started + 1) / nt;
            th.emplace_back([=] { std::memcpy(dst + a, src + a, b - a); });
        }
    } catch (...) {
        const size_t a = n * started / nt;
        std::memcpy(dst + a, src + a, n - a);
    }
    for (auto &t : th) t.join();
}

// Copies n bytes from device memory src to host memory dst, ordered after the
// work already on stream s; returns when the copy is complete.  Iteration k
// queues the DMA of chunk k into buf[k % 2] (whose previous chunk, k - 2, was
// copied out in iteration k - 1) and copies chunk k - 1 out of the other.
inline hipError_t d2h_bounced(Bounce &bb, hipStream_t s, void *dst, const void *src, size_t n)
{
    if (hipError_t e = bb.init(n)) return e;
    const size_t chunk = bb.cap, nch = (n + chunk - 1) / chunk;
    for (size_t k = 0; k <= nch; k++) {
        if (k < nch) {
            const size_t off = k * chunk, len = std::min(chunk, n - off);
            if (hipError_t e = hipMemcpyAsync(bb.buf[k & 1], static_cast<const char *>(src) + off, len,
                                              hipMemcpyDeviceToHost, s))
                return e;
            if (hipError_t e = hipEventRecord(bb.ev[k & 1], s)) return e;
        }
        if (k > 0) {
            const size_t j = k - 1, off = j * chunk, len = std::min(chunk, n - off);
            if (hipError_t e = hipEventSynchronize(bb.ev[j & 1])) return e;
            bb.pending[j & 1] = false;
            parallel_memcpy(static_cast<char *>(dst) + off, bb.buf[j & 1], len);
        }
    }
    return hipSuccess;
}

// Copies n bytes from host memory src to device memory dst on stream s (work
// queued after it on s sees the data).  Chunk k is copied by host threads
// into buf[k % 2] once the DMA of the chunk before it in that buffer has
// ended, then queued for DMA; returns when every chunk is queued, leaving the
// last DMAs pending (the next use of a buffer waits for them).
inline hipError_t h2d_bounced(Bounce &bb, hipStream_t s, void *dst, const void *src, size_t n)
{
    if (hipError_t e = bb.init(n)) return e;
    const size_t chunk = bb.cap, nch = (n + chunk - 1) / chunk;
    for (size_t k = 0; k < nch; k++) {
        const int b = int(k & 1);
        const size_t off = k * chunk, len = std::min(chunk, n - off);
        if (bb.pending[b]) {
            if (hipError_t e = hipEventSynchronize(bb.ev[b])) return e;
            bb.pending[b] = false;
        }
        parallel_memcpy(bb.buf[b], static_cast<const char *>(src) + off, len);
        if (hipError_t e = hipMemcpyAsync(static_cast<char *>(dst) + off, bb.buf[b], len, hipMemcpyHostToDevice, s))
            return e;
        if (hipError_t e = hipEventRecord(bb.ev[b], s)) return e;
        bb.pending[b] = true;
    }
    return hipSuccess;
}

}  // namespace lym
