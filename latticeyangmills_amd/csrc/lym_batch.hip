// Batched ensemble path: many independent lattices of uneven sizes flowed
// together, the whole flow loop captured in a HIP graph, the launch shape
// picked by an autotuner (declared in include/lym_batch.h; kernels in
// lym_batch_kernels.hpp).  Three launch shapes: one launch per RK stage for
// the whole batch with one workgroup per schedule slot, the same with a
// persistent grid (a multiple of 8, so a workgroup keeps its XCD) striding
// over the slots, or one persistent launch for all the steps between two
// measurements, with grid-wide barriers between RK stages (SCHED_WHOLE_FLOW).
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "lym.h"
#include "lym_batch.h"
#include "lym_launch.hpp"
#include "lym_sched.hpp"
#include "lym_xfer.hpp"

using lym::ConfDesc;
using lym::ksum;

struct lym_batch {
    int device = 0, prec = LYM_PREC_FP64, nconf = 0;
    std::vector<ConfDesc> desc;
    std::vector<std::array<int, 4>> dims;
    ConfDesc *d_desc = nullptr;
    long total_links = 0, total_sites = 0, ntiles = 0, nitems = 0;
    hipStream_t stream = nullptr;
    void *W[2] = {nullptr, nullptr};
    int cur = 0;
    bool eo = false;  // W[cur] in even-odd order (left so by lym_batch_metropolis)
    int rows = 3;     // form of the links while the flow runs (2: two-row, SCHED_TWO_ROW)
    int form = 3;     // form of W[cur] now: 3 between calls
    void *acc = nullptr;
    ksum *partials = nullptr;
    double *obs = nullptr;
    int obs_cap = 0;
    double *staging = nullptr;
    long staging_sites = 0;
    lym::Bounce bounce;  // pinned buffers of lym_batch_upload_links / lym_batch_download_links
    unsigned long long *accepted = nullptr;
    int grid = 0;  // 0: one workgroup per schedule slot, else a persistent grid
    lym::HostSchedule sched;
    int *d_order = nullptr;
    std::string tuned;
    lym::FlowSync *sync = nullptr;        // barrier state of the whole-flow launch
    unsigned long long sync_limit = 0;    // its bounded wait, in wall-clock ticks (5 s)
    hipGraphExec_t graph = nullptr;
    std::string graph_key;
    int graph_parity_flip = 0;
    std::string err;
};

namespace {

int bfail(lym_batch *b, int code, const std::string &m)
{
    if (b) b->err = m;
    return code;
}
int bhip(lym_batch *b, hipError_t e, const char *what)
{
    return bfail(b, e == hipErrorOutOfMemory ? LYM_ERR_NOMEM : LYM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define BHIP(b, call)                                    \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return bhip(b, e_, #call); \
    } while (0)
#define BLAUNCHED(b, what)                             \
    do {                                               \
        hipError_t e_ = hipGetLastError();             \
        if (e_ != hipSuccess) return bhip(b, e_, what); \
    } while (0)

unsigned nblocks(long n, int per) { return unsigned((n + per - 1) / per); }

lym::Sched dev_sched(const lym_batch *b) { return lym::Sched{b->sched.order.empty() ? nullptr : b->d_order, b->sched.per_xcd}; }

unsigned stage_grid(const lym_batch *b)
{
    const long slots = b->sched.order.empty() ? b->ntiles : long(lym::kXcds) * b->sched.per_xcd;
    return unsigned(b->grid > 0 ? std::min<long>(b->grid, slots) : slots);
}

void drop_graph(lym_batch *b)
{
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_key.clear();
}

// LDS staging mode of the batch stage kernel under schedule kind `kind`: slab
// when every lattice fits it, per lattice (4) when only some do, else whole
// x-rows or own links.
int batch_lds_mode(const lym_batch *b, int kind)
{
    if (!(kind & lym::SCHED_LDS)) return 0;
    bool xrows = true, slab = true, any_slab = false;
    for (auto &d : b->dims) {
        xrows = xrows && (lym::kSitesPerBlock % d[0] == 0);
        slab = slab && lym::slab_fits(d.data());
        any_slab = any_slab || lym::slab_fits(d.data());
    }
    return slab ? 3 : any_slab ? 4 : xrows ? 2 : 1;
}
int batch_lds_mode(const lym_batch *b) { return batch_lds_mode(b, b->sched.kind); }

// Workgroups of the whole-flow launch under schedule kind `kind`: `want`
// rounded up to a multiple of 8, or with want = 0 as many as the device holds
// at once (the kernel's workgroups per CU, from its registers and LDS, times
// the CUs) rounded down to a multiple of 8; never more than that, since every
// workgroup must be resident for the grid barriers, nor more than the
// schedule has slots (rounded up to a multiple of 8).
int whole_flow_grid(lym_batch *b, int kind, long slots, int want, int *grid)
{
    const int lds = batch_lds_mode(b, kind);
    int per_cu = 0, cus = 0;
    const hipError_t e = b->prec == LYM_PREC_FP64     ? lym::batch_flow_persistent_per_cu<double, 3>(lds, &per_cu)
                         : (kind & lym::SCHED_TWO_ROW) ? lym::batch_flow_persistent_per_cu<float, 2>(lds, &per_cu)
                                                       : lym::batch_flow_persistent_per_cu<float, 3>(lds, &per_cu);
    BHIP(b, e);
    BHIP(b, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
    const long resident = long(per_cu) * cus / lym::kXcds * lym::kXcds;
    if (resident < lym::kXcds) return bfail(b, LYM_ERR_HIP, "the whole-flow kernel does not fit the device");
    const long g = want > 0 ? (long(want) + lym::kXcds - 1) / lym::kXcds * lym::kXcds : resident;
    if (g > resident)
        return bfail(b, LYM_ERR_ARG,
                     "whole-flow grid of " + std::to_string(g) + " workgroups: the device holds " +
                         std::to_string(resident) + " at once, and every one must be resident for the grid barriers");
    *grid = int(std::min(g, (slots + lym::kXcds - 1) / lym::kXcds * lym::kXcds));
    return LYM_OK;
}

int apply_schedule(lym_batch *b, int grid, int kind, int target)
{
    std::vector<long> run0;
    for (auto &d : b->desc) run0.push_back(d.tile0);
    lym::HostSchedule h = lym::build_schedule(b->dims, run0, kind, target);
    int whole_grid = 0;
    if (kind & lym::SCHED_WHOLE_FLOW) {
        const long slots = h.order.empty() ? b->ntiles : long(lym::kXcds) * h.per_xcd;
        if (int rc = whole_flow_grid(b, kind, slots, grid, &whole_grid)) return rc;
    }
    if (b->d_order) (void)hipFree(b->d_order);
    b->d_order = nullptr;
    if (!h.order.empty()) {
        BHIP(b, hipMalloc(&b->d_order, sizeof(int) * h.order.size()));
        BHIP(b, hipMemcpy(b->d_order, h.order.data(), sizeof(int) * h.order.size(), hipMemcpyHostToDevice));
    }
    b->sched = std::move(h);
    b->grid = (kind & lym::SCHED_WHOLE_FLOW) ? whole_grid
              : grid > 0                     ? ((grid + lym::kXcds - 1) / lym::kXcds) * lym::kXcds
                                             : 0;
    b->rows = (b->sched.kind & lym::SCHED_TWO_ROW) ? 2 : 3;
    drop_graph(b);
    return LYM_OK;
}

template <typename S, int STAGE>
void launch_stage(const lym_batch *b, const lym::F2<S> *win, lym::F2<S> *wout, S *acc, double eps)
{
    if constexpr (sizeof(S) == 4)
        if (b->form == 2) {
            lym::launch_batch_stage_r2<STAGE>(batch_lds_mode(b), stage_grid(b), b->stream, win, wout, acc, b->d_desc,
                                              b->nconf, b->ntiles, dev_sched(b), eps);
            return;
        }
    lym::launch_batch_stage<S, STAGE>(batch_lds_mode(b), stage_grid(b), b->stream, win, wout, acc, b->d_desc, b->nconf,
                                      b->ntiles, dev_sched(b), eps);
}

// Every lattice of W[cur] into form `rows` (3 or 2, the mixed flow's
// SCHED_TWO_ROW) by a conversion into W[1 - cur], as the single-lattice
// to_form: a lattice's three-row field starts at link_off, its two-row field
// at link_off / 9 * 6.
int to_form(lym_batch *b, int rows)
{
    if (b->form == rows) return LYM_OK;
    auto *in = static_cast<const float2 *>(b->W[b->cur]);
    auto *out = static_cast<float2 *>(b->W[1 - b->cur]);
    for (const ConfDesc &d : b->desc) {
        const dim3 grid(nblocks(d.g.V, 256), 4);
        const long off3 = d.link_off, off2 = d.link_off / 9 * 6;
        if (rows == 2)
            lym::k_to_rows2<float><<<grid, 256, 0, b->stream>>>(in + off3, out + off2, d.g.V);
        else
            lym::k_from_rows2<float><<<grid, 256, 0, b->stream>>>(in + off2, out + off3, d.g.V);
    }
    BLAUNCHED(b, "k_to_rows2 / k_from_rows2");
    b->cur ^= 1;
    b->form = rows;
    return LYM_OK;
}

template <typename S> int enqueue_step(lym_batch *b, double eps)
{
    using F2 = lym::F2<S>;
    F2 *x = static_cast<F2 *>(b->W[b->cur]);
    F2 *y = static_cast<F2 *>(b->W[1 - b->cur]);
    S *acc = static_cast<S *>(b->acc);
    launch_stage<S, 1>(b, x, y, acc, eps);
    launch_stage<S, 2>(b, y, x, acc, eps);
    launch_stage<S, 3>(b, x, y, acc, eps);
    BLAUNCHED(b, "k_batch_flow_stage");
    b->cur = 1 - b->cur;
    return LYM_OK;
}
// nsteps steps as one whole-flow launch (SCHED_WHOLE_FLOW), its arrival
// count zeroed by a kernel just before it (graph capture records no memset).
template <typename S> int enqueue_whole(lym_batch *b, double eps, int nsteps)
{
    using F2 = lym::F2<S>;
    F2 *x = static_cast<F2 *>(b->W[b->cur]);
    F2 *y = static_cast<F2 *>(b->W[1 - b->cur]);
    S *acc = static_cast<S *>(b->acc);
    lym::k_flow_sync_reset<<<1, 256, 0, b->stream>>>(b->sync, 0);
    const int lds = batch_lds_mode(b);
    const unsigned grid = unsigned(b->grid);
    // barrier sync groups: about sqrt(grid), so that the arrivals on one
    // counter (a group's and the count of completed groups) are few
    const int ngroups = std::max(1, std::min({lym::kMaxSyncGroups, b->grid, int(std::lround(std::sqrt(b->grid)))}));
    if (sizeof(S) == 4 && b->form == 2)
        lym::launch_batch_flow_persistent<float, 2>(lds, grid, b->stream, reinterpret_cast<float2 *>(x),
                                                    reinterpret_cast<float2 *>(y), reinterpret_cast<float *>(acc),
                                                    b->d_desc, b->nconf, b->ntiles, dev_sched(b), eps, nsteps, b->sync,
                                                    ngroups, b->sync_limit);
    else
        lym::launch_batch_flow_persistent<S, 3>(lds, grid, b->stream, x, y, acc, b->d_desc, b->nconf, b->ntiles,
                                                dev_sched(b), eps, nsteps, b->sync, ngroups, b->sync_limit);
    BLAUNCHED(b, "k_batch_flow_persistent");
    if (nsteps & 1) b->cur = 1 - b->cur;
    return LYM_OK;
}
// Steps per whole-flow launch at most, so that the arrival counts (3 per step
// per workgroup at most) and the slot counters (3 per step times a queue's
// slots and workgroups, at most the XCD's) stay below 2^32.
int whole_max_steps(const lym_batch *b)
{
    const long slots = b->sched.order.empty() ? b->ntiles : long(lym::kXcds) * b->sched.per_xcd;
    const long per_xcd = (slots + lym::kXcds - 1) / lym::kXcds;
    const long per_step = 3L * std::max<long>(b->grid, per_xcd + b->grid / lym::kXcds);
    return int(std::max<long>(1, std::min<long>(1L << 20, 0xFFFFFFFFL / per_step - 1)));
}

int whole(lym_batch *b, double eps, int nsteps)
{
    for (int done = 0, rc; done < nsteps; done += whole_max_steps(b)) {
        const int n = std::min(nsteps - done, whole_max_steps(b));
        if ((rc = b->prec == LYM_PREC_FP64 ? enqueue_whole<double>(b, eps, n) : enqueue_whole<float>(b, eps, n)))
            return rc;
    }
    return LYM_OK;
}

int step(lym_batch *b, double eps)
{
    if (b->sched.kind & lym::SCHED_WHOLE_FLOW) return whole(b, eps, 1);
    return b->prec == LYM_PREC_FP64 ? enqueue_step<double>(b, eps) : enqueue_step<float>(b, eps);
}

// Whether a whole-flow launch since the flag was last cleared gave up at a
// grid barrier (its workgroups were not all resident): the links are then
// not a flow result.  Waits for the stream.
int whole_flow_aborted(lym_batch *b, bool *aborted)
{
    int flag = 0;
    BHIP(b, hipMemcpyAsync(&flag, &b->sync->abort, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    *aborted = flag != 0;
    return LYM_OK;
}

template <typename S> int enqueue_obs(lym_batch *b, int slot)
{
    const unsigned g = unsigned(std::min<long>(b->ntiles, 65535L * 4));
    if (sizeof(S) == 4 && b->form == 2)
        lym::k_batch_observables<S, 2><<<g, 192, 0, b->stream>>>(static_cast<const lym::F2<S> *>(b->W[b->cur]),
                                                                 b->d_desc, b->nconf, b->ntiles, b->partials);
    else
        lym::k_batch_observables<S><<<g, 192, 0, b->stream>>>(static_cast<const lym::F2<S> *>(b->W[b->cur]), b->d_desc,
                                                              b->nconf, b->ntiles, b->partials);
    lym::k_batch_reduce<<<b->nconf, 256, 0, b->stream>>>(b->partials, b->d_desc, b->nconf, b->ntiles,
                                                        b->obs + size_t(4) * b->nconf * slot);
    BLAUNCHED(b, "k_batch_observables");
    return LYM_OK;
}
int observables(lym_batch *b, int slot)
{
    return b->prec == LYM_PREC_FP64 ? enqueue_obs<double>(b, slot) : enqueue_obs<float>(b, slot);
}

int ensure_slots(lym_batch *b, int n)
{
    if (n <= b->obs_cap) return LYM_OK;
    BHIP(b, hipStreamSynchronize(b->stream));
    if (b->obs) (void)hipFree(b->obs);
    b->obs = nullptr;
    BHIP(b, hipMalloc(&b->obs, sizeof(double) * 4 * size_t(b->nconf) * n));
    b->obs_cap = n;
    drop_graph(b);  // the captured graph points at the old buffer
    return LYM_OK;
}

void normalise(const lym_batch *b, int c, const double *raw, double out[4])
{
    const double qn = 16 * 16 * M_PI * M_PI;
    out[0] = raw[0] / 18.0 / double(b->desc[c].g.V);
    out[1] = raw[1];
    out[2] = raw[2] / qn;
    out[3] = raw[3] / qn;
}

// enqueue the whole flow (rows-1 steps, measurements into slots after the
// steps t with t % every == 0); with SCHED_WHOLE_FLOW the steps between two
// measurements run as one launch
int enqueue_flow(lym_batch *b, double eps, int rows, int every)
{
    int slot = 0, rc;
    if ((rc = to_form(b, b->rows)) || (rc = observables(b, slot++))) return rc;
    if (b->sched.kind & lym::SCHED_WHOLE_FLOW) {
        lym::k_flow_sync_reset<<<1, 256, 0, b->stream>>>(b->sync, 1);
        for (int t = 1; t < rows;) {
            const int last = std::min((t + every - 1) / every * every, rows - 1);  // last step of the segment
            if ((rc = whole(b, eps, last - t + 1))) return rc;
            if (last % every == 0 && (rc = observables(b, slot++))) return rc;
            t = last + 1;
        }
        return to_form(b, 3);
    }
    for (int t = 1; t < rows; t++) {
        if ((rc = step(b, eps))) return rc;
        if (t % every == 0)
            if ((rc = observables(b, slot++))) return rc;
    }
    return to_form(b, 3);
}

std::string shape_key(const lym_batch *b)
{
    std::ostringstream k;
    k << "gfx950|prec" << b->prec;
    for (auto &d : b->dims) k << "|" << d[0] << "x" << d[1] << "x" << d[2] << "x" << d[3];
    return k.str();
}

}  // namespace

extern "C" {

int lym_batch_create(lym_batch **out, int nconf, const int *dims, int precision, int device)
{
    if (!out || !dims || nconf < 1 || nconf > lym::kMaxConf) return LYM_ERR_ARG;
    *out = nullptr;
    if (precision != LYM_PREC_FP64 && precision != LYM_PREC_MIXED) return LYM_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return LYM_ERR_NODEV;
    if (device < 0 || device >= ndev) return LYM_ERR_ARG;
    lym_batch *b = new lym_batch;
    b->device = device;
    b->prec = precision;
    b->nconf = nconf;
    long links = 0, tiles = 0, items = 0, maxV = 0;
    for (int c = 0; c < nconf; c++) {
        ConfDesc d{};
        long V = 1, st = 1;
        for (int k = 0; k < 4; k++) {
            const int L = dims[4 * c + k];
            if (L < 2) {
                delete b;
                return LYM_ERR_ARG;
            }
            d.g.L[k] = L;
            d.g.stride[k] = int(st);
            st *= L;
            V *= L;
        }
        if (4 * V >= (1L << 31) || 48 * V >= (1L << 31)) {  // as lym_create
            delete b;
            return LYM_ERR_ARG;
        }
        d.g.V = V;
        d.slab = lym::slab_fits(d.g.L) ? 1 : 0;
        d.link_off = links * 9;
        d.acc_off = links * 8;
        d.tile0 = tiles;
        d.item0 = items;
        links += 4 * V;
        tiles += (V + lym::kSitesPerBlock - 1) / lym::kSitesPerBlock;
        items += (V / 2 + 63) / 64 * 64;
        maxV = std::max(maxV, V);
        b->desc.push_back(d);
        b->dims.push_back({dims[4 * c], dims[4 * c + 1], dims[4 * c + 2], dims[4 * c + 3]});
    }
    b->total_links = links;
    b->total_sites = links / 4;
    b->ntiles = tiles;
    b->nitems = items;
    b->staging_sites = maxV;
    auto bail = [&](int code) {
        lym_batch_destroy(b);
        return code;
    };
    const size_t rs = precision == LYM_PREC_FP64 ? 8 : 4;
    if (hipSetDevice(device) != hipSuccess) return bail(LYM_ERR_HIP);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bail(LYM_ERR_HIP);
    for (int k = 0; k < 2; k++)
        if (hipMalloc(&b->W[k], rs * 18 * size_t(links)) != hipSuccess) return bail(LYM_ERR_NOMEM);
    if (hipMalloc(&b->acc, rs * 8 * size_t(links)) != hipSuccess) return bail(LYM_ERR_NOMEM);
    if (hipMalloc(&b->partials, sizeof(ksum) * 4 * size_t(tiles)) != hipSuccess) return bail(LYM_ERR_NOMEM);
    if (hipMalloc(&b->staging, sizeof(double) * 18 * size_t(maxV)) != hipSuccess) return bail(LYM_ERR_NOMEM);
    if (hipMalloc(&b->accepted, sizeof(unsigned long long) * nconf) != hipSuccess) return bail(LYM_ERR_NOMEM);
    if (hipMalloc(&b->d_desc, sizeof(ConfDesc) * nconf) != hipSuccess) return bail(LYM_ERR_NOMEM);
    if (hipMalloc(&b->sync, sizeof(lym::FlowSync)) != hipSuccess) return bail(LYM_ERR_NOMEM);
    if (hipMemset(b->sync, 0, sizeof(lym::FlowSync)) != hipSuccess) return bail(LYM_ERR_HIP);
    {
        int khz = 0;  // the constant wall clock of the barrier waits
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || khz <= 0)
            return bail(LYM_ERR_HIP);
        b->sync_limit = 5000ull * (unsigned long long)khz;
    }
    if (hipMemcpy(b->d_desc, b->desc.data(), sizeof(ConfDesc) * nconf, hipMemcpyHostToDevice) != hipSuccess)
        return bail(LYM_ERR_HIP);
    if (ensure_slots(b, 4) != LYM_OK) return bail(LYM_ERR_NOMEM);
    // as the single-lattice default: the mixed flow in two-row form
    if (apply_schedule(b, 0, lym::default_schedule_kind(precision == LYM_PREC_MIXED), lym::kDefaultColumnTarget) !=
        LYM_OK)
        return bail(LYM_ERR_NOMEM);
    // cold start
    for (int c = 0; c < nconf; c++) {
        const long V = b->desc[c].g.V;
        if (precision == LYM_PREC_FP64)
            lym::k_cold<double><<<nblocks(4 * V, 256), 256, 0, b->stream>>>(
                static_cast<double2 *>(b->W[0]) + b->desc[c].link_off, V);
        else
            lym::k_cold<float><<<nblocks(4 * V, 256), 256, 0, b->stream>>>(
                static_cast<float2 *>(b->W[0]) + b->desc[c].link_off, V);
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess) return bail(LYM_ERR_HIP);
    *out = b;
    return LYM_OK;
}

int lym_batch_destroy(lym_batch *b)
{
    if (!b) return LYM_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->bounce.release();
    for (void *p : {b->W[0], b->W[1], b->acc, (void *)b->partials, (void *)b->obs, (void *)b->staging,
                    (void *)b->accepted, (void *)b->d_desc, (void *)b->d_order, (void *)b->sync})
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return LYM_OK;
}

const char *lym_batch_last_error(const lym_batch *b) { return b ? b->err.c_str() : "null batch"; }

int lym_batch_info(const lym_batch *b, int *nconf, int64_t *total_sites, int *grid, int *kind, int *target)
{
    if (!b) return LYM_ERR_ARG;
    if (nconf) *nconf = b->nconf;
    if (total_sites) *total_sites = b->total_sites;
    if (grid) *grid = b->grid;
    if (kind) *kind = b->sched.kind;
    if (target) *target = b->sched.target;
    return LYM_OK;
}

// Copies W[cur] into W[1 - cur] reordered to (to_eo) or from the even-odd
// order of the Metropolis sweeps, every lattice in one launch.
static void batch_reorder_eo(lym_batch *b, bool to_eo)
{
    const dim3 grid(nblocks(b->total_sites, 256), 36);
    if (b->prec == LYM_PREC_FP64) {
        auto *in = static_cast<const double2 *>(b->W[b->cur]);
        auto *out = static_cast<double2 *>(b->W[1 - b->cur]);
        if (to_eo)
            lym::k_batch_reorder_eo<double, true><<<grid, 256, 0, b->stream>>>(in, out, b->d_desc, b->nconf, b->total_sites);
        else
            lym::k_batch_reorder_eo<double, false><<<grid, 256, 0, b->stream>>>(in, out, b->d_desc, b->nconf, b->total_sites);
    } else {
        auto *in = static_cast<const float2 *>(b->W[b->cur]);
        auto *out = static_cast<float2 *>(b->W[1 - b->cur]);
        if (to_eo)
            lym::k_batch_reorder_eo<float, true><<<grid, 256, 0, b->stream>>>(in, out, b->d_desc, b->nconf, b->total_sites);
        else
            lym::k_batch_reorder_eo<float, false><<<grid, 256, 0, b->stream>>>(in, out, b->d_desc, b->nconf, b->total_sites);
    }
}

// As the single-lattice context: the links stay in even-odd order after
// lym_batch_metropolis until an entry point reads them site by site.
static int to_order(lym_batch *b, bool eo)
{
    if (b->eo == eo) return LYM_OK;
    batch_reorder_eo(b, eo);
    BLAUNCHED(b, "k_batch_reorder_eo");
    b->cur ^= 1;
    b->eo = eo;
    return LYM_OK;
}
static int to_lex(lym_batch *b) { return to_order(b, false); }

static int check_conf(lym_batch *b, int c)
{
    if (c < 0 || c >= b->nconf) return bfail(b, LYM_ERR_ARG, "configuration index out of range");
    return LYM_OK;
}

int lym_batch_upload_links(lym_batch *b, int c, const double *const by_mu[4])
{
    if (!b || !by_mu) return LYM_ERR_ARG;
    if (check_conf(b, c)) return LYM_ERR_ARG;
    BHIP(b, hipSetDevice(b->device));
    if (int rc = to_lex(b)) return rc;  // the other lattices keep their links
    const ConfDesc &d = b->desc[c];
    const long V = d.g.V;
    for (int mu = 0; mu < 4; mu++) {
        if (!by_mu[mu]) return bfail(b, LYM_ERR_ARG, "null link array");
        BHIP(b, lym::h2d_bounced(b->bounce, b->stream, b->staging, by_mu[mu], sizeof(double) * 18 * V));
        if (b->prec == LYM_PREC_FP64)
            lym::k_aos_to_soa<double><<<nblocks(V, 256), 256, 0, b->stream>>>(
                b->staging, static_cast<double2 *>(b->W[b->cur]) + d.link_off + size_t(mu) * 9 * V, V);
        else
            lym::k_aos_to_soa<float><<<nblocks(V, 256), 256, 0, b->stream>>>(
                b->staging, static_cast<float2 *>(b->W[b->cur]) + d.link_off + size_t(mu) * 9 * V, V);
        BLAUNCHED(b, "k_aos_to_soa");
    }
    BHIP(b, hipStreamSynchronize(b->stream));
    return LYM_OK;
}

int lym_batch_download_links(lym_batch *b, int c, double *const by_mu[4])
{
    if (!b || !by_mu) return LYM_ERR_ARG;
    if (check_conf(b, c)) return LYM_ERR_ARG;
    BHIP(b, hipSetDevice(b->device));
    if (int rc = to_lex(b)) return rc;
    const ConfDesc &d = b->desc[c];
    const long V = d.g.V;
    for (int mu = 0; mu < 4; mu++) {
        if (!by_mu[mu]) return bfail(b, LYM_ERR_ARG, "null link array");
        if (b->prec == LYM_PREC_FP64)
            lym::k_soa_to_aos<double><<<nblocks(V, 256), 256, 0, b->stream>>>(
                static_cast<const double2 *>(b->W[b->cur]) + d.link_off + size_t(mu) * 9 * V, b->staging, V);
        else
            lym::k_soa_to_aos<float><<<nblocks(V, 256), 256, 0, b->stream>>>(
                static_cast<const float2 *>(b->W[b->cur]) + d.link_off + size_t(mu) * 9 * V, b->staging, V);
        BLAUNCHED(b, "k_soa_to_aos");
        BHIP(b, lym::d2h_bounced(b->bounce, b->stream, by_mu[mu], b->staging, sizeof(double) * 18 * V));
    }
    return LYM_OK;
}

int lym_batch_set_hot(lym_batch *b, const uint64_t *seeds)
{
    if (!b || !seeds) return LYM_ERR_ARG;
    BHIP(b, hipSetDevice(b->device));
    for (int c = 0; c < b->nconf; c++) {
        const long V = b->desc[c].g.V;
        if (b->prec == LYM_PREC_FP64)
            lym::k_hot<double><<<nblocks(4 * V, 256), 256, 0, b->stream>>>(
                static_cast<double2 *>(b->W[b->cur]) + b->desc[c].link_off, V, seeds[c]);
        else
            lym::k_hot<float><<<nblocks(4 * V, 256), 256, 0, b->stream>>>(
                static_cast<float2 *>(b->W[b->cur]) + b->desc[c].link_off, V, seeds[c]);
        BLAUNCHED(b, "k_hot");
    }
    b->eo = false;
    BHIP(b, hipStreamSynchronize(b->stream));
    return LYM_OK;
}

extern "C++" template <typename S>
static void launch_batch_metropolis(lym_batch *b, unsigned nb, int mu, const lym::MetroParams &p)
{
    lym::launch_batch_metropolis<S>(lym::metro_kernel_kind(p.candidate == LYM_CAND_EXPH, p.eps), (nb + 7) / 8, b->stream,
                                    static_cast<lym::F2<S> *>(b->W[b->cur]), b->d_desc, b->nconf, b->nitems, mu, p,
                                    b->accepted);
}

int lym_batch_metropolis(lym_batch *b, double beta, double eps, int nhit, int nsweeps, int first_sweep, int candidate,
                         const uint64_t *seeds, int64_t *accepted)
{
    if (!b || !seeds) return LYM_ERR_ARG;
    if (nhit < 1 || nsweeps < 0 || first_sweep < 0 || (candidate != LYM_CAND_SU2 && candidate != LYM_CAND_EXPH) ||
        !(eps > 0 && eps < 1))
        return bfail(b, LYM_ERR_ARG, "bad Metropolis arguments");
    for (auto &d : b->dims)
        for (int k = 0; k < 4; k++)
            if (d[k] % 2) return bfail(b, LYM_ERR_ARG, "checkerboard needs even extents");
    BHIP(b, hipSetDevice(b->device));
    for (int c = 0; c < b->nconf; c++) b->desc[c].seed = seeds[c];
    BHIP(b, hipMemcpyAsync(b->d_desc, b->desc.data(), sizeof(ConfDesc) * b->nconf, hipMemcpyHostToDevice, b->stream));
    lym::k_zero_u64<<<nblocks(b->nconf, 64), 64, 0, b->stream>>>(b->accepted, b->nconf);
    const int nu = candidate == LYM_CAND_SU2 ? 13 : 9;
    // the sweeps run on the links in even-odd order, which they are left in
    if (nsweeps > 0)
        if (int rc = to_order(b, true)) return rc;
    const unsigned nb = nblocks(b->nitems, 256);
    for (int sw = 0; sw < nsweeps; sw++) {
        lym::MetroParams p{beta, eps, nhit, candidate, nu, uint32_t(first_sweep + sw), 0, 0, nullptr};
        for (int mu = 0; mu < 4; mu++)
            for (int par = 0; par < 2; par++) {
                p.parity = par;
                if (b->prec == LYM_PREC_FP64)
                    launch_batch_metropolis<double>(b, nb, mu, p);
                else
                    launch_batch_metropolis<float>(b, nb, mu, p);
                BLAUNCHED(b, "k_batch_metropolis");
            }
    }
    std::vector<unsigned long long> acc(b->nconf);
    BHIP(b, hipMemcpyAsync(acc.data(), b->accepted, sizeof(unsigned long long) * b->nconf, hipMemcpyDeviceToHost,
                           b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    if (accepted)
        for (int c = 0; c < b->nconf; c++) accepted[c] = int64_t(acc[c]);
    return LYM_OK;
}

int lym_batch_observables(lym_batch *b, double *out)
{
    if (!b || !out) return LYM_ERR_ARG;
    BHIP(b, hipSetDevice(b->device));
    int rc = to_lex(b);
    if (!rc) rc = observables(b, 0);
    if (rc) return rc;
    std::vector<double> raw(4 * size_t(b->nconf));
    BHIP(b, hipMemcpyAsync(raw.data(), b->obs, sizeof(double) * raw.size(), hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    for (int c = 0; c < b->nconf; c++) normalise(b, c, &raw[4 * size_t(c)], out + 4 * size_t(c));
    return LYM_OK;
}

int lym_batch_set_schedule(lym_batch *b, int grid, int kind, int target)
{
    const int order = kind & 3;
    if (!b || grid < 0 || (kind & ~(lym::SCHED_LDS | lym::SCHED_TWO_ROW | lym::SCHED_WHOLE_FLOW | 3)) ||
        order > lym::SCHED_COLUMN ||
        (order == lym::SCHED_COLUMN && target < 1))
        return LYM_ERR_ARG;
    if ((kind & lym::SCHED_TWO_ROW) && b->prec != LYM_PREC_MIXED)
        return bfail(b, LYM_ERR_ARG, "the two-row form is for the mixed path (FP64 multiplies out all rows)");
    BHIP(b, hipSetDevice(b->device));
    BHIP(b, hipStreamSynchronize(b->stream));
    return apply_schedule(b, grid, kind, target);
}

int lym_batch_set_barrier_timeout(lym_batch *b, double seconds)
{
    if (!b || !(seconds >= 0) || seconds > 3600) return b ? bfail(b, LYM_ERR_ARG, "timeout outside [0, 3600] s") : LYM_ERR_ARG;
    int khz = 0;
    BHIP(b, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device));
    b->sync_limit = (unsigned long long)(seconds * 1000.0 * khz);
    drop_graph(b);  // the graph holds the old bound
    return LYM_OK;
}

int lym_batch_flow(lym_batch *b, double eps, double tau, int measure_every, int use_graph, double *series,
                   int max_rows, int *rows_out)
{
    if (!b || !series || measure_every < 1 || !(eps > 0) || !(tau > 0))
        return b ? bfail(b, LYM_ERR_ARG, "bad flow arguments") : LYM_ERR_ARG;
    const int rows = int(tau / eps);  // src/Apps/wilsonflow.cpp:117
    if (rows_out) *rows_out = rows;
    if (rows < 1 || rows > max_rows) return bfail(b, LYM_ERR_ARG, "series too small for int(tau/eps) rows");
    BHIP(b, hipSetDevice(b->device));
    const int nmeas = 1 + (rows - 1) / measure_every;
    int rc = to_lex(b);
    if (rc || (rc = ensure_slots(b, nmeas))) return rc;
    const int nsteps = rows - 1;
    if (use_graph) {
        // the graph is recorded for fixed buffers: W[cur] -> ... ; the final
        // field lands in W[cur ^ (nsteps & 1)], so one graph per start parity
        // key: every parameter baked into the launches, eps exactly (hex float)
        char key[160];
        snprintf(key, sizeof key, "%a|%d|%d|%d|%d|%d|%d", eps, rows, measure_every, b->cur, b->grid, b->sched.kind,
                 b->sched.target);
        if (!b->graph || b->graph_key != key) {
            if (b->graph) (void)hipGraphExecDestroy(b->graph);
            b->graph = nullptr;
            const int cur0 = b->cur;
            BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_flow(b, eps, rows, measure_every);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            b->cur = cur0;
            b->form = 3;
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return bhip(b, e, "hipStreamEndCapture");
            e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return bhip(b, e, "hipGraphInstantiate");
            b->graph_key = key;
        }
        BHIP(b, hipGraphLaunch(b->graph, b->stream));
        // nsteps steps, plus the two form conversions of the two-row flow
        if ((nsteps + (b->rows == 2 ? 2 : 0)) & 1) b->cur = 1 - b->cur;
    } else {
        if ((rc = enqueue_flow(b, eps, rows, measure_every))) return rc;
    }
    std::vector<double> raw(size_t(4) * b->nconf * nmeas);
    BHIP(b, hipMemcpyAsync(raw.data(), b->obs, sizeof(double) * raw.size(), hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    if (b->sched.kind & lym::SCHED_WHOLE_FLOW) {
        bool aborted = false;
        if ((rc = whole_flow_aborted(b, &aborted))) return rc;
        if (aborted)
            return bfail(b, LYM_ERR_HIP,
                         "whole-flow launch: a workgroup waited past the limit at a grid barrier (the grid was not "
                         "resident at once); the links are not a flow result");
    }
    for (int c = 0; c < b->nconf; c++) {
        double tcur = 0.0;
        int slot = 0;
        for (int t = 0; t < rows; t++) {
            if (t > 0) tcur += eps;
            double *row = series + (size_t(c) * max_rows + t) * 4;
            row[0] = tcur;
            if (t % measure_every == 0) {
                double o[4];
                normalise(b, c, &raw[(size_t(slot) * b->nconf + c) * 4], o);
                row[1] = o[0];
                row[2] = o[2];
                row[3] = o[1];
                slot++;
            } else {
                row[1] = row[2] = row[3] = NAN;
            }
        }
    }
    return LYM_OK;
}

// Times candidate launch shapes of the RK stage kernel over flow steps of the
// batch itself (its fields saved first and put back afterwards) and keeps the
// fastest (lym::tune_rounds).  Candidates:
// every schedule of lym::schedule_candidates() with one workgroup per slot,
// then the best of those with persistent grids of {2, 4, 8} x 256 workgroups.
// Results are cached in `cache_path` (one line per batch shape) when it is
// non-null.
int lym_batch_autotune(lym_batch *b, const char *cache_path, char *summary, int summary_len)
{
    if (!b) return LYM_ERR_ARG;
    BHIP(b, hipSetDevice(b->device));
    if (int rc = to_lex(b)) return rc;
    const std::string key = shape_key(b);
    std::map<std::string, std::string> cache;
    if (cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            auto p = line.find('\t');
            if (p != std::string::npos) cache[line.substr(0, p)] = line.substr(p + 1);
        }
    }
    int best_grid = 0, best_kind = lym::SCHED_COLUMN, best_target = 4096;
    double best_ms = -1;
    auto it = cache.find(key);
    if (it != cache.end() &&
        sscanf(it->second.c_str(), "%d %d %d %lf", &best_grid, &best_kind, &best_target, &best_ms) == 4) {
        b->tuned = "cached";
    } else {
        struct Events {
            hipEvent_t e[2] = {nullptr, nullptr};
            ~Events()
            {
                for (hipEvent_t x : e)
                    if (x) (void)hipEventDestroy(x);
            }
        } ev;
        BHIP(b, hipEventCreate(&ev.e[0]));
        BHIP(b, hipEventCreate(&ev.e[1]));
        hipEvent_t &e0 = ev.e[0], &e1 = ev.e[1];
        const bool mixed = b->prec == LYM_PREC_MIXED;
        // Each candidate is timed over whole RK steps of the batch flow itself
        // (its three stage launches, in the form the candidate keeps the links
        // in), so that the pick holds for a flow.  The fields, three rows in
        // lexicographic order after to_lex() above, are saved first and put
        // back afterwards.
        const size_t bytes = (b->prec == LYM_PREC_FP64 ? sizeof(double) : sizeof(float)) * 18 * size_t(b->total_links);
        struct Save {
            void *p = nullptr;
            ~Save()
            {
                if (p) (void)hipFree(p);
            }
        } save;
        if (int rc = to_form(b, 3)) return rc;
        BHIP(b, hipMalloc(&save.p, bytes));
        BHIP(b, hipMemcpyAsync(save.p, b->W[b->cur], bytes, hipMemcpyDeviceToDevice, b->stream));
        // an untimed, then a timed flow step under the schedule just applied
        auto time_step = [&](float *ms) -> int {
            if (int rc = to_form(b, b->rows)) return rc;
            if (int rc = step(b, lym::kTuneEps)) return rc;
            BHIP(b, hipEventRecord(e0, b->stream));
            if (int rc = step(b, lym::kTuneEps)) return rc;
            BHIP(b, hipEventRecord(e1, b->stream));
            BHIP(b, hipEventSynchronize(e1));
            BHIP(b, hipEventElapsedTime(ms, e0, e1));
            return LYM_OK;
        };
        int rc, best = 0;
        // one workgroup per slot, every order ...
        const auto cands = lym::schedule_candidates(mixed);
        auto time_order = [&](int i, float *ms) -> int {
            if (int rc = apply_schedule(b, 0, cands[size_t(i)].first, cands[size_t(i)].second)) return rc;
            return time_step(ms);
        };
        rc = lym::tune_rounds(int(cands.size()), 3, time_order, &best, &best_ms);
        if (!rc) {
            best_grid = 0;
            best_kind = cands[size_t(best)].first;
            best_target = cands[size_t(best)].second;
            // ... then the best order again with persistent grids of {2, 4, 8} x 256 workgroups
            const int grids[4] = {0, 512, 1024, 2048};
            auto time_grid = [&](int i, float *ms) -> int {
                if (int rc = apply_schedule(b, grids[i], best_kind, best_target)) return rc;
                return time_step(ms);
            };
            rc = lym::tune_rounds(4, 3, time_grid, &best, &best_ms);
            best_grid = grids[best];
        }
        if (!rc) {
            // ... and that shape against one whole-flow launch per flow segment
            // (timed over one step, i.e. three stages and two grid barriers);
            // a whole-flow launch that gave up waiting at a barrier, or a
            // kernel that does not fit the device, does not count
            const int kinds[2] = {best_kind, best_kind | lym::SCHED_WHOLE_FLOW}, wgrids[2] = {best_grid, 0};
            auto time_whole = [&](int i, float *ms) -> int {
                if (int rc = apply_schedule(b, wgrids[i], kinds[i], best_target)) {
                    if (i == 0) return rc;
                    *ms = 1e30f;
                    return apply_schedule(b, wgrids[0], kinds[0], best_target);
                }
                if (i == 1) lym::k_flow_sync_reset<<<1, 256, 0, b->stream>>>(b->sync, 1);
                if (int rc = time_step(ms)) return rc;
                bool aborted = false;
                if (i == 1)
                    if (int rc = whole_flow_aborted(b, &aborted)) return rc;
                if (aborted) *ms = 1e30f;
                return LYM_OK;
            };
            rc = lym::tune_rounds(2, 3, time_whole, &best, &best_ms);
            if (!rc) {
                best_kind = kinds[best];
                best_grid = wgrids[best];
            }
        }
        // the saved fields back, whatever the outcome
        const hipError_t e = hipMemcpyAsync(b->W[b->cur], save.p, bytes, hipMemcpyDeviceToDevice, b->stream);
        b->form = 3;
        if (rc) return rc;
        BHIP(b, e);
        BHIP(b, hipStreamSynchronize(b->stream));
        b->tuned = "measured";
        if (cache_path) {
            std::ofstream o(cache_path, std::ios::app);
            o << key << "\t" << best_grid << " " << best_kind << " " << best_target << " " << best_ms << "\n";
        }
    }
    int rc = apply_schedule(b, best_grid, best_kind, best_target);
    if (rc) return rc;
    if (summary && summary_len > 0)
        snprintf(summary, size_t(summary_len), "grid=%d schedule=%s%s target=%d step_ms=%.4f source=%s tiles=%ld",
                 b->grid, lym::schedule_name(best_kind & ~lym::SCHED_WHOLE_FLOW),
                 (best_kind & lym::SCHED_WHOLE_FLOW) ? "+whole" : "", best_target, best_ms, b->tuned.c_str(),
                 b->ntiles);
    return LYM_OK;
}

}  // extern "C"
