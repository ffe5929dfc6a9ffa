// C ABI of the single-lattice path (declared in include/lym.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lym.h"
#include "lym_launch.hpp"
#include "lym_sched.hpp"
#include "lym_xfer.hpp"

#include <fstream>
#include <map>
#include <sstream>

using lym::Geom;
using lym::ksum;

namespace lym {
const char *noalloc_last_error();  // lym_noalloc.hip
}

// Metropolis blocks in column order (lym::metro_column_schedule): each XCD
// marches (z, y) columns of about this many sites of one parity along t, so
// the neighbouring time slices of a column stay in its 4 MB L2.  A time
// slice of 32^3 x 64 in even-odd order is 9.4 MB per parity (FP64), which
// the XCD-contiguous order refetches for every slice it updates.
constexpr int kMetroColumnSites = 4096;

struct lym_ctx {
    int device = 0;
    int prec = LYM_PREC_FP64;
    int dims[4] = {0, 0, 0, 0};
    Geom g{};
    hipStream_t stream = nullptr;
    void *W[2] = {nullptr, nullptr};  // double-buffered link field (SoA)
    int cur = 0;                      // W[cur] holds the current links
    int rows = 3;                     // form of the links while the flow runs (2: two-row, SCHED_TWO_ROW)
    int form = 3;                     // form of W[cur] now: 3 between calls
    bool eo = false;                  // ... in even-odd order (left so by lym_metropolis)
    void *acc = nullptr;              // RK accumulator, 8 reals per link
    ksum *partials = nullptr;         // per-workgroup observables partials
    int nblk_obs = 0;
    double *obs = nullptr;  // measurement slots, 4 doubles each
    int obs_cap = 0;
    double *staging = nullptr;  // V x 18 doubles (reference layout of one direction)
    lym::Bounce bounce;         // pinned buffers of lym_upload_links / lym_download_links
    unsigned long long *accepted = nullptr;
    unsigned long long *check = nullptr;  // lym_check_links: max deviation bits, non-finite links
    double *predrawn = nullptr;
    size_t predrawn_cap = 0;
    std::vector<hipEvent_t> events;
    lym::HostSchedule sched;  // work schedule of the stencil kernels
    int *d_order = nullptr;
    int metro_target = kMetroColumnSites;  // Metropolis block order: 0 = XCD-contiguous, else columns
    int *d_morder = nullptr;               // its table (built on first use)
    int morder_per_xcd = 0;
    // lym_flow as a captured graph (lym_set_graph), one per start buffer
    // W[cur]; a graph holds the buffer pointers of its capture, so every
    // reallocation of d_order, partials or obs drops both
    bool use_graph = false;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    std::string graph_key[2];
    std::string err;
};

namespace {

int fail(lym_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

int hipfail(lym_ctx *c, hipError_t e, const char *what)
{
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(c, e == hipErrorOutOfMemory ? LYM_ERR_NOMEM : LYM_ERR_HIP, m);
}

#define LYM_HIP(c, call)                                   \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return hipfail(c, e_, #call); \
    } while (0)

#define LYM_LAUNCHED(c, what)                              \
    do {                                                   \
        hipError_t e_ = hipGetLastError();                 \
        if (e_ != hipSuccess) return hipfail(c, e_, what); \
    } while (0)

size_t real_size(const lym_ctx *c) { return c->prec == LYM_PREC_FP64 ? sizeof(double) : sizeof(float); }

unsigned blocks_for(long n, int per) { return unsigned((n + per - 1) / per); }

void drop_graphs(lym_ctx *c)
{
    for (int k = 0; k < 2; k++) {
        if (c->graph[k]) (void)hipGraphExecDestroy(c->graph[k]);
        c->graph[k] = nullptr;
        c->graph_key[k].clear();
    }
}

lym::Sched dev_sched(const lym_ctx *c) { return lym::Sched{c->sched.order.empty() ? nullptr : c->d_order, c->sched.per_xcd}; }

// workgroups of a stencil launch under the current schedule
unsigned stencil_blocks(const lym_ctx *c)
{
    return c->sched.order.empty() ? blocks_for(c->g.V, lym::kSitesPerBlock) : unsigned(lym::kXcds * c->sched.per_xcd);
}

// Copies W[cur] into W[1 - cur] reordered to (to_eo) or from the even-odd
// order of the Metropolis sweeps (lym::EOLinks).
void reorder_eo(lym_ctx *c, bool to_eo)
{
    const dim3 grid(blocks_for(c->g.V, 256), 36);
    if (c->prec == LYM_PREC_FP64) {
        auto *in = static_cast<const double2 *>(c->W[c->cur]);
        auto *out = static_cast<double2 *>(c->W[1 - c->cur]);
        if (to_eo)
            lym::k_reorder_eo<double, true><<<grid, 256, 0, c->stream>>>(in, out, c->g);
        else
            lym::k_reorder_eo<double, false><<<grid, 256, 0, c->stream>>>(in, out, c->g);
    } else {
        auto *in = static_cast<const float2 *>(c->W[c->cur]);
        auto *out = static_cast<float2 *>(c->W[1 - c->cur]);
        if (to_eo)
            lym::k_reorder_eo<float, true><<<grid, 256, 0, c->stream>>>(in, out, c->g);
        else
            lym::k_reorder_eo<float, false><<<grid, 256, 0, c->stream>>>(in, out, c->g);
    }
}

// The links in W[cur] are in lexicographic order, except after
// lym_metropolis, which leaves them in even-odd order so that consecutive
// calls (one sweep each, as the reference's MCUpdate) skip both reorders.
// Every entry point that reads them site by site calls to_lex first; those
// that overwrite them all clear the flag.
int to_order(lym_ctx *c, bool eo)
{
    if (c->eo == eo) return LYM_OK;
    reorder_eo(c, eo);
    LYM_LAUNCHED(c, "k_reorder_eo");
    c->cur ^= 1;
    c->eo = eo;
    return LYM_OK;
}
int to_lex(lym_ctx *c) { return to_order(c, false); }

int apply_schedule(lym_ctx *c, int kind, int target)
{
    std::vector<std::array<int, 4>> d{{c->dims[0], c->dims[1], c->dims[2], c->dims[3]}};
    lym::HostSchedule h = lym::build_schedule(d, {0L}, kind, target);
    drop_graphs(c);
    if (c->d_order) (void)hipFree(c->d_order);
    c->d_order = nullptr;
    if (!h.order.empty()) {
        LYM_HIP(c, hipMalloc(&c->d_order, sizeof(int) * h.order.size()));
        LYM_HIP(c, hipMemcpy(c->d_order, h.order.data(), sizeof(int) * h.order.size(), hipMemcpyHostToDevice));
    }
    // partials: one slot per workgroup of the observables launch
    const long nblk = h.order.empty() ? long(blocks_for(c->g.V, lym::kObsSites)) : long(lym::kXcds) * h.per_xcd;
    if (nblk > c->nblk_obs) {
        if (c->partials) (void)hipFree(c->partials);
        c->partials = nullptr;
        LYM_HIP(c, hipMalloc(&c->partials, sizeof(ksum) * 4 * size_t(nblk)));
    }
    c->nblk_obs = int(nblk);
    c->sched = std::move(h);
    c->rows = (c->sched.kind & lym::SCHED_TWO_ROW) ? 2 : 3;
    return LYM_OK;
}

// LDS staging mode of the stage kernel for this lattice under the schedule
// (flow_block): slab where whole x-rows never cross a z-plane, whole x-rows,
// 16 x 4 tiles of every (z, t) plane, else the run's own links.
int stage_lds_mode(const lym_ctx *c)
{
    if (!(c->sched.kind & lym::SCHED_LDS)) return 0;
    if (lym::slab_fits(c->dims)) return 3;
    if (lym::kSitesPerBlock % c->dims[0] == 0) return 2;
    if (lym::tile_fits(c->dims)) return 5;
    return 1;
}

// one RK stage over the lattice with the context's schedule and kernel variant
template <typename S, int STAGE>
void launch_stage(const lym_ctx *c, const lym::F2<S> *win, lym::F2<S> *wout, S *acc, double eps)
{
    if constexpr (sizeof(S) == 4)
        if (c->form == 2) {
            lym::launch_flow_stage_r2<STAGE>(stage_lds_mode(c), stencil_blocks(c), c->stream, win, wout, acc, c->g,
                                             dev_sched(c), eps);
            return;
        }
    lym::launch_flow_stage<S, STAGE>(stage_lds_mode(c), stencil_blocks(c), c->stream, win, wout, acc, c->g,
                                     dev_sched(c), eps);
}

// Puts W[cur] into form `rows` (3: three-row, 2: two-row, the mixed flow's
// SCHED_TWO_ROW) by a conversion into W[1 - cur].  The flow entry points take
// the field into the schedule's form and leave it there, so a caller that
// alternates one flow step and one measurement (the reference's loop) pays
// no conversion; lym_observables reads either form, and the entry points
// that read links as three rows (download, staples, force, Metropolis, link
// check, autotune) convert back first (readable).  Two-row form implies
// lexicographic order: Metropolis leaves even-odd order in three-row form.
int to_form(lym_ctx *c, int rows)
{
    if (c->form == rows) return LYM_OK;
    const dim3 grid(blocks_for(c->g.V, 256), 4);
    auto *in = static_cast<const float2 *>(c->W[c->cur]);
    auto *out = static_cast<float2 *>(c->W[1 - c->cur]);
    if (rows == 2)
        lym::k_to_rows2<float><<<grid, 256, 0, c->stream>>>(in, out, c->g.V);
    else
        lym::k_from_rows2<float><<<grid, 256, 0, c->stream>>>(in, out, c->g.V);
    LYM_LAUNCHED(c, "k_to_rows2 / k_from_rows2");
    c->cur ^= 1;
    c->form = rows;
    return LYM_OK;
}

// the links as three rows in lexicographic order
int readable(lym_ctx *c)
{
    if (int rc = to_form(c, 3)) return rc;
    return to_lex(c);
}

template <typename S> int enqueue_step(lym_ctx *c, double eps, hipEvent_t *ev)
{
    using F2 = lym::F2<S>;
    F2 *a = static_cast<F2 *>(c->W[c->cur]);
    F2 *b = static_cast<F2 *>(c->W[1 - c->cur]);
    S *acc = static_cast<S *>(c->acc);
    if (ev) LYM_HIP(c, hipEventRecord(ev[0], c->stream));
    launch_stage<S, 1>(c, a, b, acc, eps);
    if (ev) LYM_HIP(c, hipEventRecord(ev[1], c->stream));
    launch_stage<S, 2>(c, b, a, acc, eps);
    if (ev) LYM_HIP(c, hipEventRecord(ev[2], c->stream));
    launch_stage<S, 3>(c, a, b, acc, eps);
    if (ev) LYM_HIP(c, hipEventRecord(ev[3], c->stream));
    LYM_LAUNCHED(c, "k_flow_stage");
    c->cur = 1 - c->cur;
    return LYM_OK;
}

int step(lym_ctx *c, double eps, hipEvent_t *ev = nullptr)
{
    return c->prec == LYM_PREC_FP64 ? enqueue_step<double>(c, eps, ev) : enqueue_step<float>(c, eps, ev);
}

// observables of the current field into measurement slot `slot`
template <typename S> int enqueue_observables(lym_ctx *c, int slot, hipEvent_t *ev)
{
    const lym::F2<S> *f = static_cast<const lym::F2<S> *>(c->W[c->cur]);
    if (ev) LYM_HIP(c, hipEventRecord(ev[0], c->stream));
    if (sizeof(S) == 4 && c->form == 2)
        lym::k_observables<S, 2><<<c->nblk_obs, 192, 0, c->stream>>>(f, c->g, dev_sched(c), c->partials);
    else
        lym::k_observables<S><<<c->nblk_obs, 192, 0, c->stream>>>(f, c->g, dev_sched(c), c->partials);
    if (ev) LYM_HIP(c, hipEventRecord(ev[1], c->stream));
    lym::k_reduce_partials<<<1, lym::kReduceThreads, 0, c->stream>>>(c->partials, c->nblk_obs, c->obs + 4 * slot);
    LYM_LAUNCHED(c, "k_observables");
    return LYM_OK;
}

int observables(lym_ctx *c, int slot, hipEvent_t *ev = nullptr)
{
    return c->prec == LYM_PREC_FP64 ? enqueue_observables<double>(c, slot, ev)
                                    : enqueue_observables<float>(c, slot, ev);
}

int ensure_slots(lym_ctx *c, int n)
{
    if (n <= c->obs_cap) return LYM_OK;
    LYM_HIP(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    if (c->obs) (void)hipFree(c->obs);
    c->obs = nullptr;
    LYM_HIP(c, hipMalloc(&c->obs, sizeof(double) * 4 * size_t(n)));
    c->obs_cap = n;
    return LYM_OK;
}

// raw sums -> reference normalisation (src/Observables/superobs.cpp:125-126)
void normalise(const lym_ctx *c, const double *raw, double out[4])
{
    const double qn = 16 * 16 * M_PI * M_PI;
    out[0] = raw[0] / 18.0 / double(c->g.V);
    out[1] = raw[1];
    out[2] = raw[2] / qn;
    out[3] = raw[3] / qn;
}

int ensure_events(lym_ctx *c, size_t n)
{
    while (c->events.size() < n) {
        hipEvent_t e;
        LYM_HIP(c, hipEventCreate(&e));
        c->events.push_back(e);
    }
    return LYM_OK;
}

}  // namespace

extern "C" {

int lym_create(lym_ctx **out, const int dims[4], int precision, int device)
{
    if (!out || !dims) return LYM_ERR_ARG;
    *out = nullptr;
    if (precision != LYM_PREC_FP64 && precision != LYM_PREC_MIXED) return LYM_ERR_ARG;
    long V = 1;
    for (int d = 0; d < 4; d++) {
        if (dims[d] < 2) return LYM_ERR_ARG;
        V *= dims[d];
    }
    // 48 V bytes per buffer descriptor (FP64: a row, FP32: two rows); it also
    // keeps the 32-bit Philox link ids (4 V) in range
    if (48 * V >= (1L << 31)) return LYM_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return LYM_ERR_NODEV;
    if (device < 0 || device >= ndev) return LYM_ERR_ARG;
    lym_ctx *c = new lym_ctx;
    c->device = device;
    c->prec = precision;
    long st = 1;
    for (int d = 0; d < 4; d++) {
        c->dims[d] = dims[d];
        c->g.L[d] = dims[d];
        c->g.stride[d] = int(st);
        st *= dims[d];
    }
    c->g.V = V;
    c->nblk_obs = int(blocks_for(V, lym::kObsSites));
    auto bail = [&](int code) {
        lym_destroy(c);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(LYM_ERR_HIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(LYM_ERR_HIP);
    const size_t rs = real_size(c);
    for (int k = 0; k < 2; k++)
        if (hipMalloc(&c->W[k], rs * 2 * 9 * 4 * size_t(V)) != hipSuccess) return bail(LYM_ERR_NOMEM);
    if (hipMalloc(&c->acc, rs * 8 * 4 * size_t(V)) != hipSuccess) return bail(LYM_ERR_NOMEM);
    if (hipMalloc(&c->partials, sizeof(ksum) * 4 * size_t(c->nblk_obs)) != hipSuccess) return bail(LYM_ERR_NOMEM);
    if (hipMalloc(&c->staging, sizeof(double) * 18 * size_t(V)) != hipSuccess) return bail(LYM_ERR_NOMEM);
    if (hipMalloc(&c->accepted, sizeof(unsigned long long)) != hipSuccess) return bail(LYM_ERR_NOMEM);
    if (hipMalloc(&c->check, 2 * sizeof(unsigned long long)) != hipSuccess) return bail(LYM_ERR_NOMEM);
    if (ensure_slots(c, 16) != LYM_OK) return bail(LYM_ERR_NOMEM);
    // the default schedule (the mixed flow in two-row form), the autotuner's first candidate
    if (apply_schedule(c, lym::default_schedule_kind(precision == LYM_PREC_MIXED), lym::kDefaultColumnTarget) != LYM_OK)
        return bail(LYM_ERR_NOMEM);
    int rc = lym_set_cold(c);
    if (rc != LYM_OK) return bail(rc);
    *out = c;
    return LYM_OK;
}

int lym_destroy(lym_ctx *c)
{
    if (!c) return LYM_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    c->bounce.release();
    for (void *p : {c->W[0], c->W[1], c->acc, (void *)c->partials, (void *)c->obs, (void *)c->staging,
                    (void *)c->accepted, (void *)c->check, (void *)c->predrawn, (void *)c->d_order,
                    (void *)c->d_morder})
        if (p) (void)hipFree(p);
    for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LYM_OK;
}

const char *lym_last_error(const lym_ctx *c)
{
    if (c) return c->err.c_str();
    const char *m = lym::noalloc_last_error();  // context-free calls (lym_observables_noalloc)
    return m ? m : "null context";
}

int lym_info(const lym_ctx *c, int dims[4], int64_t *volume, int *precision)
{
    if (!c) return LYM_ERR_ARG;
    if (dims)
        for (int d = 0; d < 4; d++) dims[d] = c->dims[d];
    if (volume) *volume = c->g.V;
    if (precision) *precision = c->prec;
    return LYM_OK;
}

int lym_upload_links(lym_ctx *c, const double *const by_mu[4])
{
    if (!c || !by_mu) return LYM_ERR_ARG;
    for (int mu = 0; mu < 4; mu++)
        if (!by_mu[mu]) return fail(c, LYM_ERR_ARG, "null link array");
    LYM_HIP(c, hipSetDevice(c->device));
    const long V = c->g.V;
    // a mixed context whose flow keeps two-row links takes them in that form
    // at once (the flow then converts nothing)
    const int form = c->prec == LYM_PREC_MIXED ? c->rows : 3;
    for (int mu = 0; mu < 4; mu++) {
        LYM_HIP(c, lym::h2d_bounced(c->bounce, c->stream, c->staging, by_mu[mu], sizeof(double) * 18 * V));
        if (c->prec == LYM_PREC_FP64)
            lym::k_aos_to_soa<double><<<blocks_for(V, 256), 256, 0, c->stream>>>(
                c->staging, static_cast<double2 *>(c->W[c->cur]) + size_t(mu) * 9 * V, V);
        else if (form == 2)
            lym::k_aos_to_rows2<float><<<blocks_for(V, 256), 256, 0, c->stream>>>(
                c->staging, static_cast<float2 *>(c->W[c->cur]) + size_t(mu) * 6 * V, V);
        else
            lym::k_aos_to_soa<float><<<blocks_for(V, 256), 256, 0, c->stream>>>(
                c->staging, static_cast<float2 *>(c->W[c->cur]) + size_t(mu) * 9 * V, V);
        LYM_LAUNCHED(c, "k_aos_to_soa");
    }
    c->eo = false;
    c->form = form;
    LYM_HIP(c, hipStreamSynchronize(c->stream));
    return LYM_OK;
}

int lym_download_links(lym_ctx *c, double *const by_mu[4])
{
    if (!c || !by_mu) return LYM_ERR_ARG;
    for (int mu = 0; mu < 4; mu++)
        if (!by_mu[mu]) return fail(c, LYM_ERR_ARG, "null link array");
    LYM_HIP(c, hipSetDevice(c->device));
    if (int rc = readable(c)) return rc;
    const long V = c->g.V;
    for (int mu = 0; mu < 4; mu++) {
        if (c->prec == LYM_PREC_FP64)
            lym::k_soa_to_aos<double><<<blocks_for(V, 256), 256, 0, c->stream>>>(
                static_cast<const double2 *>(c->W[c->cur]) + size_t(mu) * 9 * V, c->staging, V);
        else
            lym::k_soa_to_aos<float><<<blocks_for(V, 256), 256, 0, c->stream>>>(
                static_cast<const float2 *>(c->W[c->cur]) + size_t(mu) * 9 * V, c->staging, V);
        LYM_LAUNCHED(c, "k_soa_to_aos");
        LYM_HIP(c, lym::d2h_bounced(c->bounce, c->stream, by_mu[mu], c->staging, sizeof(double) * 18 * V));
    }
    return LYM_OK;
}

int lym_set_cold(lym_ctx *c)
{
    if (!c) return LYM_ERR_ARG;
    LYM_HIP(c, hipSetDevice(c->device));
    const long V = c->g.V;
    if (c->prec == LYM_PREC_FP64)
        lym::k_cold<double><<<blocks_for(4 * V, 256), 256, 0, c->stream>>>(static_cast<double2 *>(c->W[c->cur]), V);
    else
        lym::k_cold<float><<<blocks_for(4 * V, 256), 256, 0, c->stream>>>(static_cast<float2 *>(c->W[c->cur]), V);
    LYM_LAUNCHED(c, "k_cold");
    c->eo = false;
    c->form = 3;
    LYM_HIP(c, hipStreamSynchronize(c->stream));
    return LYM_OK;
}

int lym_set_hot(lym_ctx *c, uint64_t seed)
{
    if (!c) return LYM_ERR_ARG;
    LYM_HIP(c, hipSetDevice(c->device));
    const long V = c->g.V;
    if (c->prec == LYM_PREC_FP64)
        lym::k_hot<double><<<blocks_for(4 * V, 256), 256, 0, c->stream>>>(static_cast<double2 *>(c->W[c->cur]), V, seed);
    else
        lym::k_hot<float><<<blocks_for(4 * V, 256), 256, 0, c->stream>>>(static_cast<float2 *>(c->W[c->cur]), V, seed);
    LYM_LAUNCHED(c, "k_hot");
    c->eo = false;
    c->form = 3;
    LYM_HIP(c, hipStreamSynchronize(c->stream));
    return LYM_OK;
}

static int export_force(lym_ctx *c, int mu, int what, double *out)
{
    if (!c || !out || mu < 0 || mu > 3) return c ? fail(c, LYM_ERR_ARG, "bad mu or null output") : LYM_ERR_ARG;
    LYM_HIP(c, hipSetDevice(c->device));
    if (int rc = readable(c)) return rc;
    const long V = c->g.V;
    if (c->prec == LYM_PREC_FP64)
        lym::k_export_force<double><<<blocks_for(V, 256), 256, 0, c->stream>>>(
            static_cast<const double2 *>(c->W[c->cur]), c->g, mu, what, c->staging);
    else
        lym::k_export_force<float><<<blocks_for(V, 256), 256, 0, c->stream>>>(
            static_cast<const float2 *>(c->W[c->cur]), c->g, mu, what, c->staging);
    LYM_LAUNCHED(c, "k_export_force");
    LYM_HIP(c, lym::d2h_bounced(c->bounce, c->stream, out, c->staging, sizeof(double) * 18 * V));
    return LYM_OK;
}

int lym_staples(lym_ctx *c, int mu, double *omega_out) { return export_force(c, mu, 0, omega_out); }
int lym_derivative(lym_ctx *c, int mu, double *z_out) { return export_force(c, mu, 1, z_out); }

int lym_observables(lym_ctx *c, double out[4])
{
    if (!c || !out) return LYM_ERR_ARG;
    LYM_HIP(c, hipSetDevice(c->device));
    int rc = to_lex(c);
    if (rc || (rc = observables(c, 0))) return rc;
    if (rc) return rc;
    double raw[4];
    LYM_HIP(c, hipMemcpyAsync(raw, c->obs, sizeof raw, hipMemcpyDeviceToHost, c->stream));
    LYM_HIP(c, hipStreamSynchronize(c->stream));
    normalise(c, raw, out);
    return LYM_OK;
}

int lym_flow_steps(lym_ctx *c, double eps, int nsteps)
{
    if (!c || nsteps < 0 || !(eps > 0)) return c ? fail(c, LYM_ERR_ARG, "bad eps or nsteps") : LYM_ERR_ARG;
    LYM_HIP(c, hipSetDevice(c->device));
    if (int rc = to_lex(c)) return rc;
    if (nsteps == 0) return LYM_OK;
    if (int rc = to_form(c, c->rows)) return rc;
    for (int i = 0; i < nsteps; i++) {
        int rc = step(c, eps);
        if (rc) return rc;
    }
    return LYM_OK;
}

int lym_flow(lym_ctx *c, double eps, double tau, int measure_every, double *series, int max_rows, int *rows_out)
{
    if (!c || !series || measure_every < 1 || !(eps > 0) || !(tau > 0))
        return c ? fail(c, LYM_ERR_ARG, "bad flow arguments") : LYM_ERR_ARG;
    const int rows = int(tau / eps);  // src/Apps/wilsonflow.cpp:117
    if (rows_out) *rows_out = rows;
    if (rows < 1 || rows > max_rows) return fail(c, LYM_ERR_ARG, "series too small for int(tau/eps) rows");
    LYM_HIP(c, hipSetDevice(c->device));
    const int nmeas = 1 + (rows - 1) / measure_every;
    int rc = to_lex(c);
    if (rc || (rc = ensure_slots(c, nmeas))) return rc;
    auto enqueue = [&]() -> int {
        int slot = 0, r = to_form(c, c->rows);
        if (!r) r = observables(c, slot++);
        for (int t = 1; t < rows && !r; t++) {
            r = step(c, eps);
            if (!r && t % measure_every == 0) r = observables(c, slot++);
        }
        return r;
    };
    if (c->use_graph) {
        // key: the flow parameters and the form the field starts in (the
        // buffers are fixed while the graph lives; a schedule change drops it)
        char key[96];
        const int form0 = c->form;
        snprintf(key, sizeof key, "%a|%d|%d|%d", eps, rows, measure_every, form0);
        const int k = c->cur;
        if (!c->graph[k] || c->graph_key[k] != key) {
            if (c->graph[k]) (void)hipGraphExecDestroy(c->graph[k]);
            c->graph[k] = nullptr;
            c->graph_key[k].clear();
            LYM_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue();
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(c->stream, &g);
            c->cur = k;  // capture enqueued nothing: the field is where it was
            c->form = form0;
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return hipfail(c, e, "hipStreamEndCapture");
            const hipError_t ei = hipGraphInstantiate(&c->graph[k], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) return hipfail(c, ei, "hipGraphInstantiate");
            c->graph_key[k] = key;
        }
        LYM_HIP(c, hipGraphLaunch(c->graph[k], c->stream));
        // rows - 1 steps, plus the form conversion at the start if the field
        // was not in the schedule's form
        if (((rows - 1) + (form0 != c->rows ? 1 : 0)) & 1) c->cur = 1 - k;
        c->form = c->rows;
    } else if ((rc = enqueue())) {
        return rc;
    }
    std::vector<double> raw(4 * size_t(nmeas));
    LYM_HIP(c, hipMemcpyAsync(raw.data(), c->obs, sizeof(double) * raw.size(), hipMemcpyDeviceToHost, c->stream));
    LYM_HIP(c, hipStreamSynchronize(c->stream));
    double tcur = 0.0;
    int slot = 0;
    for (int t = 0; t < rows; t++) {
        if (t > 0) tcur += eps;  // src/Apps/wilsonflow.cpp:126
        double *row = series + 4 * size_t(t);
        row[0] = tcur;
        if (t % measure_every == 0) {
            double o[4];
            normalise(c, &raw[4 * size_t(slot++)], o);
            row[1] = o[0];
            row[2] = o[2];
            row[3] = o[1];
        } else {
            row[1] = row[2] = row[3] = NAN;
        }
    }
    return LYM_OK;
}

int lym_set_graph(lym_ctx *c, int on)
{
    if (!c) return LYM_ERR_ARG;
    LYM_HIP(c, hipSetDevice(c->device));
    LYM_HIP(c, hipStreamSynchronize(c->stream));
    c->use_graph = on != 0;
    if (!c->use_graph) drop_graphs(c);
    return LYM_OK;
}

// One Metropolis sub-sweep launch; the candidate kind and the test mode pick
// the kernel instantiation, so each carries only its own registers.
extern "C++" template <typename S>
static void launch_metropolis(lym_ctx *c, unsigned nb, int mu, const lym::MetroParams &p)
{
    const unsigned per_xcd = c->d_morder ? unsigned(c->morder_per_xcd) : (nb + 7) / 8;
    lym::launch_metropolis<S>(lym::metro_kernel_kind(p.candidate == LYM_CAND_EXPH, p.eps), p.predrawn != nullptr,
                              per_xcd, c->stream, static_cast<lym::F2<S> *>(c->W[c->cur]), c->g, mu, p, c->d_morder,
                              c->accepted);
}

int lym_metropolis(lym_ctx *c, double beta, double eps, int nhit, int nsweeps, int first_sweep, int candidate,
                   uint64_t seed, const double *predrawn, int64_t *accepted)
{
    if (!c) return LYM_ERR_ARG;
    if (nhit < 1 || nsweeps < 0 || first_sweep < 0 || (candidate != LYM_CAND_SU2 && candidate != LYM_CAND_EXPH) ||
        !(eps > 0 && eps < 1))
        return fail(c, LYM_ERR_ARG, "bad Metropolis arguments");
    for (int d = 0; d < 4; d++)
        if (c->dims[d] % 2) return fail(c, LYM_ERR_ARG, "checkerboard needs even extents");
    LYM_HIP(c, hipSetDevice(c->device));
    const long V = c->g.V;
    const int nu = candidate == LYM_CAND_SU2 ? 13 : 9;
    const size_t per_sweep = size_t(4) * V * nhit * nu;
    if (predrawn) {
        const size_t need = per_sweep * nsweeps;
        if (need > c->predrawn_cap) {
            if (c->predrawn) (void)hipFree(c->predrawn);
            c->predrawn = nullptr;
            c->predrawn_cap = 0;
            LYM_HIP(c, hipMalloc(&c->predrawn, sizeof(double) * need));
            c->predrawn_cap = need;
        }
        LYM_HIP(c, hipMemcpyAsync(c->predrawn, predrawn, sizeof(double) * need, hipMemcpyHostToDevice, c->stream));
    }
    lym::k_zero_u64<<<1, 64, 0, c->stream>>>(c->accepted, 1);
    LYM_LAUNCHED(c, "k_zero_u64");
    // The sweeps run on the links in even-odd order, which they are left in
    // (to_lex reorders them when another entry point reads them).
    if (nsweeps > 0) {
        if (int rc = to_form(c, 3)) return rc;
        if (int rc = to_order(c, true)) return rc;
    }
    if (c->metro_target > 0 && !c->d_morder) {
        const lym::HostSchedule h =
            lym::metro_column_schedule({c->dims[0], c->dims[1], c->dims[2], c->dims[3]}, c->metro_target);
        LYM_HIP(c, hipMalloc(&c->d_morder, sizeof(int) * h.order.size()));
        LYM_HIP(c, hipMemcpy(c->d_morder, h.order.data(), sizeof(int) * h.order.size(), hipMemcpyHostToDevice));
        c->morder_per_xcd = h.per_xcd;
    }
    const unsigned nb = blocks_for(V / 2, 256);
    for (int sw = 0; sw < nsweeps; sw++) {
        lym::MetroParams p{beta, eps, nhit, candidate, nu, uint32_t(first_sweep + sw), 0, seed,
                           predrawn ? c->predrawn + per_sweep * sw : nullptr};
        for (int mu = 0; mu < 4; mu++)
            for (int par = 0; par < 2; par++) {
                p.parity = par;
                if (c->prec == LYM_PREC_FP64)
                    launch_metropolis<double>(c, nb, mu, p);
                else
                    launch_metropolis<float>(c, nb, mu, p);
                LYM_LAUNCHED(c, "k_metropolis");
            }
    }
    unsigned long long acc = 0;
    LYM_HIP(c, hipMemcpyAsync(&acc, c->accepted, sizeof acc, hipMemcpyDeviceToHost, c->stream));
    LYM_HIP(c, hipStreamSynchronize(c->stream));
    if (accepted) *accepted = int64_t(acc);
    return LYM_OK;
}

int lym_check_links(lym_ctx *c, double *max_dev, int64_t *nonfinite)
{
    if (!c) return LYM_ERR_ARG;
    LYM_HIP(c, hipSetDevice(c->device));
    // a maximum and a count over all links of the three-row field: either
    // site order will do
    if (int rc = to_form(c, 3)) return rc;
    unsigned long long *d = c->check;
    lym::k_zero_u64<<<1, 64, 0, c->stream>>>(d, 2);
    LYM_LAUNCHED(c, "k_zero_u64");
    const unsigned nb = blocks_for(4 * c->g.V, 256);
    if (c->prec == LYM_PREC_FP64)
        lym::k_check_links<double><<<nb, 256, 0, c->stream>>>(static_cast<double2 *>(c->W[c->cur]), c->g.V, d);
    else
        lym::k_check_links<float><<<nb, 256, 0, c->stream>>>(static_cast<float2 *>(c->W[c->cur]), c->g.V, d);
    LYM_LAUNCHED(c, "k_check_links");
    unsigned long long h[2];
    LYM_HIP(c, hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, c->stream));
    LYM_HIP(c, hipStreamSynchronize(c->stream));
    double v;
    std::memcpy(&v, &h[0], sizeof v);
    if (max_dev) *max_dev = v;
    if (nonfinite) *nonfinite = int64_t(h[1]);
    return LYM_OK;
}

int lym_synchronize(lym_ctx *c)
{
    if (!c) return LYM_ERR_ARG;
    LYM_HIP(c, hipStreamSynchronize(c->stream));
    return LYM_OK;
}

int lym_profile_flow(lym_ctx *c, double eps, int nsteps, int measure_every, double *total_ms, double stage_ms[4])
{
    if (!c || nsteps < 1 || measure_every < 0 || !stage_ms) return c ? fail(c, LYM_ERR_ARG, "bad profile args") : LYM_ERR_ARG;
    LYM_HIP(c, hipSetDevice(c->device));
    const int nmeas = measure_every ? nsteps / measure_every : 0;
    const size_t nev = 2 + 4 * size_t(nsteps) + 2 * size_t(nmeas);
    int rc = to_lex(c);
    if (rc || (rc = ensure_events(c, nev)) || (rc = to_form(c, c->rows))) return rc;
    hipEvent_t *ev = c->events.data();
    LYM_HIP(c, hipEventRecord(ev[0], c->stream));
    size_t k = 2;
    std::vector<size_t> obs_at;
    for (int i = 1; i <= nsteps; i++) {
        if ((rc = step(c, eps, ev + k))) return rc;
        k += 4;
        if (measure_every && i % measure_every == 0) {
            obs_at.push_back(k);
            if ((rc = observables(c, 0, ev + k))) return rc;
            k += 2;
        }
    }
    LYM_HIP(c, hipEventRecord(ev[1], c->stream));
    LYM_HIP(c, hipEventSynchronize(ev[1]));
    float ms = 0;
    LYM_HIP(c, hipEventElapsedTime(&ms, ev[0], ev[1]));
    if (total_ms) *total_ms = ms;
    double st[3] = {0, 0, 0};
    k = 2;
    for (int i = 1; i <= nsteps; i++) {
        for (int j = 0; j < 3; j++) {
            LYM_HIP(c, hipEventElapsedTime(&ms, ev[k + j], ev[k + j + 1]));
            st[j] += ms;
        }
        k += 4;
        if (measure_every && i % measure_every == 0) k += 2;
    }
    double so = 0;
    for (size_t at : obs_at) {
        LYM_HIP(c, hipEventElapsedTime(&ms, ev[at], ev[at + 1]));
        so += ms;
    }
    for (int j = 0; j < 3; j++) stage_ms[j] = st[j] / nsteps;
    stage_ms[3] = obs_at.empty() ? 0.0 : so / double(obs_at.size());
    return LYM_OK;
}

int lym_set_schedule(lym_ctx *c, int kind, int target)
{
    if (!c) return LYM_ERR_ARG;
    const int order = kind & 3;
    if ((kind & ~(lym::SCHED_LDS | lym::SCHED_TWO_ROW | 3)) || order > lym::SCHED_COLUMN ||
        (order == lym::SCHED_COLUMN && target < 1))
        return fail(c, LYM_ERR_ARG, "bad schedule");
    if ((kind & lym::SCHED_TWO_ROW) && c->prec != LYM_PREC_MIXED)
        return fail(c, LYM_ERR_ARG, "the two-row form is for the mixed path (FP64 multiplies out all rows)");
    LYM_HIP(c, hipSetDevice(c->device));
    LYM_HIP(c, hipStreamSynchronize(c->stream));
    return apply_schedule(c, kind, target);
}

int lym_autotune(lym_ctx *c, const char *cache_path, char *summary, int summary_len)
{
    if (!c) return LYM_ERR_ARG;
    LYM_HIP(c, hipSetDevice(c->device));
    if (int rc = readable(c)) return rc;
    std::ostringstream key;
    key << "gfx950|single|prec" << c->prec << "|" << c->dims[0] << "x" << c->dims[1] << "x" << c->dims[2] << "x"
        << c->dims[3];
    std::map<std::string, std::string> cache;
    if (cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            auto p = line.find('\t');
            if (p != std::string::npos) cache[line.substr(0, p)] = line.substr(p + 1);
        }
    }
    int best_kind = -1, best_target = 0;
    double best_ms = -1;
    const char *source = "cached";
    auto it = cache.find(key.str());
    if (it == cache.end() || sscanf(it->second.c_str(), "%d %d %lf", &best_kind, &best_target, &best_ms) != 3) {
        source = "measured";
        int rc = ensure_events(c, 2);
        if (rc) return rc;
        const bool mixed = c->prec == LYM_PREC_MIXED;
        // Each candidate is timed over whole RK steps of the flow itself (its
        // three stage launches, in the form the candidate keeps the links in),
        // so that the pick holds for a flow.  The field, three rows in
        // lexicographic order after readable() above, is saved first and put
        // back afterwards.
        const size_t bytes = 2 * real_size(c) * 9 * 4 * size_t(c->g.V);
        struct Save {
            void *p = nullptr;
            ~Save()
            {
                if (p) (void)hipFree(p);
            }
        } save;
        LYM_HIP(c, hipMalloc(&save.p, bytes));
        LYM_HIP(c, hipMemcpyAsync(save.p, c->W[c->cur], bytes, hipMemcpyDeviceToDevice, c->stream));
        const auto cands = lym::schedule_candidates(mixed);
        auto time_one = [&](int i, float *ms) -> int {  // an untimed, then a timed flow step
            if (int rc = apply_schedule(c, cands[size_t(i)].first, cands[size_t(i)].second)) return rc;
            if (int rc = to_form(c, c->rows)) return rc;
            if (int rc = step(c, lym::kTuneEps)) return rc;
            LYM_HIP(c, hipEventRecord(c->events[0], c->stream));
            if (int rc = step(c, lym::kTuneEps)) return rc;
            LYM_HIP(c, hipEventRecord(c->events[1], c->stream));
            LYM_HIP(c, hipEventSynchronize(c->events[1]));
            LYM_HIP(c, hipEventElapsedTime(ms, c->events[0], c->events[1]));
            return LYM_OK;
        };
        int best = 0;
        rc = lym::tune_rounds(int(cands.size()), 3, time_one, &best, &best_ms);
        // the saved field back, whatever the outcome
        const hipError_t e = hipMemcpyAsync(c->W[c->cur], save.p, bytes, hipMemcpyDeviceToDevice, c->stream);
        c->form = 3;
        if (rc) return rc;
        LYM_HIP(c, e);
        LYM_HIP(c, hipStreamSynchronize(c->stream));
        best_kind = cands[size_t(best)].first;
        best_target = cands[size_t(best)].second;
        if (cache_path) {
            std::ofstream o(cache_path, std::ios::app);
            o << key.str() << "\t" << best_kind << " " << best_target << " " << best_ms << "\n";
        }
    }
    int rc = apply_schedule(c, best_kind, best_target);
    if (rc) return rc;
    if (summary && summary_len > 0)
        snprintf(summary, size_t(summary_len), "schedule=%s target=%d step_ms=%.4f source=%s",
                 lym::schedule_name(best_kind), best_target, best_ms, source);
    return LYM_OK;
}

int lym_copy_bandwidth(int device, int64_t bytes, int reps, double *gbs)
{
    if (bytes < 4096 || reps < 1 || !gbs) return LYM_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return LYM_ERR_NODEV;
    if (device < 0 || device >= ndev || hipSetDevice(device) != hipSuccess) return LYM_ERR_ARG;
    const long n = long(bytes / 16);
    double2 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LYM_OK;
    if (hipMalloc(&a, 16 * size_t(n)) != hipSuccess || hipMalloc(&b, 16 * size_t(n)) != hipSuccess)
        rc = LYM_ERR_NOMEM;
    else if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = LYM_ERR_HIP;
    float ms = 0;
    if (rc == LYM_OK) {
        const unsigned nb = unsigned(std::min<long>((n + 255) / 256, 1L << 20));
        lym::k_stream_copy<<<nb, 256>>>(a, b, n);  // first touch
        bool ok = hipEventRecord(e0, nullptr) == hipSuccess;
        for (int r = 0; r < reps; r++) lym::k_stream_copy<<<nb, 256>>>(a, b, n);
        ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(e1, nullptr) == hipSuccess &&
             hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        if (!ok) rc = LYM_ERR_HIP;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (rc == LYM_OK) *gbs = 2.0 * 16.0 * double(n) * reps / (ms * 1e-3) / 1e9;
    return rc;
}

}  // extern "C"
