// Kernels of the batched ensemble path (many independent lattices of uneven
// sizes in one launch); the host side is lym_batch.hip.
//
// Work list: every lattice c is cut into tiles of 64 consecutive sites; tile
// t of the whole batch belongs to the lattice whose first tile index
// tile0[c] <= t < tile0[c+1].  A workgroup (4 waves = 4 directions) takes a
// tile at a time, following the XCD-aware schedule table of lym_sched.hpp
// (lattices stay whole, so small lattices each live in one XCD's L2).
#pragma once
#include "lym_kernels.hpp"
#include "lym_sched.hpp"

namespace lym {

struct ConfDesc {
    Geom g;
    long link_off;  // offset in complexes of the lattice's field in the batch buffers
    long acc_off;   // offset in reals of its RK accumulator
    long tile0;     // first tile of the lattice
    long item0;     // first Metropolis work item of the lattice: V/2 sites per sub-sweep,
                    // rounded up to whole waves (64) so that no wave spans two lattices
    uint64_t seed;
    int slab;       // the lattice fits the slab staging (slab_fits)
};

constexpr int kMaxConf = 1024;

LYM_DI int find_conf(const ConfDesc *__restrict__ d, int n, long tile)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (d[mid].tile0 <= tile)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// One schedule slot of a batch RK stage: the tile it holds (none for padding
// entries), its lattice, and the lattice's staging mode.  LDS as flow_block,
// and 4: per lattice, slab staging where the lattice fits it and the run's
// own links elsewhere (the lattice of a tile, and so the choice, is uniform
// over the workgroup); both modes stage into one array.  ROWS = 2: the fields
// are in two-row form, 6 complexes per link, so a lattice's field starts at
// link_off / 9 * 6.
template <int ROWS> LYM_DI long field_off(const ConfDesc &d) { return ROWS == 2 ? d.link_off / 9 * 6 : d.link_off; }

template <typename S, int STAGE, bool REUNIT, int LDS, int ROWS>
LYM_DI void batch_slot(long k, const F2<S> *__restrict__ win, F2<S> *__restrict__ wout, S *__restrict__ acc,
                       const ConfDesc *__restrict__ desc, int nconf, Sched sc, double eps, F2<S> *lds,
                       unsigned tid = threadIdx.x)
{
    const long tile = sched_run(sc, unsigned(k));  // uniform over the workgroup
    if (tile < 0) return;
    const int c = find_conf(desc, nconf, tile);
    const ConfDesc d = desc[c];
    const long site0 = (tile - d.tile0) * kSitesPerBlock, off = field_off<ROWS>(d);
    if constexpr (LDS == 4) {
        if (d.slab)
            flow_block<S, STAGE, REUNIT, 3, ROWS>(win + off, wout + off, acc + d.acc_off, d.g, site0, eps, lds, tid);
        else
            flow_block<S, STAGE, REUNIT, 1, ROWS>(win + off, wout + off, acc + d.acc_off, d.g, site0, eps, lds, tid);
    } else {
        flow_block<S, STAGE, REUNIT, LDS, ROWS>(win + off, wout + off, acc + d.acc_off, d.g, site0, eps, lds, tid);
    }
}

// PERSIST = false: one workgroup per schedule slot.  PERSIST = true: a
// persistent grid (fewer workgroups, a multiple of 8 so that a workgroup keeps
// its XCD) strides over the slots.  The one-slot form is a kernel of its own,
// so that no loop state weighs on its registers; the units of these kernels
// are built without machine-level loop-invariant code motion, which held the
// stage body's constants across the grid-stride loop (145-156 VGPRs instead
// of 93-95).
template <typename S, int STAGE, bool REUNIT, int LDS, int ROWS = 3, bool PERSIST = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PERSIST ? 1 : stage_min_waves<S, LDS>())))
void k_batch_flow_stage(const F2<S> *__restrict__ win, F2<S> *__restrict__ wout, S *__restrict__ acc,
                        const ConfDesc *__restrict__ desc, int nconf, long ntiles, Sched sc, double eps)
{
    const long slots = sc.order ? long(kXcds) * sc.per_xcd : ntiles;
    __shared__ F2<S> lds[stage_lds_elems<LDS == 4 ? 3 : LDS>()];
    if constexpr (PERSIST) {
        for (long k = blockIdx.x; k < slots; k += gridDim.x)
            batch_slot<S, STAGE, REUNIT, LDS, ROWS>(k, win, wout, acc, desc, nconf, sc, eps, lds);
    } else {
        if (long(blockIdx.x) < slots)
            batch_slot<S, STAGE, REUNIT, LDS, ROWS>(blockIdx.x, win, wout, acc, desc, nconf, sc, eps, lds);
    }
}

// ------------------------------------------- persistent whole-flow launch ---

// State of a whole-flow launch, in device memory.  Every counter sits on a
// 256-byte line of its own: atomics on one address are served one after
// another (about 0.13 us each, measured through a one-counter barrier of
// 1024 workgroups), so the barrier and the slot queues spread theirs.
//   group[g], arrivals: arrival counts of the barrier's sync groups and of
//     the groups that completed;
//   release[g]: the number of the last barrier that group g may leave;
//   next[q]: schedule slots handed out by slot queue q in this launch;
//   abort: raised by a workgroup that gave up waiting.
// k_flow_sync_reset zeroes everything but abort before each launch, and
// abort too once per flow call.
constexpr int kSyncLine = 64;         // unsigneds per 256-byte line
constexpr int kMaxSyncGroups = 128;   // sync groups of a barrier at most
constexpr int kQueuesPerXcd = 8;      // slot queues of one XCD
constexpr int kSlotQueues = kXcds * kQueuesPerXcd;
struct FlowSync {
    unsigned arrivals;
    int abort;
    unsigned pad[kSyncLine - 2];
    unsigned group[kMaxSyncGroups * kSyncLine];
    unsigned release[kMaxSyncGroups * kSyncLine];
    unsigned next[kSlotQueues * kSyncLine];
};

static __global__ __launch_bounds__(256) void k_flow_sync_reset(FlowSync *s, int clear_abort)
{
    const int i = threadIdx.x;
    if (i == 0) s->arrivals = 0u;
    if (i == 0 && clear_abort) s->abort = 0;
    if (i < kMaxSyncGroups) {
        s->group[i * kSyncLine] = 0u;
        s->release[i * kSyncLine] = 0u;
    }
    if (i < kSlotQueues) s->next[i * kSyncLine] = 0u;
}

// Whether no workgroup has given up, the same answer for the whole workgroup.
LYM_DI bool flow_sync_ok(FlowSync *sync)
{
    __shared__ int ok;
    if (threadIdx.x == 0) ok = !__hip_atomic_load(&sync->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const bool r = ok != 0;
    __syncthreads();
    return r;
}

// Grid-wide barrier of the whole-flow launch, the n-th of the launch.  Every
// workgroup of the launch is resident at once (the host sizes the grid from
// the kernel's occupancy), so each may wait for all the others.  Workgroup b
// belongs to sync group b % ngroups (ngroups <= kMaxSyncGroups, about the
// square root of the grid): it arrives on its group's counter; the group's
// last arrival arrives on the count of completed groups; the last group's
// last arrival sets every group's release flag to n; each workgroup waits for
// its own group's flag.  Memory order: the workgroup's barrier orders its
// waves' stores before thread 0's agent-scope fence and arrival; a group's
// last arrival fences again before it passes the group on, and the final one
// before it releases; a waiter fences after it sees its flag, then the
// workgroup's barrier orders that before every wave's next reads.  So the
// next RK stage reads the links the other workgroups (on every XCD) wrote.
// (Fences in every wave instead of one per workgroup cost the 64-lattice
// ensemble 17 %.)  The wait is bounded: past `limit` ticks of the
// constant-rate wall clock, after kMaxSyncPolls polls, or as soon as another
// workgroup has given up, it raises the abort flag and returns false for the
// whole workgroup; the caller then returns, so the launch always ends.
constexpr unsigned kMaxSyncPolls = 1u << 26;
LYM_DI bool grid_sync(FlowSync *sync, unsigned n, int ngroups, unsigned long long limit)
{
    __shared__ int go;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned g = blockIdx.x % unsigned(ngroups);
        const unsigned members = (gridDim.x - g + unsigned(ngroups) - 1) / unsigned(ngroups);
        unsigned *flag = &sync->release[g * kSyncLine];
        __threadfence();
        if (__hip_atomic_fetch_add(&sync->group[g * kSyncLine], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 ==
            n * members) {
            __threadfence();
            if (__hip_atomic_fetch_add(&sync->arrivals, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 ==
                n * unsigned(ngroups)) {
                __threadfence();
                for (int h = 0; h < ngroups; h++)
                    __hip_atomic_store(&sync->release[h * kSyncLine], n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        const unsigned long long t0 = (unsigned long long)wall_clock64();
        int ok = 1;
        unsigned polls = 0;
        while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < n) {
            if (__hip_atomic_load(&sync->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                (unsigned long long)wall_clock64() - t0 > limit || ++polls > kMaxSyncPolls) {
                __hip_atomic_store(&sync->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();
        go = ok;
    }
    __syncthreads();
    return go != 0;
}

// Waves per SIMD the registers of the whole-flow kernel must allow: five for
// the FP32 slab modes (3, 4), as the one-slot stage kernel of mode 3 (one
// VGPR spills, against 14-50 before the LDS reads of the stage were kept in
// program order); the others take what they need.
template <typename S, int LDS> constexpr int persist_min_waves()
{
    return sizeof(S) == 4 && (LDS == 3 || LDS == 4) ? 5 : 1;
}

// One RK stage of the whole-flow launch: the schedule slots of XCD x (slot
// 8 e + x for entry e of its row, the slots the one-launch-per-stage kernels
// give that XCD) are split into kQueuesPerXcd contiguous ranges, one per slot
// queue of the XCD, and workgroup b (on XCD x = b % 8) takes slots one at a
// time from queue x + 8 ((b / 8) % kQueuesPerXcd), so that a workgroup that
// finishes early takes more, as the hardware dispatcher does between
// launches.  Every workgroup of a queue takes one value past its range per
// stage, so stage j of a launch starts the queue's counter at
// j (range + users).  Thread 0 takes the next value while the workgroup works
// on the current slot (two LDS words alternate, so one workgroup barrier per
// slot suffices).
template <typename S, int STAGE, bool REUNIT, int LDS, int ROWS>
LYM_DI void whole_flow_stage(unsigned j, const F2<S> *win, F2<S> *wout, S *__restrict__ acc,
                             const ConfDesc *__restrict__ desc, int nconf, long slots, long per_xcd, Sched sc,
                             double eps, F2<S> *lds, FlowSync *sync, unsigned tid)
{
    __shared__ unsigned taken[2];
    const unsigned x = blockIdx.x % kXcds, wx = gridDim.x / kXcds;  // workgroups per XCD
    const unsigned nq = wx < unsigned(kQueuesPerXcd) ? wx : unsigned(kQueuesPerXcd);  // queues in use
    const unsigned qi = (blockIdx.x / kXcds) % nq;
    const unsigned users = (wx - qi + nq - 1) / nq;                // workgroups on this queue
    const long e0 = per_xcd * long(qi) / long(nq), e1 = per_xcd * long(qi + 1) / long(nq);  // its entries
    unsigned *counter = &sync->next[(x + kXcds * qi) * kSyncLine];
    const unsigned base = j * unsigned(e1 - e0 + users);
    unsigned pending = 0;
    if (threadIdx.x == 0) pending = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int i = 0;; i ^= 1) {
        if (threadIdx.x == 0) taken[i] = pending;
        __syncthreads();
        const long e = e0 + long(taken[i]) - long(base);
        if (e >= e1) break;
        if (threadIdx.x == 0)
            pending = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const long k = kXcds * e + x;
        if (k < slots) batch_slot<S, STAGE, REUNIT, LDS, ROWS>(k, win, wout, acc, desc, nconf, sc, eps, lds, tid);
    }
}

// nsteps whole RK3 steps of every lattice of the batch in one launch: the
// three stages of each step as in k_batch_flow_stage, with the slots handed
// out by whole_flow_stage and a grid-wide barrier between consecutive stages;
// fields alternate w0 -> w1 -> w0 -> w1 within a step and swap roles after
// it (the result of nsteps steps is in w0 for nsteps even, w1 for odd).  Each
// link is computed by the same code as in the per-stage launches, so the two
// give the same bits whichever workgroup takes a slot.  gridDim.x is a
// multiple of 8; ngroups, the barrier's sync groups, is 1 .. kMaxSyncGroups
// and at most gridDim.x.
template <typename S, bool REUNIT, int LDS, int ROWS = 3>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(persist_min_waves<S, LDS>())))
void k_batch_flow_persistent(F2<S> *w0, F2<S> *w1, S *__restrict__ acc, const ConfDesc *__restrict__ desc, int nconf,
                             long ntiles, Sched sc, double eps, int nsteps, FlowSync *sync, int ngroups,
                             unsigned long long limit)
{
    __shared__ F2<S> lds[stage_lds_elems<LDS == 4 ? 3 : LDS>()];
    const long slots = sc.order ? long(kXcds) * sc.per_xcd : ntiles;
    const long per_xcd = (slots + kXcds - 1) / kXcds;
    if (!flow_sync_ok(sync)) return;  // an earlier launch of this flow gave up
    F2<S> *x = w0, *y = w1;
    unsigned n = 0;  // stages done = grid barriers passed
    // the thread index through an opaque copy per stage: what the stage body
    // derives from it is formed there instead of being held across the loops
    auto tid = []() {
        unsigned t = threadIdx.x;
        asm volatile("" : "+v"(t));
        return t;
    };
    for (int t = 0; t < nsteps; t++) {
        whole_flow_stage<S, 1, REUNIT, LDS, ROWS>(n, x, y, acc, desc, nconf, slots, per_xcd, sc, eps, lds, sync, tid());
        if (!grid_sync(sync, ++n, ngroups, limit)) return;
        whole_flow_stage<S, 2, REUNIT, LDS, ROWS>(n, y, x, acc, desc, nconf, slots, per_xcd, sc, eps, lds, sync, tid());
        if (!grid_sync(sync, ++n, ngroups, limit)) return;
        whole_flow_stage<S, 3, REUNIT, LDS, ROWS>(n, x, y, acc, desc, nconf, slots, per_xcd, sc, eps, lds, sync, tid());
        if (t + 1 < nsteps && !grid_sync(sync, ++n, ngroups, limit)) return;
        F2<S> *const z = x;
        x = y;
        y = z;
    }
}

// observables: one partial per tile (192 threads = 64 sites x 3 plane pairs)
template <typename S, int ROWS = 3>
__global__ __launch_bounds__(192) void k_batch_observables(const F2<S> *__restrict__ f, const ConfDesc *__restrict__ desc,
                                                           int nconf, long ntiles, ksum *__restrict__ partials)
{
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int c = find_conf(desc, nconf, tile);
        const ConfDesc d = desc[c];
        const long s = (tile - d.tile0) * kObsSites + (threadIdx.x & 63);
        const int pair = threadIdx.x >> 6;
        ksum v[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
        if (s < d.g.V) {
            double o[3];
            const F2<S> *fc = f + field_off<ROWS>(d);
            if (pair == 0)
                pair_observables<S, S, 0, ROWS>(fc, d.g, s, o);
            else if (pair == 1)
                pair_observables<S, S, 1, ROWS>(fc, d.g, s, o);
            else
                pair_observables<S, S, 2, ROWS>(fc, d.g, s, o);
            v[0].s = o[0];
            v[1].s = o[1];
            v[2].s = o[2];
            v[3].s = fabs(o[2]);
        }
        block_ksum4<192>(v, partials + 4 * tile);
        __syncthreads();
    }
}

// per-lattice sum of its tiles' partials; one workgroup per lattice
static __global__ __launch_bounds__(256) void k_batch_reduce(const ksum *__restrict__ partials, const ConfDesc *__restrict__ desc,
                                                      int nconf, long ntiles, double *__restrict__ dst)
{
    const int c = blockIdx.x;
    const long t0 = desc[c].tile0, t1 = (c + 1 < nconf) ? desc[c + 1].tile0 : ntiles;
    ksum v[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (long i = t0 + threadIdx.x; i < t1; i += 256)
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = kadd(v[k], partials[4 * i + k]);
    __shared__ ksum tot[4];
    block_ksum4<256>(v, tot);
    __syncthreads();
    if (threadIdx.x < 4) dst[4 * c + threadIdx.x] = tot[threadIdx.x].s + tot[threadIdx.x].c;
}

// k_reorder_eo over the batch: thread j of plane blockIdx.y handles even-odd
// position j - site0 of lattice c, whose V sites start at site0 = link_off / 36
// in the batch's site numbering.
template <typename S, bool TO_EO>
__global__ __launch_bounds__(256) void k_batch_reorder_eo(const F2<S> *__restrict__ in, F2<S> *__restrict__ out,
                                                          const ConfDesc *__restrict__ desc, int nconf, long nsites)
{
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= nsites) return;
    int lo = 0, hi = nconf - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (desc[mid].link_off <= 36 * j)
            lo = mid;
        else
            hi = mid - 1;
    }
    const ConfDesc d = desc[lo];
    const long jl = j - d.link_off / 36, half = d.g.V / 2;
    const int par = jl >= half;
    const long s = parity_site(d.g, jl - par * half, par);
    const long base = d.link_off + (long)blockIdx.y * d.g.V;
    if constexpr (TO_EO)
        out[base + jl] = in[base + s];
    else
        out[base + s] = in[base + jl];
}

// Checkerboard Metropolis over the batch, links in even-odd order: work
// item i of sub-sweep (mu, parity) is half-site i - item0[c] of lattice c.
// Workgroup b takes block (b % 8) * per_xcd + b / 8, so each XCD sweeps a
// contiguous range (as k_metropolis).  A wave never spans two lattices, so
// the lattice is found from the wave's first item and its descriptor sits in
// scalar registers; acceptances are added with one atomic per wave.
template <typename S, int CAND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CAND == 1 ? 1 : 2))) void k_batch_metropolis(
    F2<S> *__restrict__ f, const ConfDesc *__restrict__ desc, int nconf, long nitems, int mu, MetroParams p,
    unsigned per_xcd, unsigned long long *__restrict__ accepted)
{
    const unsigned blk = (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3);
    const long iw = (long)blk * 256 + long(__builtin_amdgcn_readfirstlane(threadIdx.x & ~63u));
    if (iw >= nitems) return;
    int lo = 0, hi = nconf - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (desc[mid].item0 <= iw)
            lo = mid;
        else
            hi = mid - 1;
    }
    const ConfDesc d = desc[lo];
    MetroParams q = p;
    q.seed = d.seed;
    const long il = iw - d.item0 + (threadIdx.x & 63);
    int acc = 0;
    if (il < d.g.V / 2)
        acc = metro_link_mu<S, CAND, false>(f + d.link_off, d.g, mu, parity_site(d.g, il, p.parity), il, q);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0 && acc) atomicAdd(accepted + lo, (unsigned long long)acc);
}

}  // namespace lym
