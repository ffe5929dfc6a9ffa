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
                       const ConfDesc *__restrict__ desc, int nconf, Sched sc, double eps, F2<S> *lds)
{
    const long tile = sched_run(sc, unsigned(k));  // uniform over the workgroup
    if (tile < 0) return;
    const int c = find_conf(desc, nconf, tile);
    const ConfDesc d = desc[c];
    const long site0 = (tile - d.tile0) * kSitesPerBlock, off = field_off<ROWS>(d);
    if constexpr (LDS == 4) {
        if (d.slab)
            flow_block<S, STAGE, REUNIT, 3, ROWS>(win + off, wout + off, acc + d.acc_off, d.g, site0, eps, lds);
        else
            flow_block<S, STAGE, REUNIT, 1, ROWS>(win + off, wout + off, acc + d.acc_off, d.g, site0, eps, lds);
    } else {
        flow_block<S, STAGE, REUNIT, LDS, ROWS>(win + off, wout + off, acc + d.acc_off, d.g, site0, eps, lds);
    }
}

// PERSIST = false: one workgroup per schedule slot.  PERSIST = true: a
// persistent grid (fewer workgroups, a multiple of 8 so that a workgroup keeps
// its XCD) strides over the slots.  The loop keeps values live across
// iterations and costs the mixed kernel some 50 registers (3 instead of 5
// waves per SIMD), so the one-slot form is a kernel of its own.
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
