// Work schedules of the stencil kernels (host side).
//
// A lattice is cut into runs of 64 consecutive sites (x fastest); one
// workgroup handles one run.  Workgroups are dispatched round-robin over the
// 8 XCDs of an MI355X (workgroup b runs on XCD b % 8), and each XCD has its own
// 4 MB L2, so which runs an XCD sees decides how often a link is fetched from
// HBM: every link is read by ~19 neighbouring link updates.  The schedule is
// a table order[xcd * per_xcd + j] = run handled by the j-th workgroup of
// that XCD (-1 = none), launched as 8 * per_xcd workgroups:
//   LINEAR  no table: workgroup b takes run b (each XCD sees every 8th run);
//   CONTIG  the run list split into 8 contiguous parts, one per XCD;
//   COLUMN  runs ordered by (column, t, z, y, x) where a column is a block of
//           Zc x Yc rows of the (z, y) plane of about `target` sites (Yc | Ly,
//           Zc | Lz), so an XCD marches a narrow column along t and the
//           previous and next time slices of the column stay in its L2; then
//           split as CONTIG.
// Any schedule visits every run once, so results do not depend on it.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

namespace lym {

enum SchedKind { SCHED_LINEAR = 0, SCHED_CONTIG = 1, SCHED_COLUMN = 2 };
// flag or-ed into a kind: the stage kernel stages each run's own links in LDS
constexpr int SCHED_LDS = 16;
// flag or-ed into a kind (mixed precision, single lattice): the flow keeps the
// links in two-row form while it runs
constexpr int SCHED_TWO_ROW = 32;
// flag or-ed into a batch kind: the flow steps between two measurements run
// as one persistent launch with grid-wide barriers between RK stages
constexpr int SCHED_WHOLE_FLOW = 64;
constexpr int kXcds = 8;
constexpr int kRunSites = 64;

struct HostSchedule {
    int kind = SCHED_LINEAR, target = 0;
    int per_xcd = 0;         // table row length
    std::vector<int> order;  // kXcds * per_xcd entries (empty for LINEAR)
};

// Column cross-section (yc, zc) in (y, z) rows: divisors of Ly and Lz, so
// that every column is whole, with yc * zc closest to `rows` (in ratio) and,
// between equals, nearest to square (the least edge per site).
inline std::pair<int, int> column_shape(int Ly, int Lz, int rows)
{
    int by = 1, bz = 1;
    double best = 1e300;
    for (int yc = 1; yc <= Ly; yc++)
        for (int zc = 1; zc <= Lz; zc++) {
            if (Ly % yc || Lz % zc) continue;
            const double cost = std::fabs(std::log(double(yc) * zc / rows)) + 0.05 * std::fabs(std::log(double(yc) / zc));
            if (cost < best) {
                best = cost;
                by = yc;
                bz = zc;
            }
        }
    return {by, bz};
}

// runs of one lattice in COLUMN order (run indices relative to the lattice)
inline std::vector<int> column_order(const std::array<int, 4> &L, int target)
{
    const long V = long(L[0]) * L[1] * L[2] * L[3];
    const long nruns = (V + kRunSites - 1) / kRunSites;
    const int rows = std::max(1, target / L[0]);  // (y, z) rows per column
    const auto [yc, zc] = column_shape(L[1], L[2], rows);
    const int ny = L[1] / yc;
    const size_t n = static_cast<size_t>(nruns);
    std::vector<std::pair<uint64_t, int>> key(n);
    for (long r = 0; r < nruns; r++) {
        long s = r * kRunSites;
        const int x = int(s % L[0]);
        s /= L[0];
        const int y = int(s % L[1]);
        s /= L[1];
        const int z = int(s % L[2]);
        const int t = int(s / L[2]);
        const uint64_t col = uint64_t(z / zc) * ny + uint64_t(y / yc);
        const uint64_t k = ((((col * L[3] + t) * L[2] + z) * L[1] + y) * L[0]) + x;
        key[size_t(r)] = {k, int(r)};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> out(n);
    for (long r = 0; r < nruns; r++) out[size_t(r)] = key[size_t(r)].second;
    return out;
}

// Schedule over several lattices whose runs are numbered consecutively
// (lattice c owns runs [run0[c], run0[c] + nruns(c))).  Lattices stay whole
// and in order, so small lattices land on one XCD each.
inline HostSchedule build_schedule(const std::vector<std::array<int, 4>> &dims, const std::vector<long> &run0,
                                   int kind, int target)
{
    HostSchedule h;
    h.kind = kind;
    h.target = target;
    kind &= ~(SCHED_LDS | SCHED_TWO_ROW | SCHED_WHOLE_FLOW);
    if (kind == SCHED_LINEAR) return h;
    std::vector<int> all;
    for (size_t c = 0; c < dims.size(); c++) {
        const long V = long(dims[c][0]) * dims[c][1] * dims[c][2] * dims[c][3];
        const long nruns = (V + kRunSites - 1) / kRunSites;
        if (kind == SCHED_COLUMN) {
            for (int r : column_order(dims[c], target)) all.push_back(int(run0[c] + r));
        } else {
            for (long r = 0; r < nruns; r++) all.push_back(int(run0[c] + r));
        }
    }
    const long n = long(all.size());
    h.per_xcd = int((n + kXcds - 1) / kXcds);
    h.order.assign(size_t(kXcds) * h.per_xcd, -1);
    for (int x = 0; x < kXcds; x++) {
        const long b = n * x / kXcds, e = n * (x + 1) / kXcds;
        for (long i = b; i < e; i++) h.order[size_t(x) * h.per_xcd + size_t(i - b)] = all[size_t(i)];
    }
    return h;
}

// Metropolis blocks (256 consecutive sites of one parity in even-odd order,
// block q = sites [256 q, 256 q + 256) of the parity) in COLUMN order: a
// column is a block of zc x yc rows of the (z, y) plane holding about
// `target` sites of one parity, marched along t; then split into 8
// contiguous parts, one per XCD, as build_schedule.
inline HostSchedule metro_column_schedule(const std::array<int, 4> &L, int target)
{
    const long half_x = L[0] / 2, nsites = long(L[0]) * L[1] * L[2] * L[3] / 2;
    const long nblk = (nsites + 255) / 256;
    const int rows = std::max(1, int(target / half_x));
    const auto [yc, zc] = column_shape(L[1], L[2], rows);
    const int ny = L[1] / yc;
    std::vector<std::pair<uint64_t, int>> key(static_cast<size_t>(nblk));
    for (long q = 0; q < nblk; q++) {
        const long r = q * 256 / half_x;
        const int y = int(r % L[1]), z = int(r / L[1] % L[2]), t = int(r / (long(L[1]) * L[2]));
        const uint64_t col = uint64_t(z / zc) * ny + uint64_t(y / yc);
        key[size_t(q)] = {((col * L[3] + t) * L[2] + z) * uint64_t(nblk) + uint64_t(q), int(q)};
    }
    std::sort(key.begin(), key.end());
    HostSchedule h;
    h.kind = SCHED_COLUMN;
    h.target = target;
    h.per_xcd = int((nblk + kXcds - 1) / kXcds);
    h.order.assign(size_t(kXcds) * h.per_xcd, -1);
    for (int x = 0; x < kXcds; x++) {
        const long b = nblk * x / kXcds, e = nblk * (x + 1) / kXcds;
        for (long i = b; i < e; i++) h.order[size_t(x) * h.per_xcd + size_t(i - b)] = key[size_t(i)].second;
    }
    return h;
}

// The schedule a context takes untuned: columns of 4096 sites with LDS
// staging, in two-row form on the mixed path.
constexpr int kDefaultColumnTarget = 4096;
inline int default_schedule_kind(bool two_row) { return SCHED_COLUMN | SCHED_LDS | (two_row ? SCHED_TWO_ROW : 0); }

// the candidates the autotuners time: every order, with and without LDS
// staging, and (two_row: the mixed path) each of these again with the links
// in two-row form; the default schedule comes first
inline std::vector<std::pair<int, int>> schedule_candidates(bool two_row = false)
{
    const std::pair<int, int> def{default_schedule_kind(two_row), kDefaultColumnTarget};
    std::vector<std::pair<int, int>> c{def};
    const std::vector<std::pair<int, int>> orders{{SCHED_LINEAR, 0},    {SCHED_CONTIG, 0},    {SCHED_COLUMN, 256},
                                                  {SCHED_COLUMN, 512},  {SCHED_COLUMN, 1024}, {SCHED_COLUMN, 2048},
                                                  {SCHED_COLUMN, 4096}, {SCHED_COLUMN, 8192}};
    for (int form : {0, SCHED_TWO_ROW}) {
        if (form && !two_row) break;
        for (int lds : {0, SCHED_LDS})
            for (auto k : orders)
                if (std::make_pair(k.first | lds | form, k.second) != def) c.push_back({k.first | lds | form, k.second});
    }
    return c;
}

// A timed candidate replaces the best so far only when it is faster by more
// than this fraction: candidates within the timing noise of each other (the
// stage times of the column and linear orders often differ by under 1 %) then
// resolve to the one timed first, so that repeated tunings of a shape pick
// the same schedule.
constexpr double kTuneMargin = 0.015;
// Step size of the flow steps the autotuners time: a typical one, so that
// the exponential takes the branches it takes in a flow.
constexpr double kTuneEps = 0.02;
inline bool tune_better(double ms, double best_ms) { return best_ms < 0 || ms < best_ms * (1.0 - kTuneMargin); }

// Picks one of n candidates by timing.  time_one(i, &ms) applies candidate i
// and times one flow step after an untimed one (returns 0 or an error code);
// the untimed step takes the first touch of the schedule table that applying
// a schedule reallocates, which would otherwise weigh on every order but the
// linear one (no table).  After one untimed call for candidate 0 (the device
// warms up), every candidate is timed once per round for `rounds` rounds and
// keeps its fastest time, so that a drift of the clocks over the tuning does
// not favour the candidates timed last; then the first candidate in order is
// kept unless a later one is faster by more than kTuneMargin.
template <typename F> int tune_rounds(int n, int rounds, F &&time_one, int *best, double *best_ms)
{
    std::vector<float> t(size_t(n), 1e30f);
    float ms = 0;
    if (int rc = time_one(0, &ms)) return rc;
    for (int r = 0; r < rounds; r++)
        for (int i = 0; i < n; i++) {
            if (int rc = time_one(i, &ms)) return rc;
            t[size_t(i)] = std::min(t[size_t(i)], ms);
        }
    *best = 0;
    *best_ms = t[0];
    for (int i = 1; i < n; i++)
        if (tune_better(t[size_t(i)], *best_ms)) {
            *best = i;
            *best_ms = t[size_t(i)];
        }
    return 0;
}

inline const char *schedule_name(int kind)
{
    if (kind & SCHED_TWO_ROW) {
        switch (kind & ~SCHED_TWO_ROW) {
        case SCHED_LINEAR: return "linear+2row";
        case SCHED_CONTIG: return "contig+2row";
        case SCHED_COLUMN: return "column+2row";
        case SCHED_LINEAR | SCHED_LDS: return "linear+lds+2row";
        case SCHED_CONTIG | SCHED_LDS: return "contig+lds+2row";
        default: return "column+lds+2row";
        }
    }
    switch (kind) {
    case SCHED_LINEAR: return "linear";
    case SCHED_CONTIG: return "contig";
    case SCHED_COLUMN: return "column";
    case SCHED_LINEAR | SCHED_LDS: return "linear+lds";
    case SCHED_CONTIG | SCHED_LDS: return "contig+lds";
    default: return "column+lds";
    }
}

}  // namespace lym
