// Host-side launchers of the heavy kernel families: the RK stage kernels of
// the single-lattice and batch paths and the Metropolis sub-sweep kernels.
// Each family is instantiated for one link precision per translation unit
// (lym_stage_*.hip: single lattice, lym_bstage_*.hip: batch, the *_f32r2
// units their two-row forms, lym_metro_*.hip: Metropolis; f64 and f32), so the device code of the
// library compiles in parallel; every other translation unit sees the extern
// declarations below.
#pragma once
#include "lym_batch_kernels.hpp"

namespace lym {

// One RK stage over one lattice: nb workgroups of 256 threads on stream st;
// lds = the flow_block staging mode (0, 1, 2, 3 or 5).
template <typename S, int STAGE>
void launch_flow_stage(int lds, unsigned nb, hipStream_t st, const F2<S> *win, F2<S> *wout, S *acc, Geom g, Sched sc,
                       double eps)
{
    constexpr bool re = sizeof(S) == 4;  // mixed path: reunitarise before the store
    switch (lds) {
    case 0: k_flow_stage<S, STAGE, re, 0><<<nb, 256, 0, st>>>(win, wout, acc, g, sc, eps); break;
    case 1: k_flow_stage<S, STAGE, re, 1><<<nb, 256, 0, st>>>(win, wout, acc, g, sc, eps); break;
    case 2: k_flow_stage<S, STAGE, re, 2><<<nb, 256, 0, st>>>(win, wout, acc, g, sc, eps); break;
    case 3: k_flow_stage<S, STAGE, re, 3><<<nb, 256, 0, st>>>(win, wout, acc, g, sc, eps); break;
    default: k_flow_stage<S, STAGE, re, 5><<<nb, 256, 0, st>>>(win, wout, acc, g, sc, eps); break;
    }
}

// One RK stage of the mixed path over a field in two-row form (ROWS = 2);
// lds as launch_flow_stage.
template <int STAGE>
void launch_flow_stage_r2(int lds, unsigned nb, hipStream_t st, const F2<float> *win, F2<float> *wout, float *acc,
                          Geom g, Sched sc, double eps)
{
    switch (lds) {
    case 0: k_flow_stage<float, STAGE, true, 0, 2><<<nb, 256, 0, st>>>(win, wout, acc, g, sc, eps); break;
    case 1: k_flow_stage<float, STAGE, true, 1, 2><<<nb, 256, 0, st>>>(win, wout, acc, g, sc, eps); break;
    case 2: k_flow_stage<float, STAGE, true, 2, 2><<<nb, 256, 0, st>>>(win, wout, acc, g, sc, eps); break;
    case 3: k_flow_stage<float, STAGE, true, 3, 2><<<nb, 256, 0, st>>>(win, wout, acc, g, sc, eps); break;
    default: k_flow_stage<float, STAGE, true, 5, 2><<<nb, 256, 0, st>>>(win, wout, acc, g, sc, eps); break;
    }
}

// One RK stage over every lattice of a batch: one workgroup per schedule slot
// when grid covers them, else a persistent grid of that many workgroups
// striding over the slots (k_batch_flow_stage<..., PERSIST>); lds = 0, 1, 2, 3
// or 4 (per-lattice slab or own links).
template <typename S, int STAGE, bool REUNIT, int LDS, int ROWS>
void launch_batch_stage_lds(unsigned grid, hipStream_t st, const F2<S> *win, F2<S> *wout, S *acc, const ConfDesc *desc,
                            int nconf, long ntiles, Sched sc, double eps)
{
    const long slots = sc.order ? long(kXcds) * sc.per_xcd : ntiles;
    if (long(grid) < slots)
        k_batch_flow_stage<S, STAGE, REUNIT, LDS, ROWS, true><<<grid, 256, 0, st>>>(win, wout, acc, desc, nconf, ntiles,
                                                                                   sc, eps);
    else
        k_batch_flow_stage<S, STAGE, REUNIT, LDS, ROWS, false><<<unsigned(slots), 256, 0, st>>>(win, wout, acc, desc,
                                                                                               nconf, ntiles, sc, eps);
}

template <typename S, int STAGE, int ROWS>
void launch_batch_stage_rows(int lds, unsigned grid, hipStream_t st, const F2<S> *win, F2<S> *wout, S *acc,
                             const ConfDesc *desc, int nconf, long ntiles, Sched sc, double eps)
{
    constexpr bool re = sizeof(S) == 4;  // mixed path: reunitarise before the store
    switch (lds) {
    case 0: launch_batch_stage_lds<S, STAGE, re, 0, ROWS>(grid, st, win, wout, acc, desc, nconf, ntiles, sc, eps); break;
    case 1: launch_batch_stage_lds<S, STAGE, re, 1, ROWS>(grid, st, win, wout, acc, desc, nconf, ntiles, sc, eps); break;
    case 2: launch_batch_stage_lds<S, STAGE, re, 2, ROWS>(grid, st, win, wout, acc, desc, nconf, ntiles, sc, eps); break;
    case 3: launch_batch_stage_lds<S, STAGE, re, 3, ROWS>(grid, st, win, wout, acc, desc, nconf, ntiles, sc, eps); break;
    default: launch_batch_stage_lds<S, STAGE, re, 4, ROWS>(grid, st, win, wout, acc, desc, nconf, ntiles, sc, eps); break;
    }
}

template <typename S, int STAGE>
void launch_batch_stage(int lds, unsigned grid, hipStream_t st, const F2<S> *win, F2<S> *wout, S *acc,
                        const ConfDesc *desc, int nconf, long ntiles, Sched sc, double eps)
{
    launch_batch_stage_rows<S, STAGE, 3>(lds, grid, st, win, wout, acc, desc, nconf, ntiles, sc, eps);
}

// The mixed batch stage over fields in two-row form (ROWS = 2; lym_bstage_f32r2.hip).
template <int STAGE>
void launch_batch_stage_r2(int lds, unsigned grid, hipStream_t st, const F2<float> *win, F2<float> *wout, float *acc,
                           const ConfDesc *desc, int nconf, long ntiles, Sched sc, double eps)
{
    launch_batch_stage_rows<float, STAGE, 2>(lds, grid, st, win, wout, acc, desc, nconf, ntiles, sc, eps);
}

// The whole-flow persistent launch over a batch (k_batch_flow_persistent):
// nsteps RK3 steps of every lattice by grid workgroups that are all resident
// at once, in ngroups barrier sync groups; lds as launch_batch_stage, ROWS
// the form of the fields.
template <typename S, int ROWS>
void launch_batch_flow_persistent(int lds, unsigned grid, hipStream_t st, F2<S> *w0, F2<S> *w1, S *acc,
                                  const ConfDesc *desc, int nconf, long ntiles, Sched sc, double eps, int nsteps,
                                  FlowSync *sync, int ngroups, unsigned long long limit)
{
    constexpr bool re = sizeof(S) == 4;  // mixed path: reunitarise before the store
#define LYM_PERSIST_CASE(L)                                                                                   \
    k_batch_flow_persistent<S, re, L, ROWS><<<grid, 256, 0, st>>>(w0, w1, acc, desc, nconf, ntiles, sc, eps, \
                                                                   nsteps, sync, ngroups, limit)
    switch (lds) {
    case 0: LYM_PERSIST_CASE(0); break;
    case 1: LYM_PERSIST_CASE(1); break;
    case 2: LYM_PERSIST_CASE(2); break;
    case 3: LYM_PERSIST_CASE(3); break;
    default: LYM_PERSIST_CASE(4); break;
    }
#undef LYM_PERSIST_CASE
}
// Workgroups of that kernel one CU holds at once (its registers and LDS).
template <typename S, int ROWS> hipError_t batch_flow_persistent_per_cu(int lds, int *per_cu)
{
    constexpr bool re = sizeof(S) == 4;
    switch (lds) {
    case 0: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_batch_flow_persistent<S, re, 0, ROWS>, 256, 0);
    case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_batch_flow_persistent<S, re, 1, ROWS>, 256, 0);
    case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_batch_flow_persistent<S, re, 2, ROWS>, 256, 0);
    case 3: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_batch_flow_persistent<S, re, 3, ROWS>, 256, 0);
    default: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_batch_flow_persistent<S, re, 4, ROWS>, 256, 0);
    }
}
// Candidate kind of a Metropolis launch: 0 = R S T product, 1 = exp(i eps H)
// with the general exponential, 2 = exp(i eps H) with 8 eps^2 <= kSeriesC1Max
// (series-only exponential).
inline int metro_kernel_kind(int candidate_is_exph, double eps)
{
    if (!candidate_is_exph) return 0;
    return 8.0 * eps * eps > kSeriesC1Max ? 1 : 2;
}

// One Metropolis sub-sweep over one lattice in even-odd order (k_metropolis).
template <typename S>
void launch_metropolis(int kind, bool predrawn, unsigned per_xcd, hipStream_t st, F2<S> *f, Geom g, int mu,
                       const MetroParams &p, const int *order, unsigned long long *accepted)
{
    const unsigned nb = 8 * per_xcd;
    if (kind == 0) {
        if (predrawn) k_metropolis<S, 0, true><<<nb, 256, 0, st>>>(f, g, mu, p, order, per_xcd, accepted);
        else k_metropolis<S, 0, false><<<nb, 256, 0, st>>>(f, g, mu, p, order, per_xcd, accepted);
    } else if (kind == 1) {
        if (predrawn) k_metropolis<S, 1, true><<<nb, 256, 0, st>>>(f, g, mu, p, order, per_xcd, accepted);
        else k_metropolis<S, 1, false><<<nb, 256, 0, st>>>(f, g, mu, p, order, per_xcd, accepted);
    } else {
        if (predrawn) k_metropolis<S, 2, true><<<nb, 256, 0, st>>>(f, g, mu, p, order, per_xcd, accepted);
        else k_metropolis<S, 2, false><<<nb, 256, 0, st>>>(f, g, mu, p, order, per_xcd, accepted);
    }
}

// One Metropolis sub-sweep over every lattice of a batch (k_batch_metropolis).
template <typename S>
void launch_batch_metropolis(int kind, unsigned per_xcd, hipStream_t st, F2<S> *f, const ConfDesc *desc, int nconf,
                             long nitems, int mu, const MetroParams &p, unsigned long long *accepted)
{
    const unsigned nb = 8 * per_xcd;
    if (kind == 0) k_batch_metropolis<S, 0><<<nb, 256, 0, st>>>(f, desc, nconf, nitems, mu, p, per_xcd, accepted);
    else if (kind == 1) k_batch_metropolis<S, 1><<<nb, 256, 0, st>>>(f, desc, nconf, nitems, mu, p, per_xcd, accepted);
    else k_batch_metropolis<S, 2><<<nb, 256, 0, st>>>(f, desc, nconf, nitems, mu, p, per_xcd, accepted);
}

#define LYM_STAGE_EXTERN(S, STAGE)                                                                                 \
    extern template void launch_flow_stage<S, STAGE>(int, unsigned, hipStream_t, const F2<S> *, F2<S> *, S *, Geom, \
                                                     Sched, double);                                               \
    extern template void launch_batch_stage<S, STAGE>(int, unsigned, hipStream_t, const F2<S> *, F2<S> *, S *,    \
                                                      const ConfDesc *, int, long, Sched, double);
#define LYM_METRO_EXTERN(S)                                                                                        \
    extern template void launch_metropolis<S>(int, bool, unsigned, hipStream_t, F2<S> *, Geom, int,               \
                                              const MetroParams &, const int *, unsigned long long *);             \
    extern template void launch_batch_metropolis<S>(int, unsigned, hipStream_t, F2<S> *, const ConfDesc *, int,  \
                                                    long, int, const MetroParams &, unsigned long long *);

#ifndef LYM_INSTANTIATE_STAGES
LYM_STAGE_EXTERN(double, 1)
LYM_STAGE_EXTERN(double, 2)
LYM_STAGE_EXTERN(double, 3)
LYM_STAGE_EXTERN(float, 1)
LYM_STAGE_EXTERN(float, 2)
LYM_STAGE_EXTERN(float, 3)
#endif
#ifndef LYM_INSTANTIATE_STAGES_R2
extern template void launch_flow_stage_r2<1>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *, Geom,
                                             Sched, double);
extern template void launch_flow_stage_r2<2>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *, Geom,
                                             Sched, double);
extern template void launch_flow_stage_r2<3>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *, Geom,
                                             Sched, double);
#endif
#ifndef LYM_INSTANTIATE_BSTAGES_R2
extern template void launch_batch_stage_r2<1>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *,
                                              const ConfDesc *, int, long, Sched, double);
extern template void launch_batch_stage_r2<2>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *,
                                              const ConfDesc *, int, long, Sched, double);
extern template void launch_batch_stage_r2<3>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *,
                                              const ConfDesc *, int, long, Sched, double);
#endif
#ifndef LYM_INSTANTIATE_BFLOW
#define LYM_BFLOW_EXTERN(S, ROWS)                                                                                   \
    extern template void launch_batch_flow_persistent<S, ROWS>(int, unsigned, hipStream_t, F2<S> *, F2<S> *, S *,   \
                                                               const ConfDesc *, int, long, Sched, double, int,     \
                                                               FlowSync *, int, unsigned long long);                \
    extern template hipError_t batch_flow_persistent_per_cu<S, ROWS>(int, int *);
LYM_BFLOW_EXTERN(double, 3)
LYM_BFLOW_EXTERN(float, 3)
LYM_BFLOW_EXTERN(float, 2)
#undef LYM_BFLOW_EXTERN
#endif
#ifndef LYM_INSTANTIATE_METRO
LYM_METRO_EXTERN(double)
LYM_METRO_EXTERN(float)
#endif

}  // namespace lym
