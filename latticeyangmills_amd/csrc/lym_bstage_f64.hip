// RK stage kernels of the batched ensemble path with double links (the
// launchers of lym_launch.hpp, instantiated here).
#define LYM_INSTANTIATE_STAGES
#include "lym_launch.hpp"

namespace lym {
template void launch_batch_stage<double, 1>(int, unsigned, hipStream_t, const F2<double> *, F2<double> *, double *, const ConfDesc *,
                                        int, long, Sched, double);
template void launch_batch_stage<double, 2>(int, unsigned, hipStream_t, const F2<double> *, F2<double> *, double *, const ConfDesc *,
                                        int, long, Sched, double);
template void launch_batch_stage<double, 3>(int, unsigned, hipStream_t, const F2<double> *, F2<double> *, double *, const ConfDesc *,
                                        int, long, Sched, double);
}  // namespace lym
