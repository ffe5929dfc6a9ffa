// RK stage kernels of the batched ensemble path with float links (the
// launchers of lym_launch.hpp, instantiated here).
#define LYM_INSTANTIATE_STAGES
#include "lym_launch.hpp"

namespace lym {
template void launch_batch_stage<float, 1>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *, const ConfDesc *,
                                        int, long, Sched, double);
template void launch_batch_stage<float, 2>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *, const ConfDesc *,
                                        int, long, Sched, double);
template void launch_batch_stage<float, 3>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *, const ConfDesc *,
                                        int, long, Sched, double);
}  // namespace lym
