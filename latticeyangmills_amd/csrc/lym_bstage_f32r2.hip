// RK stage kernels of the mixed-precision batch flow over links in two-row
// form (the launcher of lym_launch.hpp, instantiated here).
#define LYM_INSTANTIATE_BSTAGES_R2
#include "lym_launch.hpp"

namespace lym {
template void launch_batch_stage_r2<1>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *,
                                       const ConfDesc *, int, long, Sched, double);
template void launch_batch_stage_r2<2>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *,
                                       const ConfDesc *, int, long, Sched, double);
template void launch_batch_stage_r2<3>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *,
                                       const ConfDesc *, int, long, Sched, double);
}  // namespace lym
