// RK stage kernels of the mixed-precision flow over links in two-row form
// (the launcher of lym_launch.hpp, instantiated here).
#define LYM_INSTANTIATE_STAGES_R2
#include "lym_launch.hpp"

namespace lym {
template void launch_flow_stage_r2<1>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *, Geom, Sched,
                                      double);
template void launch_flow_stage_r2<2>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *, Geom, Sched,
                                      double);
template void launch_flow_stage_r2<3>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *, Geom, Sched,
                                      double);
}  // namespace lym
