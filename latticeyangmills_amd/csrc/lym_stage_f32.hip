// RK stage kernels of the Wilson flow with float links, single lattice and
// batch (the launchers of lym_launch.hpp, instantiated here).
#define LYM_INSTANTIATE_STAGES
#include "lym_launch.hpp"

namespace lym {
template void launch_flow_stage<float, 1>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *, Geom, Sched, double);
template void launch_flow_stage<float, 2>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *, Geom, Sched, double);
template void launch_flow_stage<float, 3>(int, unsigned, hipStream_t, const F2<float> *, F2<float> *, float *, Geom, Sched, double);
}  // namespace lym
