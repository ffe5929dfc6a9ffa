// RK stage kernels of the Wilson flow with double links, single lattice and
// batch (the launchers of lym_launch.hpp, instantiated here).
#define LYM_INSTANTIATE_STAGES
#include "lym_launch.hpp"

namespace lym {
template void launch_flow_stage<double, 1>(int, unsigned, hipStream_t, const F2<double> *, F2<double> *, double *, Geom, Sched, double);
template void launch_flow_stage<double, 2>(int, unsigned, hipStream_t, const F2<double> *, F2<double> *, double *, Geom, Sched, double);
template void launch_flow_stage<double, 3>(int, unsigned, hipStream_t, const F2<double> *, F2<double> *, double *, Geom, Sched, double);
}  // namespace lym
