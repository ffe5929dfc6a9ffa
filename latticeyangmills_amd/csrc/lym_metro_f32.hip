// Checkerboard Metropolis kernels with float links, single lattice and batch
// (the launchers of lym_launch.hpp, instantiated here).
#define LYM_INSTANTIATE_METRO
#include "lym_launch.hpp"

namespace lym {
template void launch_metropolis<float>(int, bool, unsigned, hipStream_t, F2<float> *, Geom, int, const MetroParams &,
                                    const int *, unsigned long long *);
template void launch_batch_metropolis<float>(int, unsigned, hipStream_t, F2<float> *, const ConfDesc *, int, long, int,
                                          const MetroParams &, unsigned long long *);
}  // namespace lym
