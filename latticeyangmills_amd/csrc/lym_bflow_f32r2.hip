// Whole-flow persistent kernel of the batched ensemble path, float links, fields
// of 2 rows (the launcher of lym_launch.hpp, instantiated here).
#define LYM_INSTANTIATE_BFLOW
#include "lym_launch.hpp"

namespace lym {
template void launch_batch_flow_persistent<float, 2>(int, unsigned, hipStream_t, F2<float> *, F2<float> *, float *,
                                                    const ConfDesc *, int, long, Sched, double, int, FlowSync *, int,
                                                    unsigned long long);
template hipError_t batch_flow_persistent_per_cu<float, 2>(int, int *);
}  // namespace lym
