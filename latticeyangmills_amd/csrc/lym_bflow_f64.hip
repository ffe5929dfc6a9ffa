// Whole-flow persistent kernel of the batched ensemble path, double links, fields
// of 3 rows (the launcher of lym_launch.hpp, instantiated here).
#define LYM_INSTANTIATE_BFLOW
#include "lym_launch.hpp"

namespace lym {
template void launch_batch_flow_persistent<double, 3>(int, unsigned, hipStream_t, F2<double> *, F2<double> *, double *,
                                                    const ConfDesc *, int, long, Sched, double, int, FlowSync *, int,
                                                    unsigned long long);
template hipError_t batch_flow_persistent_per_cu<double, 3>(int, int *);
}  // namespace lym
