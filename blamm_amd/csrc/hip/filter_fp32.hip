// Instances of the FP32 filter kernel.
#define BLAMM_INSTANTIATE
#include "instances.hpp"

namespace blamm {
#define BLAMM_DEF_FILTER(B, M, L) template __global__ void filter_persistent_kernel<B, M, L>(FilterArgs, WorkQueue, ListArgs);
BLAMM_DEF_FILTER(false, false, false)
BLAMM_DEF_FILTER(false, false, true)
BLAMM_DEF_FILTER(false, true, false)
BLAMM_DEF_FILTER(false, true, true)
}  // namespace blamm
