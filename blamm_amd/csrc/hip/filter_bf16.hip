// Instances of the split-BF16 filter kernels (hi-only and lazy-lo forms).
#define BLAMM_INSTANTIATE
#include "instances.hpp"

namespace blamm {
#define BLAMM_DEF_FILTER(B, M, L) template __global__ void filter_persistent_kernel<B, M, L>(FilterArgs, WorkQueue, ListArgs);
BLAMM_DEF_FILTER(true, false, false)
BLAMM_DEF_FILTER(true, false, true)
BLAMM_DEF_FILTER(true, true, false)
BLAMM_DEF_FILTER(true, true, true)
#define BLAMM_DEF_LO(M, L) template __global__ void filter_lo_kernel<M, L>(FilterArgs, WorkQueue, ListArgs);
BLAMM_FILTER_LO_INSTANCES(BLAMM_DEF_LO)
}  // namespace blamm
