// Instances of the empirical-histogram kernel.
#define BLAMM_INSTANTIATE
#include "instances.hpp"

namespace blamm {
#define BLAMM_DEF_HIST(M, L, W) template __global__ void hist_filter_kernel<M, L, W>(HistArgs);
BLAMM_HIST_INSTANCES(BLAMM_DEF_HIST)
}  // namespace blamm
