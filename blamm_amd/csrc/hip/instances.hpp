// The filter and histogram kernels are compiled in translation units of their
// own (filter_bf16.hip, filter_fp32.hip, hist_filter.hip) so that the library
// builds in parallel; the launching code sees these declarations and leaves
// the instantiation to those units.
#pragma once

#include "scan_kernels.hpp"
#include "hist_kernels.hpp"

namespace blamm {

#define BLAMM_FILTER_INSTANCES(X)                                         \
    X(true, false, false) X(true, false, true) X(true, true, false) X(true, true, true) \
    X(false, false, false) X(false, false, true) X(false, true, false) X(false, true, true)
#define BLAMM_FILTER_LO_INSTANCES(X) X(false, false) X(false, true) X(true, false) X(true, true)
#define BLAMM_HIST_INSTANCES(X)                                                             \
    X(false, false, false) X(false, true, false) X(true, false, false) X(true, true, false) \
    X(false, false, true) X(false, true, true) X(true, false, true) X(true, true, true)

#ifndef BLAMM_INSTANTIATE
#define BLAMM_EXTERN_FILTER(B, M, L) \
    extern template __global__ void filter_persistent_kernel<B, M, L>(FilterArgs, WorkQueue, ListArgs);
#define BLAMM_EXTERN_LO(M, L) extern template __global__ void filter_lo_kernel<M, L>(FilterArgs, WorkQueue, ListArgs);
#define BLAMM_EXTERN_HIST(M, L, W) extern template __global__ void hist_filter_kernel<M, L, W>(HistArgs);
BLAMM_FILTER_INSTANCES(BLAMM_EXTERN_FILTER)
BLAMM_FILTER_LO_INSTANCES(BLAMM_EXTERN_LO)
BLAMM_HIST_INSTANCES(BLAMM_EXTERN_HIST)
#undef BLAMM_EXTERN_FILTER
#undef BLAMM_EXTERN_LO
#undef BLAMM_EXTERN_HIST
#endif

}  // namespace blamm
