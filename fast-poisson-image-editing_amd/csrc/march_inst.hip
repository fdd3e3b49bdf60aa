// One (sweeps, columns per lane) variant of the sweep kernel, for every
// schedule and neighbour order.  Compiled once per entry of PIE_TILES with
// -DPIE_S=<sweeps> -DPIE_C=<columns per lane>.
#include "march.h"

#ifndef PIE_S
#error "compile with -DPIE_S=<sweeps per launch> -DPIE_C=<columns per lane>"
#endif

#define PIE_NAME2(s, c) launch_march_##s##_##c
#define PIE_NAME(s, c) PIE_NAME2(s, c)

namespace pie {

hipError_t PIE_NAME(PIE_S, PIE_C)(int sched, int order, const MarchArgs& a,
                                  hipStream_t st) {
  using namespace detail;
  const dim3 grid((a.ntask + kBlockWaves - 1) / kBlockWaves), block(kBlock);
#define PIE_L(SC, OR)                                                     \
  if (sched == SC && order == OR) {                                      \
    hipLaunchKernelGGL((march_kernel<PIE_S, PIE_C, SC, OR>), grid, block, \
                       0, st, a);                                         \
    return hipGetLastError();                                            \
  }
  PIE_L(SCHED_JACOBI, ORDER_UDLR)
  PIE_L(SCHED_JACOBI, ORDER_ULRD)
  PIE_L(SCHED_GS_LEX, ORDER_UDLR)
  PIE_L(SCHED_GS_LEX, ORDER_ULRD)
  PIE_L(SCHED_REDBLACK, ORDER_UDLR)
  PIE_L(SCHED_REDBLACK, ORDER_ULRD)
#undef PIE_L
  return hipErrorInvalidValue;
}

}  // namespace pie
