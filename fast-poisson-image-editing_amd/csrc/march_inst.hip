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

#define PIE_BNAME2(s, c) launch_batch_##s##_##c
#define PIE_BNAME(s, c) PIE_BNAME2(s, c)
#define PIE_ONAME2(s, c) batch_blocks_per_cu_##s##_##c
#define PIE_ONAME(s, c) PIE_ONAME2(s, c)

hipError_t PIE_BNAME(PIE_S, PIE_C)(int sched, int order, const BatchArgs& b,
                                   int blocks, hipStream_t st) {
  using namespace detail;
#define PIE_L(SC, OR)                                                          \
  if (sched == SC && order == OR) {                                           \
    hipLaunchKernelGGL((march_batch_kernel<PIE_S, PIE_C, SC, OR>),             \
                       dim3(blocks), dim3(kBlock), 0, st, b);                  \
    return hipGetLastError();                                                 \
  }
  PIE_L(SCHED_JACOBI, ORDER_UDLR)
  PIE_L(SCHED_JACOBI, ORDER_ULRD)
#undef PIE_L
  return hipErrorInvalidValue;
}

int PIE_ONAME(PIE_S, PIE_C)(int sched, int order) {
  using namespace detail;
  int n = 0;
  hipError_t e = hipErrorInvalidValue;
#define PIE_L(SC, OR)                                                       \
  if (sched == SC && order == OR)                                          \
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(                       \
        &n, (march_batch_kernel<PIE_S, PIE_C, SC, OR>), kBlock, 0);
  PIE_L(SCHED_JACOBI, ORDER_UDLR)
  PIE_L(SCHED_JACOBI, ORDER_ULRD)
#undef PIE_L
  return e == hipSuccess && n > 0 ? n : 1;
}

}  // namespace pie
