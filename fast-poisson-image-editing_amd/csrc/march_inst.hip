// One (sweeps, columns per lane, channels per task) variant of the sweep
// kernel, for every schedule and neighbour order.  Compiled once per entry
// of PIE_TILES with -DPIE_S=<sweeps> -DPIE_C=<columns per lane>
// -DPIE_N=<channels per task>.
#include "march.h"

#if !defined(PIE_S) || !defined(PIE_C) || !defined(PIE_N)
#error "compile with -DPIE_S=<sweeps> -DPIE_C=<columns> -DPIE_N=<channels>"
#endif

#define PIE_NAME2(s, c, n) launch_march_##s##_##c##_##n
#define PIE_NAME(s, c, n) PIE_NAME2(s, c, n)

namespace pie {

hipError_t PIE_NAME(PIE_S, PIE_C, PIE_N)(int sched, int order,
                                         const MarchArgs& a, hipStream_t st) {
  using namespace detail;
  const dim3 grid((a.ntask + kBlockWaves - 1) / kBlockWaves), block(kBlock);
#define PIE_L(SC, OR)                                                     \
  if (sched == SC && order == OR) {                                      \
    hipLaunchKernelGGL((march_kernel<PIE_S, PIE_C, PIE_N, SC, OR>), grid, block, \
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

#define PIE_MNAME2(s, c, n) march_blocks_per_cu_##s##_##c##_##n
#define PIE_MNAME(s, c, n) PIE_MNAME2(s, c, n)

int PIE_MNAME(PIE_S, PIE_C, PIE_N)(int sched, int order) {
  using namespace detail;
  int n = 0;
  hipError_t e = hipErrorInvalidValue;
#define PIE_L(SC, OR)                                                       \
  if (sched == SC && order == OR)                                          \
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(                       \
        &n, (march_kernel<PIE_S, PIE_C, PIE_N, SC, OR>), kBlock, 0);
  PIE_L(SCHED_JACOBI, ORDER_UDLR)
  PIE_L(SCHED_JACOBI, ORDER_ULRD)
  PIE_L(SCHED_GS_LEX, ORDER_UDLR)
  PIE_L(SCHED_GS_LEX, ORDER_ULRD)
  PIE_L(SCHED_REDBLACK, ORDER_UDLR)
  PIE_L(SCHED_REDBLACK, ORDER_ULRD)
#undef PIE_L
  return e == hipSuccess && n > 0 ? n : 1;
}

#define PIE_BNAME2(s, c, n) launch_batch_##s##_##c##_##n
#define PIE_BNAME(s, c, n) PIE_BNAME2(s, c, n)
#define PIE_ONAME2(s, c, n) batch_blocks_per_cu_##s##_##c##_##n
#define PIE_ONAME(s, c, n) PIE_ONAME2(s, c, n)

hipError_t PIE_BNAME(PIE_S, PIE_C, PIE_N)(int sched, int order,
                                          const BatchArgs& b, int blocks,
                                          hipStream_t st) {
  using namespace detail;
#define PIE_L(SC, OR)                                                          \
  if (sched == SC && order == OR) {                                           \
    hipLaunchKernelGGL((march_batch_kernel<PIE_S, PIE_C, PIE_N, SC, OR>),             \
                       dim3(blocks), dim3(kBatchBlock), 0, st, b);             \
    return hipGetLastError();                                                 \
  }
  PIE_L(SCHED_JACOBI, ORDER_UDLR)
  PIE_L(SCHED_JACOBI, ORDER_ULRD)
#undef PIE_L
  return hipErrorInvalidValue;
}

int PIE_ONAME(PIE_S, PIE_C, PIE_N)(int sched, int order) {
  using namespace detail;
  int n = 0;
  hipError_t e = hipErrorInvalidValue;
#define PIE_L(SC, OR)                                                       \
  if (sched == SC && order == OR)                                          \
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(                       \
        &n, (march_batch_kernel<PIE_S, PIE_C, PIE_N, SC, OR>), kBatchBlock, 0);
  PIE_L(SCHED_JACOBI, ORDER_UDLR)
  PIE_L(SCHED_JACOBI, ORDER_ULRD)
#undef PIE_L
  return e == hipSuccess && n > 0 ? n : 1;
}

}  // namespace pie
