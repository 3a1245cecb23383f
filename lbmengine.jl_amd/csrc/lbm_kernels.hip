// lbm_kernels.hip — launchers of the step and auxiliary kernels for
// one storage kind.  Compiled three times with -DLBM_STORE=0 (FP64),
// 1 (FP32) and 2 (FP16).
#include "lbm_dispatch.hpp"

#ifndef LBM_STORE
#error "LBM_STORE must be 0 (FP64), 1 (FP32) or 2 (FP16)"
#endif

namespace lbm {

constexpr int S = LBM_STORE;

#define LBM_DISPATCH_QCF(KERNEL, ARGS)                                                   \
  do {                                                                                   \
    const int key = (Q == 19 ? 100 : 0) + coll * 10 + fm;                                \
    switch (key) {                                                                       \
      case 0: hipLaunchKernelGGL((KERNEL<9, S, 0, 0>), l.grid, l.block, l.shmem, l.stream, ARGS); break;    \
      case 1: hipLaunchKernelGGL((KERNEL<9, S, 0, 1>), l.grid, l.block, l.shmem, l.stream, ARGS); break;    \
      case 2: hipLaunchKernelGGL((KERNEL<9, S, 0, 2>), l.grid, l.block, l.shmem, l.stream, ARGS); break;    \
      case 10: hipLaunchKernelGGL((KERNEL<9, S, 1, 0>), l.grid, l.block, l.shmem, l.stream, ARGS); break;   \
      case 11: hipLaunchKernelGGL((KERNEL<9, S, 1, 1>), l.grid, l.block, l.shmem, l.stream, ARGS); break;   \
      case 12: hipLaunchKernelGGL((KERNEL<9, S, 1, 2>), l.grid, l.block, l.shmem, l.stream, ARGS); break;   \
      case 100: hipLaunchKernelGGL((KERNEL<19, S, 0, 0>), l.grid, l.block, l.shmem, l.stream, ARGS); break; \
      case 101: hipLaunchKernelGGL((KERNEL<19, S, 0, 1>), l.grid, l.block, l.shmem, l.stream, ARGS); break; \
      case 102: hipLaunchKernelGGL((KERNEL<19, S, 0, 2>), l.grid, l.block, l.shmem, l.stream, ARGS); break; \
      case 110: hipLaunchKernelGGL((KERNEL<19, S, 1, 0>), l.grid, l.block, l.shmem, l.stream, ARGS); break; \
      case 111: hipLaunchKernelGGL((KERNEL<19, S, 1, 1>), l.grid, l.block, l.shmem, l.stream, ARGS); break; \
      case 112: hipLaunchKernelGGL((KERNEL<19, S, 1, 2>), l.grid, l.block, l.shmem, l.stream, ARGS); break; \
      default: return hipErrorInvalidValue;                                              \
    }                                                                                    \
  } while (0)

template <>
hipError_t launch_step<S>(int Q, int coll, int fm, const StepArgs &a, const Launch &l) {
  LBM_DISPATCH_QCF(step_kernel, a);
  return hipGetLastError();
}

#define LBM_DISPATCH_Q(KERNEL, ARGS)                                                \
  do {                                                                              \
    if (Q == 9)                                                                     \
      hipLaunchKernelGGL((KERNEL<9, S>), l.grid, l.block, l.shmem, l.stream, ARGS);       \
    else                                                                            \
      hipLaunchKernelGGL((KERNEL<19, S>), l.grid, l.block, l.shmem, l.stream, ARGS);      \
  } while (0)

template <>
hipError_t launch_init_uniform<S>(int Q, const AuxArgs &a, const Launch &l) {
  LBM_DISPATCH_Q(init_uniform_kernel, a);
  return hipGetLastError();
}

template <>
hipError_t launch_set_populations<S>(int Q, const AuxArgs &a, const Launch &l) {
  LBM_DISPATCH_Q(set_populations_kernel, a);
  return hipGetLastError();
}

template <>
hipError_t launch_get_populations<S>(int Q, const AuxArgs &a, const Launch &l) {
  LBM_DISPATCH_Q(get_populations_kernel, a);
  return hipGetLastError();
}

template <>
hipError_t launch_hydro<S>(int Q, const AuxArgs &a, const Launch &l) {
  LBM_DISPATCH_Q(hydro_kernel, a);
  return hipGetLastError();
}

}  // namespace lbm
