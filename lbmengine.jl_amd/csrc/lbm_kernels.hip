// lbm_kernels.hip — launchers of the step and auxiliary kernels for
// one storage kind.  Compiled three times with -DLBM_STORE=0 (FP64),
// 1 (FP32) and 2 (FP16).
#include "lbm_dispatch.hpp"

#ifndef LBM_STORE
#error "LBM_STORE must be 0 (FP64), 1 (FP32) or 2 (FP16)"
#endif

namespace lbm {

constexpr int S = LBM_STORE;

// (internal linkage: each storage's object file has its own)
template <int Q, int COLL, int FM>
static hipError_t launch_step_qcf(int mode, const StepArgs &a, const Launch &l) {
  switch (mode) {
    case kCollide: hipLaunchKernelGGL((step_kernel<Q, S, COLL, FM, kCollide>), l.grid, l.block, l.shmem, l.stream, a); break;
    case kFused: hipLaunchKernelGGL((step_kernel<Q, S, COLL, FM, kFused>), l.grid, l.block, l.shmem, l.stream, a); break;
    case kGather: hipLaunchKernelGGL((step_kernel<Q, S, COLL, FM, kGather>), l.grid, l.block, l.shmem, l.stream, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <>
hipError_t launch_step<S>(int Q, int coll, int fm, int mode, const StepArgs &a, const Launch &l) {
  const int key = (Q == 19 ? 100 : 0) + coll * 10 + fm;
  switch (key) {
    case 0: return launch_step_qcf<9, 0, 0>(mode, a, l);
    case 1: return launch_step_qcf<9, 0, 1>(mode, a, l);
    case 2: return launch_step_qcf<9, 0, 2>(mode, a, l);
    case 10: return launch_step_qcf<9, 1, 0>(mode, a, l);
    case 11: return launch_step_qcf<9, 1, 1>(mode, a, l);
    case 12: return launch_step_qcf<9, 1, 2>(mode, a, l);
    case 100: return launch_step_qcf<19, 0, 0>(mode, a, l);
    case 101: return launch_step_qcf<19, 0, 1>(mode, a, l);
    case 102: return launch_step_qcf<19, 0, 2>(mode, a, l);
    case 110: return launch_step_qcf<19, 1, 0>(mode, a, l);
    case 111: return launch_step_qcf<19, 1, 1>(mode, a, l);
    case 112: return launch_step_qcf<19, 1, 2>(mode, a, l);
    default: return hipErrorInvalidValue;
  }
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
