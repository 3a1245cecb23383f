// lbm_kernels.hip — launchers of the step and auxiliary kernels for
// one storage kind.  Compiled once per storage, -DLBM_STORE=0 (FP64),
// 1 (FP32), 2 (FP16) or 3 (custom 16-bit), and per part: -DLBM_PART=0 builds
// the one-lattice step of D2Q9/D3Q19 and the auxiliary kernels of every
// lattice, -DLBM_PART=1 the batched step (D2Q9/D3Q19), -DLBM_PART=2 the
// one-lattice step of D3Q15/D3Q27, -DLBM_PART=3 the D3Q19 kernels that keep
// the moments (separate objects compile in parallel).
#include "lbm_dispatch.hpp"

#ifndef LBM_STORE
#error "LBM_STORE must be 0 (FP64), 1 (FP32), 2 (FP16) or 3 (custom 16-bit)"
#endif
#ifndef LBM_PART
#error "LBM_PART must be 0 (one lattice and auxiliary kernels), 1 (batched step), 2 (D3Q15/D3Q27 step) or 3 (keeping modes)"
#endif

namespace lbm {

constexpr int S = LBM_STORE;

// (lattice, collision, force) -> the launcher instantiated for them, for
// the two lattices QA, QB of this part
#define LBM_KEY_SWITCH(QA, QB, FN, ...)                   \
  switch ((Q == QB ? 100 : (Q == QA ? 0 : 1000)) + coll * 10 + fm) { \
    case 0: return FN<QA, 0, 0>(__VA_ARGS__);             \
    case 1: return FN<QA, 0, 1>(__VA_ARGS__);             \
    case 2: return FN<QA, 0, 2>(__VA_ARGS__);             \
    case 10: return FN<QA, 1, 0>(__VA_ARGS__);            \
    case 11: return FN<QA, 1, 1>(__VA_ARGS__);            \
    case 12: return FN<QA, 1, 2>(__VA_ARGS__);            \
    case 100: return FN<QB, 0, 0>(__VA_ARGS__);           \
    case 101: return FN<QB, 0, 1>(__VA_ARGS__);           \
    case 102: return FN<QB, 0, 2>(__VA_ARGS__);           \
    case 110: return FN<QB, 1, 0>(__VA_ARGS__);           \
    case 111: return FN<QB, 1, 1>(__VA_ARGS__);           \
    case 112: return FN<QB, 1, 2>(__VA_ARGS__);           \
    default: return hipErrorInvalidValue;                 \
  }

// (the collide-only mode has no box path: one occupancy variant)
#define LBM_MODE_SWITCH(KERNEL, ARGS)                                                                        \
  switch (mode * 2 + (occ ? 1 : 0)) {                                                                        \
    case 2 * kCollide:                                                                                       \
    case 2 * kCollide + 1: hipLaunchKernelGGL((KERNEL<Q, S, COLL, FM, kCollide, 0>), l.grid, l.block, l.shmem, l.stream, ARGS); break; \
    case 2 * kFused: hipLaunchKernelGGL((KERNEL<Q, S, COLL, FM, kFused, 0>), l.grid, l.block, l.shmem, l.stream, ARGS); break;         \
    case 2 * kFused + 1: hipLaunchKernelGGL((KERNEL<Q, S, COLL, FM, kFused, 1>), l.grid, l.block, l.shmem, l.stream, ARGS); break;     \
    case 2 * kGather: hipLaunchKernelGGL((KERNEL<Q, S, COLL, FM, kGather, 0>), l.grid, l.block, l.shmem, l.stream, ARGS); break;       \
    case 2 * kGather + 1: hipLaunchKernelGGL((KERNEL<Q, S, COLL, FM, kGather, 1>), l.grid, l.block, l.shmem, l.stream, ARGS); break;   \
    default: return hipErrorInvalidValue;                                                                    \
  }                                                                                                          \
  return hipGetLastError();

#if LBM_PART == 1

// (internal linkage: each storage's object file has its own)
template <int Q, int COLL, int FM>
static hipError_t launch_batch_qcf(int mode, int occ, const BatchArgs &a, const Launch &l) {
  LBM_MODE_SWITCH(batch_step_kernel, a)
}

template <>
hipError_t launch_batch_step<S>(int Q, int coll, int fm, int mode, int occ, const BatchArgs &a, const Launch &l) {
  LBM_KEY_SWITCH(9, 19, launch_batch_qcf, mode, occ, a, l)
}

#elif LBM_PART == 3

template <int COLL, int FM>
static hipError_t launch_keep_cf(int mode, const StepArgs &a, const Launch &l) {
  constexpr int Q = 19;
  switch (mode * 2 + (a.t.occ ? 1 : 0)) {
    case 2 * kCollideKeep:
    case 2 * kCollideKeep + 1:
      hipLaunchKernelGGL((step_kernel<Q, S, COLL, FM, kCollideKeep, 0>), l.grid, l.block, l.shmem, l.stream, a);
      break;
    case 2 * kFusedKeep:
      hipLaunchKernelGGL((step_kernel<Q, S, COLL, FM, kFusedKeep, 0>), l.grid, l.block, l.shmem, l.stream, a);
      break;
    case 2 * kFusedKeep + 1:
      hipLaunchKernelGGL((step_kernel<Q, S, COLL, FM, kFusedKeep, 1>), l.grid, l.block, l.shmem, l.stream, a);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <>
hipError_t launch_step_keep<S>(int coll, int fm, int mode, const StepArgs &a, const Launch &l) {
  switch (coll * 10 + fm) {
    case 0: return launch_keep_cf<0, 0>(mode, a, l);
    case 1: return launch_keep_cf<0, 1>(mode, a, l);
    case 2: return launch_keep_cf<0, 2>(mode, a, l);
    case 10: return launch_keep_cf<1, 0>(mode, a, l);
    case 11: return launch_keep_cf<1, 1>(mode, a, l);
    case 12: return launch_keep_cf<1, 2>(mode, a, l);
    default: return hipErrorInvalidValue;
  }
}

#elif LBM_PART == 2

template <int Q, int COLL, int FM>
static hipError_t launch_step_qcf(int mode, const StepArgs &a, const Launch &l) {
  const int occ = a.t.occ;
  LBM_MODE_SWITCH(step_kernel, a)
}

template <>
hipError_t launch_step_q15_q27<S>(int Q, int coll, int fm, int mode, const StepArgs &a, const Launch &l) {
  LBM_KEY_SWITCH(15, 27, launch_step_qcf, mode, a, l)
}

#else

template <int Q, int COLL, int FM>
static hipError_t launch_step_qcf(int mode, const StepArgs &a, const Launch &l) {
  const int occ = a.t.occ;
  LBM_MODE_SWITCH(step_kernel, a)
}

template <>
hipError_t launch_step<S>(int Q, int coll, int fm, int mode, const StepArgs &a, const Launch &l) {
  if (Q == 15 || Q == 27) return launch_step_q15_q27<S>(Q, coll, fm, mode, a, l);
  LBM_KEY_SWITCH(9, 19, launch_step_qcf, mode, a, l)
}

#define LBM_DISPATCH_Q(KERNEL, ...)                                                          \
  do {                                                                                       \
    switch (Q) {                                                                             \
      case 9: hipLaunchKernelGGL((KERNEL<9, S>), l.grid, l.block, l.shmem, l.stream, __VA_ARGS__); break;   \
      case 15: hipLaunchKernelGGL((KERNEL<15, S>), l.grid, l.block, l.shmem, l.stream, __VA_ARGS__); break; \
      case 19: hipLaunchKernelGGL((KERNEL<19, S>), l.grid, l.block, l.shmem, l.stream, __VA_ARGS__); break; \
      case 27: hipLaunchKernelGGL((KERNEL<27, S>), l.grid, l.block, l.shmem, l.stream, __VA_ARGS__); break; \
      default: return hipErrorInvalidValue;                                                   \
    }                                                                                        \
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

template <>
hipError_t launch_stress<S>(int Q, const AuxArgs &a, const Launch &l) {
  LBM_DISPATCH_Q(stress_kernel, a);
  return hipGetLastError();
}

template <>
hipError_t launch_scan<S>(int Q, const AuxArgs &a, unsigned long long *counts, const Launch &l) {
  LBM_DISPATCH_Q(scan_populations_kernel, a, counts);
  return hipGetLastError();
}

#endif

}  // namespace lbm
