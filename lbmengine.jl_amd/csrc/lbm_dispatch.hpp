// lbm_dispatch.hpp — host-callable launchers for the kernels of
// liblbm_hip.so.  lbm_kernels.hip is compiled once per storage kind S
// (FP64, FP32, FP16) and defines the templated launchers for that S.
#pragma once

#include "lbm_step.hpp"

namespace lbm {

struct Launch {
  dim3 grid, block;
  hipStream_t stream;
  size_t shmem = 0;  // dynamic LDS bytes
};

template <int S>
hipError_t launch_step(int Q, int coll, int fm, int mode, const StepArgs &a, const Launch &l);
template <int S>
hipError_t launch_step_q15_q27(int Q, int coll, int fm, int mode, const StepArgs &a, const Launch &l);
template <int S>
hipError_t launch_batch_step(int Q, int coll, int fm, int mode, int occ, const BatchArgs &a, const Launch &l);
// kCollideKeep or kFusedKeep on a D3Q19 lattice (the adjoint's recomputation)
template <int S>
hipError_t launch_step_keep(int coll, int fm, int mode, const StepArgs &a, const Launch &l);
template <int S>
hipError_t launch_init_uniform(int Q, const AuxArgs &a, const Launch &l);
template <int S>
hipError_t launch_set_populations(int Q, const AuxArgs &a, const Launch &l);
template <int S>
hipError_t launch_get_populations(int Q, const AuxArgs &a, const Launch &l);
template <int S>
hipError_t launch_hydro(int Q, const AuxArgs &a, const Launch &l);
template <int S>
hipError_t launch_stress(int Q, const AuxArgs &a, const Launch &l);
template <int S>
hipError_t launch_scan(int Q, const AuxArgs &a, unsigned long long *counts, const Launch &l);

hipError_t launch_link_mask(int Q, const LinkMaskArgs &a, const Launch &l);
hipError_t launch_finalize(const StepParticles &sp, hipStream_t stream);
hipError_t launch_move(const StepParticles &sp, hipStream_t stream);
hipError_t launch_finalize_batch(const StepArgs *boxes, int n, hipStream_t stream);
hipError_t launch_copy(const void *src, void *dst, int64_t bytes, hipStream_t stream);
// the D3Q19 step's gathers and stores alone (elements of 2, 4 or 8 bytes)
hipError_t launch_pull(int elem_bytes, const void *src, void *dst, int nx, int ny, int nz, hipStream_t stream);

struct AdjArgs;
// near = false: the bulk kernel (nodes outside the particles' near boxes);
// true: the near kernel (their nodes, with the particle terms and sums)
hipError_t launch_adjoint(int storage, int coll, int fm, bool near, const AdjArgs &a, const Launch &l);
struct AdjMoveArgs;
struct AdjTailArgs;
hipError_t launch_adjoint_move(const AdjMoveArgs &a, int np, hipStream_t stream);
hipError_t launch_adjoint_tail(const AdjTailArgs &a, hipStream_t stream);
hipError_t launch_adjoint_fold(float *src, double *dst, int64_t n, hipStream_t stream);

}  // namespace lbm
