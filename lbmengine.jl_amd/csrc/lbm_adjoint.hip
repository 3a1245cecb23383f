// lbm_adjoint.hip — launchers of the adjoint tick kernel (D3Q19; every
// storage, collision and force kind) and of its fixed-order reduction.
#include "lbm_adjoint.hpp"
#include "lbm_dispatch.hpp"

namespace lbm {

// Fixed-order sum of the blocks' partials: out[p][k] = sum_b partial[b][p][k].
__global__ void __launch_bounds__(256) adjoint_reduce_kernel(const double *partial, int64_t nblocks, int np,
                                                            double *out) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = __lane_id();
  for (int pk = wave; pk < np * kAdjSums; pk += (int)(blockDim.x >> 6)) {
    DD s{0.0, 0.0};
    for (int64_t b = lane; b < nblocks; b += 64) s = dd_merge(s, DD{partial[b * np * kAdjSums + pk], 0.0});
    s = wave_sum(s);
    if (lane == kWaveSumLane) out[pk] = s.hi + s.lo;
  }
}

template <int S, int COLL>
static hipError_t launch_adjoint_sc(int fm, const AdjArgs &a, const Launch &l) {
  switch (fm) {
    case LBM_FORCE_NONE: hipLaunchKernelGGL((adjoint_kernel<19, S, COLL, LBM_FORCE_NONE>), l.grid, l.block, 0, l.stream, a); break;
    case LBM_FORCE_UNIFORM: hipLaunchKernelGGL((adjoint_kernel<19, S, COLL, LBM_FORCE_UNIFORM>), l.grid, l.block, 0, l.stream, a); break;
    case LBM_FORCE_FIELD: hipLaunchKernelGGL((adjoint_kernel<19, S, COLL, LBM_FORCE_FIELD>), l.grid, l.block, 0, l.stream, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int S>
static hipError_t launch_adjoint_s(int coll, int fm, const AdjArgs &a, const Launch &l) {
  return coll == LBM_COLLISION_TRT ? launch_adjoint_sc<S, LBM_COLLISION_TRT>(fm, a, l)
                                   : launch_adjoint_sc<S, LBM_COLLISION_BGK>(fm, a, l);
}

hipError_t launch_adjoint(int storage, int coll, int fm, const AdjArgs &a, const Launch &l) {
  switch (storage) {
    case LBM_STORAGE_F64: return launch_adjoint_s<kF64>(coll, fm, a, l);
    case LBM_STORAGE_F32: return launch_adjoint_s<kF32>(coll, fm, a, l);
    default: return launch_adjoint_s<kF16>(coll, fm, a, l);
  }
}

hipError_t launch_adjoint_reduce(const double *partial, int64_t nblocks, int np, double *out, hipStream_t stream) {
  hipLaunchKernelGGL(adjoint_reduce_kernel, dim3(1), dim3(256), 0, stream, partial, nblocks, np, out);
  return hipGetLastError();
}

}  // namespace lbm
