// lbm_adjoint.hip — launchers of the adjoint tick kernel (D3Q19; every
// storage, collision and force kind) and of its fixed-order reduction.
#include "lbm_adjoint.hpp"
#include "lbm_dispatch.hpp"

namespace lbm {

// Fixed-order sum of the blocks' partials: out[p][k] = sum_b partial[b][p][k],
// one block per (p, k): each thread sums a strided share in double-double,
// then the waves and the block combine in a fixed order.
__global__ void __launch_bounds__(256) adjoint_reduce_kernel(const double *partial, int64_t nblocks, int np,
                                                            double *out) {
  const int pk = blockIdx.x;
  const int tid = threadIdx.x, wave = tid >> 6, lane = __lane_id();
  const int64_t stride = (int64_t)np * kAdjSums;
  DD s{0.0, 0.0};
#pragma unroll 4
  for (int64_t b = tid; b < nblocks; b += 256) s = dd_merge(s, DD{partial[b * stride + pk], 0.0});
  s = wave_sum(s);
  __shared__ DD part[4];
  if (lane == kWaveSumLane) part[wave] = s;
  __syncthreads();
  if (tid == 0) {
    DD t = part[0];
    for (int w = 1; w < 4; w++) t = dd_merge(t, part[w]);
    out[pk] = t.hi + t.lo;
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
  hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((unsigned)(np * kAdjSums)), dim3(256), 0, stream, partial, nblocks, np,
                     out);
  return hipGetLastError();
}

}  // namespace lbm
