// lbm_adjoint.hip — launchers of the adjoint tick kernel (D3Q19; every
// storage, collision and force kind) and of its fixed-order reduction.
#include "lbm_adjoint.hpp"
#include "lbm_adjoint_move.hpp"
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

__global__ void __launch_bounds__(64) adjoint_move_kernel(AdjMoveArgs a) {
#pragma clang fp contract(off)
  const int p = blockIdx.x, j = threadIdx.x;
  if (j >= kDualN) return;
  const DevParticle P = a.recs_k[p];
  const DevParticle &P1 = a.recs_k1[p];
  const double F[3] = {P1.force[0], P1.force[1], P1.force[2]};
  const double Tq[3] = {P1.torque[0], P1.torque[1], P1.torque[2]};
  DualStateT<1> s = move_dual_t<1>(P, F, Tq, a.dt, j);
  double g = 0.0;
  for (int o = 0; o < kPState; o++) {
    const double m = a.mu[p * kPState + o];
    if (m != 0.0) g += m * s.slot(o)->d[0];
  }
  if (j < kPState)
    a.mu_in[p * kPState + j] = g;
  else
    a.phitau[p * 6 + (j - kPState)] = g;
}

__global__ void __launch_bounds__(256) adjoint_finish_kernel(AdjFinishArgs a) {
#pragma clang fp contract(off)
  const int t = threadIdx.x;
  if (t >= a.np * kPState) return;
  const int p = t / kPState, i = t % kPState;
  const double *sm = a.sums + p * kAdjSums;
  double v = a.mu_in[t];
  int si = -1;
  if (i < kStR + 3)
    si = kAdjR + i;
  else if (i >= kStNR && i < kStNR + 3)
    si = kAdjNR + (i - kStNR);
  else if (i >= kStNV && i < kStNV + 3)
    si = kAdjNV + (i - kStNV);
  else if (i >= kStNW && i < kStNW + 3)
    si = kAdjNW + (i - kStNW);
  else if (i >= kStNE && i < kStNE + 3)
    si = kAdjNE + (i - kStNE);
  if (si >= 0) v = v + sm[si];
  if (a.direct && p == a.sw && i >= kStV && i < kStV + 3) {
    const double *V = a.recs_k[p].V;
    const double sp = sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
    if (sp > 0.0) v = v + (0.0 + a.gscale * V[i - kStV] / sp);
  }
  a.mu[t] = v;
  if (i == 0) {
    a.gB[p * 2] += sm[kAdjB1];
    a.gB[p * 2 + 1] += sm[kAdjB2];
  }
}

template <int S, int COLL, bool NEAR>
static hipError_t launch_adjoint_sc(int fm, const AdjArgs &a, const Launch &l) {
  switch (fm) {
    case LBM_FORCE_NONE: hipLaunchKernelGGL((adjoint_kernel<19, S, COLL, LBM_FORCE_NONE, NEAR>), l.grid, l.block, 0, l.stream, a); break;
    case LBM_FORCE_UNIFORM: hipLaunchKernelGGL((adjoint_kernel<19, S, COLL, LBM_FORCE_UNIFORM, NEAR>), l.grid, l.block, 0, l.stream, a); break;
    case LBM_FORCE_FIELD: hipLaunchKernelGGL((adjoint_kernel<19, S, COLL, LBM_FORCE_FIELD, NEAR>), l.grid, l.block, 0, l.stream, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int S, bool NEAR>
static hipError_t launch_adjoint_s(int coll, int fm, const AdjArgs &a, const Launch &l) {
  return coll == LBM_COLLISION_TRT ? launch_adjoint_sc<S, LBM_COLLISION_TRT, NEAR>(fm, a, l)
                                   : launch_adjoint_sc<S, LBM_COLLISION_BGK, NEAR>(fm, a, l);
}

template <bool NEAR>
static hipError_t launch_adjoint_part(int storage, int coll, int fm, const AdjArgs &a, const Launch &l) {
  switch (storage) {
    case LBM_STORAGE_F64: return launch_adjoint_s<kF64, NEAR>(coll, fm, a, l);
    case LBM_STORAGE_F32: return launch_adjoint_s<kF32, NEAR>(coll, fm, a, l);
    case LBM_STORAGE_F16C: return launch_adjoint_s<kF16C, NEAR>(coll, fm, a, l);
    default: return launch_adjoint_s<kF16, NEAR>(coll, fm, a, l);
  }
}

hipError_t launch_adjoint(int storage, int coll, int fm, bool near, const AdjArgs &a, const Launch &l) {
  return near ? launch_adjoint_part<true>(storage, coll, fm, a, l) : launch_adjoint_part<false>(storage, coll, fm, a, l);
}

hipError_t launch_adjoint_move(const AdjMoveArgs &a, int np, hipStream_t stream) {
  hipLaunchKernelGGL(adjoint_move_kernel, dim3((unsigned)np), dim3(64), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_adjoint_finish(const AdjFinishArgs &a, hipStream_t stream) {
  hipLaunchKernelGGL(adjoint_finish_kernel, dim3(1), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_adjoint_reduce(const double *partial, int64_t nblocks, int np, double *out, hipStream_t stream) {
  hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((unsigned)(np * kAdjSums)), dim3(256), 0, stream, partial, nblocks, np,
                     out);
  return hipGetLastError();
}

}  // namespace lbm
