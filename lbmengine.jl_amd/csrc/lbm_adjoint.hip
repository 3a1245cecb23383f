// lbm_adjoint.hip — launchers of the adjoint tick's kernels (D3Q19; every
// storage, collision and force kind): the near and bulk kernels, the move's
// reverse and the tick's tail.
#include "lbm_adjoint.hpp"
#include "lbm_adjoint_move.hpp"
#include "lbm_dispatch.hpp"

namespace lbm {

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

// Reverse tick k's tail in one block (AdjTailArgs): the near kernel's
// per-particle sums in a fixed order (thread g ns + s sums partial s of the
// blocks b = g mod G in double-double, then the G groups in order); the
// particles' adjoint state at the start of tick k, mu = mu_in (the move's
// reverse) + the sums' terms of the state + dL/dV_k of the speed term when
// V_k is averaged, and the B1, B2 gradients; then, when asked, the reverse
// of tick k - 1's move from that state (wave p for particle p, lane j the
// tangent of input j, as adjoint_move_kernel).
__global__ void __launch_bounds__(kAdjTailThreads) adjoint_tail_kernel(AdjTailArgs a) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x;
  const AdjFinishArgs &fa = a.fin;
  const int ns = fa.np * kAdjSums;
  const int G = kAdjTailThreads / ns;
  __shared__ DD part[kAdjTailThreads];
  __shared__ double sums[kAdjParticles * kAdjSums];
  __shared__ double mu[kAdjParticles * kPState];
  if (tid < G * ns) {
    const int s = tid % ns, gi = tid / ns;
    DD acc{0.0, 0.0};
    for (int64_t b = gi; b < a.nblocks; b += G) acc = dd_merge(acc, DD{a.partial[b * ns + s], 0.0});
    part[tid] = acc;
  }
  __syncthreads();
  if (tid < ns) {
    DD t = part[tid];
    for (int gi = 1; gi < G; gi++) t = dd_merge(t, part[gi * ns + tid]);
    sums[tid] = t.hi + t.lo;
  }
  __syncthreads();
  if (tid < fa.np * kPState) {
    const int p = tid / kPState, i = tid % kPState;
    const double *sm = sums + p * kAdjSums;
    double v = fa.mu_in[tid];
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
    if (fa.direct && p == fa.sw && i >= kStV && i < kStV + 3) {
      const double *V = fa.recs_k[p].V;
      const double sp = sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
      if (sp > 0.0) v = v + (0.0 + fa.gscale * V[i - kStV] / sp);
    }
    fa.mu[tid] = v;
    mu[tid] = v;
    if (i == 0) {
      fa.gB[p * 2] += sm[kAdjB1];
      fa.gB[p * 2 + 1] += sm[kAdjB2];
    }
  }
  if (!a.do_move) return;
  __syncthreads();
  const AdjMoveArgs &m = a.mv;
  const int p = tid >> 6, j = tid & 63;
  if (p >= fa.np || j >= kDualN) return;
  const DevParticle P = m.recs_k[p];
  const DevParticle &P1 = m.recs_k1[p];
  const double F[3] = {P1.force[0], P1.force[1], P1.force[2]};
  const double Tq[3] = {P1.torque[0], P1.torque[1], P1.torque[2]};
  DualStateT<1> st = move_dual_t<1>(P, F, Tq, m.dt, j);
  double g = 0.0;
  for (int o = 0; o < kPState; o++) {
    const double mo = mu[p * kPState + o];
    if (mo != 0.0) g += mo * st.slot(o)->d[0];
  }
  if (j < kPState)
    m.mu_in[p * kPState + j] = g;
  else
    m.phitau[p * 6 + (j - kPState)] = g;
}

template <int S, int COLL, bool NEAR>
static hipError_t launch_adjoint_sc(int fm, const AdjArgs &a, const Launch &l) {
#define LBM_ADJ_LAUNCH(FM)                                                                       \
  if (NEAR)                                                                                      \
    hipLaunchKernelGGL((adjoint_near_kernel<19, S, COLL, FM>), l.grid, l.block, 0, l.stream, a); \
  else                                                                                           \
    hipLaunchKernelGGL((adjoint_kernel<19, S, COLL, FM>), l.grid, l.block, 0, l.stream, a);
  switch (fm) {
    case LBM_FORCE_NONE: LBM_ADJ_LAUNCH(LBM_FORCE_NONE) break;
    case LBM_FORCE_UNIFORM: LBM_ADJ_LAUNCH(LBM_FORCE_UNIFORM) break;
    case LBM_FORCE_FIELD: LBM_ADJ_LAUNCH(LBM_FORCE_FIELD) break;
    default: return hipErrorInvalidValue;
  }
#undef LBM_ADJ_LAUNCH
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

// dst += src, src = 0 (n values): a segment's FP32 sum of dL/dF folded into
// the FP64 total.
__global__ void __launch_bounds__(256) adjoint_fold_kernel(float *src, double *dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dst[i] += (double)src[i];
  src[i] = 0.0f;
}

hipError_t launch_adjoint_fold(float *src, double *dst, int64_t n, hipStream_t stream) {
  hipLaunchKernelGGL(adjoint_fold_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, src, dst, n);
  return hipGetLastError();
}

hipError_t launch_adjoint_tail(const AdjTailArgs &a, hipStream_t stream) {
  hipLaunchKernelGGL(adjoint_tail_kernel, dim3(1), dim3(kAdjTailThreads), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_adjoint_move(const AdjMoveArgs &a, int np, hipStream_t stream) {
  hipLaunchKernelGGL(adjoint_move_kernel, dim3((unsigned)np), dim3(64), 0, stream, a);
  return hipGetLastError();
}

}  // namespace lbm
