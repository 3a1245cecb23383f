// lbm_step.hpp — the time step of LBMengine.jl as HIP kernels for gfx950.
//
// tick! (gpu/fluids.jl:132-226) = collide, stream with halfway bounce-back
// and the Ladd term, momentum exchange, move the particles.  Here a run of n
// ticks is n + 1 launches of step_kernel in the pull form:
//   kCollide  P_0 = C(S_0)                      (S: the reference's
//   kFused    S_k = G(P_{k-1}), P_k = C(S_k)     model.distributions,
//   kGather   S_n = G(P_{n-1})                   P: after collision)
// G gathers f_q(x) = f*_q(x - c_q) (one read per population, each node
// writes only its own populations), with halfway bounce-back where x - c_q
// is a wall or particle node, the Ladd moving-wall term with the squirmer
// slip, and the momentum exchange; C is BGK/TRT with Guo forcing.  The
// nodes around the particles are handled by the leading "box" blocks of each
// gathering kernel (see box_work); the others stream the rest.
//
// Populations are stored row-blocked: the Q planes of each x-row are
// contiguous, f[(row Q + q) nx + x] with row = y + ny z.  FP64 storage holds
// f; FP32 holds f - w_i rho0; FP16 holds (f - w_i rho0) * 2^14 / w_i; the
// custom 16-bit format (kF16C) holds (f - w_i rho0) / w_i.
#pragma once

#include "lbm_common.hpp"
#include "lbm_integrate.hpp"

namespace lbm {

enum StoreKind { kF64 = 0, kF32 = 1, kF16 = 2, kF16C = 3 };

template <int S>
struct Store;
template <>
struct Store<kF64> {
  using T = double;  // storage element
  using R = double;  // compute type
};
template <>
struct Store<kF32> {
  using T = float;
  using R = float;
};
template <>
struct Store<kF16> {
  using T = _Float16;
  using R = float;
};
template <>
struct Store<kF16C> {
  using T = int16_t;  // loaded sign-extended (see f16c_scaled)
  using R = float;
};

// kF16C element -> h * 2^-108 as an FP32 value (exact): with the element
// sign-extended, shifting it left by 12 puts the exponent and mantissa bits
// in place and copies of the sign in bits 31..27; the mask keeps bit 31.
__device__ __forceinline__ float f16c_scaled(int16_t v) {
  return __builtin_bit_cast(float, (int32_t)((uint32_t)(int32_t)v << 12) & (int32_t)0x87FFF000);
}

// h * 2^-108 (FP32, gradual underflow kept) -> kF16C element, rounded to the
// nearest (ties away from zero) and clamped to the largest magnitude.
__device__ __forceinline__ int16_t f16c_from_scaled(float s) {
  const float lim = __builtin_bit_cast(float, 0x07FFF7FFu);  // rounds down to 0x07FFF000
  s = __builtin_amdgcn_fmed3f(s, -lim, lim);
  const uint32_t b = __builtin_bit_cast(uint32_t, s) + 0x800u;  // round the 12 dropped bits
  return (int16_t)(uint16_t)(((b >> 12) & 0x7FFFu) | ((b >> 16) & 0x8000u));
}

// Every population a step kernel writes is read next by the following
// kernel, long after the caches would have dropped it, and nothing reads it
// in between: stores are non-temporal (streamed), which took the 192^3 FP16
// fused kernel from 0.1066 to 0.1004 ms in a timing build.  (Non-temporal
// loads of the gathered populations were slower: 0.1122 ms.)
template <typename T>
__device__ __forceinline__ void store_pop(T *p, T v) {
  __builtin_nontemporal_store(v, p);
}

template <int S, int Q>
__device__ __forceinline__ typename Store<S>::R decode(typename Store<S>::T v, int i) {
  if constexpr (S == kF16)
    return static_cast<float>(v) * static_cast<float>(Lattice<Q>::w(i) / kF16Scale);
  else if constexpr (S == kF16C)
    return f16c_scaled(v) * static_cast<float>(Lattice<Q>::w(i) * kF16CUp);
  else
    return v;
}

template <int S, int Q>
__device__ __forceinline__ typename Store<S>::T encode(typename Store<S>::R v, int i) {
  if constexpr (S == kF16)
    return static_cast<_Float16>(v * static_cast<float>(kF16Scale * Lattice<Q>::invw(i)));
  else if constexpr (S == kF16C)
    return f16c_from_scaled(v * static_cast<float>(Lattice<Q>::invw(i) * kF16CDown));
  else
    return v;
}

// Population f_i in FP64 from a stored element.
template <int S, int Q>
__device__ __forceinline__ double to_f64(typename Store<S>::T v, int i, double rho0) {
  if constexpr (S == kF64)
    return v;
  else
    return Lattice<Q>::w(i) * rho0 + static_cast<double>(decode<S, Q>(v, i));
}

// Stored element from a population f_i in FP64.
template <int S, int Q>
__device__ __forceinline__ typename Store<S>::T from_f64(double f, int i, double rho0) {
  if constexpr (S == kF64)
    return f;
  else
    return encode<S, Q>(static_cast<float>(f - Lattice<Q>::w(i) * rho0), i);
}

struct StepArgs {
  const void *fin;
  void *fout;
  const void *Ffield;  // (N, d): double for FP64 storage, float otherwise
  void *rho_side;      // (N): rho of the last collision at link-box nodes (compute type)
  void *hydro;         // the moments the keeping modes write (see StepMode), else null
  BoxGeom g;
  CollParams c;
  Tiling t;
  StepParticles sp;    // particles (sp.np == 0: none)
};

// ------------------------------------------------------------ collision
// FP64 storage: the reference's formulas in the reference's order of
// operations, without FMA contraction, so the step reproduces the FP64
// CPU restatement to the last bit wherever both sum in the same order.
template <int Q, int COLL, int FM>
__host__ __device__ inline void collide_ref(const double (&f)[Q], const double *Fn,
                                            const CollParams &c, double (&fs)[Q],
                                            double &rho_out, double *u_out = nullptr) {
#pragma clang fp contract(off)
  using L = Lattice<Q>;
  constexpr int D = L::D;
  // hydroVariablesUpdate! (gpu/fluids.jl:7-37)
  double rho = f[0];
#pragma unroll
  for (int q = 1; q < Q; q++) rho = rho + f[q];
  // Terms whose velocity component is 0 are left out of the sums here and
  // below: each would add a signed zero to a sum that starts at +0 and so
  // changes no bit of it (the compiler may not drop them itself, since f *
  // 0.0 is not 0 for every f).
  double mom[3];
#pragma unroll
  for (int k = 0; k < D; k++) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int ck = k == 0 ? L::cx[q] : (k == 1 ? L::cy[q] : L::cz[q]);
      if (ck != 0) acc = acc + f[q] * static_cast<double>(ck);
    }
    mom[k] = acc;
  }
  double F[3] = {0.0, 0.0, 0.0};
  if constexpr (FM != LBM_FORCE_NONE) {
#pragma unroll
    for (int k = 0; k < D; k++) F[k] = Fn[k];
    const double half_dt = 0.5 * c.dt;
#pragma unroll
    for (int k = 0; k < D; k++) mom[k] = mom[k] + half_dt * F[k];
  }
  double u[3];
#pragma unroll
  for (int k = 0; k < D; k++) u[k] = rho != 0.0 ? mom[k] / rho : 0.0;
  rho_out = rho;
  if (u_out)
    for (int k = 0; k < 3; k++) u_out[k] = k < D ? u[k] : 0.0;
  // equilibriumDistribution (gpu/fluids.jl:39-52)
  double uu = 0.0;
#pragma unroll
  for (int k = 0; k < D; k++) uu = uu + u[k] * u[k];
  double fe[Q], udc_[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int cq[3] = {L::cx[q], L::cy[q], L::cz[q]};
    const double ci[3] = {cq[0] * c.ls, cq[1] * c.ls, cq[2] * c.ls};
    double udc = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++)
      if (cq[k] != 0) udc = udc + u[k] * ci[k];
    udc_[q] = udc;
    const double first = udc * c.invC2s + (udc * udc) * (0.5 * c.invC4s);
    const double second = first - uu * (0.5 * c.invC2s);
    // compressible: w ((second + 1) rho); incompressible: w rho + w (rho0 second)
    fe[q] = c.incompressible ? L::w(q) * rho + L::w(q) * (c.rho0 * second) : L::w(q) * ((second + 1.0) * rho);
  }
  // collisionStep (gpu/fluids.jl:58-101) with guoForcingTerm (:103-127)
  const double cb = -c.dt / c.tau;
  const double cm = -c.dt / c.tau_minus;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    double om;
    if constexpr (COLL == LBM_COLLISION_BGK) {
      om = cb * (f[q] - fe[q]);
    } else {
      const int o = L::opp[q];
      const double fp = 0.5 * (f[q] + f[o]), fm = 0.5 * (f[q] - f[o]);
      const double ep = 0.5 * (fe[q] + fe[o]), em = 0.5 * (fe[q] - fe[o]);
      om = cb * (fp - ep) + cm * (fm - em);
    }
    if constexpr (FM != LBM_FORCE_NONE) {
      // (with c_qk = 0, the c_qk and (u.c) c_qk terms are signed zeros added
      // to nonzero or +0 values: left out)
      const int cq[3] = {L::cx[q], L::cy[q], L::cz[q]};
      const double ci[3] = {cq[0] * c.ls, cq[1] * c.ls, cq[2] * c.ls};
      const double udc = udc_[q];
      double second[3];
#pragma unroll
      for (int k = 0; k < D; k++) second[k] = (udc * ci[k]) * c.invC4s;
      double S;
      if constexpr (COLL == LBM_COLLISION_BGK) {
        double inter = 0.0;
#pragma unroll
        for (int k = 0; k < D; k++) {
          const double a = cq[k] != 0 ? (-u[k] + ci[k]) * c.invC2s + second[k] : (-u[k]) * c.invC2s;
          inter = inter + a * F[k];
        }
        S = L::w(q) * c.dt * (1.0 - c.dt / (2.0 * c.tau)) * inter;
      } else {
        double inter = 0.0;
#pragma unroll
        for (int k = 0; k < D; k++) {
          const double a = cq[k] != 0 ? (-u[k]) * c.invC2s + second[k] : (-u[k]) * c.invC2s;
          inter = inter + a * F[k];
        }
        const double plus = (1.0 - c.dt / (2.0 * c.tau)) * inter;
        double inter2 = 0.0;
#pragma unroll
        for (int k = 0; k < D; k++)
          if (cq[k] != 0) inter2 = inter2 + F[k] * (ci[k] * c.invC2s);
        const double minus = (1.0 - c.dt / (2.0 * c.tau_minus)) * inter2;
        S = L::w(q) * c.dt * (plus + minus);
      }
      om = om + S;
    }
    fs[q] = om + f[q];
  }
}

// FP32/FP16 storage: the same collision in deviation variables d_i =
// f_i - w_i rho0, computed in FP32 (Sum_i w_i = 1, Sum_i w_i c_i = 0 make
// the shift drop out of every moment).
template <int Q, int COLL, int FM>
__device__ __forceinline__ void collide_dev(const float (&d)[Q], const float *Fn,
                                            const CollParams &c, float (&ds)[Q],
                                            float &rho_out, float *u_out = nullptr) {
  using L = Lattice<Q>;
  constexpr int D = L::D;
  const float ls = c.f_ls;
  const float invC2s = c.f_invC2s;
  const float invC4s = c.f_invC4s;
  float drho = 0.0f;
#pragma unroll
  for (int q = 0; q < Q; q++) drho += d[q];
  float j[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int q = 0; q < Q; q++) {
    if (L::cx[q]) j[0] += L::cx[q] > 0 ? d[q] : -d[q];
    if (L::cy[q]) j[1] += L::cy[q] > 0 ? d[q] : -d[q];
    if (L::cz[q]) j[2] += L::cz[q] > 0 ? d[q] : -d[q];
  }
  float F[3] = {0.0f, 0.0f, 0.0f};
  if constexpr (FM != LBM_FORCE_NONE) {
#pragma unroll
    for (int k = 0; k < D; k++) F[k] = Fn[k];
    const float half_dt = c.f_half_dt;
#pragma unroll
    for (int k = 0; k < D; k++) j[k] += half_dt * F[k];
  }
  const float rho = c.f_rho0 + drho;
  const float inv = rho != 0.0f ? 1.0f / rho : 0.0f;
  float u[3];
#pragma unroll
  for (int k = 0; k < 3; k++) u[k] = j[k] * inv;
  rho_out = rho;
  if (u_out)
    for (int k = 0; k < 3; k++) u_out[k] = k < D ? u[k] : 0.0f;
  // density multiplying the velocity terms of the equilibrium
  const float rv = c.incompressible ? c.f_rho0 : rho;
  float uu = 0.0f;
#pragma unroll
  for (int k = 0; k < D; k++) uu += u[k] * u[k];
  const float dt = c.f_dt;
  const float wp = c.f_wp;
  const float wm = COLL == LBM_COLLISION_TRT ? c.f_wm : wp;
  const float gp = 1.0f - 0.5f * wp, gm = 1.0f - 0.5f * wm;
  float uF = 0.0f;
  if constexpr (FM != LBM_FORCE_NONE) {
#pragma unroll
    for (int k = 0; k < D; k++) uF += u[k] * F[k];
  }
  const float base = -0.5f * invC2s * uu;  // -u.u/(2 c_s^2)
  // Rest population (self-conjugate)
  {
    const float w0 = static_cast<float>(L::w(0));
    const float dfe = w0 * (drho + rv * base);
    float S = 0.0f;
    if constexpr (FM != LBM_FORCE_NONE) S = w0 * dt * gp * (-uF * invC2s);
    ds[0] = d[0] - wp * (d[0] - dfe) + S;
  }
  // Conjugate pairs (q, q̄) with q < q̄: symmetric and antisymmetric parts.
#pragma unroll
  for (int q = 1; q < Q; q++) {
    const int o = L::opp[q];
    if (o < q) continue;
    const float w = static_cast<float>(L::w(q));
    const float cu = ls * (L::cx[q] * u[0] + L::cy[q] * u[1] + (D == 3 ? L::cz[q] * u[2] : 0.0f));
    const float esym = w * (drho + rv * (0.5f * invC4s * cu * cu + base));
    const float eanti = w * rv * invC2s * cu;
    float Ssym = 0.0f, Santi = 0.0f;
    if constexpr (FM != LBM_FORCE_NONE) {
      const float cF =
          ls * (L::cx[q] * F[0] + L::cy[q] * F[1] + (D == 3 ? L::cz[q] * F[2] : 0.0f));
      Ssym = w * dt * gp * (-uF * invC2s + cu * cF * invC4s);
      Santi = w * dt * gm * (cF * invC2s);
    }
    const float psym = 0.5f * (d[q] + d[o]) - esym;
    const float panti = 0.5f * (d[q] - d[o]) - eanti;
    const float dsym = -wp * psym + Ssym;
    const float danti = -wm * panti + Santi;
    ds[q] = d[q] + dsym + danti;
    ds[o] = d[o] + dsym - danti;
  }
}

// 16-bit storage (FP16 or the custom format): the same collision on the
// stored values h_q = d_q / s_q directly (FP16: s_q = w_q / kF16Scale, read
// by mixed-precision FMAs as they are; custom format: h_q * 2^-108 as FP32
// from the bit pattern, s_q = w_q 2^108): the moments are sums of the stored
// values, and each population is updated as h_q + delta_q / s_q and rounded
// once to the storage format.  The FMAs are written out and contraction is
// off, so the result is the same bits wherever the routine is inlined.
template <int S>
__device__ __forceinline__ float h16_value(typename Store<S>::T v) {
  if constexpr (S == kF16C)
    return f16c_scaled(v);
  else
    return static_cast<float>(v);
}
template <int S>
__device__ __forceinline__ typename Store<S>::T h16_store(float v) {
  if constexpr (S == kF16C)
    return f16c_from_scaled(v);
  else
    return static_cast<_Float16>(v);
}

// Whether some velocity of the set has weight class k.
template <int Q>
constexpr bool lattice_has_class(int k) {
  for (int i = 0; i < Q; i++)
    if (Lattice<Q>::wclass[i] == k) return true;
  return false;
}

template <int Q, int COLL, int FM, int SK = kF16>
__device__ __forceinline__ void collide_h16(const typename Store<SK>::T (&h)[Q], const float *Fn, const CollParams &c,
                                            typename Store<SK>::T (&hs)[Q], float &rho_out, float *u_out = nullptr) {
#pragma clang fp contract(off)
  using L = Lattice<Q>;
  constexpr int D = L::D, NC = L::NC;
  // stored unit -> deviation (s) and back (is), per weight class
  constexpr double kUnit = SK == kF16C ? kF16CUp : 1.0 / kF16Scale;
  constexpr float sc[4] = {static_cast<float>(L::wc(0) * kUnit), static_cast<float>(L::wc(1) * kUnit),
                           static_cast<float>(L::wc(NC > 2 ? 2 : 1) * kUnit),
                           static_cast<float>(L::wc(NC > 3 ? 3 : 1) * kUnit)};
  constexpr float isc[4] = {static_cast<float>(L::invwc(0) / kUnit), static_cast<float>(L::invwc(1) / kUnit),
                            static_cast<float>(L::invwc(NC > 2 ? 2 : 1) / kUnit),
                            static_cast<float>(L::invwc(NC > 3 ? 3 : 1) / kUnit)};
  const float ls = c.f_ls;
  const float invC2s = c.f_invC2s;
  const float invC4s = c.f_invC4s;
  // class sums of the stored values and of their first moments
  float Sc[4] = {0.0f, 0.0f, 0.0f, 0.0f}, Ac[4][3] = {};
#pragma unroll
  for (int q = 1; q < Q; q++) {
    const float v = h16_value<SK>(h[q]);
    const int cq[3] = {L::cx[q], L::cy[q], L::cz[q]};
    const int k = L::wclass[q];
    Sc[k] = fmaf(v, 1.0f, Sc[k]);
#pragma unroll
    for (int a = 0; a < D; a++)
      if (cq[a]) Ac[k][a] = fmaf(v, cq[a] > 0 ? 1.0f : -1.0f, Ac[k][a]);
  }
  const float h0 = h16_value<SK>(h[0]);
  float drho = sc[0] * h0;
#pragma unroll
  for (int k = 1; k < 4; k++)
    if (lattice_has_class<Q>(k)) drho = fmaf(sc[k], Sc[k], drho);
  float j[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int a = 0; a < D; a++) {
    bool first = true;
#pragma unroll
    for (int k = 1; k < 4; k++)
      if (lattice_has_class<Q>(k)) {
        j[a] = first ? sc[k] * Ac[k][a] : fmaf(sc[k], Ac[k][a], j[a]);
        first = false;
      }
  }
  float F[3] = {0.0f, 0.0f, 0.0f};
  if constexpr (FM != LBM_FORCE_NONE) {
    const float half_dt = c.f_half_dt;
#pragma unroll
    for (int k = 0; k < D; k++) {
      F[k] = Fn[k];
      j[k] = fmaf(half_dt, F[k], j[k]);
    }
  }
  const float rho = c.f_rho0 + drho;
  const float inv = rho != 0.0f ? 1.0f / rho : 0.0f;
  float u[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int k = 0; k < D; k++) u[k] = j[k] * inv;
  rho_out = rho;
  if (u_out)
    for (int k = 0; k < 3; k++) u_out[k] = u[k];
  // density multiplying the velocity terms of the equilibrium
  const float rv = c.incompressible ? c.f_rho0 : rho;
  float uu = 0.0f;
#pragma unroll
  for (int k = 0; k < D; k++) uu = fmaf(u[k], u[k], uu);
  const float dt = c.f_dt;
  const float wp = c.f_wp;
  const float wm = COLL == LBM_COLLISION_TRT ? c.f_wm : wp;
  const float gp = 1.0f - 0.5f * wp, gm = 1.0f - 0.5f * wm;
  float uF = 0.0f;
  if constexpr (FM != LBM_FORCE_NONE) {
#pragma unroll
    for (int k = 0; k < D; k++) uF = fmaf(u[k], F[k], uF);
  }
  const float base = (-0.5f * invC2s) * uu;  // -u.u/(2 c_s^2)
  {
    // rest population (self-conjugate)
    const float w0 = static_cast<float>(L::w(0));
    const float dfe = w0 * fmaf(rv, base, drho);
    float S = 0.0f;
    if constexpr (FM != LBM_FORCE_NONE) S = (w0 * dt * gp) * (-uF * invC2s);
    const float delta = fmaf(-wp, fmaf(sc[0], h0, -dfe), S);
    hs[0] = h16_store<SK>(fmaf(delta, isc[0], h0));
  }
  // conjugate pairs (q, q̄), q < q̄: symmetric and antisymmetric parts
#pragma unroll
  for (int q = 1; q < Q; q++) {
    const int o = L::opp[q];
    if (o < q) continue;
    const int k = L::wclass[q];
    const float w = static_cast<float>(L::w(q));
    const float sq = sc[k], isq = isc[k];
    float cu = 0.0f;
    if (L::cx[q]) cu = L::cx[q] > 0 ? cu + u[0] : cu - u[0];
    if (L::cy[q]) cu = L::cy[q] > 0 ? cu + u[1] : cu - u[1];
    if (D == 3 && L::cz[q]) cu = L::cz[q] > 0 ? cu + u[2] : cu - u[2];
    cu = ls * cu;
    const float esym = w * fmaf(rv, fmaf(0.5f * invC4s, cu * cu, base), drho);
    const float eanti = (w * rv * invC2s) * cu;
    float Ssym = 0.0f, Santi = 0.0f;
    if constexpr (FM != LBM_FORCE_NONE) {
      float cF = 0.0f;
      if (L::cx[q]) cF = L::cx[q] > 0 ? cF + F[0] : cF - F[0];
      if (L::cy[q]) cF = L::cy[q] > 0 ? cF + F[1] : cF - F[1];
      if (D == 3 && L::cz[q]) cF = L::cz[q] > 0 ? cF + F[2] : cF - F[2];
      cF = ls * cF;
      Ssym = (w * dt * gp) * fmaf(cu * cF, invC4s, -uF * invC2s);
      Santi = (w * dt * gm) * (cF * invC2s);
    }
    const float hq = h16_value<SK>(h[q]), ho = h16_value<SK>(h[o]);
    const float psym = fmaf(0.5f * sq, hq + ho, -esym);
    const float panti = fmaf(0.5f * sq, hq - ho, -eanti);
    const float dsym = fmaf(-wp, psym, Ssym);
    const float danti = fmaf(-wm, panti, Santi);
    hs[q] = h16_store<SK>(fmaf(dsym + danti, isq, hq));
    hs[o] = h16_store<SK>(fmaf(dsym - danti, isq, ho));
  }
}

// Collision from stored values to stored values (the tiles' path); u_out,
// when given, receives the velocity the equilibrium was taken at.
template <int Q, int S, int COLL, int FM>
__device__ __forceinline__ void collide_stored(const typename Store<S>::T (&raw)[Q], const typename Store<S>::R *Fn,
                                               const CollParams &c, typename Store<S>::T (&out)[Q],
                                               typename Store<S>::R &rho, typename Store<S>::R *u_out = nullptr) {
  using R = typename Store<S>::R;
  if constexpr (S == kF16 || S == kF16C) {
    collide_h16<Q, COLL, FM, S>(raw, Fn, c, out, rho, u_out);
  } else {
    R f[Q], fs[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) f[q] = decode<S, Q>(raw[q], q);
    if constexpr (S == kF64)
      collide_ref<Q, COLL, FM>(f, Fn, c, fs, rho, u_out);
    else
      collide_dev<Q, COLL, FM>(f, Fn, c, fs, rho, u_out);
#pragma unroll
    for (int q = 0; q < Q; q++) out[q] = encode<S, Q>(fs[q], q);
  }
}

// ------------------------------------------------------- node geometry
struct NodeIdx {
  int x, y, z;
};

__device__ __forceinline__ int wrapi(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// wallNodes (gpu/misc.jl:7-31): the first and last layer of each walled
// dimension; with a wall mask, its bit 0 (which includes those layers).
__device__ __forceinline__ bool is_wall_node(const BoxGeom &g, int x, int y, int z) {
  if (g.walled & kWallMask) return g.wmask[(int64_t)x + g.nx * ((int64_t)y + (int64_t)g.ny * z)] & 1u;
  return ((g.walled & 1) && (x == 0 || x == g.nx - 1)) ||
         ((g.walled & 2) && (y == 0 || y == g.ny - 1)) ||
         ((g.walled & 4) && (z == 0 || z == g.nz - 1));
}

// The wall bits of node (x, y, z) in the layout of BoxGeom::wmask: bit 0 the
// node is a wall, bit q population q bounces back (x - c_q is a wall and x
// is not), kMovingLinkBit a moving wall among them.  Without a mask, from
// the walled layers (halfway bounce-back next to them).
template <int Q>
__device__ __forceinline__ uint32_t wall_bits(const BoxGeom &g, int x, int y, int z) {
  using L = Lattice<Q>;
  if (g.walled & kWallMask) return g.wmask[(int64_t)x + g.nx * ((int64_t)y + (int64_t)g.ny * z)];
  if (is_wall_node(g, x, y, z)) return 1u;
  const bool xl = (g.walled & 1) && x == 1, xh = (g.walled & 1) && x == g.nx - 2;
  const bool yl = (g.walled & 2) && y == 1, yh = (g.walled & 2) && y == g.ny - 2;
  const bool zl = (g.walled & 4) && z == 1, zh = (g.walled & 4) && z == g.nz - 2;
  uint32_t m = 0;
#pragma unroll
  for (int q = 1; q < Q; q++) {
    const int cx = L::cx[q], cy = L::cy[q], cz = L::cz[q];
    const bool bb = (cx > 0 && xl) || (cx < 0 && xh) || (cy > 0 && yl) || (cy < 0 && yh) || (cz > 0 && zl) ||
                    (cz < 0 && zh);
    m |= (uint32_t)bb << q;
  }
  return m;
}

// The moving-wall term of a bounced population q at fluid node (x, y, z):
// 2 w_q / c_s^2 rho(x) (u_w(x - c_q) . c_q) (gpu/fluids.jl:168-174), in the
// reference's order of operations; rho is the node's massDensity.
template <int Q, typename R>
__device__ __forceinline__ R moving_wall_term(const BoxGeom &g, const CollParams &c, int q, int x, int y, int z,
                                             R rho) {
#pragma clang fp contract(off)
  using L = Lattice<Q>;
  constexpr int D = L::D;
  const int xs = wrapi(x - L::cx[q], g.nx), ys = wrapi(y - L::cy[q], g.ny), zs = wrapi(z - L::cz[q], g.nz);
  const int64_t nw = (int64_t)xs + g.nx * ((int64_t)ys + (int64_t)g.ny * zs);
  const R *uw = static_cast<const R *>(g.uwall);
  const int ck[3] = {L::cx[q], L::cy[q], L::cz[q]};
  double udc = 0.0;
#pragma unroll
  for (int k = 0; k < D; k++) udc = udc + static_cast<double>(uw[k * g.N + nw]) * (ck[k] * c.ls);
  const double lc = 2.0 * L::w(q) * c.invC2s;
  return static_cast<R>((lc * static_cast<double>(rho)) * udc);
}

// (X + (-R))_1^2 + ... at node (x, y, z) (gpu/kernels.jl:344-353)
template <int D>
__device__ __forceinline__ double dist2_at(const double *const coords[3], const double Rg[3], int x,
                                           int y, int z) {
#pragma clang fp contract(off)
  const double a = coords[0][x] + (-Rg[0]);
  const double b = coords[1][y] + (-Rg[1]);
  double s = a * a + b * b;
  if constexpr (D == 3) {
    const double c = coords[2][z] + (-Rg[2]);
    s = s + c * c;
  }
  return s;
}

// getSphere: (X + (-R))_1^2 + ... < r^2
template <int D>
__device__ __forceinline__ bool solid_at(const double *const coords[3], const double Rg[3],
                                         double r2, int x, int y, int z) {
#pragma clang fp contract(off)
  const double a = coords[0][x] + (-Rg[0]);
  const double b = coords[1][y] + (-Rg[1]);
  double s = a * a + b * b;
  if constexpr (D == 3) {
    const double c = coords[2][z] + (-Rg[2]);
    s = s + c * c;
  }
  return s < r2;
}

// ------------------------------------------------------------ particles
// Node velocity of particle p at the solid node with coordinates xc:
// getNodeVelocity (gpu/aux.jl:229-263) or bulkVelocity
// (cpu/particles.jl:20-43), from the node-velocity snapshot.
template <int D>
__host__ __device__ inline void node_velocity(const double (&xc)[3], const DevParticle &p, int semantics,
                                              double cutoff, double (&out)[3]) {
#pragma clang fp contract(off)
  double r[3] = {0.0, 0.0, 0.0};
  if (semantics == LBM_SEMANTICS_GPU) {
#pragma unroll
    for (int k = 0; k < D; k++) r[k] = xc[k] + (-p.nR[k]);
  } else {
#pragma unroll
    for (int k = 0; k < D; k++) r[k] = xc[k] - p.nR[k];
  }
  double nn = 0.0;
#pragma unroll
  for (int k = 0; k < D; k++) nn = nn + r[k] * r[k];
  nn = sqrt(nn);
  const double *V = p.nV, *W = p.nW, *e = p.ne;
  double bulk[3];
  if (semantics == LBM_SEMANTICS_GPU) {
    if constexpr (D == 2) {
      bulk[0] = V[0] - W[0] * r[1];
      bulk[1] = V[1] + W[0] * r[0];
    } else {
      bulk[0] = V[0] + W[1] * r[2] - W[2] * r[1];
      bulk[1] = V[1] - W[0] * r[2] + W[2] * r[0];
      bulk[2] = V[2] + W[0] * r[1] - W[1] * r[0];
    }
    if (nn == 0.0) nn = 1.0;
    if (p.kind == LBM_PARTICLE_BEAD) {
#pragma unroll
      for (int k = 0; k < D; k++) out[k] = bulk[k];
      return;
    }
    const double inv = 1.0 / nn;
    double unit[3], edr = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) unit[k] = inv * r[k];
#pragma unroll
    for (int k = 0; k < D; k++) edr = edr + unit[k] * e[k];
    const double first = p.B1 + p.B2 * edr;
#pragma unroll
    for (int k = 0; k < D; k++) out[k] = first * (edr * unit[k] + (-e[k])) + bulk[k];
  } else {
#pragma unroll
    for (int k = 0; k < D; k++) out[k] = 0.0;
    // nodes farther than 2 dx from the surface keep zero velocity
    // (cpu/particles.jl:25); link nodes never are
    if (fabs(nn - p.radius) > cutoff) return;
    if constexpr (D == 2) {
      bulk[0] = V[0] + (0.0 * 0.0 - W[0] * r[1]);
      bulk[1] = V[1] + (-0.0 * 0.0 + W[0] * r[0]);
    } else {
      bulk[0] = V[0] + (W[1] * r[2] - W[2] * r[1]);
      bulk[1] = V[1] + (-W[0] * r[2] + W[2] * r[0]);
      bulk[2] = V[2] + (W[0] * r[1] - W[1] * r[0]);
    }
    if (p.kind == LBM_PARTICLE_BEAD || nn == 0.0) {
#pragma unroll
      for (int k = 0; k < D; k++) out[k] = bulk[k];
      return;
    }
    double fr[3], edr = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) fr[k] = r[k] / nn;
#pragma unroll
    for (int k = 0; k < D; k++) edr = edr + e[k] * fr[k];
    const double first = p.B1 + p.B2 * edr;
#pragma unroll
    for (int k = 0; k < D; k++) out[k] = first * (edr * fr[k] - e[k]) + bulk[k];
  }
}

__device__ __forceinline__ bool in_box(const DevParticle &P, int x, int y, int z) {
  return x >= P.lo[0] && x <= P.hi[0] && y >= P.lo[1] && y <= P.hi[1] && z >= P.lo[2] && z <= P.hi[2];
}

// Range along axis k of the nodes the box blocks own for particle P (its
// link box grown by kBoxMargin, clamped to the lattice).
__host__ __device__ inline void owned_range(const DevParticle &P, const BoxGeom &g, int k, int32_t &lo,
                                            int32_t &hi) {
  const int n = k == 0 ? g.nx : (k == 1 ? g.ny : g.nz);
  lo = P.lo[k] - kBoxMargin < 0 ? 0 : P.lo[k] - kBoxMargin;
  hi = P.hi[k] + kBoxMargin > n - 1 ? n - 1 : P.hi[k] + kBoxMargin;
}

__device__ __forceinline__ bool owned(const DevParticle &P, const BoxGeom &g, int x, int y, int z) {
  const int c[3] = {x, y, z};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int32_t lo, hi;
    owned_range(P, g, k, lo, hi);
    if (c[k] < lo || c[k] > hi) return false;
  }
  return true;
}

// Band of nodes the box blocks take for a particle: the nodes of its owned
// region whose distance from its centre as read at the kernel's start (in
// node units) lies in [r - kBandIn, r + kBandOut]; the tiles take every other
// node.  Link fluid nodes lie within [r, r + |c|) of their tick's centre (|c|
// at most sqrt(2), or sqrt(3) on the sets with body diagonals) and a move is
// below half a node, so the band holds this kernel's links and the nodes
// whose rho the next kernel's links read, with a margin either side.  Tiles and box blocks evaluate the same arithmetic on
// the same record, so they agree on every node.  A particle whose record is
// marked whole gives its box blocks its whole owned region instead (radii
// whose band rows would not fit the blocks' LDS).
constexpr double kBandIn = 1.25, kBandOut = 2.75;

struct Band {
  double c[3];         // centre in node units
  double rin2, rout2;  // squared band radii in node units
  int32_t lo[3], hi[3];  // owned region (inclusive)
  int32_t whole;
};

__host__ __device__ inline void band_of(const DevParticle &P, const StepParticles &sp, const BoxGeom &g, Band &b) {
#pragma clang fp contract(off)
  for (int k = 0; k < 3; k++) {
    b.c[k] = k < sp.dims ? (P.Rg[k] - sp.x0[k]) / sp.dx : 0.0;
    owned_range(P, g, k, b.lo[k], b.hi[k]);
  }
  const double r = P.radius / sp.dx;
  const double ri = r - kBandIn > 0.0 ? r - kBandIn : 0.0, ro = r + kBandOut;
  b.rin2 = ri * ri;
  b.rout2 = ro * ro;
  b.whole = P.whole;
}

// The band's nodes on row (y, z): x in [a0, a0 + na) and [b0, b0 + nb).
__host__ __device__ inline void band_row(const Band &b, int y, int z, int &a0, int &na, int &b0, int &nb) {
#pragma clang fp contract(off)
  a0 = b0 = 0;
  na = nb = 0;
  if (y < b.lo[1] || y > b.hi[1] || z < b.lo[2] || z > b.hi[2]) return;
  if (b.whole) {
    a0 = b.lo[0];
    na = b.hi[0] - b.lo[0] + 1;
    return;
  }
  const double dy = y - b.c[1], dz = z - b.c[2];
  const double rho2 = dy * dy + dz * dz;
  if (rho2 > b.rout2) return;
  // (the ends are clamped to the owned row in double precision before they
  // become integers, so that a centre of any value, NaN included, gives a
  // range inside the row)
  const double ao = sqrt(b.rout2 - rho2);
  const double rl = b.lo[0], rh = b.hi[0];
  double lo = ceil(b.c[0] - ao), hi = floor(b.c[0] + ao);
  lo = lo >= rl ? (lo <= rh + 1.0 ? lo : rh + 1.0) : rl;
  hi = hi <= rh ? (hi >= rl - 1.0 ? hi : rl - 1.0) : rh;
  if (rho2 < b.rin2) {
    // minus the hole |x - c| < ai
    const double ai = sqrt(b.rin2 - rho2);
    double l1 = floor(b.c[0] - ai), h0 = ceil(b.c[0] + ai);
    l1 = l1 <= hi ? (l1 >= lo - 1.0 ? l1 : lo - 1.0) : hi;
    h0 = h0 >= lo ? (h0 <= hi + 1.0 ? h0 : hi + 1.0) : lo;
    a0 = (int)lo;
    na = (int)(l1 - lo) + 1;
    b0 = (int)h0;
    nb = (int)(hi - h0) + 1;  // (l1 < c0 < h0: the two never overlap)
  } else {
    a0 = (int)lo;
    na = (int)(hi - lo) + 1;
  }
  na = na > 0 ? na : 0;
  nb = nb > 0 ? nb : 0;
}

__host__ __device__ inline bool in_band(const Band &b, int x, int y, int z) {
  int a0, na, b0, nb;
  band_row(b, y, z, a0, na, b0, nb);
  return (x >= a0 && x < a0 + na) || (x >= b0 && x < b0 + nb);
}

// Node gi of a band region of ey x ez rows, from the tables: plane_pre[k] =
// band nodes of planes below k (ez + 1 entries), row_off[k ey + j] = band
// nodes of plane k's rows below j.  gi must be below plane_pre[ez].
__host__ __device__ inline void band_node(const Band &b, const int32_t *plane_pre, const uint16_t *row_off, int ey,
                                          int ez, int gi, int &x, int &y, int &z) {
  int zl = 0, zh = ez - 1;  // last plane starting at or before gi
  while (zl < zh) {
    const int mid = (zl + zh + 1) >> 1;
    if (plane_pre[mid] <= gi)
      zl = mid;
    else
      zh = mid - 1;
  }
  const int off = gi - plane_pre[zl];
  const uint16_t *ro = row_off + zl * ey;
  int yl = 0, yh = ey - 1;  // last row starting at or before off
  while (yl < yh) {
    const int mid = (yl + yh + 1) >> 1;
    if (ro[mid] <= off)
      yl = mid;
    else
      yh = mid - 1;
  }
  const int k = off - ro[yl];
  y = b.lo[1] + yl;
  z = b.lo[2] + zl;
  int a0, na, b0, nb;
  band_row(b, y, z, a0, na, b0, nb);
  x = k < na ? a0 + k : b0 + (k - na);
}

// Index of population q of node (x, row) in the row-blocked layout, row =
// y + ny z: the Q planes of one x-row are contiguous.
__device__ __forceinline__ int64_t pop_index(const BoxGeom &g, int Q, int q, int x, int row) {
  return ((int64_t)row * Q + q) * g.nx + x;
}

// Row offsets (in elements) of the rows a node gathers from, for every
// (cy, cz) in {-1, 0, 1}^2: source row of (y - cy, z - cz).
struct GatherRows {
  int64_t r[3][3];  // [cy + 1][cz + 1]
};

__device__ __forceinline__ GatherRows gather_rows(const BoxGeom &g, int Q, int y, int z) {
  GatherRows o;
  const int ys[3] = {y == g.ny - 1 ? 0 : y + 1, y, y == 0 ? g.ny - 1 : y - 1};  // y - cy, cy = -1, 0, 1
  const int zs[3] = {z == g.nz - 1 ? 0 : z + 1, z, z == 0 ? g.nz - 1 : z - 1};
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) o.r[a][b] = (int64_t)(ys[a] + g.ny * zs[b]) * Q * g.nx;
  return o;
}

// Kernel modes of one run of ticks (S: the reference's distributions, P:
// post-collision populations): kCollide P_0 = C(S_0); kFused S_k = G(P_{k-1})
// then P_k = C(S_k); kGather S_n = G(P_{n-1}).  G is streaming with halfway
// bounce-back, the Ladd term and the momentum exchange (tick n-1's, which
// also moves the particles), C the collision.  kCollideKeep and kFusedKeep
// are kCollide and kFused that also keep each fluid node's moments (rho and
// the velocity of its equilibrium) in StepArgs::hydro: the adjoint's
// recomputation, whose reverse collision needs nothing else of S_k.
enum StepMode { kCollide = 0, kFused = 1, kGather = 2, kCollideKeep = 3, kFusedKeep = 4 };
constexpr int base_mode(int m) { return m == kCollideKeep ? kCollide : (m == kFusedKeep ? kFused : m); }
constexpr bool keeps_moments(int m) { return m == kCollideKeep || m == kFusedKeep; }

// Collision of a node's populations f (S), into fs (P), with rho.
// The Guo force at node n (zero without forcing).
template <int Q, int S, int FM>
__device__ __forceinline__ void node_force(const StepArgs &a, int64_t n, typename Store<S>::R (&Fn)[3]) {
  using R = typename Store<S>::R;
  constexpr int D = Lattice<Q>::D;
  Fn[0] = Fn[1] = Fn[2] = R(0);
  if constexpr (FM == LBM_FORCE_UNIFORM) {
#pragma unroll
    for (int k = 0; k < D; k++) Fn[k] = static_cast<R>(a.c.F[k]);
  } else if constexpr (FM == LBM_FORCE_FIELD) {
    const R *Ff = static_cast<const R *>(a.Ffield);
#pragma unroll
    for (int k = 0; k < D; k++) Fn[k] = Ff[k * a.g.N + n];
  }
}

template <int Q, int S, int COLL, int FM>
__device__ __forceinline__ void collide_node(const StepArgs &a, int64_t n,
                                             const typename Store<S>::R (&f)[Q],
                                             typename Store<S>::R (&fs)[Q], typename Store<S>::R &rho,
                                             typename Store<S>::R *u_out = nullptr) {
  using R = typename Store<S>::R;
  R Fn[3];
  node_force<Q, S, FM>(a, n, Fn);
  if constexpr (S == kF64)
    collide_ref<Q, COLL, FM>(f, Fn, a.c, fs, rho, u_out);
  else
    collide_dev<Q, COLL, FM>(f, Fn, a.c, fs, rho, u_out);
}

// The moments of node (x, row) into StepArgs::hydro (row-blocked like the
// populations: plane k of row r is rho, u_x, u_y, u_z for k = 0..3).
template <int S>
__device__ __forceinline__ void keep_node_moments(const StepArgs &a, int x, int row, typename Store<S>::R rho,
                                                  const typename Store<S>::R (&u)[3]) {
  using R = typename Store<S>::R;
  R *H = static_cast<R *>(a.hydro) + (int64_t)row * 4 * a.g.nx + x;
  store_pop(&H[0], rho);
#pragma unroll
  for (int k = 0; k < 3; k++) store_pop(&H[(int64_t)(k + 1) * a.g.nx], u[k]);
}

// The stored value of a population that is zero (wall layers).
template <int Q, int S>
__device__ __forceinline__ typename Store<S>::T zero_pop(const CollParams &c, int q) {
  using R = typename Store<S>::R;
  if constexpr (S == kF64)
    return 0.0;
  else
    return encode<S, Q>(static_cast<R>(-Lattice<Q>::w(q) * c.rho0), q);
}

// -------------------------------------------------------------- boxes
// The nodes of every particle's band (see Band) are handled by the first
// blocks of a gathering kernel's grid (dispatched first, spread over all
// XCDs), while the other blocks do the rest of the lattice.  Every box block
// first applies the move left pending
// by the previous box kernel (moveParticles!, from that kernel's partials,
// in a fixed order, so all blocks agree).  A box node then gathers like any
// other, except at particle links: a fluid node x whose neighbour x - c_q is
// solid takes f*_{q̄}(x) (halfway bounce-back) plus the Ladd terms of the
// particles the link belongs to, in particle order (gpu/fluids.jl:157-207);
// each particle's momentum exchange is taken from the running value and
// reduced over the block into one double-double partial per block, which
// the next box kernel (or the finalize kernel after a run) turns into the
// move.
//
// Box block bid of particle pb covers nodes [j T, (j + 1) T) of pb's band
// listed row by row in x-fastest order (j = bid - first_block[pb], T threads
// per block); a node in the bands of several particles belongs to the first.
template <int Q, int S, int COLL, int FM, int MODE>
__device__ __forceinline__ void box_work(const StepArgs &a, int bid) {
  using L = Lattice<Q>;
  using T = typename Store<S>::T;
  using R = typename Store<S>::R;
  constexpr int D = L::D;
  const BoxGeom &g = a.g;
  const StepParticles &sp = a.sp;
  const int nthreads = blockDim.x * blockDim.y * blockDim.z;
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);

  // The records of the particles this block works with (its local set), in
  // LDS (read throughout), with the regions this kernel's box blocks own
  // (from the records as read).  With at most kLocalParticles particles the
  // local set is all of them, and the previous box kernel's move is applied
  // here by every box block alike; with more, the move kernel has moved them
  // already, and the local set is the particles whose owned regions meet
  // this block's particle's, in id order (local index j < lpb: id < pb).
  __shared__ DevParticle prec[kLocalParticles];
  __shared__ Band bands[kLocalParticles];
  __shared__ int32_t lid[kLocalParticles];  // id of local particle j
  __shared__ int32_t nloc_s, lpb_s;
  int pb = 0;
  while (pb + 1 < sp.np && bid >= sp.first_block[pb + 1]) pb++;
  // the records and the previous kernel's partials load together
  const DD pre = sp.many ? DD{0.0, 0.0} : prefetch_partial(sp);
  if (!sp.many) {
    if (tid < sp.np) lid[tid] = tid;
    if (tid == 0) {
      nloc_s = sp.np;
      lpb_s = pb;
    }
  } else if (tid < 64) {
    // (one wave decides, over at most 64 particles)
    bool meet = false;
    if (tid < sp.np) {
      meet = true;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        int32_t lo, hi, lo_b, hi_b;
        owned_range(sp.parts_in[tid], g, k, lo, hi);
        owned_range(sp.parts_in[pb], g, k, lo_b, hi_b);
        meet = meet && lo <= hi_b && lo_b <= hi;
      }
    }
    const uint64_t m = __ballot(meet);
    const int j = __popcll(m & ((uint64_t(1) << tid) - 1));
    if (meet && j < kLocalParticles) lid[j] = tid;
    if (tid == 0) {
      const int cnt = __popcll(m);
      const int at = __popcll(m & ((uint64_t(1) << pb) - 1));
      nloc_s = cnt < kLocalParticles ? cnt : kLocalParticles;
      // too many particles reach this block: the host reports it after the
      // run (the block still works, on the first kLocalParticles of them)
      if (cnt > kLocalParticles) atomicOr(sp.flags, 1);
      lpb_s = at < nloc_s ? at : 0;
    }
  }
  __syncthreads();
  const int nloc = nloc_s, lpb = lpb_s;
  {
    static_assert(sizeof(DevParticle) % 8 == 0, "records are copied as 8-byte words");
    constexpr int kWords = sizeof(DevParticle) / 8;
    uint64_t *dst = reinterpret_cast<uint64_t *>(prec);
    for (int w = tid; w < nloc * kWords; w += nthreads)
      dst[w] = reinterpret_cast<const uint64_t *>(sp.parts_in + lid[w / kWords])[w % kWords];
  }
  __syncthreads();
  // pb's owned region (from the records as read) and the coordinates to
  // stage for it, loaded now so that they arrive during the move
  int bl[3], ext[3];
  for (int k = 0; k < 3; k++) {
    int32_t lo, hi;
    owned_range(prec[lpb], g, k, lo, hi);
    bl[k] = lo;
    ext[k] = hi - lo + 1;
  }
  constexpr int kCoordRegs = 4;
  double cval[kCoordRegs];
  {
    const int len[3] = {ext[0] + 2, ext[1] + 2, D == 3 ? ext[2] + 2 : 1};
    const int nn[3] = {g.nx, g.ny, g.nz};
    const int tot3 = len[0] + len[1] + len[2];
#pragma unroll
    for (int j = 0; j < kCoordRegs; j++) {
      const int i = tid + j * nthreads;
      cval[j] = 0.0;
      if (i < tot3) {
        const int k = i < len[0] ? 0 : (i < len[0] + len[1] ? 1 : 2);
        const int ii = i - (k == 0 ? 0 : (k == 1 ? len[0] : len[0] + len[1]));
        int gi = bl[k] - 1 + ii;
        gi = gi < 0 ? gi + nn[k] : (gi >= nn[k] ? gi - nn[k] : gi);
        cval[j] = sp.coords[k][gi];
      }
    }
  }
  if (tid < nloc) band_of(prec[tid], sp, g, bands[tid]);  // (before the move rewrites prec)
  if (!sp.many) apply_pending_moves(sp, prec, &pre, bid == 0);  // (synchronises the block)
  else __syncthreads();

  const Band &B = bands[lpb];
  const int ex = ext[0], ey = ext[1], ez = ext[2];

  // Dynamic LDS: coordinate stage (3 axes of coord_stride doubles), link
  // list (2 T slots of f*, value and key), rho and node of each lane, and the
  // band's row offsets (coord_stride^2) and plane offsets (coord_stride + 1).
  extern __shared__ double lds_dyn[];
  const int cstride = sp.coord_stride;
  const int cap = 2 * nthreads;
  double *cst = lds_dyn;
  R *fs_l = reinterpret_cast<R *>(cst + 3 * cstride);  // f*_{q̄}(x) of the slot
  R *v_l = fs_l + cap;                               // streamed value of the slot so far
  R *rho_l = v_l + cap;                              // rho(x) of each lane's node
  uint32_t *pk_l = reinterpret_cast<uint32_t *>(rho_l + nthreads);
  uint32_t *node_l = pk_l + cap;                     // lane's node - bl, 10 bits per axis
  int32_t *plane_pre = reinterpret_cast<int32_t *>(node_l + nthreads);
  uint16_t *row_off = reinterpret_cast<uint16_t *>(plane_pre + cstride + 1);

  // Band nodes per row of pb's region, then offsets within each plane and of
  // each plane (a whole region needs no table).
  const int nrows = B.whole ? 0 : ey * ez;
  for (int r = tid; r < nrows; r += nthreads) {
    int a0, na, b0, nb;
    band_row(B, bl[1] + r % ey, bl[2] + r / ey, a0, na, b0, nb);
    row_off[r] = (uint16_t)(na + nb);
  }
  __syncthreads();
  if (!B.whole && tid < ez) {
    int sum = 0;
    for (int j = 0; j < ey; j++) {
      const int c = row_off[tid * ey + j];
      row_off[tid * ey + j] = (uint16_t)sum;
      sum += c;
    }
    plane_pre[tid + 1] = sum;
  }
  __syncthreads();
  if (!B.whole && tid == 0) {
    plane_pre[0] = 0;
    for (int k = 1; k <= ez; k++) plane_pre[k] += plane_pre[k - 1];
  }
  __syncthreads();
  // this lane's band node: plane, then row, then place in the row
  const int gi = (bid - sp.first_block[pb]) * nthreads + tid;
  int x = bl[0], y = bl[1], z = bl[2];
  bool mine = gi < (B.whole ? ex * ey * ez : plane_pre[ez]);
  if (mine && B.whole) {
    x = bl[0] + gi % ex;
    y = bl[1] + (gi / ex) % ey;
    z = bl[2] + gi / (ex * ey);
    for (int j = 0; j < lpb; j++)
      if (in_band(bands[j], x, y, z)) mine = false;
  } else if (mine) {
    band_node(B, plane_pre, row_off, ey, ez, gi, x, y, z);
    // (a node off its region would mean inconsistent tables: never touch it)
    if (x < bl[0] || x > B.hi[0]) {
      x = bl[0];
      mine = false;
    }
    for (int j = 0; j < lpb; j++)
      if (in_band(bands[j], x, y, z)) mine = false;
  }
  node_l[tid] = (uint32_t)(x - bl[0]) | (uint32_t)(y - bl[1]) << 10 | (uint32_t)(z - bl[2]) << 20;
  // box block 0 hands the moved records to the next kernel
  if (!sp.many && bid == 0) {
    constexpr int kWords = sizeof(DevParticle) / 8;
    for (int w = tid; w < sp.np * kWords; w += nthreads) {
      uint64_t v = reinterpret_cast<const uint64_t *>(prec)[w];
      reinterpret_cast<uint64_t *>(sp.parts_out)[w] = v;
    }
  }

  // Coordinates of pb's owned region and one node around it, in LDS: axis k
  // from index bl_k - 1 (periodic) at cst[k][0] (the loads above; a region
  // too long for them loads the rest here).
  {
    const int len[3] = {ex + 2, ey + 2, D == 3 ? ez + 2 : 1};
    const int nn[3] = {g.nx, g.ny, g.nz};
    const int tot3 = len[0] + len[1] + len[2];
    for (int i = tid; i < tot3; i += nthreads) {
      const int k = i < len[0] ? 0 : (i < len[0] + len[1] ? 1 : 2);
      const int ii = i - (k == 0 ? 0 : (k == 1 ? len[0] : len[0] + len[1]));
      const int j = (i - tid) / nthreads;
      double v;
      if (j < kCoordRegs) {
        v = cval[0];
#pragma unroll
        for (int jj = 1; jj < kCoordRegs; jj++)
          if (j == jj) v = cval[jj];
      } else {
        int gi = bl[k] - 1 + ii;
        gi = gi < 0 ? gi + nn[k] : (gi >= nn[k] ? gi - nn[k] : gi);
        v = sp.coords[k][gi];
      }
      cst[k * cstride + ii] = v;
    }
  }
  __syncthreads();
  // this kernel's partials are to be applied by the next one
  if (!sp.many && bid == 0 && tid < sp.np) sp.parts_out[tid].pending = 1;
  // coordinate k of the node at unwrapped index i (bl_k - 1 <= i <= bl_k + extent)
  auto crd = [&](int k, int i) { return cst[k * cstride + (i - bl[k] + 1)]; };
  const uint32_t wb = mine ? wall_bits<Q>(g, x, y, z) : 0u;
  const bool wall = wb & 1u;
  const bool node = mine && !wall;
  const int row = y + g.ny * z;
  const T *__restrict__ fin = static_cast<const T *>(a.fin);
  T *__restrict__ fout = static_cast<T *>(a.fout);
  R *rho_side = static_cast<R *>(a.rho_side);

  if (wall) {
#pragma unroll
    for (int q = 0; q < Q; q++) store_pop(&fout[pop_index(g, Q, q, x, row)], zero_pop<Q, S>(a.c, q));
  }

  const int64_t own_idx = (int64_t)row * Q * g.nx + x;  // population 0 of this node

  // Geometry of this tick (bounceBackPrep with the getSphere test): only nodes
  // within the shell (sp.shell) of a sphere surface can be a link's fluid end.
  uint32_t own_any = 0;
  uint32_t own[kLocalParticles];
#pragma unroll
  for (int p = 0; p < kLocalParticles; p++) own[p] = 0;
  if (node) {
    // coordinates of the node and its neighbours, [1 - c] for offset -c
    double cxv[3], cyv[3], czv[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; i++) {
      cxv[i] = crd(0, x - 1 + i);
      cyv[i] = crd(1, y - 1 + i);
      if (D == 3) czv[i] = crd(2, z - 1 + i);
    }
    for (int p = 0; p < nloc; p++) {
      const DevParticle &P = prec[p];
      if (!in_box(P, x, y, z)) continue;
      // (X + (-R))_1^2 + ... in the getSphere order (gpu/kernels.jl:344-353),
      // from the per-axis squares
      double sa[3], sb[3], sc[3];
      {
#pragma clang fp contract(off)
#pragma unroll
        for (int i = 0; i < 3; i++) {
          const double u = cxv[i] + (-P.Rg[0]), v = cyv[i] + (-P.Rg[1]), w = czv[i] + (-P.Rg[2]);
          sa[i] = u * u;
          sb[i] = v * v;
          sc[i] = w * w;
        }
      }
      auto d2at = [&](int i, int j, int k) {
#pragma clang fp contract(off)
        const double s2 = sa[i] + sb[j];
        return D == 3 ? s2 + sc[k] : s2;
      };
      const double d2 = d2at(1, 1, 1);
      const double r2 = P.radius * P.radius;
      const double ro = P.radius + sp.shell;
      if (d2 < r2 || d2 >= ro * ro) continue;  // solid, or too far from the surface
      uint32_t op = 0;
#pragma unroll
      for (int q = 1; q < Q; q++) {
        // (x, q) is a link slot of P when x - c_q is solid
        if (d2at(1 - L::cx[q], 1 - L::cy[q], 1 - L::cz[q]) < r2) op |= 1u << q;
      }
#pragma unroll
      for (int i = 0; i < kLocalParticles; i++)
        if (i == p) own[i] = op;
      own_any |= op;
    }
  }

  // Link slots, one lane per slot: the block's slots are listed in LDS in a
  // fixed order (node, then q), each starting from the bounce-back value
  // f*_{q̄}(x); the particles then add their Ladd terms in particle order
  // and take their momentum exchange from the running value.  The list holds
  // 2 T slots at a time; more are taken in further rounds.
  // final values of the block's link slots, in list order
  R *linkval = static_cast<R *>(sp.linkval) + (int64_t)bid * nthreads * (Q - 1);
  // pk: lane | q << 9 | (c + 1) << 14, 16, 18 | weight class << 20 | owners << 22
  __shared__ uint32_t wave_cnt[8];
  __shared__ DD red[8][6];
  __shared__ DD blk[kLocalParticles][6];
  const int wave = tid >> 6, lane = __lane_id();
  const int nwaves = (nthreads + 63) >> 6;

  const uint32_t cnt = __popc(own_any);
  uint32_t incl = cnt;  // inclusive scan over the wave
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(incl, d);
    if (lane >= d) incl += o;
  }
  if (lane == 63 || tid == nthreads - 1) wave_cnt[wave] = incl;
  // rho of this node's populations of the previous tick (massDensity), kept
  // by the collision that produced them
  rho_l[tid] = (own_any || (wb & kMovingLinkBit)) ? rho_side[(int64_t)x + g.nx * (int64_t)row] : R(0);
  if (tid < 6 * kLocalParticles) blk[tid / 6][tid % 6] = DD{0.0, 0.0};
  __syncthreads();
  uint32_t off = incl - cnt, total = 0;
  for (int w = 0; w < nwaves; w++) {
    if (w < wave) off += wave_cnt[w];
    total += wave_cnt[w];
  }
  const int lx0 = bl[0], ly0 = bl[1], lz0 = bl[2];
  // The node's Q loads at once: f*_{q̄}(x) at its link slots (bounce-back),
  // f*_q(x - c_q) elsewhere (halfway bounce-back off a wall layer, gpu/
  // fluids.jl:165); the link slots go through the list below, the others
  // wait in registers for the node's update.  (w_q = w_q̄, so either index
  // decodes the bounced value.)
  R bbv[Q];
  {
    const GatherRows gr = gather_rows(g, Q, y, z);
    const int xs[3] = {x == g.nx - 1 ? 0 : x + 1, x, x == 0 ? g.nx - 1 : x - 1};
    const uint32_t bbm = own_any | wb;  // particle links and wall links
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int cx = L::cx[q], cy = L::cy[q], cz = L::cz[q];
      const bool bb = (bbm >> q) & 1u;
      const int64_t src = bb ? own_idx + (int64_t)L::opp[q] * g.nx
                             : gr.r[cy + 1][cz + 1] + (int64_t)q * g.nx + xs[cx + 1];
      bbv[q] = node ? decode<S, Q>(fin[src], q) : R(0);
    }
    if (node && (wb & kMovingLinkBit)) {
      // moving walls at the wall links (gpu/fluids.jl:168-174)
#pragma unroll
      for (int q = 1; q < Q; q++)
        if ((wb >> q) & 1u) bbv[q] = bbv[q] + moving_wall_term<Q, R>(g, a.c, q, x, y, z, rho_l[tid]);
    }
  }
  for (uint32_t c0 = 0; c0 < total; c0 += cap) {
    const int m = (int)min((uint32_t)cap, total - c0);
    uint32_t idx = off;
#pragma unroll
    for (int q = 1; q < Q; q++) {
      if ((own_any >> q) & 1) {
        if (idx >= c0 && idx < c0 + cap) {
          uint32_t owners = 0;
#pragma unroll
          for (int i = 0; i < kLocalParticles; i++) owners |= ((own[i] >> q) & 1u) << i;
          fs_l[idx - c0] = bbv[q];
          v_l[idx - c0] = bbv[q];
          pk_l[idx - c0] = (uint32_t)tid | (uint32_t)q << 9 | (uint32_t)(L::cx[q] + 1) << 14 |
                           (uint32_t)(L::cy[q] + 1) << 16 | (uint32_t)(L::cz[q] + 1) << 18 |
                           (uint32_t)L::wclass[q] << 20 | owners << 22;
        }
        idx++;
      }
    }
    __syncthreads();
    for (int p = 0; p < nloc; p++) {
      const DevParticle &P = prec[p];
      const bool coupled = P.couple_forces || P.couple_torques;
      // Each pass of the block over its slots is reduced over every wave at
      // once (lanes without a slot of p add zero), so no accumulator has to
      // live across the pass.
      if (lane == kWaveSumLane) {
#pragma unroll
        for (int k = 0; k < 6; k++) red[wave][k] = DD{0.0, 0.0};
      }
      const int passes = (m + nthreads - 1) / nthreads;
      for (int it = 0; it < passes; it++) {
        const int e = it * nthreads + tid;
        // double-double sums: the result does not depend on how the slots
        // fall into blocks and passes (any block shape gives the same bits)
        DD acc[6];
#pragma unroll
        for (int k = 0; k < 6; k++) acc[k] = DD{0.0, 0.0};
        auto add = [](DD &a, double x) { dd_add(a, x); };
        const uint32_t pk = e < m ? pk_l[e] : 0u;
        if ((pk >> (22 + p)) & 1) {
        const int t = pk & 511;
        const uint32_t nd = node_l[t];
        const int lx = lx0 + (int)(nd & 1023), ly = ly0 + (int)((nd >> 10) & 1023), lz = lz0 + (int)(nd >> 20);
        const int c0 = (int)((pk >> 14) & 3) - 1, c1 = (int)((pk >> 16) & 3) - 1,
                  c2 = (int)((pk >> 18) & 3) - 1;
        const double wq = L::wc((int)((pk >> 20) & 3));
        const double ci[3] = {c0 * a.c.ls, c1 * a.c.ls, c2 * a.c.ls};
        const double lc = 2.0 * wq * a.c.invC2s;
        const double shift = S == kF64 ? 0.0 : 2.0 * wq * a.c.rho0;
        const double xs[3] = {crd(0, lx - c0), crd(1, ly - c1), D == 3 ? crd(2, lz - c2) : 0.0};
        double ub[3] = {0.0, 0.0, 0.0};
        node_velocity<D>(xs, P, sp.semantics, sp.cutoff, ub);
        R v = v_l[e];
        {
#pragma clang fp contract(off)
          double udc = 0.0;
#pragma unroll
          for (int k = 0; k < D; k++) udc = udc + ub[k] * ci[k];
          // Ladd: streamed += 2 w_q / c_s^2 rho(x) (u_b . c_q) (gpu/fluids.jl:188-189)
          if constexpr (S == kF64)
            v = v + lc * static_cast<double>(rho_l[t]) * udc;
          else
            v = v + static_cast<R>(lc * static_cast<double>(rho_l[t]) * udc);
        }
        v_l[e] = v;
        if (coupled) {
#pragma clang fp contract(off)
          // s = f*_{q̄}(x) + f_q(x) (gpu/fluids.jl:193); the loads are
          // -coeff * sum(s) * c_q and -coeff * cross(sum(s (x - R)), c_q),
          // formed from these sums by the integrator
          const double s = static_cast<double>(fs_l[e]) + static_cast<double>(v) + shift;
          const double X[3] = {crd(0, lx), crd(1, ly), D == 3 ? crd(2, lz) : 0.0};
          double pr[3] = {0.0, 0.0, 0.0};
#pragma unroll
          for (int k = 0; k < D; k++) pr[k] = s * (X[k] + (-P.R[k]));
          if (P.couple_forces) {
            if (c0) add(acc[0], c0 > 0 ? s : -s);
            if (c1) add(acc[1], c1 > 0 ? s : -s);
            if (D == 3 && c2) add(acc[2], c2 > 0 ? s : -s);
          }
          if (P.couple_torques) {
            if constexpr (D == 2) {
              if (c1) add(acc[3], c1 > 0 ? pr[0] : -pr[0]);
              if (c0) add(acc[3], c0 > 0 ? -pr[1] : pr[1]);
            } else {
              if (c2) add(acc[3], c2 > 0 ? pr[1] : -pr[1]);
              if (c1) add(acc[3], c1 > 0 ? -pr[2] : pr[2]);
              if (c2) add(acc[4], c2 > 0 ? -pr[0] : pr[0]);
              if (c0) add(acc[4], c0 > 0 ? pr[2] : -pr[2]);
              if (c1) add(acc[5], c1 > 0 ? pr[0] : -pr[0]);
              if (c0) add(acc[5], c0 > 0 ? -pr[1] : pr[1]);
            }
          }
        }
        }
#pragma unroll
        for (int k = 0; k < 6; k++) {
          const DD v = wave_sum(acc[k]);
          if (lane == kWaveSumLane) red[wave][k] = dd_merge(red[wave][k], v);
        }
      }
      // this round's part of the block's partial of particle p
      __syncthreads();
      if (tid < 6) {
        DD s = red[0][tid];
        for (int w = 1; w < nwaves; w++) s = dd_merge(s, red[w][tid]);
        blk[p][tid] = dd_merge(blk[p][tid], s);
      }
      __syncthreads();
    }
    // the link slots' streamed values, kept for their nodes' collision (in
    // v_l itself when one round takes them all)
    if (total > (uint32_t)cap)
      for (int e = tid; e < m; e += nthreads) linkval[c0 + e] = v_l[e];
    __syncthreads();  // the next round reuses the list
  }
  if (sp.many && tid < kLocalParticles) sp.ids_out[(int64_t)bid * kLocalParticles + tid] = tid < nloc ? lid[tid] : -1;
  if (tid < 6 * nloc) {
    const DD s = blk[tid / 6][tid % 6];
    double *dst = sp.partial_out + ((int64_t)bid * kLocalParticles * 6 + tid) * 2;
    dst[0] = s.hi;
    dst[1] = s.lo;
  }

  // this node: S (gather only) or its collision P, and rho for the next tick's links
  if (node) {
    R fg[Q];
    uint32_t idx = off;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      if ((own_any >> q) & 1) {
        fg[q] = total > (uint32_t)cap ? linkval[idx] : v_l[idx];  // particle link
        idx++;
      } else {
        fg[q] = bbv[q];
      }
    }
    if constexpr (base_mode(MODE) == kGather) {
#pragma unroll
      for (int q = 0; q < Q; q++) store_pop(&fout[pop_index(g, Q, q, x, row)], encode<S, Q>(fg[q], q));
    } else {
      R fs[Q], rho, u[3];
      collide_node<Q, S, COLL, FM>(a, (int64_t)x + g.nx * (int64_t)row, fg, fs, rho,
                                   keeps_moments(MODE) ? u : nullptr);
#pragma unroll
      for (int q = 0; q < Q; q++) store_pop(&fout[pop_index(g, Q, q, x, row)], encode<S, Q>(fs[q], q));
      rho_side[(int64_t)x + g.nx * (int64_t)row] = rho;
      if constexpr (keeps_moments(MODE)) keep_node_moments<S>(a, x, row, rho, u);
    }
  }
}

// -------------------------------------------------------------- step
// Block b of one lattice's part of a step kernel (see StepMode): blocks
// [0, nbox_pad) work on the particles' link boxes (gathering modes only),
// the others each do one tile outside them.  FLAT: the block is
// one-dimensional (the batched kernel, whose lattices have tiles of their
// own shapes with the block's thread count) and thread t takes node
// (t mod bx, (t / bx) mod by, t / (bx by)) of the tile, the node a
// three-dimensional block's thread of the same linear index takes.
template <int Q, int S, int COLL, int FM, int MODE, bool FLAT>
__device__ __forceinline__ void step_block(const StepArgs &a, int b) {
  using L = Lattice<Q>;
  using T = typename Store<S>::T;
  using R = typename Store<S>::R;
  const BoxGeom &g = a.g;
  constexpr int M = base_mode(MODE);
  constexpr bool KEEP = keeps_moments(MODE);
  if constexpr (M != kCollide) {
    if (b < a.sp.nbox_pad) {
      if (b < a.sp.nbox) box_work<Q, S, COLL, FM, MODE>(a, b);
      return;
    }
    b -= a.sp.nbox_pad;
  }
  const int tile = a.t.remap ? xcd_remap(b, a.t.ntiles) : b;
  const uint32_t tq = tile_div((uint32_t)tile, a.t.tx, a.t.mtx);  // tile / tx
  const uint32_t tqq = tile_div(tq, a.t.ty, a.t.mty);              // tile / (tx ty)
  const int tpx = tile - (int)tq * a.t.tx;
  const int tpy = (int)tq - (int)tqq * a.t.ty;
  const int tpz = (int)tqq;
  int ix, iy, iz;
  if constexpr (FLAT) {
    const uint32_t t = threadIdx.x;
    const uint32_t r = tile_div(t, a.t.bx, a.t.mbx);   // t / bx
    const uint32_t zz = tile_div(r, a.t.by, a.t.mby);  // t / (bx by)
    ix = (int)(t - r * (uint32_t)a.t.bx);
    iy = (int)(r - zz * (uint32_t)a.t.by);
    iz = (int)zz;
  } else {
    ix = threadIdx.x;
    iy = threadIdx.y;
    iz = threadIdx.z;
  }
  const int x = tpx * a.t.bx + ix;
  const int y = tpy * a.t.by + iy;
  const int z = tpz * a.t.bz + iz;
  if (x >= g.nx || y >= g.ny || z >= g.nz) return;
  // whether this tile overlaps particle P's owned region: the six bounds
  // in one go, compared without short-circuits (which would make six
  // dependent loads of them)
  auto tile_touches = [&](const DevParticle &P) {
    const int32_t lo0 = P.lo[0], lo1 = P.lo[1], lo2 = P.lo[2], hi0 = P.hi[0], hi1 = P.hi[1], hi2 = P.hi[2];
    return (tpx * a.t.bx <= hi0 + kBoxMargin) & ((tpx + 1) * a.t.bx > lo0 - kBoxMargin) &
           (tpy * a.t.by <= hi1 + kBoxMargin) & ((tpy + 1) * a.t.by > lo1 - kBoxMargin) &
           (tpz * a.t.bz <= hi2 + kBoxMargin) & ((tpz + 1) * a.t.bz > lo2 - kBoxMargin);
  };
  // a band node of some particle belongs to the box blocks
  auto band_node_of_any = [&]() {
    bool in = false;
    for (int p = 0; p < a.sp.np; p++) {
      const DevParticle &P = a.sp.parts_in[p];
      if (tile_touches(P)) {
        Band bp;
        band_of(P, a.sp, g, bp);
        if (in_band(bp, x, y, z)) in = true;
      }
    }
    return in;
  };
  // whether no node of this tile is a wall or next to one: walled layers
  // only, and the tile inside planes 2 .. n - 3 of every walled dimension
  auto clear_of_walls = [&]() {
    if (!g.walled) return true;
    if (g.walled & (kWallMask | kMovingWalls)) return false;
    const int lo[3] = {tpx * a.t.bx, tpy * a.t.by, tpz * a.t.bz}, ext[3] = {a.t.bx, a.t.by, a.t.bz};
    const int nn[3] = {g.nx, g.ny, g.nz};
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 3; k++)
      if ((g.walled >> k) & 1) ok = ok & (lo[k] >= 2) & (lo[k] + ext[k] - 1 <= nn[k] - 3);
    return ok;
  };
  if constexpr (M == kFused) {
    if (clear_of_walls()) {
      // The run's repeated kernel away from walls: the node's Q gathers go
      // out first and the particles' regions (scalar, one load each) are
      // checked while they are in flight.  A tile that touches none of
      // them finishes here; one that does takes the general path below.
      // Offsets in I: 32 bits when every population's byte offset fits
      // (loads and stores then take the buffer's address from scalar
      // registers and one 32-bit offset), 64 otherwise.
      auto fast = [&](auto zero) -> bool {
        using I = decltype(zero);
        // byte offsets: row (y', z') of population q at x' is
        // y' Qb + z' ny Qb + q nxb + x' Tb (nxb, Qb uniform)
        const I Tb = (I)sizeof(T), nxb = (I)g.nx * Tb, Qb = (I)Q * nxb, Zs = (I)g.ny * Qb;
        const char *__restrict__ fin = static_cast<const char *>(a.fin);
        char *__restrict__ fout = static_cast<char *>(a.fout);
        const I ya[3] = {(I)(y == g.ny - 1 ? 0 : y + 1) * Qb, (I)y * Qb, (I)(y == 0 ? g.ny - 1 : y - 1) * Qb};
        const I zb[3] = {(I)(z == g.nz - 1 ? 0 : z + 1) * Zs, (I)z * Zs, (I)(z == 0 ? g.nz - 1 : z - 1) * Zs};
        const I xb[3] = {(I)(x == g.nx - 1 ? 0 : x + 1) * Tb, (I)x * Tb, (I)(x == 0 ? g.nx - 1 : x - 1) * Tb};
        I rb[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int j = 0; j < 3; j++) rb[i][j] = ya[i] + zb[j];
        T raw[Q];
#pragma unroll
        for (int q = 0; q < Q; q++)
          raw[q] = *reinterpret_cast<const T *>(fin + (rb[L::cy[q] + 1][L::cz[q] + 1] + xb[L::cx[q] + 1] + (I)q * nxb));
        bool near = false;
        for (int p = 0; p < a.sp.np; p++) near = near | tile_touches(a.sp.parts_in[p]);
        if (near) return false;
        R Fn[3], rho, u[3];
        node_force<Q, S, FM>(a, (int64_t)x + g.nx * (int64_t)(y + g.ny * z), Fn);
        T out[Q];
        collide_stored<Q, S, COLL, FM>(raw, Fn, a.c, out, rho, KEEP ? u : nullptr);
        const I own = rb[1][1] + xb[1];
#pragma unroll
        for (int q = 0; q < Q; q++) store_pop(reinterpret_cast<T *>(fout + (own + (I)q * nxb)), out[q]);
        if constexpr (KEEP) keep_node_moments<S>(a, x, y + g.ny * z, rho, u);
        return true;
      };
      if (g.off32 ? fast(uint32_t(0)) : fast(int64_t(0))) return;
    }
  }
  bool boxed = false;  // owned by the box blocks
  boxed = band_node_of_any();
  if (M != kCollide && boxed) return;  // the box blocks' node
  const int row = y + g.ny * z;
  const int64_t own = (int64_t)row * Q * g.nx + x;
  const int64_t n = (int64_t)x + g.nx * (int64_t)row;
  const T *__restrict__ fin = static_cast<const T *>(a.fin);
  T *__restrict__ fout = static_cast<T *>(a.fout);

  const uint32_t wb = wall_bits<Q>(g, x, y, z);
  if (wb & 1u) {
    // wall nodes hold f = 0 (`streamed[wallRegion] .= 0`)
#pragma unroll
    for (int q = 0; q < Q; q++) store_pop(&fout[own + (int64_t)q * g.nx], zero_pop<Q, S>(a.c, q));
    return;
  }

  T raw[Q];
  if constexpr (M == kCollide) {
#pragma unroll
    for (int q = 0; q < Q; q++) raw[q] = fin[own + (int64_t)q * g.nx];
  } else {
    // gather f_q(x) = f*_q(x - c_q); where x - c_q is a wall node, halfway
    // bounce-back f_q(x) = f*_{q̄}(x) (`streamed[cIR] = f*_{q̄}[cIR]`,
    // gpu/fluids.jl:165)
    const GatherRows gr = gather_rows(g, Q, y, z);
    const int xs[3] = {x == g.nx - 1 ? 0 : x + 1, x, x == 0 ? g.nx - 1 : x - 1};
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int cx = L::cx[q], cy = L::cy[q], cz = L::cz[q];
      const int64_t src = ((wb >> q) & 1u) ? own + (int64_t)L::opp[q] * g.nx
                                            : gr.r[cy + 1][cz + 1] + (int64_t)q * g.nx + xs[cx + 1];
      raw[q] = fin[src];
    }
  }

  R *rho_side = static_cast<R *>(a.rho_side);
  if (M != kCollide && (wb & kMovingLinkBit)) {
    // moving walls: the bounced populations also take the wall's momentum
    // (gpu/fluids.jl:168-174), with rho(x) of the node's last collision
    const R rx = rho_side[n];
    R f[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
      f[q] = decode<S, Q>(raw[q], q);
      if ((wb >> q) & 1u) f[q] = f[q] + moving_wall_term<Q, R>(g, a.c, q, x, y, z, rx);
    }
    if constexpr (M == kGather) {
#pragma unroll
      for (int q = 0; q < Q; q++) store_pop(&fout[own + (int64_t)q * g.nx], encode<S, Q>(f[q], q));
    } else {
      R fs[Q], rho, u[3];
      collide_node<Q, S, COLL, FM>(a, n, f, fs, rho, KEEP ? u : nullptr);
#pragma unroll
      for (int q = 0; q < Q; q++) store_pop(&fout[own + (int64_t)q * g.nx], encode<S, Q>(fs[q], q));
      rho_side[n] = rho;
      if constexpr (KEEP) keep_node_moments<S>(a, x, row, rho, u);
    }
    return;
  }

  if constexpr (M == kGather) {
#pragma unroll
    for (int q = 0; q < Q; q++) store_pop(&fout[own + (int64_t)q * g.nx], raw[q]);
  } else {
    R Fn[3], rho, u[3];
    node_force<Q, S, FM>(a, n, Fn);
    T out[Q];
    collide_stored<Q, S, COLL, FM>(raw, Fn, a.c, out, rho, KEEP ? u : nullptr);
#pragma unroll
    for (int q = 0; q < Q; q++) store_pop(&fout[own + (int64_t)q * g.nx], out[q]);
    // rho of the next gather's links: particle links (only the collide-only
    // kernel reaches box nodes) and moving walls
    if ((M == kCollide && boxed) || (wb & kMovingLinkBit)) rho_side[n] = rho;
    if constexpr (KEEP) keep_node_moments<S>(a, x, row, rho, u);
  }
}

// Waves per SIMD a step kernel is built for.  OCC = 0: registers capped at
// the tiles' needs (FP32/FP16: 64, eight waves; FP64: 128, four), so the
// wider box path spills rather than lowering the occupancy of every tile.
// OCC = 1: twice the registers and half the waves; the box path then runs
// without spills, which pays where the box blocks' chain of dependent steps
// outlasts the tiles (lattices of up to about 10^6 nodes with particles:
// 96^3 FP32 with a squirmer r = 10, 50 -> 34 us per fused kernel; 192^3
// FP16 r = 12, 94 -> 110 us), and so do the tiles of FP64 with a force,
// whose collision needs 157 registers.  The autotuner times both.
template <int S, int OCC>
constexpr int kWavesPerEU = OCC ? (S == kF64 ? 2 : 4) : (S == kF64 ? 4 : 8);

// One step kernel over one lattice.
template <int Q, int S, int COLL, int FM, int MODE, int OCC>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<S, OCC>)))
step_kernel(StepArgs a) {
  step_block<Q, S, COLL, FM, MODE, false>(a, blockIdx.x);
}

// One step kernel over a batch of independent lattices (the same lattice,
// storage, collision and force kinds; any sizes, parameters and particles),
// in one-dimensional blocks: each lattice has tiles of its own shape with
// the block's thread count.
// The grid is a list of segments, each a multiple of 8 blocks: segment s
// covers blocks [start[s], start[s + 1]) and holds either the box blocks
// (gathering modes) or the tiles, followed by idle blocks, of lattice
// seg[s] >> 1 (seg[s] & 1: tiles).  A lattice's box segment may come well
// before its tiles, so that the box blocks' chain of dependent steps starts
// early in the launch rather than behind the last tiles.
constexpr int kMaxBatch = 256;
struct BatchArgs {
  const StepArgs *boxes;  // device array of the lattices' arguments
  int32_t n;              // segments
  int32_t start[2 * kMaxBatch + 1];
  uint16_t seg[2 * kMaxBatch];
};

template <int Q, int S, int COLL, int FM, int MODE, int OCC>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<S, OCC>)))
batch_step_kernel(BatchArgs ba) {
  const int b = blockIdx.x;
  int lo = 0, hi = ba.n - 1;  // the last segment starting at or before b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ba.start[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  const StepArgs &a = ba.boxes[ba.seg[lo] >> 1];
  const int nbox = base_mode(MODE) != kCollide ? a.sp.nbox_pad : 0;
  const int local = b - ba.start[lo] + ((ba.seg[lo] & 1) ? nbox : 0);
  if (local >= nbox + a.t.ntiles) return;
  step_block<Q, S, COLL, FM, MODE, true>(a, local);
}

// -------------------------------------------------------- aux kernels
struct AuxArgs {
  void *f;              // populations (storage)
  const double *f64;    // FP64 host-layout buffer (set)
  double *out64;        // FP64 output (get populations / rho)
  double *uout;         // FP64 velocity output
  const void *Ffield;
  BoxGeom g;
  CollParams c;
  double rho_init, u_init[3];
};

// findInitialConditions (gpu/initMethods.jl:417-438) with uniform rho, u.
template <int Q, int S>
__global__ void init_uniform_kernel(AuxArgs a) {
#pragma clang fp contract(off)
  using L = Lattice<Q>;
  using T = typename Store<S>::T;
  constexpr int D = L::D;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.g.N) return;
  const int x = n % a.g.nx, y = (n / a.g.nx) % a.g.ny, z = n / ((int64_t)a.g.nx * a.g.ny);
  const bool wall = is_wall_node(a.g, x, y, z);
  const double rho = wall ? 0.0 : a.rho_init;
  double u[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < D; k++) u[k] = wall ? 0.0 : a.u_init[k];
  if (a.c.force_mode != LBM_FORCE_NONE) {
#pragma unroll
    for (int k = 0; k < D; k++) {
      double Fk;
      if (a.c.force_mode == LBM_FORCE_UNIFORM)
        Fk = wall ? 0.0 : a.c.F[k];
      else if constexpr (S == kF64)
        Fk = static_cast<const double *>(a.Ffield)[k * a.g.N + n];
      else
        Fk = static_cast<const float *>(a.Ffield)[k * a.g.N + n];
      const double ct = rho != 0.0 ? Fk / rho : 0.0;
      u[k] = u[k] - 0.5 * a.c.dt * ct;
    }
  }
  T *f = static_cast<T *>(a.f);
  const int row = y + a.g.ny * z;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const double ci[3] = {L::cx[q] * a.c.ls, L::cy[q] * a.c.ls, L::cz[q] * a.c.ls};
    double udc = 0.0, uu = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) udc = udc + u[k] * ci[k];
#pragma unroll
    for (int k = 0; k < D; k++) uu = uu + u[k] * u[k];
    const double first = udc * a.c.invC2s + (udc * udc) * (0.5 * a.c.invC4s);
    const double second = first - uu * (0.5 * a.c.invC2s) + 1.0;
    f[pop_index(a.g, Q, q, x, row)] = from_f64<S, Q>(second * (L::w(q) * rho), q, a.c.rho0);
  }
}

template <int Q, int S>
__global__ void set_populations_kernel(AuxArgs a) {
  using T = typename Store<S>::T;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.g.N) return;
  T *f = static_cast<T *>(a.f);
  const int x = n % a.g.nx, row = n / a.g.nx;
#pragma unroll
  for (int q = 0; q < Q; q++) f[pop_index(a.g, Q, q, x, row)] = from_f64<S, Q>(a.f64[q * a.g.N + n], q, a.c.rho0);
}

template <int Q, int S>
__global__ void get_populations_kernel(AuxArgs a) {
  using T = typename Store<S>::T;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.g.N) return;
  const T *f = static_cast<const T *>(a.f);
  const int x = n % a.g.nx, y = (n / a.g.nx) % a.g.ny, z = n / ((int64_t)a.g.nx * a.g.ny);
  const bool wall = S != kF64 && is_wall_node(a.g, x, y, z);
  const int row = y + a.g.ny * z;
#pragma unroll
  for (int q = 0; q < Q; q++)
    a.out64[q * a.g.N + n] = wall ? 0.0 : to_f64<S, Q>(f[pop_index(a.g, Q, q, x, row)], q, a.c.rho0);
}

// Health of the stored populations: counts[0] += the non-finite ones,
// counts[1] += those at the storage's clamp (f16c's largest magnitude),
// over the fluid nodes (wall nodes hold no meaningful value).  Grid-stride
// over the row-blocked buffer, one pair of atomics per block.
template <int Q, int S>
__global__ void scan_populations_kernel(AuxArgs a, unsigned long long *counts) {
  using T = typename Store<S>::T;
  const T *f = static_cast<const T *>(a.f);
  const int64_t total = a.g.N * Q;
  unsigned long long bad = 0, sat = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(e % a.g.nx);
    const int64_t rq = e / a.g.nx;
    const int q = (int)(rq % Q);
    const int64_t row = rq / Q;
    const int y = (int)(row % a.g.ny), z = (int)(row / a.g.ny);
    if (is_wall_node(a.g, x, y, z)) continue;
    const T v = f[e];
    if constexpr (S == kF16C) {
      sat += ((uint32_t)(uint16_t)v & 0x7FFFu) == 0x7FFFu;
    } else {
      if (!isfinite(static_cast<double>(decode<S, Q>(v, q)))) bad++;
    }
  }
  __shared__ unsigned long long sb[2];
  if (threadIdx.x == 0) sb[0] = sb[1] = 0;
  __syncthreads();
  if (bad) atomicAdd(&sb[0], bad);
  if (sat) atomicAdd(&sb[1], sat);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (sb[0]) atomicAdd(&counts[0], sb[0]);
    if (sb[1]) atomicAdd(&counts[1], sb[1]);
  }
}

// hydroVariablesUpdate! (gpu/fluids.jl:7-37, cpu/fluids.jl:7-38) at node n
// from its populations f (FP64, read from storage): rho and u.
template <int Q, int S>
__device__ inline void node_hydro(const AuxArgs &a, int64_t n, int x, int y, int z, double (&f)[Q], double &rho,
                                  double (&u)[3]) {
#pragma clang fp contract(off)
  using L = Lattice<Q>;
  using T = typename Store<S>::T;
  constexpr int D = L::D;
  const T *fp = static_cast<const T *>(a.f);
#pragma unroll
  for (int q = 0; q < Q; q++) f[q] = to_f64<S, Q>(fp[pop_index(a.g, Q, q, x, y + a.g.ny * z)], q, a.c.rho0);
  if (S != kF64 && is_wall_node(a.g, x, y, z)) {
#pragma unroll
    for (int q = 0; q < Q; q++) f[q] = 0.0;
  }
  rho = f[0];
#pragma unroll
  for (int q = 1; q < Q; q++) rho = rho + f[q];
  // Terms whose velocity component is 0 are left out of the sums here and
  // below: each would add a signed zero to a sum that starts at +0 and so
  // changes no bit of it (the compiler may not drop them itself, since f *
  // 0.0 is not 0 for every f).
  double mom[3];
#pragma unroll
  for (int k = 0; k < D; k++) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int ck = k == 0 ? L::cx[q] : (k == 1 ? L::cy[q] : L::cz[q]);
      if (ck != 0) acc = acc + f[q] * static_cast<double>(ck);
    }
    mom[k] = acc;
  }
  if (a.c.force_mode != LBM_FORCE_NONE) {
    const double half_dt = 0.5 * a.c.dt;
    const bool wall = is_wall_node(a.g, x, y, z);
#pragma unroll
    for (int k = 0; k < D; k++) {
      double Fk;
      if (a.c.force_mode == LBM_FORCE_UNIFORM)
        Fk = wall ? 0.0 : a.c.F[k];
      else if constexpr (S == kF64)
        Fk = static_cast<const double *>(a.Ffield)[k * a.g.N + n];
      else
        Fk = static_cast<const float *>(a.Ffield)[k * a.g.N + n];
      mom[k] = mom[k] + half_dt * Fk;
    }
  }
  u[0] = u[1] = u[2] = 0.0;
#pragma unroll
  for (int k = 0; k < D; k++) u[k] = rho != 0.0 ? mom[k] / rho : 0.0;
}

// hydroVariablesUpdate! (gpu/fluids.jl:7-37) on the current populations,
// evaluated in FP64 whatever the storage.
template <int Q, int S>
__global__ void hydro_kernel(AuxArgs a) {
  constexpr int D = Lattice<Q>::D;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.g.N) return;
  const int x = n % a.g.nx, y = (n / a.g.nx) % a.g.ny, z = n / ((int64_t)a.g.nx * a.g.ny);
  double f[Q], rho, u[3];
  node_hydro<Q, S>(a, n, x, y, z, f, rho, u);
  if (a.out64) a.out64[n] = rho;
  if (a.uout) {
#pragma unroll
    for (int k = 0; k < D; k++) a.uout[k * a.g.N + n] = u[k];
  }
}

// viscousStressTensor (cpu/fluids.jl:507-520) on the current populations, in
// FP64: sigma_ab = -(1 - dt/(2 tau)) sum_i c_ia c_ib (f_i - feq_i), with
// feq from rho and u as equilibriumDistribution builds it
// (cpu/fluids.jl:46-63, either form) and the integer c_i in the sum.  a.out64 receives
// sigma as (N, D, D), column major: sigma_ab of node n at n + N (a + D b).
template <int Q, int S>
__global__ void stress_kernel(AuxArgs a) {
#pragma clang fp contract(off)
  using L = Lattice<Q>;
  constexpr int D = L::D;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.g.N) return;
  const int x = n % a.g.nx, y = (n / a.g.nx) % a.g.ny, z = n / ((int64_t)a.g.nx * a.g.ny);
  double f[Q], rho, u[3];
  node_hydro<Q, S>(a, n, x, y, z, f, rho, u);
  double uu = u[0] * u[0];
#pragma unroll
  for (int k = 1; k < D; k++) uu = uu + u[k] * u[k];
  const double coeff = 1.0 - 0.5 * a.c.dt / a.c.tau;
  double sig[D][D];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int c[3] = {L::cx[q], L::cy[q], L::cz[q]};
    double udc = u[0] * (static_cast<double>(c[0]) * a.c.ls);
#pragma unroll
    for (int k = 1; k < D; k++) udc = udc + u[k] * (static_cast<double>(c[k]) * a.c.ls);
    const double first = udc * a.c.invC2s + udc * udc * (0.5 * a.c.invC4s);
    const double second = first - uu * (0.5 * a.c.invC2s);
    const double feq = a.c.incompressible ? L::w(q) * rho + L::w(q) * (a.c.rho0 * second)
                                          : L::w(q) * ((second + 1.0) * rho);
    const double neq = f[q] - feq;
#pragma unroll
    for (int al = 0; al < D; al++)
#pragma unroll
      for (int be = 0; be < D; be++) {
        const double t = static_cast<double>(c[al] * c[be]) * neq;
        sig[al][be] = q == 0 ? t : sig[al][be] + t;
      }
  }
#pragma unroll
  for (int al = 0; al < D; al++)
#pragma unroll
    for (int be = 0; be < D; be++) a.out64[n + a.g.N * (al + D * be)] = -coeff * sig[al][be];
}

// Link bits of one particle over its box (for lbm_get_links):
// bit j of mask[local node] set when node is fluid and node + c_j solid.
struct LinkMaskArgs {
  BoxGeom g;
  const DevParticle *part;
  const double *coords[3];
  uint32_t *mask;
  int32_t lo[3], ext[3];
};

template <int Q>
__global__ void link_mask_kernel(LinkMaskArgs a) {
  using L = Lattice<Q>;
  constexpr int D = L::D;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tot = (int64_t)a.ext[0] * a.ext[1] * a.ext[2];
  if (t >= tot) return;
  const int x = wrapi(a.lo[0] + (int)(t % a.ext[0]), a.g.nx);
  const int y = wrapi(a.lo[1] + (int)((t / a.ext[0]) % a.ext[1]), a.g.ny);
  const int z = D == 3 ? wrapi(a.lo[2] + (int)(t / ((int64_t)a.ext[0] * a.ext[1])), a.g.nz) : 0;
  const DevParticle &P = *a.part;
  const double r2 = P.radius * P.radius;
  uint32_t m = 0;
  if (!solid_at<D>(a.coords, P.Rg, r2, x, y, z)) {
#pragma unroll
    for (int j = 1; j < Q; j++) {
      const int sx = wrapi(x + L::cx[j], a.g.nx), sy = wrapi(y + L::cy[j], a.g.ny),
                sz = D == 3 ? wrapi(z + L::cz[j], a.g.nz) : 0;
      if (solid_at<D>(a.coords, P.Rg, r2, sx, sy, sz)) m |= 1u << j;
    }
  }
  a.mask[t] = m;
}

}  // namespace lbm
