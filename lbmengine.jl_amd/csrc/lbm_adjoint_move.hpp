// lbm_adjoint_move.hpp — derivative of one particle move (moveParticles!,
// GPU semantics: move_gpu plus the node-velocity snapshot refresh of
// move_record) by forward-mode dual numbers, for the adjoint's
// vector-Jacobian product: on the host with every tangent at once, on the
// device with one tangent per lane.  The branches (moved, snapshot refresh,
// zero angular speed) are taken from the primal values.
#pragma once

#include <cmath>

#include <hip/hip_runtime.h>

#include "lbm_integrate.hpp"

namespace lbm {

// A number with N tangent components: the host carries all kDualN tangents
// at once; the device gives each lane one tangent (N = 1).
constexpr int kPState = 24;              // R, V, W, e, nR, nV, nW, ne
constexpr int kDualN = kPState + 6;      // plus the force and torque inputs
// Indices of the particle state in the adjoint's vectors.
enum { kStR = 0, kStV = 3, kStW = 6, kStE = 9, kStNR = 12, kStNV = 15, kStNW = 18, kStNE = 21 };

template <int N>
struct DualT {
  double v;
  double d[N];
  __host__ __device__ DualT() : v(0.0) {
    for (double &x : d) x = 0.0;
  }
  __host__ __device__ DualT(double x) : v(x) {  // NOLINT: constants lift implicitly
    for (double &t : d) t = 0.0;
  }
};
template <int N>
__host__ __device__ inline DualT<N> operator+(const DualT<N> &a, const DualT<N> &b) {
#pragma clang fp contract(off)
  DualT<N> r(a.v + b.v);
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N>
__host__ __device__ inline DualT<N> operator-(const DualT<N> &a, const DualT<N> &b) {
#pragma clang fp contract(off)
  DualT<N> r(a.v - b.v);
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N>
__host__ __device__ inline DualT<N> operator-(const DualT<N> &a) {
  DualT<N> r(-a.v);
  for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
  return r;
}
template <int N>
__host__ __device__ inline DualT<N> operator*(const DualT<N> &a, const DualT<N> &b) {
#pragma clang fp contract(off)
  DualT<N> r(a.v * b.v);
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
__host__ __device__ inline DualT<N> operator/(const DualT<N> &a, const DualT<N> &b) {
#pragma clang fp contract(off)
  DualT<N> r(a.v / b.v);
  for (int i = 0; i < N; i++) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v);
  return r;
}
template <int N>
__host__ __device__ inline DualT<N> dsqrt(const DualT<N> &a) {
#pragma clang fp contract(off)
  DualT<N> r(sqrt(a.v));
  for (int i = 0; i < N; i++) r.d[i] = a.v > 0.0 ? a.d[i] / (2.0 * r.v) : 0.0;
  return r;
}
template <int N>
__host__ __device__ inline DualT<N> dcos(const DualT<N> &a) {
  DualT<N> r(cos(a.v));
  const double s = -sin(a.v);
  for (int i = 0; i < N; i++) r.d[i] = s * a.d[i];
  return r;
}
template <int N>
__host__ __device__ inline DualT<N> dsin(const DualT<N> &a) {
  DualT<N> r(sin(a.v));
  const double c = cos(a.v);
  for (int i = 0; i < N; i++) r.d[i] = c * a.d[i];
  return r;
}
using Dual = DualT<kDualN>;

// One particle's differentiable state, in the order of the tangents.
template <int N>
struct DualStateT {
  DualT<N> R[3], V[3], W[3], e[3], nR[3], nV[3], nW[3], ne[3];
  __host__ __device__ DualT<N> *slot(int i) {
    DualT<N> *g[8] = {R, V, W, e, nR, nV, nW, ne};
    return &g[i / 3][i % 3];
  }
};
using DualState = DualStateT<kDualN>;

// Move of record p (3D) with tangents seeded on its state and on (F, T):
// input i (state 0..23, force 24..26, torque 27..29) carries tangent
// i - base when that lies in [0, N).  Returns the new state; the tangents
// hold the Jacobian's columns base .. base + N - 1.
template <int N>
__host__ __device__ inline DualStateT<N> move_dual_t(const DevParticle &p, const double F[3], const double Tq[3],
                                                     double dt, int base) {
  using D = DualT<N>;
  DualStateT<N> s;
  const double *src[8] = {p.R, p.V, p.W, p.e, p.nR, p.nV, p.nW, p.ne};
  auto seed = [&](D &x, int i) {
    if (i - base >= 0 && i - base < N) x.d[i - base] = 1.0;
  };
  for (int i = 0; i < kPState; i++) {
    D &x = *s.slot(i);
    x = D(src[i / 3][i % 3]);
    seed(x, i);
  }
  D Fd[3], Td[3];
  for (int k = 0; k < 3; k++) {
    Fd[k] = D(F[k]);
    seed(Fd[k], kPState + k);
    Td[k] = D(Tq[k]);
    seed(Td[k], kPState + 3 + k);
  }
  // primal branches, from the device routine itself
  DevParticle q = p;
  bool moved = false, find_nv = false;
  move_gpu(q, F, Tq, 3, dt, moved, find_nv);
  if (p.couple_forces) {
    for (int k = 0; k < 3; k++) s.V[k] = s.V[k] + D(dt) * (Fd[k] * D(p.inv_mass));
    for (int k = 0; k < 3; k++) s.R[k] = s.R[k] + D(dt) * s.V[k];
  }
  if (p.couple_torques) {
    for (int k = 0; k < 3; k++) s.W[k] = s.W[k] + D(dt) * (Td[k] * D(p.inv_I));
    if (p.kind == LBM_PARTICLE_SQUIRMER) {
      D th = dsqrt(s.W[0] * s.W[0] + s.W[1] * s.W[1] + s.W[2] * s.W[2]);
      if (!(th.v == 0.0)) {
        const D kk[3] = {s.W[0] / th, s.W[1] / th, s.W[2] / th};
        th = th * D(dt);
        const D c = dcos(th), sn = dsin(th);
        const D *e = s.e;
        const D cr[3] = {kk[1] * e[2] - kk[2] * e[1], -kk[0] * e[2] + kk[2] * e[0], kk[0] * e[1] - kk[1] * e[0]};
        D vr[3];
        for (int k = 0; k < 3; k++) vr[k] = e[k] * c + cr[k] * sn;
        const D kd = kk[0] * e[0] + kk[1] * e[1] + kk[2] * e[2];
        for (int k = 0; k < 3; k++) vr[k] = vr[k] + kk[k] * kd * (D(1.0) - c);
        const D nv = dsqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
        for (int k = 0; k < 3; k++) s.e[k] = vr[k] / nv;
      }
    }
  }
  if (find_nv)
    for (int k = 0; k < 3; k++) {
      s.nR[k] = s.R[k];
      s.nV[k] = s.V[k];
      s.nW[k] = s.W[k];
      s.ne[k] = s.e[k];
    }
  return s;
}

// All kDualN tangents at once (host).
inline DualState move_dual(const DevParticle &p, const double F[3], const double Tq[3], double dt) {
  return move_dual_t<kDualN>(p, F, Tq, dt, 0);
}

// mu_in = (d state'/d state)^T mu_out, phi = (d state'/dF)^T mu_out,
// tau = (d state'/dT)^T mu_out.
inline void move_vjp(const DevParticle &p, const double F[3], const double Tq[3], double dt,
                     const double (&mu_out)[kPState], double (&mu_in)[kPState], double (&phi)[3],
                     double (&tau)[3]) {
  DualState s = move_dual(p, F, Tq, dt);
  double g[kDualN];
  for (int i = 0; i < kDualN; i++) g[i] = 0.0;
  for (int o = 0; o < kPState; o++) {
    const Dual &y = *s.slot(o);
    if (mu_out[o] == 0.0) continue;
    for (int i = 0; i < kDualN; i++) g[i] += mu_out[o] * y.d[i];
  }
  for (int i = 0; i < kPState; i++) mu_in[i] = g[i];
  for (int k = 0; k < 3; k++) {
    phi[k] = g[kPState + k];
    tau[k] = g[kPState + 3 + k];
  }
}

}  // namespace lbm
