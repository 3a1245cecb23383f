// lbm_adjoint_move.hpp — host-side derivative of one particle move
// (moveParticles!, GPU semantics: move_gpu plus the node-velocity snapshot
// refresh of move_record) by forward-mode dual numbers, for the adjoint's
// vector-Jacobian product.  The branches (moved, snapshot refresh, zero
// angular speed) are taken from the primal values.
#pragma once

#include <cmath>

#include "lbm_integrate.hpp"

namespace lbm {

// A number with kDualN tangent components.
constexpr int kPState = 24;              // R, V, W, e, nR, nV, nW, ne
constexpr int kDualN = kPState + 6;      // plus the force and torque inputs

struct Dual {
  double v;
  double d[kDualN];
  Dual() : v(0.0) {
    for (double &x : d) x = 0.0;
  }
  Dual(double x) : v(x) {  // NOLINT: constants lift implicitly
    for (double &t : d) t = 0.0;
  }
};
inline Dual operator+(const Dual &a, const Dual &b) {
  Dual r(a.v + b.v);
  for (int i = 0; i < kDualN; i++) r.d[i] = a.d[i] + b.d[i];
  return r;
}
inline Dual operator-(const Dual &a, const Dual &b) {
  Dual r(a.v - b.v);
  for (int i = 0; i < kDualN; i++) r.d[i] = a.d[i] - b.d[i];
  return r;
}
inline Dual operator-(const Dual &a) {
  Dual r(-a.v);
  for (int i = 0; i < kDualN; i++) r.d[i] = -a.d[i];
  return r;
}
inline Dual operator*(const Dual &a, const Dual &b) {
  Dual r(a.v * b.v);
  for (int i = 0; i < kDualN; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
inline Dual operator/(const Dual &a, const Dual &b) {
  Dual r(a.v / b.v);
  for (int i = 0; i < kDualN; i++) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v);
  return r;
}
inline Dual dsqrt(const Dual &a) {
  Dual r(std::sqrt(a.v));
  for (int i = 0; i < kDualN; i++) r.d[i] = a.v > 0.0 ? a.d[i] / (2.0 * r.v) : 0.0;
  return r;
}
inline Dual dcos(const Dual &a) {
  Dual r(std::cos(a.v));
  const double s = -std::sin(a.v);
  for (int i = 0; i < kDualN; i++) r.d[i] = s * a.d[i];
  return r;
}
inline Dual dsin(const Dual &a) {
  Dual r(std::sin(a.v));
  const double c = std::cos(a.v);
  for (int i = 0; i < kDualN; i++) r.d[i] = c * a.d[i];
  return r;
}

// One particle's differentiable state, in the order of the tangents.
struct DualState {
  Dual R[3], V[3], W[3], e[3], nR[3], nV[3], nW[3], ne[3];
  Dual *slot(int i) {
    Dual *g[8] = {R, V, W, e, nR, nV, nW, ne};
    return &g[i / 3][i % 3];
  }
};

// Move of record p (3D) with tangents seeded on its state and on (F, T):
// returns the new state; Jacobian columns in the tangents.
inline DualState move_dual(const DevParticle &p, const double F[3], const double Tq[3], double dt) {
  DualState s;
  const double *src[8] = {p.R, p.V, p.W, p.e, p.nR, p.nV, p.nW, p.ne};
  for (int i = 0; i < kPState; i++) {
    Dual &x = *s.slot(i);
    x = Dual(src[i / 3][i % 3]);
    x.d[i] = 1.0;
  }
  Dual Fd[3], Td[3];
  for (int k = 0; k < 3; k++) {
    Fd[k] = Dual(F[k]);
    Fd[k].d[kPState + k] = 1.0;
    Td[k] = Dual(Tq[k]);
    Td[k].d[kPState + 3 + k] = 1.0;
  }
  // primal branches, from the device routine itself
  DevParticle q = p;
  bool moved = false, find_nv = false;
  move_gpu(q, F, Tq, 3, dt, moved, find_nv);
  if (p.couple_forces) {
    for (int k = 0; k < 3; k++) s.V[k] = s.V[k] + Dual(dt) * (Fd[k] * Dual(p.inv_mass));
    for (int k = 0; k < 3; k++) s.R[k] = s.R[k] + Dual(dt) * s.V[k];
  }
  if (p.couple_torques) {
    for (int k = 0; k < 3; k++) s.W[k] = s.W[k] + Dual(dt) * (Td[k] * Dual(p.inv_I));
    if (p.kind == LBM_PARTICLE_SQUIRMER) {
      Dual th = dsqrt(s.W[0] * s.W[0] + s.W[1] * s.W[1] + s.W[2] * s.W[2]);
      if (!(th.v == 0.0)) {
        const Dual kk[3] = {s.W[0] / th, s.W[1] / th, s.W[2] / th};
        th = th * Dual(dt);
        const Dual c = dcos(th), sn = dsin(th);
        const Dual *e = s.e;
        const Dual cr[3] = {kk[1] * e[2] - kk[2] * e[1], -kk[0] * e[2] + kk[2] * e[0], kk[0] * e[1] - kk[1] * e[0]};
        Dual vr[3];
        for (int k = 0; k < 3; k++) vr[k] = e[k] * c + cr[k] * sn;
        const Dual kd = kk[0] * e[0] + kk[1] * e[1] + kk[2] * e[2];
        for (int k = 0; k < 3; k++) vr[k] = vr[k] + kk[k] * kd * (Dual(1.0) - c);
        const Dual nv = dsqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
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
