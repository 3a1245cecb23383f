// Host-side checks of the adjoint's building blocks against central finite
// differences of the forward routines they reverse (all FP64):
//  - collide_vjp against collide_ref (BGK and TRT; no force and a force
//    field), for dL/dS and dL/dF;
//  - node_velocity_vjp against node_velocity (squirmer and bead), for
//    B1, B2 and the node-velocity snapshot;
//  - move_vjp (dual numbers) against move_gpu with the snapshot refresh.
// Prints "ok" and exits 0, or names the first failure and exits 1.
#include <cmath>
#include <cstdio>
#include <random>

#include "lbm_adjoint.hpp"
#include "lbm_adjoint_move.hpp"

using namespace lbm;

static std::mt19937_64 rng(7);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

// |a - b| <= tol max(|b|, floor): floor is the size of the vector's entries,
// below which a difference quotient's cancellation error dominates
static bool close(double a, double b, double tol, const char *what, int i, double floor = 1e-12) {
  const double scale = std::max(std::fabs(b), floor);
  if (std::fabs(a - b) <= tol * scale) return true;
  std::printf("%s[%d]: adjoint %.12e finite difference %.12e\n", what, i, a, b);
  return false;
}

template <int COLL, int FM>
static bool check_collision() {
  constexpr int Q = 19;
  using L = Lattice<Q>;
  CollParams c{};
  c.dt = 1.0;
  c.tau = 0.8;
  c.tau_minus = COLL == LBM_COLLISION_TRT ? 0.5 + (3.0 / 16.0) / (0.8 - 0.5) : 0.8;
  c.ls = 1.0;
  c.invC2s = 3.0;
  c.invC4s = 9.0;
  c.rho0 = 1.0;
  set_float_params(c);
  for (int trial = 0; trial < 20; trial++) {
    double f[Q], Fn[3], lam[Q];
    for (int q = 0; q < Q; q++) f[q] = L::w(q) * (1.0 + U(-0.05, 0.05)), lam[q] = U(-1, 1);
    for (int k = 0; k < 3; k++) Fn[k] = FM == LBM_FORCE_NONE ? 0.0 : U(-1e-3, 1e-3);
    auto loss = [&](const double (&ff)[Q], const double *FF) {
      double fs[Q], rho;
      collide_ref<Q, COLL, FM>(ff, FF, c, fs, rho);
      double s = 0.0;
      for (int q = 0; q < Q; q++) s += lam[q] * fs[q];
      return s;
    };
    double rho = 0.0, j[3] = {0, 0, 0};
    for (int q = 0; q < Q; q++) {
      rho += f[q];
      j[0] += L::cx[q] * f[q];
      j[1] += L::cy[q] * f[q];
      j[2] += L::cz[q] * f[q];
    }
    double u[3], F3[3] = {Fn[0], Fn[1], Fn[2]};
    for (int k = 0; k < 3; k++) u[k] = (j[k] + (FM == LBM_FORCE_NONE ? 0.0 : 0.5 * c.dt * Fn[k])) / rho;
    double ls_[Q], gF[3] = {0, 0, 0};
    collide_vjp<Q, COLL, FM, double>(rho, u, F3, c, lam, ls_, gF);
    const double h = 1e-6;
    for (int q = 0; q < Q; q++) {
      double fp[Q], fm[Q];
      for (int i = 0; i < Q; i++) fp[i] = fm[i] = f[i];
      fp[q] += h;
      fm[q] -= h;
      const double fd = (loss(fp, Fn) - loss(fm, Fn)) / (2 * h);
      if (!close(ls_[q], fd, 1e-7, "collision dL/dS", q, 0.1)) return false;
    }
    if (FM == LBM_FORCE_FIELD)
      for (int k = 0; k < 3; k++) {
        double Fp[3] = {Fn[0], Fn[1], Fn[2]}, Fm[3] = {Fn[0], Fn[1], Fn[2]};
        const double hf = 1e-7;
        Fp[k] += hf;
        Fm[k] -= hf;
        const double fd = (loss(f, Fp) - loss(f, Fm)) / (2 * hf);
        if (!close(gF[k], fd, 1e-6, "collision dL/dF", k, 0.1)) return false;
      }
  }
  return true;
}

static DevParticle random_particle(int kind) {
  DevParticle p{};
  p.kind = kind;
  p.couple_forces = p.couple_torques = 1;
  p.radius = 6.0;
  p.B1 = 1e-3;
  p.B2 = U(-5e-3, 5e-3);
  p.inv_mass = 1.0 / 900.0;
  p.inv_I = 1.0 / 20000.0;
  double en = 0.0;
  for (int k = 0; k < 3; k++) {
    p.R[k] = U(-1, 1);
    p.V[k] = U(-1e-3, 1e-3);
    p.W[k] = U(-1e-4, 1e-4);
    p.e[k] = U(-1, 1);
    en += p.e[k] * p.e[k];
  }
  for (int k = 0; k < 3; k++) {
    p.e[k] /= std::sqrt(en);
    p.nR[k] = p.R[k] + U(-0.01, 0.01);
    p.nV[k] = p.V[k] + U(-1e-4, 1e-4);
    p.nW[k] = p.W[k] + U(-1e-5, 1e-5);
    p.ne[k] = p.e[k] + U(-0.01, 0.01);
    p.Rg[k] = p.R[k];
  }
  return p;
}

static bool check_node_velocity() {
  for (int trial = 0; trial < 40; trial++) {
    const int kind = trial % 4 == 0 ? LBM_PARTICLE_BEAD : LBM_PARTICLE_SQUIRMER;
    DevParticle p = random_particle(kind);
    double xc[3], ub[3];
    for (int k = 0; k < 3; k++) xc[k] = p.nR[k] + U(-6, 6), ub[k] = U(-1, 1);
    double acc[kAdjSums] = {0};
    node_velocity_vjp(xc, p, ub, acc);
    auto loss = [&](const DevParticle &pp) {
      double out[3];
      node_velocity<3>(xc, pp, LBM_SEMANTICS_GPU, 0.0, out);
      return ub[0] * out[0] + ub[1] * out[1] + ub[2] * out[2];
    };
    struct {
      int idx;
      double *field;
      double h;
    } vars[] = {{kAdjB1, &p.B1, 1e-7}, {kAdjB2, &p.B2, 1e-7}};
    for (auto &v : vars) {
      const double x0 = *v.field;
      *v.field = x0 + v.h;
      const double lp = loss(p);
      *v.field = x0 - v.h;
      const double lm = loss(p);
      *v.field = x0;
      if (!close(acc[v.idx], (lp - lm) / (2 * v.h), 1e-6, "node velocity dB", v.idx)) return false;
    }
    double *fields[4] = {p.nR, p.nV, p.nW, p.ne};
    const int base[4] = {kAdjNR, kAdjNV, kAdjNW, kAdjNE};
    for (int f = 0; f < 4; f++)
      for (int k = 0; k < 3; k++) {
        const double h = 1e-6, x0 = fields[f][k];
        fields[f][k] = x0 + h;
        const double lp = loss(p);
        fields[f][k] = x0 - h;
        const double lm = loss(p);
        fields[f][k] = x0;
        const double fd = (lp - lm) / (2 * h);
        if (std::fabs(fd) < 1e-14 && std::fabs(acc[base[f] + k]) < 1e-12) continue;
        if (!close(acc[base[f] + k], fd, 1e-6, "node velocity d(snapshot)", base[f] + k)) return false;
      }
  }
  return true;
}

static void state_of(const DevParticle &p, double (&s)[kPState]) {
  const double *src[8] = {p.R, p.V, p.W, p.e, p.nR, p.nV, p.nW, p.ne};
  for (int i = 0; i < kPState; i++) s[i] = src[i / 3][i % 3];
}

static void set_state(DevParticle &p, const double (&s)[kPState]) {
  double *dst[8] = {p.R, p.V, p.W, p.e, p.nR, p.nV, p.nW, p.ne};
  for (int i = 0; i < kPState; i++) dst[i / 3][i % 3] = s[i];
}

static bool check_move() {
  const double dt = 1.0;
  for (int trial = 0; trial < 40; trial++) {
    DevParticle p = random_particle(trial % 5 == 0 ? LBM_PARTICLE_BEAD : LBM_PARTICLE_SQUIRMER);
    double F[3], T[3], mu[kPState];
    for (int k = 0; k < 3; k++) F[k] = U(-1e-2, 1e-2), T[k] = U(-1e-1, 1e-1);
    for (int i = 0; i < kPState; i++) mu[i] = U(-1, 1);
    // primal of the dual routine against move_gpu + the snapshot refresh
    DevParticle q = p;
    bool moved, find_nv;
    move_gpu(q, F, T, 3, dt, moved, find_nv);
    if (find_nv)
      for (int k = 0; k < 3; k++) q.nR[k] = q.R[k], q.nV[k] = q.V[k], q.nW[k] = q.W[k], q.ne[k] = q.e[k];
    DualState ds = move_dual(p, F, T, dt);
    double sq[kPState];
    state_of(q, sq);
    for (int i = 0; i < kPState; i++)
      if (ds.slot(i)->v != sq[i]) {
        std::printf("move primal[%d]: dual %.17g device %.17g\n", i, ds.slot(i)->v, sq[i]);
        return false;
      }
    double mu_in[kPState], phi[3], tau[3];
    move_vjp(p, F, T, dt, mu, mu_in, phi, tau);
    auto loss = [&](const double (&s)[kPState], const double *FF, const double *TT) {
      DevParticle r = p;
      set_state(r, s);
      bool m2, f2;
      move_gpu(r, FF, TT, 3, dt, m2, f2);
      if (f2)
        for (int k = 0; k < 3; k++) r.nR[k] = r.R[k], r.nV[k] = r.V[k], r.nW[k] = r.W[k], r.ne[k] = r.e[k];
      double o[kPState];
      state_of(r, o);
      double l = 0.0;
      for (int i = 0; i < kPState; i++) l += mu[i] * o[i];
      return l;
    };
    double s0[kPState];
    state_of(p, s0);
    for (int i = 0; i < kPState; i++) {
      const double h = 1e-6 * std::max(1.0, std::fabs(s0[i]));
      double sp[kPState], sm[kPState];
      for (int j = 0; j < kPState; j++) sp[j] = sm[j] = s0[j];
      sp[i] += h;
      sm[i] -= h;
      const double fd = (loss(sp, F, T) - loss(sm, F, T)) / (2 * h);
      if (!close(mu_in[i], fd, 1e-6, "move d(state)", i, 1e-3)) return false;
    }
    for (int k = 0; k < 6; k++) {
      double Fp[3] = {F[0], F[1], F[2]}, Fm[3] = {F[0], F[1], F[2]};
      double Tp[3] = {T[0], T[1], T[2]}, Tm[3] = {T[0], T[1], T[2]};
      const double h = 1e-6;
      if (k < 3)
        Fp[k] += h, Fm[k] -= h;
      else
        Tp[k - 3] += h, Tm[k - 3] -= h;
      const double fd = (loss(s0, Fp, Tp) - loss(s0, Fm, Tm)) / (2 * h);
      if (!close(k < 3 ? phi[k] : tau[k - 3], fd, 1e-6, "move d(force, torque)", k, 1e-3)) return false;
    }
  }
  return true;
}

int main() {
  if (!check_collision<LBM_COLLISION_BGK, LBM_FORCE_NONE>()) return 1;
  if (!check_collision<LBM_COLLISION_TRT, LBM_FORCE_NONE>()) return 1;
  if (!check_collision<LBM_COLLISION_BGK, LBM_FORCE_FIELD>()) return 1;
  if (!check_collision<LBM_COLLISION_TRT, LBM_FORCE_FIELD>()) return 1;
  if (!check_node_velocity()) return 1;
  if (!check_move()) return 1;
  std::printf("ok\n");
  return 0;
}
