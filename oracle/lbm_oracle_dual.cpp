/*
 * lbm_oracle_dual.cpp — forward-mode (dual-number) restatement of the
 * oracle's tick, for the derivative of a swimming-speed loss along one
 * direction in (B1, B2, force field).
 *
 * TEST INFRASTRUCTURE, like lbm_oracle.c: only tests/ load it, as the
 * checker of the HIP adjoint (lbm_adjoint_run).  Nothing in the product
 * links, loads or calls it.
 *
 * Every value is a dual number v + t eps.  The value parts follow
 * lbm_oracle.c operation by operation (GPU semantics; built with
 * -ffp-contract=off), so they reproduce the FP64 oracle bit for bit — the
 * tests check this — and the tangent parts carry the directional derivative:
 *   tick!                 gpu/fluids.jl:132-226     (lbm_oracle.c tick)
 *   collisionStep         gpu/fluids.jl:58-101      (collision)
 *   equilibriumDistribution gpu/fluids.jl:39-52     (feq_one)
 *   guoForcingTerm        gpu/fluids.jl:103-127     (guo_one)
 *   hydroVariablesUpdate! gpu/fluids.jl:7-37        (hydro_update)
 *   getNodeVelocity       gpu/aux.jl:229-263        (node_velocity_gpu)
 *   moveParticles!        gpu/particles.jl:7-92     (move_particle_gpu)
 * The geometry (getSphere, bounceBackPrep) is taken from the value parts:
 * the link set is piecewise constant and has no derivative.
 *
 * Build: g++ -O2 -std=c++17 -ffp-contract=off -shared -fPIC (oracle/Makefile).
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/lbm_hip.h"

namespace {

struct D {
  double v, t;
  D() : v(0.0), t(0.0) {}
  D(double x) : v(x), t(0.0) {}  // NOLINT: constants lift implicitly
  D(double x, double dx) : v(x), t(dx) {}
};
inline D operator+(D a, D b) { return D(a.v + b.v, a.t + b.t); }
inline D operator-(D a, D b) { return D(a.v - b.v, a.t - b.t); }
inline D operator-(D a) { return D(-a.v, -a.t); }
inline D operator*(D a, D b) { return D(a.v * b.v, a.t * b.v + a.v * b.t); }
inline D operator/(D a, D b) { return D(a.v / b.v, (a.t * b.v - a.v * b.t) / (b.v * b.v)); }
inline D dsqrt(D a) {
  const double r = std::sqrt(a.v);
  return D(r, a.v > 0.0 ? a.t / (2.0 * r) : 0.0);
}
inline D dcos(D a) { return D(std::cos(a.v), -std::sin(a.v) * a.t); }
inline D dsin(D a) { return D(std::sin(a.v), std::cos(a.v) * a.t); }

/* double-double value, plain tangent: the oracle's link sums */
struct DDD {
  double hi = 0.0, lo = 0.0, t = 0.0;
  void add(D x) {
    const double s = hi + x.v;
    const double bb = s - hi;
    const double err = (hi - (s - bb)) + (x.v - bb);
    hi = s;
    lo = lo + err;
    t = t + x.t;
  }
  D value() const { return D(hi + lo, t); }
};

struct Part {
  lbm_particle p;  // flags, radius, masses (value parts)
  D R[3], V[3], W[3], e[3], B1, B2;
  std::vector<uint8_t> solid, sir;
  std::vector<D> node_vel;  // d x N
  D force[3], torque[3];
  DDD fsum[3], tsum[3];
};

struct Model {
  lbm_config cfg;
  int d, Q;
  int64_t n[3], N;
  std::vector<double> X;  // d x N
  std::vector<D> f, fs, fn, rho, u, F;
  std::vector<uint8_t> wall, sir_wall;
  int guo = 0, bounce_back = 0;
  std::vector<Part> parts;
  int c[19][3];
  double w[19];
  int opp[19];
  double ls, invC2s, invC4s;
};

void tables(Model &m) {
  static const int c2[9][2] = {{0, 0}, {-1, 0}, {1, 0}, {0, -1}, {0, 1}, {-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
  static const int c3[19][3] = {{0, 0, 0},  {-1, 0, 0}, {1, 0, 0},  {0, -1, 0},  {0, 1, 0},
                                {0, 0, -1}, {0, 0, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1},
                                {0, 1, 1},  {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {1, 0, 1},
                                {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0}};
  std::memset(m.c, 0, sizeof m.c);
  for (int i = 0; i < m.Q; i++) {
    for (int k = 0; k < m.d; k++) m.c[i][k] = m.Q == 9 ? c2[i][k] : c3[i][k];
    m.w[i] = m.Q == 9 ? (i == 0 ? 4.0 / 9.0 : (i < 5 ? 1.0 / 9.0 : 1.0 / 36.0))
                      : (i == 0 ? 1.0 / 3.0 : (i < 7 ? 1.0 / 18.0 : 1.0 / 36.0));
  }
  for (int i = 0; i < m.Q; i++)
    for (int j = 0; j < m.Q; j++)
      if (m.c[j][0] == -m.c[i][0] && m.c[j][1] == -m.c[i][1] && m.c[j][2] == -m.c[i][2]) {
        m.opp[i] = j;
        break;
      }
}

inline int64_t wrap(int64_t a, int64_t n) {
  a %= n;
  return a < 0 ? a + n : a;
}

int64_t shifted(const Model &m, int64_t idx, const int s[3]) {
  int64_t x = idx % m.n[0], y = (idx / m.n[0]) % m.n[1], z = idx / (m.n[0] * m.n[1]);
  x = wrap(x - s[0], m.n[0]);
  y = wrap(y - s[1], m.n[1]);
  z = wrap(z - s[2], m.n[2]);
  return x + m.n[0] * (y + m.n[1] * z);
}

void bounce_back_prep(const Model &m, const std::vector<uint8_t> &region, std::vector<uint8_t> &sir) {
  sir.assign((size_t)m.Q * m.N, 0);
  for (int i = 0; i < m.Q; i++) {
    const int s[3] = {-m.c[i][0], -m.c[i][1], -m.c[i][2]};
    for (int64_t x = 0; x < m.N; x++) {
      const uint8_t a = region[shifted(m, x, s)], b = region[x];
      sir[(size_t)i * m.N + x] = (uint8_t)((a || b) ^ b);
    }
  }
}

void hydro_update(Model &m) {
  const double half_dt = 0.5 * m.cfg.dt;
  for (int64_t x = 0; x < m.N; x++) {
    D rho = m.f[x];
    for (int q = 1; q < m.Q; q++) rho = rho + m.f[(size_t)q * m.N + x];
    D mom[3];
    for (int k = 0; k < m.d; k++) {
      D acc = m.f[x] * D((double)m.c[0][k]);
      for (int q = 1; q < m.Q; q++) acc = acc + m.f[(size_t)q * m.N + x] * D((double)m.c[q][k]);
      mom[k] = acc;
    }
    if (m.guo)
      for (int k = 0; k < m.d; k++) mom[k] = mom[k] + D(half_dt) * m.F[(size_t)k * m.N + x];
    m.rho[x] = rho;
    for (int k = 0; k < m.d; k++) m.u[(size_t)k * m.N + x] = rho.v != 0.0 ? mom[k] / rho : D(0.0);
  }
}

D feq_one(const Model &m, int q, D rho, const D *u) {
  D ci[3], udc = 0.0, uu = 0.0;
  for (int k = 0; k < m.d; k++) ci[k] = D((double)m.c[q][k] * m.ls);
  for (int k = 0; k < m.d; k++) udc = udc + u[k] * ci[k];
  for (int k = 0; k < m.d; k++) uu = uu + u[k] * u[k];
  const D first = udc * D(m.invC2s) + (udc * udc) * D(0.5 * m.invC4s);
  const D second = first - uu * D(0.5 * m.invC2s);
  return D(m.w[q]) * ((second + D(1.0)) * rho);
}

D guo_one(const Model &m, int q, const D *U, const D *F) {
  const int d = m.d;
  D ci[3], udc = 0.0;
  for (int k = 0; k < d; k++) ci[k] = D((double)m.c[q][k] * m.ls);
  for (int k = 0; k < d; k++) udc = udc + U[k] * ci[k];
  D second[3];
  for (int k = 0; k < d; k++) second[k] = (udc * ci[k]) * D(m.invC4s);
  const double dt = m.cfg.dt, wi = m.w[q];
  if (m.cfg.collision == LBM_COLLISION_BGK) {
    D inter = 0.0;
    for (int k = 0; k < d; k++) inter = inter + ((-U[k] + ci[k]) * D(m.invC2s) + second[k]) * F[k];
    return D(wi * dt * (1.0 - dt / (2.0 * m.cfg.tau))) * inter;
  }
  D inter = 0.0;
  for (int k = 0; k < d; k++) inter = inter + ((-U[k]) * D(m.invC2s) + second[k]) * F[k];
  const D plus = D(1.0 - dt / (2.0 * m.cfg.tau)) * inter;
  D inter2 = 0.0;
  for (int k = 0; k < d; k++) inter2 = inter2 + F[k] * (ci[k] * D(m.invC2s));
  const D minus = D(1.0 - dt / (2.0 * m.cfg.tau_minus)) * inter2;
  return D(wi * dt) * (plus + minus);
}

void collision(Model &m) {
  const double cp = -m.cfg.dt / m.cfg.tau, cm = -m.cfg.dt / m.cfg.tau_minus;
  for (int64_t x = 0; x < m.N; x++) {
    D u[3], F[3], fe[19], fv[19];
    for (int k = 0; k < m.d; k++) u[k] = m.u[(size_t)k * m.N + x];
    if (m.guo)
      for (int k = 0; k < m.d; k++) F[k] = m.F[(size_t)k * m.N + x];
    const D rho = m.rho[x];
    for (int q = 0; q < m.Q; q++) {
      fe[q] = feq_one(m, q, rho, u);
      fv[q] = m.f[(size_t)q * m.N + x];
    }
    for (int q = 0; q < m.Q; q++) {
      D om;
      if (m.cfg.collision == LBM_COLLISION_BGK) {
        om = D(cp) * (fv[q] - fe[q]);
      } else {
        const int o = m.opp[q];
        const D fp = D(0.5) * (fv[q] + fv[o]), fm = D(0.5) * (fv[q] - fv[o]);
        const D ep = D(0.5) * (fe[q] + fe[o]), em = D(0.5) * (fe[q] - fe[o]);
        om = D(cp) * (fp - ep) + D(cm) * (fm - em);
      }
      if (m.guo) om = om + guo_one(m, q, u, F);
      m.fs[(size_t)q * m.N + x] = om + fv[q];
    }
  }
}

int sphere_member(const Model &m, int64_t x, const D *R, double r) {
  double s = 0.0;
  for (int k = 0; k < m.d; k++) {
    const double a = m.X[(size_t)k * m.N + x] + (-R[k].v);
    s = k == 0 ? a * a : s + a * a;
  }
  return s < r * r;
}

void node_velocity_gpu(const Model &m, const Part &pp, int64_t x, D *out) {
  const int d = m.d;
  D r[3];
  for (int k = 0; k < d; k++) r[k] = D(m.X[(size_t)k * m.N + x]) + (-pp.R[k]);
  D nn = 0.0;
  for (int k = 0; k < d; k++) nn = nn + r[k] * r[k];
  nn = dsqrt(nn);
  if (nn.v == 0.0) nn = D(1.0);
  D bulk[3];
  const D *V = pp.V, *W = pp.W;
  if (d == 2) {
    bulk[0] = V[0] - W[0] * r[1];
    bulk[1] = V[1] + W[0] * r[0];
  } else {
    bulk[0] = V[0] + W[1] * r[2] - W[2] * r[1];
    bulk[1] = V[1] - W[0] * r[2] + W[2] * r[0];
    bulk[2] = V[2] + W[0] * r[1] - W[1] * r[0];
  }
  if (pp.p.kind == LBM_PARTICLE_BEAD) {
    for (int k = 0; k < d; k++) out[k] = bulk[k];
    return;
  }
  const D inv = D(1.0) / nn;
  D unit[3], edr = 0.0;
  for (int k = 0; k < d; k++) unit[k] = inv * r[k];
  for (int k = 0; k < d; k++) edr = edr + unit[k] * pp.e[k];
  const D first = pp.B1 + pp.B2 * edr;
  for (int k = 0; k < d; k++) {
    const D second = edr * unit[k] + (-pp.e[k]);
    out[k] = first * second + bulk[k];
  }
}

void rebuild_geometry(Model &m, Part &pp) {
  pp.solid.assign(m.N, 0);
  for (int64_t x = 0; x < m.N; x++) pp.solid[x] = (uint8_t)sphere_member(m, x, pp.R, pp.p.radius);
  bounce_back_prep(m, pp.solid, pp.sir);
}

void rebuild_node_velocity(Model &m, Part &pp) {
  pp.node_vel.assign((size_t)m.d * m.N, D(0.0));
  for (int64_t x = 0; x < m.N; x++) {
    D v[3];
    node_velocity_gpu(m, pp, x, v);
    for (int k = 0; k < m.d; k++) pp.node_vel[(size_t)k * m.N + x] = v[k];
  }
}

bool any_nonzero(const D *v, int n) {
  for (int k = 0; k < n; k++)
    if (v[k].v != 0.0) return true;
  return false;
}

void move_particle_gpu(Model &m, Part &pp) {
  const int d = m.d;
  const double dt = m.cfg.dt;
  bool moved = false, find_nv = false;
  if (pp.p.couple_forces) {
    D A[3];
    for (int k = 0; k < d; k++) A[k] = pp.force[k] * D(pp.p.inverse_mass);
    for (int k = 0; k < d; k++) pp.V[k] = pp.V[k] + D(dt) * A[k];
    for (int k = 0; k < d; k++) pp.R[k] = pp.R[k] + D(dt) * pp.V[k];
    if (any_nonzero(pp.V, d)) moved = true;
  }
  if (pp.p.couple_torques) {
    const int nw = d == 2 ? 1 : 3;
    D Al[3];
    for (int k = 0; k < nw; k++) Al[k] = pp.torque[k] * D(pp.p.inverse_moment_of_inertia);
    for (int k = 0; k < nw; k++) pp.W[k] = pp.W[k] + D(dt) * Al[k];
    if (pp.p.kind == LBM_PARTICLE_SQUIRMER) {
      D *e = pp.e, *W = pp.W;
      D theta;
      if (d == 2) {
        theta = W[0].v < 0.0 ? -W[0] : W[0];
      } else {
        D s = 0.0;
        for (int k = 0; k < 3; k++) s = s + W[k] * W[k];
        theta = dsqrt(s);
      }
      if (!(theta.v == 0.0)) {
        D vr[3];
        if (d == 2) {
          const D kk = W[0] / theta;
          theta = theta * D(dt);
          const D c = dcos(theta), s = dsin(theta);
          const D cr[2] = {-kk * e[1], kk * e[0]};
          vr[0] = e[0] * c + cr[0] * s;
          vr[1] = e[1] * c + cr[1] * s;
          const D nv = dsqrt(vr[0] * vr[0] + vr[1] * vr[1]);
          e[0] = vr[0] / nv;
          e[1] = vr[1] / nv;
        } else {
          const D kk[3] = {W[0] / theta, W[1] / theta, W[2] / theta};
          theta = theta * D(dt);
          const D c = dcos(theta), s = dsin(theta);
          const D cr[3] = {kk[1] * e[2] - kk[2] * e[1], -kk[0] * e[2] + kk[2] * e[0], kk[0] * e[1] - kk[1] * e[0]};
          for (int k = 0; k < 3; k++) vr[k] = e[k] * c + cr[k] * s;
          D kd = 0.0;
          for (int k = 0; k < 3; k++) kd = kd + kk[k] * e[k];
          for (int k = 0; k < 3; k++) vr[k] = vr[k] + kk[k] * kd * (D(1.0) - c);
          D nv = 0.0;
          for (int k = 0; k < 3; k++) nv = nv + vr[k] * vr[k];
          nv = dsqrt(nv);
          for (int k = 0; k < 3; k++) e[k] = vr[k] / nv;
        }
      }
    }
    if (any_nonzero(Al, nw)) find_nv = true;
  }
  find_nv = moved || find_nv;
  for (int k = 0; k < 3; k++) {
    pp.force[k] = D(0.0);
    pp.torque[k] = D(0.0);
  }
  if (moved) rebuild_geometry(m, pp);
  if (find_nv) rebuild_node_velocity(m, pp);
}

void tick(Model &m) {
  const int Q = m.Q, d = m.d;
  double coeff = 1.0;
  for (int k = 0; k < d; k++) coeff *= m.cfg.dx;
  coeff = coeff / m.cfg.dt;
  collision(m);
  for (int q = 0; q < Q; q++) {
    const int o = m.opp[q];
    const D *fsq = &m.fs[(size_t)q * m.N];
    const D *fso = &m.fs[(size_t)o * m.N];
    D *st = &m.fn[(size_t)q * m.N];
    for (int64_t x = 0; x < m.N; x++) st[x] = fsq[shifted(m, x, m.c[q])];
    if (!m.bounce_back) continue;
    const uint8_t *cir = &m.sir_wall[(size_t)o * m.N];
    for (int64_t x = 0; x < m.N; x++) {
      if (m.wall[x]) st[x] = D(0.0);
      int in = cir[x];
      for (const Part &pp : m.parts) in = in || pp.sir[(size_t)o * m.N + x];
      if (in) st[x] = fso[x];
    }
    if (m.parts.empty()) continue;
    D ci[3];
    for (int k = 0; k < d; k++) ci[k] = D((double)m.c[q][k] * m.ls);
    const double lc = 2.0 * m.w[q] * m.invC2s;
    for (Part &pp : m.parts) {
      const uint8_t *cbn = &pp.sir[(size_t)o * m.N];
      int64_t cnt = 0;
      for (int64_t x = 0; x < m.N; x++) cnt += cbn[x];
      if (cnt == 0) break;
      for (int64_t x = 0; x < m.N; x++) {
        if (!cbn[x]) continue;
        const int64_t y = shifted(m, x, m.c[q]);
        D udc = 0.0;
        for (int k = 0; k < d; k++) udc = udc + pp.node_vel[(size_t)k * m.N + y] * ci[k];
        st[x] = st[x] + D(lc) * m.rho[x] * udc;
      }
      if (!(pp.p.couple_forces || pp.p.couple_torques)) continue;
      for (int64_t x = 0; x < m.N; x++) {
        if (!cbn[x]) continue;
        const D s = fso[x] + st[x];
        D p[3];
        for (int k = 0; k < d; k++) p[k] = s * (D(m.X[(size_t)k * m.N + x]) + (-pp.R[k]));
        if (pp.p.couple_forces)
          for (int k = 0; k < d; k++)
            if (m.c[q][k]) pp.fsum[k].add(m.c[q][k] > 0 ? s : -s);
        if (pp.p.couple_torques) {
          const int *c = m.c[q];
          if (d == 2) {
            if (c[1]) pp.tsum[0].add(c[1] > 0 ? p[0] : -p[0]);
            if (c[0]) pp.tsum[0].add(c[0] > 0 ? -p[1] : p[1]);
          } else {
            if (c[2]) pp.tsum[0].add(c[2] > 0 ? p[1] : -p[1]);
            if (c[1]) pp.tsum[0].add(c[1] > 0 ? -p[2] : p[2]);
            if (c[2]) pp.tsum[1].add(c[2] > 0 ? -p[0] : p[0]);
            if (c[0]) pp.tsum[1].add(c[0] > 0 ? p[2] : -p[2]);
            if (c[1]) pp.tsum[2].add(c[1] > 0 ? p[0] : -p[0]);
            if (c[0]) pp.tsum[2].add(c[0] > 0 ? -p[1] : p[1]);
          }
        }
      }
    }
  }
  for (Part &pp : m.parts) {
    const int nw = d == 2 ? 1 : 3;
    for (int k = 0; k < 3; k++) {
      const D F = k < d ? D(-coeff) * (pp.fsum[k].value() * D(m.ls)) : D(0.0);
      const D T = k < nw ? D(-coeff) * (pp.tsum[k].value() * D(m.ls)) : D(0.0);
      if (pp.p.couple_forces) pp.force[k] = F;
      if (pp.p.couple_torques) pp.torque[k] = T;
      pp.fsum[k] = DDD();
      pp.tsum[k] = DDD();
    }
  }
  std::swap(m.f, m.fn);
  hydro_update(m);
  for (Part &pp : m.parts) move_particle_gpu(m, pp);
}

}  // namespace

extern "C" {

/* Loss and its derivative along (dB1, dB2, dF) from the populations f0
 * (Q x N, Julia layout) over `ticks` ticks:
 *   L = (mean |V_swimmer| over the last `window` ticks - target)^2 + penalty sum F^2,
 * F the model's force field (zero at wall nodes).
 * dB1, dB2: one per particle; F, dF: (d x N) force field and its direction
 * (NULL: none).  speeds (optional, `window` values): the swimmer's speeds. */
int lbmd_loss_jvp(const lbm_config *cfg, const double *f0, const double *F, int32_t np, const lbm_particle *parts,
                  const double *dB1, const double *dB2, const double *dF, int64_t ticks, int64_t window,
                  int32_t swimmer, double target, double penalty, double *loss, double *dloss, double *speeds) {
  if (!cfg || !f0 || !parts || np < 1 || np > LBM_MAX_PARTICLES || ticks < 1 || window < 1 || window > ticks ||
      swimmer < 0 || swimmer >= np || !loss || !dloss)
    return LBM_EINVAL;
  if (cfg->semantics != LBM_SEMANTICS_GPU) return LBM_EINVAL;
  Model m;
  m.cfg = *cfg;
  m.d = cfg->dims;
  m.Q = cfg->q;
  m.n[0] = cfg->n[0];
  m.n[1] = cfg->n[1];
  m.n[2] = cfg->dims == 3 ? cfg->n[2] : 1;
  m.N = m.n[0] * m.n[1] * m.n[2];
  tables(m);
  m.ls = cfg->dx / cfg->dt;
  m.invC2s = 3.0 / (m.ls * m.ls);
  m.invC4s = 9.0 / (m.ls * m.ls * m.ls * m.ls);
  const int64_t N = m.N;
  m.X.assign((size_t)m.d * N, 0.0);
  m.wall.assign(N, 0);
  for (int64_t x = 0; x < N; x++) {
    const int64_t id[3] = {x % m.n[0], (x / m.n[0]) % m.n[1], x / (m.n[0] * m.n[1])};
    for (int k = 0; k < m.d; k++) m.X[(size_t)k * N + x] = cfg->coords[k][id[k]];
    for (int k = 0; k < m.d; k++)
      if (((cfg->walled >> k) & 1) && (id[k] == 0 || id[k] == m.n[k] - 1)) m.wall[x] = 1;
  }
  m.sir_wall.assign((size_t)m.Q * N, 0);
  if (cfg->walled) bounce_back_prep(m, m.wall, m.sir_wall);
  m.bounce_back = 1;
  m.f.resize((size_t)m.Q * N);
  m.fs.assign((size_t)m.Q * N, D(0.0));
  m.fn.assign((size_t)m.Q * N, D(0.0));
  m.rho.assign(N, D(0.0));
  m.u.assign((size_t)m.d * N, D(0.0));
  for (size_t i = 0; i < m.f.size(); i++) m.f[i] = D(f0[i]);
  D Fsq = 0.0;
  if (cfg->force != LBM_FORCE_NONE) {
    m.guo = 1;
    m.F.assign((size_t)m.d * N, D(0.0));
    for (int64_t x = 0; x < N; x++)
      for (int k = 0; k < m.d; k++) {
        const size_t i = (size_t)k * N + x;
        if (cfg->force == LBM_FORCE_UNIFORM)
          m.F[i] = D(m.wall[x] ? 0.0 : cfg->force_uniform[k]);
        else if (F)
          m.F[i] = m.wall[x] ? D(0.0) : D(F[i], dF ? dF[i] : 0.0);
      }
    // the penalty is on the model's field, which is zero at wall nodes
    if (cfg->force == LBM_FORCE_FIELD && F)
      for (size_t i = 0; i < m.F.size(); i++) Fsq = Fsq + m.F[i] * m.F[i];
  }
  hydro_update(m);
  for (int p = 0; p < np; p++) {
    Part pp;
    pp.p = parts[p];
    for (int k = 0; k < 3; k++) {
      pp.R[k] = D(parts[p].position[k]);
      pp.V[k] = D(parts[p].velocity[k]);
      pp.W[k] = D(parts[p].angular_velocity[k]);
      pp.e[k] = D(parts[p].swimming_direction[k]);
    }
    pp.B1 = D(parts[p].B1, dB1 ? dB1[p] : 0.0);
    pp.B2 = D(parts[p].B2, dB2 ? dB2[p] : 0.0);
    m.parts.push_back(pp);
  }
  for (Part &pp : m.parts) {
    rebuild_geometry(m, pp);
    rebuild_node_velocity(m, pp);
  }
  D mean = 0.0;
  for (int64_t t = 0; t < ticks; t++) {
    tick(m);
    if (t >= ticks - window) {
      const Part &s = m.parts[swimmer];
      D v2 = 0.0;
      for (int k = 0; k < m.d; k++) v2 = v2 + s.V[k] * s.V[k];
      const D sp = dsqrt(v2);
      if (speeds) speeds[t - (ticks - window)] = sp.v;
      mean = mean + sp;
    }
  }
  mean = mean / D((double)window);
  const D L = (mean - D(target)) * (mean - D(target)) + D(penalty) * Fsq;
  *loss = L.v;
  *dloss = L.t;
  return LBM_OK;
}

}  // extern "C"
