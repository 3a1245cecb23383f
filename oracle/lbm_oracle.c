/*
 * lbm_oracle.c — CPU restatement, in FP64 C, of LBMengine.jl's time step.
 *
 * TEST INFRASTRUCTURE.  This library is the checker for liblbm_hip.so: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * Nothing in the product (lbmengine.jl_amd/) links, loads or calls it.
 *
 * It follows the reference array by array, in the reference's order of
 * operations, so that it is independent of the node-local formulation the
 * HIP kernels use:
 *   tick!                 gpu/fluids.jl:132-226     cpu/fluids.jl:377-461
 *   collisionStep         gpu/fluids.jl:58-101      cpu/fluids.jl:69-130,204-214
 *   equilibriumDistribution gpu/fluids.jl:39-52     cpu/fluids.jl:46-63
 *   guoForcingTerm        gpu/fluids.jl:103-127     cpu/fluids.jl:217-241
 *   hydroVariablesUpdate! gpu/fluids.jl:7-37        cpu/fluids.jl:7-38
 *   circshift_gpu         gpu/aux.jl:127-137, gpu/kernels.jl:254-289
 *   bounceBackPrep        gpu/aux.jl:187-198        cpu/aux.jl:93-103
 *   getSphere             gpu/aux.jl:206-215, gpu/kernels.jl:335-353
 *   getNodeVelocity       gpu/aux.jl:229-263, gpu/kernels.jl:376-415
 *   bulkVelocity          cpu/particles.jl:20-43
 *   moveParticles!        gpu/particles.jl:7-92     cpu/particles.jl:45-100
 *   molecularDynamicsTick!, applyInteraction!  gpu/particles.jl:100-197
 *   viscousStressTensor   cpu/fluids.jl:507-520 (nonEquilibriumDistribution :65-67)
 *   findInitialConditions gpu/initMethods.jl:417-438
 *   wallNodes             gpu/misc.jl:7-31
 *   walls, movingWalls    gpu/initMethods.jl:591-626, gpu/fluids.jl:168-174
 *   D2Q9, D3Q15, D3Q19, D3Q27  gpu/misc.jl:195-272
 *
 * Summation order: where the reference sums with CUDA.jl/BLAS reductions of
 * unspecified order (sum over q for rho, the momentum-exchange sums, norm,
 * dot), this file sums sequentially in index order; that choice is
 * documented as unpinned at the ~1e-16 relative level.
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -shared -fPIC (oracle/Makefile).
 */
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/lbm_hip.h"

#define OMAXQ 27

static __thread char g_err[512];

static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char *lbmo_last_error(void) { return g_err; }

/* Double-double accumulation (Knuth's TwoSum): the momentum-exchange sums are
 * evaluated essentially exactly and rounded once, so their value does not
 * depend on the order in which links are visited. */
typedef struct {
  double hi, lo;
} dd_t;

static inline void dd_add(dd_t *a, double x) {
  double s = a->hi + x;
  double bb = s - a->hi;
  double err = (a->hi - (s - bb)) + (x - bb);
  a->hi = s;
  a->lo = a->lo + err;
}

static inline double dd_value(dd_t a) { return a.hi + a.lo; }

typedef struct {
  lbm_particle p;      /* parameters and current state                 */
  uint8_t *solid;      /* particle.boundaryConditionsParams.solidRegion */
  uint8_t *sir;        /* streamingInvasionRegions, Q x N              */
  double *node_vel;    /* particle.nodeVelocity, d x N                 */
  double force[3];     /* forceInput  (gpu)                            */
  double torque[3];    /* torqueInput (gpu); [0] is the 2D scalar      */
  double mom[OMAXQ + 1][3];  /* momentumInput per q (cpu)              */
  double amom[OMAXQ + 1][3]; /* angularMomentumInput per q (cpu)       */
  double last_force[3], last_torque[3];
  dd_t fsum[3], tsum[3]; /* sum over links of s c_i and of (s (x - R)) x c_i */
} opart;

typedef struct lbmo_model {
  lbm_config cfg;
  int d, Q;
  int64_t n[3], N;
  double *coords[3];
  double *X;          /* spaceTime.X, d x N                             */
  double *f, *fs, *fn;/* distributions, collisioned, propagated; Q x N  */
  double *rho, *u;    /* massDensity (N), fluidVelocity (d x N)         */
  double *F;          /* forceDensity (d x N) when :guo                 */
  uint8_t *wall;      /* boundaryConditionsParams.wallRegion            */
  uint8_t *sir_wall;  /* boundaryConditionsParams.streamingInvasionRegions */
  double *solid_vel;  /* boundaryConditionsParams.solidNodeVelocity, d x N (movingWalls) */
  int incompressible; /* fluidParams.isFluidCompressible == false       */
  int guo, bounce_back, ladd;
  int np;
  opart parts[LBM_MAX_PARTICLES];
  int ninter;
  lbm_interaction inter[LBM_MAX_INTERACTIONS]; /* particleInteractions, pairs expanded */
  /* a magnetic field that is a function of time: field_fn(model.time, B, ctx)
   * (DipoleDipoleInteraction.magneticField(model.time), gpu/particles.jl:161) */
  void (*field_fn)(double t, double *B, void *ctx);
  void *field_ctx;
  int64_t tick;
  double time;
  int c[OMAXQ][3];
  double w[OMAXQ];
  int opp[OMAXQ];
  double ls, invC2s, invC4s;
} lbmo_model;

/* ---------------------------------------------------------------- tables */

/* The velocity sets in the reference's order (gpu/misc.jl:195-272): rest,
 * the axis pairs, then comprehensions [[p,q] for p in [-1,1], q in [-1,1]]
 * (and [[p,q,r] ...]) run p fastest. */
static void tables(lbmo_model *m) {
  static const int c2[9][2] = {{0, 0},  {-1, 0}, {1, 0},  {0, -1}, {0, 1},
                               {-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
  /* rest, axes, edges of the yz, xz, xy planes, corners */
  static const int c3[27][3] = {
      {0, 0, 0},   {-1, 0, 0},  {1, 0, 0},  {0, -1, 0}, {0, 1, 0},
      {0, 0, -1},  {0, 0, 1},   {0, -1, -1}, {0, 1, -1}, {0, -1, 1},
      {0, 1, 1},   {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {1, 0, 1},
      {-1, -1, 0}, {1, -1, 0},  {-1, 1, 0}, {1, 1, 0},
      {-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1},
      {-1, -1, 1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1}};
  memset(m->c, 0, sizeof m->c);
  if (m->Q == 9) {
    for (int i = 0; i < 9; i++) {
      m->c[i][0] = c2[i][0];
      m->c[i][1] = c2[i][1];
      m->w[i] = i == 0 ? 4.0 / 9.0 : (i < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
    }
  } else if (m->Q == 15) { /* rest, axes, corners: 2/9, 1/9, 1/72 */
    for (int i = 0; i < 15; i++) {
      const int src = i < 7 ? i : i + 12;
      for (int k = 0; k < 3; k++) m->c[i][k] = c3[src][k];
      m->w[i] = i == 0 ? 2.0 / 9.0 : (i < 7 ? 1.0 / 9.0 : 1.0 / 72.0);
    }
  } else if (m->Q == 19) { /* 1/3, 1/18, 1/36 */
    for (int i = 0; i < 19; i++) {
      for (int k = 0; k < 3; k++) m->c[i][k] = c3[i][k];
      m->w[i] = i == 0 ? 1.0 / 3.0 : (i < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
    }
  } else { /* 8/27, 2/27, 1/54, 1/216 */
    for (int i = 0; i < 27; i++) {
      for (int k = 0; k < 3; k++) m->c[i][k] = c3[i][k];
      m->w[i] = i == 0 ? 8.0 / 27.0 : (i < 7 ? 2.0 / 27.0 : (i < 19 ? 1.0 / 54.0 : 1.0 / 216.0));
    }
  }
  /* oppositeVectorId = [findfirst(x -> x == -1 .* c, cs) for c in cs] */
  for (int i = 0; i < m->Q; i++)
    for (int j = 0; j < m->Q; j++)
      if (m->c[j][0] == -m->c[i][0] && m->c[j][1] == -m->c[i][1] &&
          m->c[j][2] == -m->c[i][2]) {
        m->opp[i] = j;
        break;
      }
}

/* ------------------------------------------------------------ utilities */

static inline int64_t wrap(int64_t a, int64_t n) {
  a %= n;
  return a < 0 ? a + n : a;
}

/* circshift: out[x] = in[x - s] with periodic wrap (gpu/kernels.jl:254-289) */
static int64_t shifted(const lbmo_model *m, int64_t idx, const int s[3]) {
  int64_t x = idx % m->n[0];
  int64_t y = (idx / m->n[0]) % m->n[1];
  int64_t z = idx / (m->n[0] * m->n[1]);
  x = wrap(x - s[0], m->n[0]);
  y = wrap(y - s[1], m->n[1]);
  z = wrap(z - s[2], m->n[2]);
  return x + m->n[0] * (y + m->n[1] * z);
}

/* bounceBackPrep: SIR_i = (circshift(region, -c_i) .|| region) .⊻ region,
 * i.e. SIR_i(x) = region(x + c_i) and not region(x)  (gpu/aux.jl:187-198) */
static void bounce_back_prep(const lbmo_model *m, const uint8_t *region,
                             uint8_t *sir) {
  for (int i = 0; i < m->Q; i++) {
    int s[3] = {-m->c[i][0], -m->c[i][1], -m->c[i][2]};
    uint8_t *out = sir + (int64_t)i * m->N;
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < m->N; x++) {
      uint8_t a = region[shifted(m, x, s)];
      uint8_t b = region[x];
      out[x] = (uint8_t)((a || b) ^ b);
    }
  }
}

/* --------------------------------------------------------------- hydro */

/* hydroVariablesUpdate! (gpu/fluids.jl:7-37): rho = sum_q f_q; rho u =
 * sum_q f_q c_q (+ dt F / 2 under :guo); u = rho u / rho where rho != 0. */
static void hydro_update(lbmo_model *m) {
  const int d = m->d, Q = m->Q;
  const double half_dt = 0.5 * m->cfg.dt;
#pragma omp parallel for schedule(static)
  for (int64_t x = 0; x < m->N; x++) {
    double rho = m->f[x];
    for (int q = 1; q < Q; q++) rho = rho + m->f[(int64_t)q * m->N + x];
    double mom[3] = {0, 0, 0};
    for (int k = 0; k < d; k++) {
      double acc = m->f[x] * (double)m->c[0][k];
      for (int q = 1; q < Q; q++)
        acc = acc + m->f[(int64_t)q * m->N + x] * (double)m->c[q][k];
      mom[k] = acc;
    }
    if (m->guo)
      for (int k = 0; k < d; k++)
        mom[k] = mom[k] + half_dt * m->F[(int64_t)k * m->N + x];
    m->rho[x] = rho;
    for (int k = 0; k < d; k++)
      m->u[(int64_t)k * m->N + x] = rho != 0.0 ? mom[k] / rho : 0.0;
  }
}

/* ----------------------------------------------------------- collision */

/* equilibriumDistribution (gpu/fluids.jl:39-52): compressible
 * w_i ((u.c_i invC2 + (u.c_i)^2 invC4/2 - u.u invC2/2) + 1) rho, or
 * incompressible w_i rho + w_i (rho0 (...)) with the initial density rho0 */
static inline double feq_one(const lbmo_model *m, int q, double rho,
                             const double *u) {
  double ci[3], udc = 0.0, uu = 0.0;
  for (int k = 0; k < m->d; k++) ci[k] = (double)m->c[q][k] * m->ls;
  for (int k = 0; k < m->d; k++) udc = udc + u[k] * ci[k];
  for (int k = 0; k < m->d; k++) uu = uu + u[k] * u[k];
  double first = udc * m->invC2s + (udc * udc) * (0.5 * m->invC4s);
  double second = first - uu * (0.5 * m->invC2s);
  if (m->incompressible) return m->w[q] * rho + m->w[q] * (m->cfg.rho0 * second);
  return m->w[q] * ((second + 1.0) * rho);
}

/* guoForcingTerm (gpu/fluids.jl:103-127) */
static inline double guo_one(const lbmo_model *m, int q, const double *U,
                             const double *F) {
  const int d = m->d;
  double ci[3], udc = 0.0;
  for (int k = 0; k < d; k++) ci[k] = (double)m->c[q][k] * m->ls;
  for (int k = 0; k < d; k++) udc = udc + U[k] * ci[k];
  double second[3];
  for (int k = 0; k < d; k++) second[k] = (udc * ci[k]) * m->invC4s;
  const double dt = m->cfg.dt, wi = m->w[q];
  if (m->cfg.collision == LBM_COLLISION_BGK) {
    double inter = 0.0;
    for (int k = 0; k < d; k++)
      inter = inter + ((-U[k] + ci[k]) * m->invC2s + second[k]) * F[k];
    return wi * dt * (1.0 - dt / (2.0 * m->cfg.tau)) * inter;
  }
  double inter = 0.0;
  for (int k = 0; k < d; k++)
    inter = inter + ((-U[k]) * m->invC2s + second[k]) * F[k];
  double plus = (1.0 - dt / (2.0 * m->cfg.tau)) * inter;
  double inter2 = 0.0;
  for (int k = 0; k < d; k++) inter2 = inter2 + F[k] * (ci[k] * m->invC2s);
  double minus = (1.0 - dt / (2.0 * m->cfg.tau_minus)) * inter2;
  return wi * dt * (plus + minus);
}

/* collisionStep (gpu/fluids.jl:58-101): fs = Omega + f, with
 * BGK Omega = -dt/tau (f - feq);  TRT Omega = -dt/tau+ (f+ - feq+)
 * - dt/tau- (f- - feq-);  Omega += S_guo.                                */
static void collision(lbmo_model *m) {
  const int Q = m->Q, d = m->d;
  const double cb = -m->cfg.dt / m->cfg.tau;
  const double cp = -m->cfg.dt / m->cfg.tau;
  const double cm = -m->cfg.dt / m->cfg.tau_minus;
#pragma omp parallel for schedule(static)
  for (int64_t x = 0; x < m->N; x++) {
    double u[3] = {0, 0, 0}, F[3] = {0, 0, 0}, fe[OMAXQ], fv[OMAXQ];
    for (int k = 0; k < d; k++) u[k] = m->u[(int64_t)k * m->N + x];
    if (m->guo)
      for (int k = 0; k < d; k++) F[k] = m->F[(int64_t)k * m->N + x];
    const double rho = m->rho[x];
    for (int q = 0; q < Q; q++) {
      fe[q] = feq_one(m, q, rho, u);
      fv[q] = m->f[(int64_t)q * m->N + x];
    }
    for (int q = 0; q < Q; q++) {
      double om;
      if (m->cfg.collision == LBM_COLLISION_BGK) {
        om = cb * (fv[q] - fe[q]);
      } else {
        const int o = m->opp[q];
        double fp = 0.5 * (fv[q] + fv[o]), fm = 0.5 * (fv[q] - fv[o]);
        double ep = 0.5 * (fe[q] + fe[o]), em = 0.5 * (fe[q] - fe[o]);
        om = cp * (fp - ep) + cm * (fm - em);
      }
      if (m->guo) om = om + guo_one(m, q, u, F);
      m->fs[(int64_t)q * m->N + x] = om + fv[q];
    }
  }
}

/* ------------------------------------------------------------ particles */

static int sphere_member(const lbmo_model *m, int64_t x, const double *R,
                         double r) {
  /* getSphere: (X+(-R))_1^2 + (X+(-R))_2^2 [+ ...]_3^2 < r^2, no wrap;
   * cpu: sum((x - R).^2) < radius^2 — the same left-to-right sum */
  double s = 0.0;
  for (int k = 0; k < m->d; k++) {
    double a = m->X[(int64_t)k * m->N + x] + (-R[k]);
    s = k == 0 ? a * a : s + a * a;
  }
  return s < r * r;
}

/* getNodeVelocity (gpu/aux.jl:229-263, kernels.jl:376-415) at node x */
static void node_velocity_gpu(const lbmo_model *m, const opart *pp, int64_t x,
                              double *out) {
  const lbm_particle *p = &pp->p;
  const int d = m->d;
  double r[3] = {0, 0, 0};
  for (int k = 0; k < d; k++) r[k] = m->X[(int64_t)k * m->N + x] + (-p->position[k]);
  double nn = 0.0;
  for (int k = 0; k < d; k++) nn = nn + r[k] * r[k];
  nn = sqrt(nn);
  if (nn == 0.0) nn = 1.0;
  double bulk[3];
  const double *V = p->velocity, *W = p->angular_velocity;
  if (d == 2) {
    bulk[0] = V[0] - W[0] * r[1];
    bulk[1] = V[1] + W[0] * r[0];
  } else {
    bulk[0] = V[0] + W[1] * r[2] - W[2] * r[1];
    bulk[1] = V[1] - W[0] * r[2] + W[2] * r[0];
    bulk[2] = V[2] + W[0] * r[1] - W[1] * r[0];
  }
  if (p->kind == LBM_PARTICLE_BEAD) {
    for (int k = 0; k < d; k++) out[k] = bulk[k];
    return;
  }
  double inv = 1.0 / nn, unit[3], edr = 0.0;
  for (int k = 0; k < d; k++) unit[k] = inv * r[k];
  for (int k = 0; k < d; k++) edr = edr + unit[k] * p->swimming_direction[k];
  double first = p->B1 + p->B2 * edr;
  for (int k = 0; k < d; k++) {
    double second = edr * unit[k] + (-p->swimming_direction[k]);
    out[k] = first * second + bulk[k];
  }
}

/* bulkVelocity (cpu/particles.jl:20-43) at a solid node x */
static void node_velocity_cpu(const lbmo_model *m, const opart *pp, int64_t x,
                              double *out) {
  const lbm_particle *p = &pp->p;
  const int d = m->d;
  double r[3] = {0, 0, 0};
  for (int k = 0; k < d; k++) r[k] = m->X[(int64_t)k * m->N + x] - p->position[k];
  double nn = 0.0;
  for (int k = 0; k < d; k++) nn = nn + r[k] * r[k];
  nn = sqrt(nn);
  for (int k = 0; k < d; k++) out[k] = 0.0;
  if (fabs(nn - p->radius) > 2.0 * m->cfg.dx) return;
  const double *V = p->velocity, *W = p->angular_velocity;
  double bulk[3];
  if (d == 2) { /* cross(omega::Real, V) = cross([0;0;w],[V;0])[1:2] */
    bulk[0] = V[0] + (0.0 * 0.0 - W[0] * r[1]);
    bulk[1] = V[1] + (-0.0 * 0.0 + W[0] * r[0]);
  } else {
    bulk[0] = V[0] + (W[1] * r[2] - W[2] * r[1]);
    bulk[1] = V[1] + (-W[0] * r[2] + W[2] * r[0]);
    bulk[2] = V[2] + (W[0] * r[1] - W[1] * r[0]);
  }
  if (p->kind == LBM_PARTICLE_BEAD || nn == 0.0) {
    for (int k = 0; k < d; k++) out[k] = bulk[k];
    return;
  }
  double fr[3], edr = 0.0;
  for (int k = 0; k < d; k++) fr[k] = r[k] / nn;
  for (int k = 0; k < d; k++) edr = edr + p->swimming_direction[k] * fr[k];
  double first = p->B1 + p->B2 * edr;
  for (int k = 0; k < d; k++) {
    double second = edr * fr[k] - p->swimming_direction[k];
    out[k] = first * second + bulk[k];
  }
}

static void rebuild_geometry(lbmo_model *m, opart *pp) {
#pragma omp parallel for schedule(static)
  for (int64_t x = 0; x < m->N; x++)
    pp->solid[x] = (uint8_t)sphere_member(m, x, pp->p.position, pp->p.radius);
  bounce_back_prep(m, pp->solid, pp->sir);
}

static void rebuild_node_velocity(lbmo_model *m, opart *pp) {
  const int d = m->d;
  if (m->cfg.semantics == LBM_SEMANTICS_GPU) {
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < m->N; x++) {
      double v[3];
      node_velocity_gpu(m, pp, x, v);
      for (int k = 0; k < d; k++) pp->node_vel[(int64_t)k * m->N + x] = v[k];
    }
  } else {
    /* nodeVelocity[solidRegion] = bulkVelocity(...), zero elsewhere */
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < m->N; x++) {
      double v[3] = {0, 0, 0};
      if (pp->solid[x]) node_velocity_cpu(m, pp, x, v);
      for (int k = 0; k < d; k++) pp->node_vel[(int64_t)k * m->N + x] = v[k];
    }
  }
}

static inline int any_nonzero(const double *v, int n) {
  for (int k = 0; k < n; k++)
    if (v[k] != 0.0) return 1;
  return 0;
}

/* moveParticles! — gpu/particles.jl:7-92 */
static void move_particle_gpu(lbmo_model *m, opart *pp) {
  lbm_particle *p = &pp->p;
  const int d = m->d;
  const double dt = m->cfg.dt;
  int moved = 0, find_nv = 0;
  if (p->couple_forces) {
    /* the half-kick branch (model.tick == 0) is unreachable: tick! has
     * already incremented the tick (gpu/fluids.jl:217) */
    double A[3];
    for (int k = 0; k < d; k++) A[k] = pp->force[k] * p->inverse_mass;
    for (int k = 0; k < d; k++) p->velocity[k] = p->velocity[k] + dt * A[k];
    for (int k = 0; k < d; k++) p->position[k] = p->position[k] + dt * p->velocity[k];
    if (any_nonzero(p->velocity, d)) moved = 1;
  }
  if (p->couple_torques) {
    const int nw = d == 2 ? 1 : 3;
    double Al[3];
    for (int k = 0; k < nw; k++) Al[k] = pp->torque[k] * p->inverse_moment_of_inertia;
    for (int k = 0; k < nw; k++)
      p->angular_velocity[k] = p->angular_velocity[k] + dt * Al[k];
    if (p->kind == LBM_PARTICLE_SQUIRMER) {
      double *e = p->swimming_direction, *W = p->angular_velocity;
      double theta;
      if (d == 2) {
        theta = fabs(W[0]);
      } else {
        double s = 0.0;
        for (int k = 0; k < 3; k++) s = s + W[k] * W[k];
        theta = sqrt(s);
      }
      if (!(theta == 0.0)) {
        double vr[3];
        if (d == 2) {
          double kk = W[0] / theta;
          theta *= dt;
          double c = cos(theta), s = sin(theta);
          double cr[2] = {-kk * e[1], kk * e[0]}; /* cross(omega::Real, V) */
          vr[0] = e[0] * c + cr[0] * s;
          vr[1] = e[1] * c + cr[1] * s;
          double nv = sqrt(vr[0] * vr[0] + vr[1] * vr[1]);
          e[0] = vr[0] / nv;
          e[1] = vr[1] / nv;
        } else {
          double kk[3] = {W[0] / theta, W[1] / theta, W[2] / theta};
          theta *= dt;
          double c = cos(theta), s = sin(theta);
          double cr[3] = {kk[1] * e[2] - kk[2] * e[1], -kk[0] * e[2] + kk[2] * e[0],
                          kk[0] * e[1] - kk[1] * e[0]};
          for (int k = 0; k < 3; k++) vr[k] = e[k] * c + cr[k] * s;
          double kd = 0.0;
          for (int k = 0; k < 3; k++) kd = kd + kk[k] * e[k];
          for (int k = 0; k < 3; k++) vr[k] = vr[k] + kk[k] * kd * (1.0 - c);
          double nv = 0.0;
          for (int k = 0; k < 3; k++) nv = nv + vr[k] * vr[k];
          nv = sqrt(nv);
          for (int k = 0; k < 3; k++) e[k] = vr[k] / nv;
        }
      }
    }
    if (any_nonzero(Al, nw)) find_nv = 1;
  }
  find_nv = moved || find_nv;
  for (int k = 0; k < 3; k++) {
    pp->last_force[k] = pp->force[k];
    pp->last_torque[k] = pp->torque[k];
    pp->force[k] = 0.0;
    pp->torque[k] = 0.0;
  }
  if (moved) rebuild_geometry(m, pp);
  if (find_nv) rebuild_node_velocity(m, pp);
}

/* moveParticles! — cpu/particles.jl:45-100 (eulerStep! :7-18) */
static void move_particle_cpu(lbmo_model *m, opart *pp) {
  lbm_particle *p = &pp->p;
  const int d = m->d, Q = m->Q, nw = d == 2 ? 1 : 3;
  double dP[3] = {0, 0, 0}, dL[3] = {0, 0, 0};
  for (int q = 0; q <= Q; q++) {
    for (int k = 0; k < d; k++) dP[k] = q == 0 ? pp->mom[q][k] : dP[k] + pp->mom[q][k];
    for (int k = 0; k < nw; k++) dL[k] = q == 0 ? pp->amom[q][k] : dL[k] + pp->amom[q][k];
  }
  for (int k = 0; k < d; k++) p->velocity[k] = p->velocity[k] + p->inverse_mass * dP[k];
  for (int k = 0; k < nw; k++)
    p->angular_velocity[k] =
        p->angular_velocity[k] + p->inverse_moment_of_inertia * dL[k];
  int moved = 0;
  if (any_nonzero(p->velocity, d)) {
    for (int k = 0; k < d; k++)
      p->position[k] = p->position[k] + m->cfg.dt * p->velocity[k];
    moved = 1;
  }
  int find_nv = any_nonzero(dL, nw) || moved;
  for (int k = 0; k < 3; k++) {
    pp->last_force[k] = dP[k];
    pp->last_torque[k] = dL[k];
  }
  memset(pp->mom, 0, sizeof pp->mom);
  memset(pp->amom, 0, sizeof pp->amom);
  if (moved) rebuild_geometry(m, pp);
  if (find_nv) rebuild_node_velocity(m, pp);
}

/* ------------------------------------------------- molecular dynamics */

/* x^n for integer n >= 1 as Julia's ^(::Float64, ::Integer) computes it:
 * Base.Math.pow_body, binary powering whose products keep their rounding
 * errors (two_mul via fma) and fold them back at the end; n == 3 is the
 * literal cube x*x*x.  The result is within about an ulp of x^n. */
static double jpow(double x, int n) {
  if (n == 3) return x * x * x;
  double y = 1.0, xnlo = 0.0, ynlo = 0.0;
  while (n > 1) {
    if (n & 1) {
      double err = fma(y, xnlo, x * ynlo);
      double hi = x * y;
      double lo = fma(x, y, -hi);
      y = hi;
      ynlo = lo + err;
    }
    double err = x * 2.0 * xnlo;
    double hi = x * x;
    double lo = fma(x, x, -hi);
    x = hi;
    xnlo = lo + err;
    n >>= 1;
  }
  double err = fma(y, xnlo, x * ynlo);
  return (isfinite(x) && isfinite(err)) ? fma(x, y, err) : x * y;
}

/* `v |> Array |> norm` (LinearAlgebra.generic_norm2 for short vectors) */
static double vnorm(const double *v, int d) {
  double s = v[0] * v[0];
  for (int k = 1; k < d; k++) s = s + v[k] * v[k];
  return sqrt(s);
}

/* cross, gpu/aux.jl:139-157 */
static void vcross3(const double *v, const double *w, double *o) {
  o[0] = v[1] * w[2] - v[2] * w[1];
  o[1] = -v[0] * w[2] + v[2] * w[0];
  o[2] = v[0] * w[1] - v[1] * w[0];
}
static double vcross2(const double *v, const double *w) { return v[0] * w[1] - v[1] * w[0]; }
static void vcross_real(double om, const double *V, double *o) { /* both argument orders */
  o[0] = -om * V[1];
  o[1] = om * V[0];
}

static void add_force(lbmo_model *m, int id, const double *f, double sgn) {
  for (int k = 0; k < m->d; k++)
    m->parts[id].force[k] = sgn > 0 ? m->parts[id].force[k] + f[k] : m->parts[id].force[k] - f[k];
}

/* applyInteraction!, gpu/particles.jl:110-197 (one expanded term) */
static void apply_interaction(lbmo_model *m, const lbm_interaction *I) {
  const int d = m->d;
  const double *R1 = m->parts[I->ids[0]].p.position, *R2 = m->parts[I->ids[1]].p.position;
  if (I->kind == LBM_INTERACTION_POLAR) { /* :136-158 */
    const double *R3 = m->parts[I->ids[2]].p.position;
    double disp12[3], disp32[3], u12[3], u32[3];
    for (int k = 0; k < d; k++) disp12[k] = R1[k] - R2[k];
    const double n12 = vnorm(disp12, d);
    for (int k = 0; k < d; k++) u12[k] = disp12[k] / n12;
    for (int k = 0; k < d; k++) disp32[k] = R3[k] - R2[k];
    const double n32 = vnorm(disp32, d);
    for (int k = 0; k < d; k++) u32[k] = disp32[k] / n32;
    double x = 0.0;
    for (int k = 0; k < d; k++) x = k == 0 ? u12[0] * u32[0] : x + u12[k] * u32[k];
    const double angle = acos(fabs(x) > 1 ? (x > 0 ? 1.0 : -1.0) : x);
    double f321[3] = {0, 0, 0}, f123[3] = {0, 0, 0}, both[3];
    if (d == 3) {
      double td[3], torque[3], c[3];
      vcross3(u32, u12, td);
      const double an = vnorm(td, 3);
      if (an > 0)
        for (int k = 0; k < 3; k++) td[k] = td[k] / an;
      for (int k = 0; k < 3; k++) torque[k] = I->params[1] * (angle - I->params[0]) * td[k];
      vcross3(torque, u12, c);
      for (int k = 0; k < 3; k++) f321[k] = -c[k] / n12;
      vcross3(torque, u32, c);
      for (int k = 0; k < 3; k++) f123[k] = c[k] / n32;
    } else {
      double td = vcross2(u32, u12), c[2];
      const double an = fabs(td);
      if (an > 0) td = td / an;
      const double torque = I->params[1] * (angle - I->params[0]) * td;
      vcross_real(torque, u12, c);
      for (int k = 0; k < 2; k++) f321[k] = -c[k] / n12;
      vcross_real(torque, u32, c);
      for (int k = 0; k < 2; k++) f123[k] = c[k] / n32;
    }
    add_force(m, I->ids[0], f321, 1);
    add_force(m, I->ids[2], f123, 1);
    for (int k = 0; k < d; k++) both[k] = f321[k] + f123[k];
    add_force(m, I->ids[1], both, -1);
    return;
  }
  double disp21[3], u[3], f21[3] = {0, 0, 0};
  for (int k = 0; k < d; k++) disp21[k] = R2[k] - R1[k];
  const double disp = vnorm(disp21, d);
  switch (I->kind) {
    case LBM_INTERACTION_LINEAR: /* :110-121 */
      for (int k = 0; k < d; k++) u[k] = disp21[k] / disp;
      for (int k = 0; k < d; k++) f21[k] = I->params[1] * (disp - I->params[0]) * u[k];
      break;
    case LBM_INTERACTION_BISTABLE: { /* :123-134 */
      for (int k = 0; k < d; k++) u[k] = disp21[k] / disp;
      const double x = disp - I->params[0];
      for (int k = 0; k < d; k++) f21[k] = (I->params[1] * jpow(x, 3) - I->params[2] * x) * u[k];
      break;
    }
    case LBM_INTERACTION_DIPOLE: { /* :160-177, one pair */
      double Bt[3] = {0, 0, 0};
      const double *B = I->field;
      if (I->params[2] != 0.0) { /* B = magneticField(model.time), after tick! advanced the time */
        if (!m->field_fn) return;
        m->field_fn(m->time, Bt, m->field_ctx);
        if (Bt[0] == 0.0 && Bt[1] == 0.0 && Bt[2] == 0.0) return; /* norm(B) ≈ 0 */
        B = Bt;
      }
      for (int k = 0; k < d; k++) u[k] = disp21[k] / disp;
      double t1[3] = {0, 0, 0}, BB, cc;
      if (d == 3) {
        double c1[3];
        vcross3(u, B, c1);
        vcross3(c1, B, t1);
        BB = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
        cc = c1[0] * c1[0] + c1[1] * c1[1] + c1[2] * c1[2];
      } else if (I->params[1] != 0.0) { /* B a number: cross(V, B) = (-B V2, B V1) */
        double c1[2];
        vcross_real(B[0], u, c1);
        vcross_real(B[0], c1, t1);
        BB = B[0] * B[0];
        cc = c1[0] * c1[0] + c1[1] * c1[1];
      } else {
        const double c1 = vcross2(u, B);
        vcross_real(c1, B, t1);
        BB = B[0] * B[0] + B[1] * B[1];
        cc = c1 * c1;
      }
      const double pre = -I->params[0] / jpow(disp, 4);
      for (int k = 0; k < d; k++) f21[k] = pre * (2 * t1[k] - 2 * BB * u[k] + 5 * u[k] * cc);
      break;
    }
    default: { /* LBM_INTERACTION_LENNARD_JONES, :179-197 */
      if (!(disp < I->params[2])) return;
      for (int k = 0; k < d; k++) u[k] = disp21[k] / disp;
      const double sd = I->params[1] / disp;
      for (int k = 0; k < d; k++)
        f21[k] = -4 * I->params[0] / disp * (12 * jpow(sd, 12) - 6 * jpow(sd, 6)) * u[k];
      break;
    }
  }
  add_force(m, I->ids[0], f21, 1);
  add_force(m, I->ids[1], f21, -1);
}

/* ----------------------------------------------------------------- tick */

/* tick! — gpu/fluids.jl:132-226 (gpu semantics), cpu/fluids.jl:377-461 */
static void tick(lbmo_model *m) {
  const int Q = m->Q, d = m->d;
  const int gpu = m->cfg.semantics == LBM_SEMANTICS_GPU;
  double coeff = 1.0;
  for (int k = 0; k < d; k++) coeff *= m->cfg.dx;
  if (gpu) coeff = coeff / m->cfg.dt;

  collision(m);

  for (int q = 0; q < Q; q++) {
    const int o = m->opp[q];
    const double *fsq = m->fs + (int64_t)q * m->N;
    const double *fso = m->fs + (int64_t)o * m->N;
    double *st = m->fn + (int64_t)q * m->N;
    /* streamedDistribution = circshift(f*_q, c_q) */
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < m->N; x++) st[x] = fsq[shifted(m, x, m->c[q])];
    if (!m->bounce_back) continue;
    /* streamed[wallRegion] .= 0 */
    const uint8_t *cir = m->sir_wall + (int64_t)o * m->N;
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < m->N; x++) {
      if (m->wall[x]) st[x] = 0.0;
      int in = cir[x];
      if (m->ladd)
        for (int p = 0; p < m->np; p++) in = in || m->parts[p].sir[(int64_t)o * m->N + x];
      /* streamed[cIR] = f*_{q̄}[cIR] */
      if (in) st[x] = fso[x];
      /* moving walls: streamed[cIR] += (2 w_q invC2) rho[cIR] .* (u_w(x - c_q) . c_q)
       * (gpu/fluids.jl:168-174) */
      if (in && m->solid_vel) {
        const int64_t y = shifted(m, x, m->c[q]);
        double udc = 0.0;
        for (int k = 0; k < d; k++) udc = udc + m->solid_vel[(int64_t)k * m->N + y] * ((double)m->c[q][k] * m->ls);
        st[x] = st[x] + ((2.0 * m->w[q] * m->invC2s) * m->rho[x]) * udc;
      }
    }
    if (!m->ladd) continue;
    double ci[3] = {0, 0, 0};
    for (int k = 0; k < d; k++) ci[k] = (double)m->c[q][k] * m->ls;
    const double lc = 2.0 * m->w[q] * m->invC2s;
    for (int p = 0; p < m->np; p++) {
      opart *pp = &m->parts[p];
      const uint8_t *cbn = pp->sir + (int64_t)o * m->N;
      int64_t cnt = 0;
      for (int64_t x = 0; x < m->N; x++) cnt += cbn[x];
      if (cnt == 0) break; /* `sum(conjugateBoundaryNodes) == 0 && break` */
      /* uw = circshift(nodeVelocity, c_q)  [gpu]
       * uw = nodeVelocity + circshift(nodeVelocity, c_q)  [cpu] */
#pragma omp parallel for schedule(static)
      for (int64_t x = 0; x < m->N; x++) {
        if (!cbn[x]) continue;
        int64_t y = shifted(m, x, m->c[q]);
        double udc = 0.0;
        for (int k = 0; k < d; k++) {
          double uwk = pp->node_vel[(int64_t)k * m->N + y];
          if (!gpu) uwk = pp->node_vel[(int64_t)k * m->N + x] + uwk;
          udc = udc + uwk * ci[k];
        }
        st[x] = st[x] + lc * m->rho[x] * udc;
      }
      if (!(pp->p.couple_forces || pp->p.couple_torques)) continue;
      /* forceInput -= coeff * sum(s) * c_q and torqueInput -= coeff *
       * cross(sum(s (x - R)), c_q) (gpu/fluids.jl:193-204): the sums over
       * links (and over q) are accumulated in double-double; the products
       * s * (x - R)_k are rounded once each, as in the reference. */
      for (int64_t x = 0; x < m->N; x++) {
        if (!cbn[x]) continue;
        const double s = fso[x] + st[x];
        double p[3] = {0, 0, 0};
        for (int k = 0; k < d; k++) p[k] = s * (m->X[(int64_t)k * m->N + x] + (-pp->p.position[k]));
        if (pp->p.couple_forces)
          for (int k = 0; k < d; k++)
            if (m->c[q][k]) dd_add(&pp->fsum[k], m->c[q][k] > 0 ? s : -s);
        if (pp->p.couple_torques) {
          /* cross(p, c): (p1 c2 - p2 c1, -p0 c2 + p2 c0, p0 c1 - p1 c0); in 2D
           * the scalar p0 c1 - p1 c0 */
          const int *c = m->c[q];
          if (d == 2) {
            if (c[1]) dd_add(&pp->tsum[0], c[1] > 0 ? p[0] : -p[0]);
            if (c[0]) dd_add(&pp->tsum[0], c[0] > 0 ? -p[1] : p[1]);
          } else {
            if (c[2]) dd_add(&pp->tsum[0], c[2] > 0 ? p[1] : -p[1]);
            if (c[1]) dd_add(&pp->tsum[0], c[1] > 0 ? -p[2] : p[2]);
            if (c[2]) dd_add(&pp->tsum[1], c[2] > 0 ? -p[0] : p[0]);
            if (c[0]) dd_add(&pp->tsum[1], c[0] > 0 ? p[2] : -p[2]);
            if (c[1]) dd_add(&pp->tsum[2], c[1] > 0 ? p[0] : -p[0]);
            if (c[0]) dd_add(&pp->tsum[2], c[0] > 0 ? -p[1] : p[1]);
          }
        }
      }
    }
  }
  /* round the exact sums once: load = -coeff * sum (ci = c * dx/dt) */
  for (int p = 0; p < m->np; p++) {
    opart *pp = &m->parts[p];
    const int nw = d == 2 ? 1 : 3;
    for (int k = 0; k < 3; k++) {
      const double F = k < d ? -coeff * (dd_value(pp->fsum[k]) * m->ls) : 0.0;
      const double T = k < nw ? -coeff * (dd_value(pp->tsum[k]) * m->ls) : 0.0;
      if (gpu) {
        if (pp->p.couple_forces) pp->force[k] = F;
        if (pp->p.couple_torques) pp->torque[k] = T;
      } else {
        if (pp->p.couple_forces) pp->mom[0][k] = F;
        if (pp->p.couple_torques) pp->amom[0][k] = T;
      }
      pp->fsum[k].hi = pp->fsum[k].lo = 0.0;
      pp->tsum[k].hi = pp->tsum[k].lo = 0.0;
    }
  }
  double *t = m->f;
  m->f = m->fn;
  m->fn = t;
  m->time += m->cfg.dt;
  m->tick += 1;
  hydro_update(m);
  if (m->ladd && gpu)
    for (int t = 0; t < m->ninter; t++) apply_interaction(m, &m->inter[t]);
  if (m->ladd)
    for (int p = 0; p < m->np; p++) {
      if (gpu)
        move_particle_gpu(m, &m->parts[p]);
      else
        move_particle_cpu(m, &m->parts[p]);
    }
}

/* ------------------------------------------------------------------ ABI */

static void *xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

int lbmo_destroy(lbmo_model *m) {
  if (!m) return LBM_OK;
  for (int k = 0; k < 3; k++) free(m->coords[k]);
  free(m->X); free(m->f); free(m->fs); free(m->fn); free(m->rho); free(m->u);
  free(m->F); free(m->wall); free(m->sir_wall); free(m->solid_vel);
  for (int p = 0; p < m->np; p++) {
    free(m->parts[p].solid); free(m->parts[p].sir); free(m->parts[p].node_vel);
  }
  free(m);
  return LBM_OK;
}

int lbmo_create(const lbm_config *cfg, lbmo_model **out) {
  if (!cfg || !out) return fail(LBM_EINVAL, "null argument");
  if (!((cfg->dims == 2 && cfg->q == 9) || (cfg->dims == 3 && (cfg->q == 15 || cfg->q == 19 || cfg->q == 27))))
    return fail(LBM_EINVAL, "lattice must be D2Q9, D3Q15, D3Q19 or D3Q27");
  lbmo_model *m = (lbmo_model *)calloc(1, sizeof *m);
  if (!m) return fail(LBM_ENOMEM, "out of memory");
  m->cfg = *cfg;
  m->d = cfg->dims;
  m->Q = cfg->q;
  m->incompressible = (cfg->flags & LBM_FLAG_INCOMPRESSIBLE) != 0;
  m->n[0] = cfg->n[0];
  m->n[1] = cfg->n[1];
  m->n[2] = cfg->dims == 3 ? cfg->n[2] : 1;
  m->N = m->n[0] * m->n[1] * m->n[2];
  tables(m);
  m->ls = cfg->dx / cfg->dt;
  m->invC2s = 3.0 / (m->ls * m->ls);
  m->invC4s = 9.0 / (m->ls * m->ls * m->ls * m->ls);
  const int64_t N = m->N;
  for (int k = 0; k < m->d; k++) {
    m->coords[k] = (double *)xcalloc((size_t)m->n[k], sizeof(double));
    memcpy(m->coords[k], cfg->coords[k], (size_t)m->n[k] * sizeof(double));
  }
  m->X = (double *)xcalloc((size_t)(m->d * N), sizeof(double));
  m->f = (double *)xcalloc((size_t)(m->Q * N), sizeof(double));
  m->fs = (double *)xcalloc((size_t)(m->Q * N), sizeof(double));
  m->fn = (double *)xcalloc((size_t)(m->Q * N), sizeof(double));
  m->rho = (double *)xcalloc((size_t)N, sizeof(double));
  m->u = (double *)xcalloc((size_t)(m->d * N), sizeof(double));
  m->wall = (uint8_t *)xcalloc((size_t)N, 1);
  m->sir_wall = (uint8_t *)xcalloc((size_t)(m->Q * N), 1);
  if (!m->X || !m->f || !m->fs || !m->fn || !m->rho || !m->u || !m->wall || !m->sir_wall) {
    lbmo_destroy(m);
    return fail(LBM_ENOMEM, "out of memory");
  }
  for (int64_t x = 0; x < N; x++) {
    int64_t id[3] = {x % m->n[0], (x / m->n[0]) % m->n[1], x / (m->n[0] * m->n[1])};
    for (int k = 0; k < m->d; k++) m->X[(int64_t)k * N + x] = m->coords[k][id[k]];
    /* wallNodes: first and last layer of each walled dimension */
    for (int k = 0; k < m->d; k++)
      if ((cfg->walled >> k) & 1)
        if (id[k] == 0 || id[k] == m->n[k] - 1) m->wall[x] = 1;
  }
  if (cfg->walled) {
    m->bounce_back = 1;
    bounce_back_prep(m, m->wall, m->sir_wall);
  }
  if (cfg->force != LBM_FORCE_NONE) {
    m->guo = 1;
    m->F = (double *)xcalloc((size_t)(m->d * N), sizeof(double));
    if (cfg->force == LBM_FORCE_UNIFORM)
      for (int64_t x = 0; x < N; x++)
        for (int k = 0; k < m->d; k++)
          m->F[(int64_t)k * N + x] = m->wall[x] ? 0.0 : cfg->force_uniform[k];
  }
  *out = m;
  return LBM_OK;
}

/* forceDensity[wallRegion] .= 0 (gpu/initMethods.jl:643-647) */
static void zero_force_on_walls(lbmo_model *m) {
  if (!m->guo) return;
  for (int64_t x = 0; x < m->N; x++)
    if (m->wall[x])
      for (int k = 0; k < m->d; k++) m->F[(int64_t)k * m->N + x] = 0.0;
}

/* `walls` (gpu/initMethods.jl:591-605): wallRegion .|| the given region;
 * the invasion regions are those of the final region (:622-627). */
int lbmo_set_walls(lbmo_model *m, const uint8_t *wall) {
  for (int64_t x = 0; x < m->N; x++)
    if (wall[x]) m->wall[x] = 1;
  m->bounce_back = 1;
  bounce_back_prep(m, m->wall, m->sir_wall);
  zero_force_on_walls(m);
  return LBM_OK;
}

/* `movingWalls` (gpu/initMethods.jl:607-626): solidNodeVelocity = the given
 * field; its non-zero nodes join the wall region. */
int lbmo_set_wall_velocity(lbmo_model *m, const double *u) {
  free(m->solid_vel);
  m->solid_vel = (double *)xcalloc((size_t)(m->d * m->N), sizeof(double));
  if (!m->solid_vel) return fail(LBM_ENOMEM, "out of memory");
  memcpy(m->solid_vel, u, (size_t)(m->d * m->N) * sizeof(double));
  for (int64_t x = 0; x < m->N; x++) {
    double nn = 0.0;
    for (int k = 0; k < m->d; k++) nn = nn + u[(int64_t)k * m->N + x] * u[(int64_t)k * m->N + x];
    if (sqrt(nn) > 0) m->wall[x] = 1;
  }
  m->bounce_back = 1;
  bounce_back_prep(m, m->wall, m->sir_wall);
  zero_force_on_walls(m);
  return LBM_OK;
}

int lbmo_set_force_field(lbmo_model *m, const double *F) {
  if (m->cfg.force != LBM_FORCE_FIELD) return fail(LBM_ESTATE, "force mode is not FIELD");
  for (int64_t x = 0; x < m->N; x++)
    for (int k = 0; k < m->d; k++)
      m->F[(int64_t)k * m->N + x] = m->wall[x] ? 0.0 : F[(int64_t)k * m->N + x];
  return LBM_OK;
}

/* findInitialConditions (gpu/initMethods.jl:417-438) for uniform rho, u:
 * rho and u are zero on walls (:511-527 via :584-586); under :guo,
 * u -= dt/2 * F/rho (F/rho = 0 where rho == 0); f_i = second .* (w_i rho) */
int lbmo_init_uniform(lbmo_model *m, double rho0, const double *u0) {
  const int d = m->d;
  for (int64_t x = 0; x < m->N; x++) {
    double rho = m->wall[x] ? 0.0 : rho0;
    double u[3] = {0, 0, 0};
    for (int k = 0; k < d; k++) u[k] = m->wall[x] ? 0.0 : u0[k];
    if (m->guo)
      for (int k = 0; k < d; k++) {
        double ct = rho != 0.0 ? m->F[(int64_t)k * m->N + x] / rho : 0.0;
        u[k] = u[k] - 0.5 * m->cfg.dt * ct;
      }
    for (int q = 0; q < m->Q; q++) {
      double ci[3], udc = 0.0, uu = 0.0;
      for (int k = 0; k < d; k++) ci[k] = (double)m->c[q][k] * m->ls;
      for (int k = 0; k < d; k++) udc = udc + u[k] * ci[k];
      for (int k = 0; k < d; k++) uu = uu + u[k] * u[k];
      double first = udc * m->invC2s + (udc * udc) * (0.5 * m->invC4s);
      double second = first - uu * (0.5 * m->invC2s) + 1.0;
      m->f[(int64_t)q * m->N + x] = second * (m->w[q] * rho);
    }
  }
  hydro_update(m);
  return LBM_OK;
}

int lbmo_set_populations(lbmo_model *m, const double *f) {
  memcpy(m->f, f, (size_t)(m->Q * m->N) * sizeof(double));
  hydro_update(m);
  return LBM_OK;
}

int lbmo_get_populations(lbmo_model *m, double *f) {
  memcpy(f, m->f, (size_t)(m->Q * m->N) * sizeof(double));
  return LBM_OK;
}

int lbmo_get_hydro(lbmo_model *m, double *rho, double *u) {
  if (rho) memcpy(rho, m->rho, (size_t)m->N * sizeof(double));
  if (u) memcpy(u, m->u, (size_t)(m->d * m->N) * sizeof(double));
  return LBM_OK;
}

/* addSquirmer!/addBead! after parameter resolution (gpu/initMethods.jl:
 * 91-222): append, switch on :ladd and :bounceBack, then
 * moveParticles!(; initialSetup = true). */
int lbmo_add_particle(lbmo_model *m, const lbm_particle *p, int32_t *id) {
  if (m->np >= LBM_MAX_PARTICLES) return fail(LBM_EINVAL, "too many particles");
  opart *pp = &m->parts[m->np];
  memset(pp, 0, sizeof *pp);
  pp->p = *p;
  pp->solid = (uint8_t *)xcalloc((size_t)m->N, 1);
  pp->sir = (uint8_t *)xcalloc((size_t)(m->Q * m->N), 1);
  pp->node_vel = (double *)xcalloc((size_t)(m->d * m->N), sizeof(double));
  if (!pp->solid || !pp->sir || !pp->node_vel) return fail(LBM_ENOMEM, "out of memory");
  if (id) *id = m->np;
  m->np++;
  m->ladd = 1;
  m->bounce_back = 1; /* an all-false wallRegion is installed if absent */
  rebuild_geometry(m, pp);
  rebuild_node_velocity(m, pp);
  return LBM_OK;
}

int lbmo_get_particle(lbmo_model *m, int32_t id, lbm_particle *out) {
  if (id < 0 || id >= m->np) return fail(LBM_EINVAL, "bad particle id");
  *out = m->parts[id].p;
  return LBM_OK;
}

int lbmo_get_particle_load(lbmo_model *m, int32_t id, double *force, double *torque) {
  if (id < 0 || id >= m->np) return fail(LBM_EINVAL, "bad particle id");
  for (int k = 0; k < 3; k++) {
    force[k] = m->parts[id].last_force[k];
    torque[k] = m->parts[id].last_torque[k];
  }
  return LBM_OK;
}

/* findall over streamingInvasionRegions (N..., Q): node fastest, j slowest */
int lbmo_get_links(lbmo_model *m, int32_t id, int64_t *node, int32_t *dir,
                   int64_t cap, int64_t *count) {
  if (id < 0 || id >= m->np) return fail(LBM_EINVAL, "bad particle id");
  int64_t c = 0;
  for (int j = 0; j < m->Q; j++)
    for (int64_t x = 0; x < m->N; x++)
      if (m->parts[id].sir[(int64_t)j * m->N + x]) {
        if (c < cap) {
          node[c] = x;
          dir[c] = j;
        }
        c++;
      }
  *count = c;
  return LBM_OK;
}

int lbmo_run(lbmo_model *m, int64_t nsteps) {
  for (int64_t s = 0; s < nsteps; s++) tick(m);
  return LBM_OK;
}

int lbmo_get_tick(lbmo_model *m, int64_t *t, double *time) {
  *t = m->tick;
  *time = m->time;
  return LBM_OK;
}

int lbmo_set_field_fn(lbmo_model *m, void (*fn)(double, double *, void *), void *ctx) {
  m->field_fn = fn;
  m->field_ctx = ctx;
  return LBM_OK;
}

int lbmo_set_interactions(lbmo_model *m, const lbm_interaction *terms, int32_t n) {
  if (n < 0 || n > LBM_MAX_INTERACTIONS) return fail(LBM_EINVAL, "interaction count out of range");
  for (int t = 0; t < n; t++)
    for (int k = 0; k < (terms[t].kind == LBM_INTERACTION_POLAR ? 3 : 2); k++)
      if (terms[t].ids[k] < 0 || terms[t].ids[k] >= m->np) return fail(LBM_EINVAL, "bad particle id");
  if (n > 0) memcpy(m->inter, terms, sizeof(lbm_interaction) * (size_t)n);
  m->ninter = n;
  return LBM_OK;
}

/* viscousStressTensor (cpu/fluids.jl:507-520) of the current populations:
 * the non-equilibrium populations f_q - feq_q(rho, u) of every q
 * (nonEquilibriumDistribution, :65-67, with equilibriumDistribution's
 * compressible form, :46-63), then for every (a, b) the sum over q of
 * (c_qa c_qb) f_neq_q, in q order, times -(1 - 0.5 dt / tau).  out is
 * (N, d, d), column major. */
int lbmo_viscous_stress(lbmo_model *m, double *out) {
  const int d = m->d, Q = m->Q;
  const int64_t N = m->N;
  hydro_update(m);
  double *neq = (double *)xcalloc((size_t)(Q * N), sizeof(double));
  if (!neq) return fail(LBM_ENOMEM, "out of memory");
  for (int q = 0; q < Q; q++)
    for (int64_t x = 0; x < N; x++) {
      double u[3];
      for (int k = 0; k < d; k++) u[k] = m->u[(int64_t)k * N + x];
      neq[(int64_t)q * N + x] = m->f[(int64_t)q * N + x] - feq_one(m, q, m->rho[x], u);
    }
  const double coeff = 1 - 0.5 * m->cfg.dt / m->cfg.tau;
  for (int a = 0; a < d; a++)
    for (int b = 0; b < d; b++) {
      double *o = out + N * (a + d * b);
      for (int64_t x = 0; x < N; x++) {
        double s = 0.0;
        for (int q = 0; q < Q; q++) {
          const double t = (double)(m->c[q][a] * m->c[q][b]) * neq[(int64_t)q * N + x];
          s = q == 0 ? t : s + t;
        }
        o[x] = -coeff * s;
      }
    }
  free(neq);
  return LBM_OK;
}

/* Solid mask of particle id (N), for geometry checks. */
int lbmo_get_solid(lbmo_model *m, int32_t id, uint8_t *out) {
  if (id < 0 || id >= m->np) return fail(LBM_EINVAL, "bad particle id");
  memcpy(out, m->parts[id].solid, (size_t)m->N);
  return LBM_OK;
}
