// lbm_adjoint.hpp — discrete adjoint of one tick of the D3Q19 step (GPU
// semantics): reverse collision, streaming, halfway bounce-back on walls and
// particles, the Ladd term and the momentum exchange, for gradients of a
// loss with respect to the populations, the particle state, the squirmer
// slip (B1, B2) and the Guo force field.
//
// The forward tick k, in the reference's order (gpu/fluids.jl:132-226):
//   P_k(x)       = C(S_k(x), F(x))                      collision + Guo
//   S_{k+1,q}(x) = P_{k,q}(x - c_q)                     streaming
//                = P_{k,q̄}(x)                           x - c_q a wall node
//                = P_{k,q̄}(x) + sum_p 2 w_q/c_s^2 rho_k(x) (u_b^p . c_q)
//                                                       (x, q) a link of p
//   F_p = -coeff ls sum s c_q, T_p = -coeff ls sum s (X - R_p) x c_q over
//   p's links, s = P_{k,q̄}(x) + S_{k+1,q}(x)           momentum exchange
//   particles    = moveParticles!(particles, F, T)       (host, by duals)
// The link set is a piecewise-constant function of the positions, so it
// carries no derivative (the gradient of the smooth piece is taken).
#pragma once

#include "lbm_step.hpp"

namespace lbm {

// Per-particle sums of one adjoint tick: dL/dB1, dL/dB2, then dL/d of the
// node-velocity snapshot (nR, nV, nW, ne) and of the position R (torque arm).
enum AdjSum { kAdjB1 = 0, kAdjB2 = 1, kAdjNR = 2, kAdjNV = 5, kAdjNW = 8, kAdjNE = 11, kAdjR = 14, kAdjSums = 17 };

// Vector-Jacobian product of the collision (collide_ref and collide_dev are
// the same function of f): given lp = dL/dP at one node, lsum = (dP/dS)^T lp
// and gF += (dP/dF)^T lp.  rho and u are the node's moments (u with the
// half-force shift), F its force.
template <int Q, int COLL, int FM, class R>
__host__ __device__ inline void collide_vjp(R rho, const R (&u)[3], const R (&F)[3], const CollParams &c,
                                            const R (&lp)[Q], R (&lsum)[Q], R (&gF)[3]) {
  using L = Lattice<Q>;
  constexpr int D = L::D;
  const R ls = (R)c.ls, i2 = (R)c.invC2s, i4 = (R)c.invC4s, dt = (R)c.dt;
  const R cb = (R)(-c.dt / c.tau);
  const R cm = COLL == LBM_COLLISION_TRT ? (R)(-c.dt / c.tau_minus) : cb;
  const R gp = (R)(1.0 - c.dt / (2.0 * c.tau));
  const R g2 = COLL == LBM_COLLISION_TRT ? (R)(1.0 - c.dt / (2.0 * c.tau_minus)) : gp;
  R uu = 0;
  for (int k = 0; k < D; k++) uu += u[k] * u[k];
  R uF = 0;
  if (FM != LBM_FORCE_NONE)
    for (int k = 0; k < D; k++) uF += u[k] * F[k];
  R prho = 0, pu[3] = {0, 0, 0}, gd[3] = {0, 0, 0};
  R lin[Q];
  for (int q = 0; q < Q; q++) {
    const int o = L::opp[q];
    // omega = cb (f+ - feq+) + cm (f- - feq-): kappa_q = cb lp+_q + cm lp-_q
    const R kap = COLL == LBM_COLLISION_TRT ? cb * (R)0.5 * (lp[q] + lp[o]) + cm * (R)0.5 * (lp[q] - lp[o])
                                            : cb * lp[q];
    lin[q] = lp[q] + kap;
    const R w = (R)L::w(q);
    const R ci[3] = {(R)L::cx[q] * ls, (R)L::cy[q] * ls, (R)L::cz[q] * ls};
    R udc = 0;
    for (int k = 0; k < D; k++) udc += u[k] * ci[k];
    // feq = w rho (1 + udc/c2 + udc^2/(2 c4) - uu/(2 c2))
    const R feq_rho = w * ((R)1 + udc * i2 + (R)0.5 * udc * udc * i4 - (R)0.5 * uu * i2);
    prho -= kap * feq_rho;
    for (int k = 0; k < D; k++) pu[k] -= kap * w * rho * (ci[k] * i2 + udc * ci[k] * i4 - u[k] * i2);
    if (FM != LBM_FORCE_NONE) {
      // S = w dt [gp (-u.F/c2 + udc cF/c4) + g2 cF/c2]
      R cF = 0;
      for (int k = 0; k < D; k++) cF += ci[k] * F[k];
      const R a = lp[q] * w * dt;
      for (int k = 0; k < D; k++) {
        pu[k] += a * gp * (-F[k] * i2 + ci[k] * cF * i4);
        gd[k] += a * (gp * (-u[k] * i2 + udc * ci[k] * i4) + g2 * ci[k] * i2);
      }
    }
  }
  // rho = sum f, u = (sum f c + F dt/2) / rho
  const R inv = rho != (R)0 ? (R)1 / rho : (R)0;
  R pus = 0;
  for (int k = 0; k < D; k++) pus += pu[k] * u[k];
  for (int q = 0; q < Q; q++) {
    const R cq[3] = {(R)L::cx[q], (R)L::cy[q], (R)L::cz[q]};
    R pc = 0;
    for (int k = 0; k < D; k++) pc += pu[k] * cq[k];
    lsum[q] = lin[q] + prho + inv * (pc - pus);
  }
  if (FM != LBM_FORCE_NONE)
    for (int k = 0; k < D; k++) gF[k] += gd[k] + pu[k] * (R)0.5 * dt * inv;
}

// Vector-Jacobian product of node_velocity (GPU semantics, 3D) at the solid
// node with coordinates xc: given ub = dL/du_b, adds dL/d(B1, B2, nR, nV,
// nW, ne) of particle p to acc.
__host__ __device__ inline void node_velocity_vjp(const double (&xc)[3], const DevParticle &p, const double (&ub)[3],
                                                  double (&acc)[kAdjSums]) {
  double r[3];
  for (int k = 0; k < 3; k++) r[k] = xc[k] + (-p.nR[k]);
  double nn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  if (nn == 0.0) nn = 1.0;
  const double *W = p.nW, *e = p.ne;
  // bulk V + W x r
  for (int k = 0; k < 3; k++) acc[kAdjNV + k] += ub[k];
  acc[kAdjNW + 0] += r[1] * ub[2] - r[2] * ub[1];
  acc[kAdjNW + 1] += r[2] * ub[0] - r[0] * ub[2];
  acc[kAdjNW + 2] += r[0] * ub[1] - r[1] * ub[0];
  double dr[3] = {ub[1] * W[2] - ub[2] * W[1], ub[2] * W[0] - ub[0] * W[2], ub[0] * W[1] - ub[1] * W[0]};
  if (p.kind == LBM_PARTICLE_SQUIRMER) {
    // (B1 + B2 a) (a r^ - e), a = e . r^, r^ = r / |r|
    double un[3];
    for (int k = 0; k < 3; k++) un[k] = r[k] / nn;
    const double a = un[0] * e[0] + un[1] * e[1] + un[2] * e[2];
    const double first = p.B1 + p.B2 * a;
    const double uu = ub[0] * un[0] + ub[1] * un[1] + ub[2] * un[2];
    const double ue = ub[0] * e[0] + ub[1] * e[1] + ub[2] * e[2];
    const double G = a * uu - ue;
    acc[kAdjB1] += G;
    acc[kAdjB2] += a * G;
    const double da = p.B2 * G + first * uu;
    double dun[3];
    for (int k = 0; k < 3; k++) {
      acc[kAdjNE + k] += -first * ub[k] + da * un[k];
      dun[k] = first * a * ub[k] + da * e[k];
    }
    const double dd = dun[0] * un[0] + dun[1] * un[1] + dun[2] * un[2];
    for (int k = 0; k < 3; k++) dr[k] += (dun[k] - dd * un[k]) / nn;
  }
  for (int k = 0; k < 3; k++) acc[kAdjNR + k] -= dr[k];
}

struct AdjArgs {
  const void *S;        // S_k, storage type, row-blocked (as the step's buffers)
  const void *lam_in;   // dL/dS_{k+1}, compute type, row-blocked
  void *lam_out;        // dL/dS_k
  double *gF;           // (N, 3) += dL/dF (Julia layout k N + n)
  const void *Ffield;   // (N, 3) force field (compute type)
  BoxGeom g;
  CollParams c;
  const DevParticle *parts;  // the particle records at the start of tick k
  const double *coords[3];
  double *partial;      // [block][np][kAdjSums]
  int32_t np;
  double coeff, ls;     // momentum-exchange coefficient and lattice speed
  double shell;         // 1.5 dx: link nodes lie within r + shell of a centre
  const double *phitau;  // [particle][6]: dL/dF_p (0..2) and dL/dT_p (3..5) of the tick's move
};

// The reverse of one tick's particle moves, on the device (block p, lane j <
// kDualN carries the tangent of input j), and the update of the particles'
// adjoint state once the tick's adjoint sums are in.
struct AdjMoveArgs {
  const DevParticle *recs_k;   // records at the start of tick k
  const DevParticle *recs_k1;  // after its move (force, torque consumed)
  const double *mu;            // [np][kPState]: dL/d(state after the move)
  double *mu_in;               // [np][kPState]: dL/d(state before it), from the move alone
  double *phitau;              // [np][6]
  double dt;
};
struct AdjFinishArgs {
  const double *sums;          // [np][kAdjSums] of tick k's adjoint kernel
  const double *mu_in;
  double *mu;                  // [np][kPState] <- dL/d(state at the start of tick k)
  double *gB;                  // [np][2] += dL/dB1, dL/dB2
  const DevParticle *recs_k;
  int32_t np, sw, direct;      // direct: V_k is one of the averaged speeds
  double gscale;               // 2 (mean - target) / window
};

// Squared distance of node (x, y, z) from Rg in getSphere's order, and the
// link predicates of the forward box blocks: (x, q) is a link slot of p when
// x lies in p's link box, r^2 <= d2(x) < (r + 1.5 dx)^2 and d2(x - c_q) < r^2.
__device__ __forceinline__ double adj_d2(const AdjArgs &a, const DevParticle &P, int x, int y, int z) {
#pragma clang fp contract(off)
  const double u = a.coords[0][x] + (-P.Rg[0]), v = a.coords[1][y] + (-P.Rg[1]), w = a.coords[2][z] + (-P.Rg[2]);
  const double su = u * u, sv = v * v, sw = w * w;
  return (su + sv) + sw;
}

// One adjoint tick: dL/dS_k from dL/dS_{k+1}, with dL/dF += and the
// per-particle sums of the block.  One thread per node (x fastest).
template <int Q, int S, int COLL, int FM>
__global__ void __launch_bounds__(256) adjoint_kernel(AdjArgs a) {
  using L = Lattice<Q>;
  using T = typename Store<S>::T;
  using R = typename Store<S>::R;
  static_assert(Q == 19, "the adjoint is built for D3Q19");
  const BoxGeom &g = a.g;
  const int tid = threadIdx.x;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + tid;
  const bool live = n < g.N;
  const int x = live ? (int)(n % g.nx) : 0;
  const int64_t row = live ? n / g.nx : 0;
  const int y = (int)(row % g.ny), z = (int)(row / g.ny);
  const T *Sk = static_cast<const T *>(a.S);
  const R *lin = static_cast<const R *>(a.lam_in);
  R *lout = static_cast<R *>(a.lam_out);
  const int64_t own_idx = row * Q * g.nx + x;
  const bool wall = live && is_wall_node(g, x, y, z);

  // neighbour x + c_q (periodic), its row offset and whether it is a wall node
  auto nb = [&](int q, int &xn, int &yn, int &zn) {
    xn = wrapi(x + L::cx[q], g.nx);
    yn = wrapi(y + L::cy[q], g.ny);
    zn = wrapi(z + L::cz[q], g.nz);
  };

  // forward collision of this node: rho, u and P_k (the link values need it)
  R f[Q], P[Q], rho = 0;
  R Fn[3] = {0, 0, 0};
  if (live && !wall) {
#pragma unroll
    for (int q = 0; q < Q; q++) f[q] = decode<S, Q>(Sk[own_idx + (int64_t)q * g.nx], q);
    if constexpr (FM == LBM_FORCE_UNIFORM) {
      for (int k = 0; k < 3; k++) Fn[k] = (R)a.c.F[k];
    } else if constexpr (FM == LBM_FORCE_FIELD) {
      const R *Ff = static_cast<const R *>(a.Ffield);
      for (int k = 0; k < 3; k++) Fn[k] = Ff[k * g.N + n];
    }
    if constexpr (S == kF64)
      collide_ref<Q, COLL, FM>(f, Fn, a.c, P, rho);
    else
      collide_dev<Q, COLL, FM>(f, Fn, a.c, P, rho);
  }

  // link geometry near the particles: own[p] = link slots (x, q) of p,
  // drop = directions q whose stream x -> x + c_q a link of some particle replaces
  uint32_t own[LBM_MAX_PARTICLES];
  uint32_t own_any = 0, drop = 0;
  for (int p = 0; p < LBM_MAX_PARTICLES; p++) own[p] = 0;
  if (live && !wall) {
    for (int p = 0; p < a.np; p++) {
      const DevParticle &Pp = a.parts[p];
      if (x < Pp.lo[0] - 1 || x > Pp.hi[0] + 1 || y < Pp.lo[1] - 1 || y > Pp.hi[1] + 1 || z < Pp.lo[2] - 1 ||
          z > Pp.hi[2] + 1)
        continue;
      const double r2 = Pp.radius * Pp.radius, ro = Pp.radius + a.shell;
      const double d0 = adj_d2(a, Pp, x, y, z);
      const bool fl0 = in_box(Pp, x, y, z) && d0 >= r2 && d0 < ro * ro;  // may hold link slots
      uint32_t op = 0;
      for (int q = 1; q < Q; q++) {
        int xn, yn, zn;
        nb(q, xn, yn, zn);
        const double dn = adj_d2(a, Pp, xn, yn, zn);
        // (x, q̄) is a link slot: x - c_q̄ = x + c_q solid
        if (fl0 && dn < r2) op |= 1u << L::opp[q];
        // (x + c_q, q) is a link slot: x + c_q fluid within the shell, x solid
        if (d0 < r2 && in_box(Pp, xn, yn, zn) && dn >= r2 && dn < ro * ro) drop |= 1u << q;
      }
      own[p] = op;
      own_any |= op;
    }
  }

  // dL/dP_k at this node
  R lp[Q];
  R grho = 0;
  double coords_x[3] = {0.0, 0.0, 0.0};
  if (live) coords_x[0] = a.coords[0][x], coords_x[1] = a.coords[1][y], coords_x[2] = a.coords[2][z];
  if (live && !wall) {
#pragma unroll
    for (int q = 0; q < Q; q++) {
      int xn, yn, zn;
      nb(q, xn, yn, zn);
      const bool wn = is_wall_node(g, xn, yn, zn);
      R v = 0;
      // streamed to x + c_q, unless that node is a wall or the slot is a link
      if (!wn && !((drop >> q) & 1))
        v = lin[((int64_t)(yn + g.ny * zn) * Q + q) * g.nx + xn];
      // bounced back into (x, q̄) off a wall layer
      if (wn) v += lin[own_idx + (int64_t)L::opp[q] * g.nx];
      lp[q] = v;
    }
    // particle link slots (x, q'): S_q'(x) = P_q̄'(x) + the owners' Ladd terms
    for (int qq = 1; qq < Q; qq++) {
      if (!((own_any >> qq) & 1)) continue;
      const int ob = L::opp[qq];
      const double lam = (double)lin[own_idx + (int64_t)qq * g.nx];
      const double cq[3] = {(double)L::cx[qq], (double)L::cy[qq], (double)L::cz[qq]};
      const double xs[3] = {a.coords[0][wrapi(x - L::cx[qq], g.nx)], a.coords[1][wrapi(y - L::cy[qq], g.ny)],
                            a.coords[2][wrapi(z - L::cz[qq], g.nz)]};
      const double lc = 2.0 * L::w(qq) * a.c.invC2s;
      const double fs = (double)P[ob];
      // owners in particle order: running value v_j, s_j = fs + v_j, shat_j
      double v = fs, shat_sum = 0.0;
      for (int p = 0; p < a.np; p++) {
        if (!((own[p] >> qq) & 1)) continue;
        const DevParticle &Pp = a.parts[p];
        double ubv[3];
        node_velocity<3>(xs, Pp, LBM_SEMANTICS_GPU, 0.0, ubv);
        const double udc = (ubv[0] * cq[0] + ubv[1] * cq[1] + ubv[2] * cq[2]) * a.ls;
        v = v + lc * (double)rho * udc;
        const double arm[3] = {coords_x[0] - Pp.R[0], coords_x[1] - Pp.R[1], coords_x[2] - Pp.R[2]};
        const double cr[3] = {arm[1] * cq[2] - arm[2] * cq[1], arm[2] * cq[0] - arm[0] * cq[2],
                              arm[0] * cq[1] - arm[1] * cq[0]};
        const double sh = -a.coeff * a.ls *
                          (a.phitau[p * 6 + 0] * cq[0] + a.phitau[p * 6 + 1] * cq[1] + a.phitau[p * 6 + 2] * cq[2] + a.phitau[p * 6 + 3] * cr[0] +
                           a.phitau[p * 6 + 4] * cr[1] + a.phitau[p * 6 + 5] * cr[2]);
        shat_sum += sh;
      }
      // dL/dfs = lam + 2 sum shat; dL/dv_j = lam + sum_{i >= j} shat_i
      lp[ob] += (R)(lam + 2.0 * shat_sum);
      double pre = 0.0;
      for (int p = 0; p < a.np; p++) {
        if (!((own[p] >> qq) & 1)) continue;
        const DevParticle &Pp = a.parts[p];
        double ubv[3];
        node_velocity<3>(xs, Pp, LBM_SEMANTICS_GPU, 0.0, ubv);
        const double udc = (ubv[0] * cq[0] + ubv[1] * cq[1] + ubv[2] * cq[2]) * a.ls;
        const double aj = lam + shat_sum - pre;
        grho += (R)(aj * lc * udc);
        const double arm[3] = {coords_x[0] - Pp.R[0], coords_x[1] - Pp.R[1], coords_x[2] - Pp.R[2]};
        const double cr[3] = {arm[1] * cq[2] - arm[2] * cq[1], arm[2] * cq[0] - arm[0] * cq[2],
                              arm[0] * cq[1] - arm[1] * cq[0]};
        pre += -a.coeff * a.ls *
               (a.phitau[p * 6 + 0] * cq[0] + a.phitau[p * 6 + 1] * cq[1] + a.phitau[p * 6 + 2] * cq[2] + a.phitau[p * 6 + 3] * cr[0] +
                a.phitau[p * 6 + 4] * cr[1] + a.phitau[p * 6 + 5] * cr[2]);
      }
    }
  }

  // dL/dS_k = collision VJP + the Ladd terms' rho (rho = sum of S_k)
  if (live) {
    if (wall) {
#pragma unroll
      for (int q = 0; q < Q; q++) store_pop(&lout[own_idx + (int64_t)q * g.nx], (R)0);
    } else {
      R u[3] = {0, 0, 0};
      {
        R j[3] = {0, 0, 0};
#pragma unroll
        for (int q = 0; q < Q; q++) {
          j[0] += (R)L::cx[q] * f[q];
          j[1] += (R)L::cy[q] * f[q];
          j[2] += (R)L::cz[q] * f[q];
        }
        if (FM != LBM_FORCE_NONE)
          for (int k = 0; k < 3; k++) j[k] += (R)(0.5 * a.c.dt) * Fn[k];
        const R inv = rho != (R)0 ? (R)1 / rho : (R)0;
        for (int k = 0; k < 3; k++) u[k] = j[k] * inv;
      }
      R ls_[Q], gFn[3] = {0, 0, 0};
      collide_vjp<Q, COLL, FM, R>(rho, u, Fn, a.c, lp, ls_, gFn);
#pragma unroll
      for (int q = 0; q < Q; q++) store_pop(&lout[own_idx + (int64_t)q * g.nx], ls_[q] + grho);
      if constexpr (FM == LBM_FORCE_FIELD)
        for (int k = 0; k < 3; k++) a.gF[k * g.N + n] += (double)gFn[k];
    }
  }

  // per-particle sums over the block (only blocks that hold link slots)
  __shared__ int any_link;
  __shared__ DD red[4][kAdjSums];
  if (tid == 0) any_link = 0;
  __syncthreads();
  if (own_any) any_link = 1;
  __syncthreads();
  const int wave = tid >> 6, lane = __lane_id();
  for (int p = 0; p < a.np; p++) {
    double acc[kAdjSums];
    for (int k = 0; k < kAdjSums; k++) acc[k] = 0.0;
    if (any_link && (own[p] != 0)) {
      const DevParticle &Pp = a.parts[p];
      for (int qq = 1; qq < Q; qq++) {
        if (!((own[p] >> qq) & 1)) continue;
        const int ob = L::opp[qq];
        const double lam = (double)lin[own_idx + (int64_t)qq * g.nx];
        const double cq[3] = {(double)L::cx[qq], (double)L::cy[qq], (double)L::cz[qq]};
        const double xs[3] = {a.coords[0][wrapi(x - L::cx[qq], g.nx)], a.coords[1][wrapi(y - L::cy[qq], g.ny)],
                              a.coords[2][wrapi(z - L::cz[qq], g.nz)]};
        const double lc = 2.0 * L::w(qq) * a.c.invC2s;
        const double shift = S == kF64 ? 0.0 : 2.0 * L::w(qq) * a.c.rho0;
        const double fs = (double)P[ob];
        // the owner chain up to and after p
        double v = fs, shat_sum = 0.0, pre = 0.0, s_p = 0.0;
        for (int i = 0; i < a.np; i++) {
          if (!((own[i] >> qq) & 1)) continue;
          const DevParticle &Pi = a.parts[i];
          double ubv[3];
          node_velocity<3>(xs, Pi, LBM_SEMANTICS_GPU, 0.0, ubv);
          const double udc = (ubv[0] * cq[0] + ubv[1] * cq[1] + ubv[2] * cq[2]) * a.ls;
          v = v + lc * (double)rho * udc;
          const double arm[3] = {coords_x[0] - Pi.R[0], coords_x[1] - Pi.R[1], coords_x[2] - Pi.R[2]};
          const double cr[3] = {arm[1] * cq[2] - arm[2] * cq[1], arm[2] * cq[0] - arm[0] * cq[2],
                                arm[0] * cq[1] - arm[1] * cq[0]};
          const double sh = -a.coeff * a.ls *
                            (a.phitau[i * 6 + 0] * cq[0] + a.phitau[i * 6 + 1] * cq[1] + a.phitau[i * 6 + 2] * cq[2] + a.phitau[i * 6 + 3] * cr[0] +
                             a.phitau[i * 6 + 4] * cr[1] + a.phitau[i * 6 + 5] * cr[2]);
          if (i < p) pre += sh;
          if (i == p) s_p = fs + v + shift;
          shat_sum += sh;
        }
        const double aj = lam + shat_sum - pre;
        // u_b of p enters through lc rho (u_b . c ls)
        const double ubar[3] = {aj * lc * (double)rho * cq[0] * a.ls, aj * lc * (double)rho * cq[1] * a.ls,
                                aj * lc * (double)rho * cq[2] * a.ls};
        node_velocity_vjp(xs, Pp, ubar, acc);
        // torque arm: d/dR of -coeff ls tau . ((X - R) x c) s = coeff ls s (c x tau)
        const double *tp = a.phitau + p * 6 + 3;
        const double cxt[3] = {cq[1] * tp[2] - cq[2] * tp[1], cq[2] * tp[0] - cq[0] * tp[2],
                               cq[0] * tp[1] - cq[1] * tp[0]};
        for (int k = 0; k < 3; k++) acc[kAdjR + k] += a.coeff * a.ls * s_p * cxt[k];
      }
    }
    double *dst = a.partial + ((int64_t)blockIdx.x * a.np + p) * kAdjSums;
    if (!any_link) {
      if (tid < kAdjSums) dst[tid] = 0.0;
      continue;
    }
    for (int k = 0; k < kAdjSums; k++) {
      const DD s = wave_sum(DD{acc[k], 0.0});
      if (lane == kWaveSumLane) red[wave][k] = s;
    }
    __syncthreads();
    if (tid < kAdjSums) {
      DD s = red[0][tid];
      for (int w = 1; w < (int)(blockDim.x >> 6); w++) s = dd_merge(s, red[w][tid]);
      dst[tid] = s.hi + s.lo;
    }
    __syncthreads();
  }
}

}  // namespace lbm
