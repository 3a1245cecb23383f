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

// The adjoint takes at most kLocalParticles particles (lbm_adjoint_run
// refuses more): its per-particle arrays live in registers, and each slot of
// them costs every thread of the near kernel registers and selects.
constexpr int kAdjParticles = kLocalParticles;

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
#pragma unroll
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
#pragma unroll
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
__host__ __device__ __forceinline__ void node_velocity_vjp(const double (&xc)[3], const DevParticle &p, const double (&ub)[3],
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
  const void *P;        // P_k = C(S_k), storage type, row-blocked (as the step's buffers)
  const void *H;        // the moments of S_k (rho, u; compute type, row-blocked by 4, as the keeping modes write them)
  const void *lam_in;   // dL/dS_{k+1}, compute type, row-blocked
  void *lam_out;        // dL/dS_k
  void *gF;             // (N, 3) += dL/dF (Julia layout k N + n), compute type: the
                        // segment's sum for reduced storage, folded into FP64 after it
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
  // the near kernel's blocks: particle p's are [near_first[p],
  // near_first[p + 1]), enough for near_ext[p]^3 nodes of its near box
  int32_t near_first[kAdjParticles + 1];
  int32_t near_ext[kAdjParticles];
  // the bulk kernel's fast path (no walls, 32-bit offsets) and its
  // tile_div multipliers for nx and ny
  int32_t fast;
  uint32_t mnx, mny;
};

// Near box of a particle: its link box grown by one node (clamped to the
// lattice).  Every node that holds a link slot, or whose stream a link
// replaces, lies inside; the near kernel takes these nodes, the bulk kernel
// every other.
__device__ __forceinline__ void near_box(const DevParticle &P, const BoxGeom &g, int lo[3], int hi[3]) {
  const int n[3] = {g.nx, g.ny, g.nz};
  for (int k = 0; k < 3; k++) {
    lo[k] = P.lo[k] - 1 < 0 ? 0 : P.lo[k] - 1;
    hi[k] = P.hi[k] + 1 > n[k] - 1 ? n[k] - 1 : P.hi[k] + 1;
  }
}
__device__ __forceinline__ bool in_near_box(const DevParticle &P, const BoxGeom &g, int x, int y, int z) {
  int lo[3], hi[3];
  near_box(P, g, lo, hi);
  return x >= lo[0] && x <= hi[0] && y >= lo[1] && y <= hi[1] && z >= lo[2] && z <= hi[2];
}

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
  const double *mu_in;
  double *mu;                  // [np][kPState] <- dL/d(state at the start of tick k)
  double *gB;                  // [np][2] += dL/dB1, dL/dB2
  const DevParticle *recs_k;
  int32_t np, sw, direct;      // direct: V_k is one of the averaged speeds
  double gscale;               // 2 (mean - target) / window
};

// The tail of reverse tick k in one block (adjoint_tail_kernel): the
// fixed-order sums of the near kernel's per-block partials, the particles'
// adjoint state at the start of tick k (as AdjFinishArgs describes), and,
// when tick k - 1 lies in the same segment, the reverse of tick k - 1's move
// (as AdjMoveArgs describes).
struct AdjTailArgs {
  const double *partial;  // [block][np][kAdjSums] of the near kernel
  int32_t nblocks;
  int32_t do_move;
  AdjFinishArgs fin;
  AdjMoveArgs mv;         // (mv.mu unused: the state comes from fin, in LDS)
};
constexpr int kAdjThreads = 256;  // threads per block of the adjoint kernels
constexpr int kAdjTailThreads = 64 * kAdjParticles;  // one wave per particle's move
constexpr int kNearSlots = 1024;  // link slots a near block stages in LDS per pass

// What the reverse collision of a fluid node needs of the forward one (x,
// row: its position; n: its Julia index): the density and the equilibrium
// velocity of S_k, kept by the recomputation, and the force.
template <int Q, int S, int FM>
__device__ __forceinline__ void adj_moments(const AdjArgs &a, int x, int64_t row, int64_t n,
                                           typename Store<S>::R &rho, typename Store<S>::R (&u)[3],
                                           typename Store<S>::R (&Fn)[3]) {
  using R = typename Store<S>::R;
  const BoxGeom &g = a.g;
  const R *H = static_cast<const R *>(a.H) + row * 4 * g.nx + x;
  rho = H[0];
#pragma unroll
  for (int k = 0; k < 3; k++) u[k] = H[(int64_t)(k + 1) * g.nx];
  Fn[0] = Fn[1] = Fn[2] = (R)0;
  if constexpr (FM == LBM_FORCE_UNIFORM) {
    for (int k = 0; k < 3; k++) Fn[k] = (R)a.c.F[k];
  } else if constexpr (FM == LBM_FORCE_FIELD) {
    const R *Ff = static_cast<const R *>(a.Ffield);
    for (int k = 0; k < 3; k++) Fn[k] = Ff[k * g.N + n];
  }
}

// dL/dP_k of a fluid node from dL/dS_{k+1}: lambda_q(x + c_q), except where
// that node is a wall (bounced back into (x, q̄) instead) or the stream is
// replaced by a particle link (drop bit q).
template <int Q, class R>
__device__ __forceinline__ void adj_gather(const AdjArgs &a, int x, int y, int z, int64_t own_idx, uint32_t drop,
                                          R (&lp)[Q]) {
  using L = Lattice<Q>;
  const BoxGeom &g = a.g;
  const R *lin = static_cast<const R *>(a.lam_in);
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int xn = wrapi(x + L::cx[q], g.nx), yn = wrapi(y + L::cy[q], g.ny), zn = wrapi(z + L::cz[q], g.nz);
    const bool wn = is_wall_node(g, xn, yn, zn);
    R v = 0;
    if (!wn && !((drop >> q) & 1)) v = lin[((int64_t)(yn + g.ny * zn) * Q + q) * g.nx + xn];
    if (wn) v += lin[own_idx + (int64_t)L::opp[q] * g.nx];
    lp[q] = v;
  }
}

// dL/dS_k of a fluid node = the collision's VJP of lp + the Ladd terms'
// dL/drho (rho = sum of S_k); dL/dF += the Guo term's part.
template <int Q, int S, int COLL, int FM>
__device__ __forceinline__ void adj_store(const AdjArgs &a, int64_t own_idx, int64_t n, typename Store<S>::R rho,
                                         const typename Store<S>::R (&u)[3], const typename Store<S>::R (&Fn)[3],
                                         const typename Store<S>::R (&lp)[Q], typename Store<S>::R grho) {
  using R = typename Store<S>::R;
  const BoxGeom &g = a.g;
  R *lout = static_cast<R *>(a.lam_out);
  R ls_[Q], gFn[3] = {0, 0, 0};
  collide_vjp<Q, COLL, FM, R>(rho, u, Fn, a.c, lp, ls_, gFn);
#pragma unroll
  for (int q = 0; q < Q; q++) store_pop(&lout[own_idx + (int64_t)q * g.nx], ls_[q] + grho);
  if constexpr (FM == LBM_FORCE_FIELD)
    for (int k = 0; k < 3; k++) static_cast<R *>(a.gF)[k * g.N + n] += gFn[k];
}

// One adjoint tick on the bulk: dL/dS_k from dL/dS_{k+1}, with dL/dF +=, at
// every node outside the particles' near boxes (block b of the bulk, one
// thread per node, x fastest over the lattice; no particle term reaches
// these nodes).
template <int Q, int S, int COLL, int FM>
__device__ __forceinline__ void adjoint_bulk_block(const AdjArgs &a, int b) {
  using R = typename Store<S>::R;
  const BoxGeom &g = a.g;
  if (a.fast) {
    // Lattices without walls whose population offsets fit in 32 bits: node
    // coordinates by multiply-high divisions, the neighbours' rows from nine
    // row bases (y and z, each -1, 0, +1, periodic), no wall tests.
    using L = Lattice<Q>;
    const uint32_t nn = (uint32_t)b * kAdjThreads + threadIdx.x;
    if (nn >= (uint32_t)g.N) return;
    const uint32_t r = tile_div(nn, g.nx, a.mnx), zz = tile_div(r, g.ny, a.mny);
    const int x = (int)(nn - r * (uint32_t)g.nx), y = (int)(r - zz * (uint32_t)g.ny), z = (int)zz;
    for (int p = 0; p < a.np; p++)
      if (in_near_box(a.parts[p], g, x, y, z)) return;
    const uint32_t Qnx = (uint32_t)Q * (uint32_t)g.nx;
    const uint32_t ya[3] = {(uint32_t)(y == g.ny - 1 ? 0 : y + 1), (uint32_t)y, (uint32_t)(y == 0 ? g.ny - 1 : y - 1)};
    const uint32_t za[3] = {(uint32_t)(z == g.nz - 1 ? 0 : z + 1), (uint32_t)z, (uint32_t)(z == 0 ? g.nz - 1 : z - 1)};
    const uint32_t xa[3] = {(uint32_t)(x == g.nx - 1 ? 0 : x + 1), (uint32_t)x, (uint32_t)(x == 0 ? g.nx - 1 : x - 1)};
    uint32_t rb[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) rb[i][j] = (ya[i] + (uint32_t)g.ny * za[j]) * Qnx;
    const R *lin = static_cast<const R *>(a.lam_in);
    R lp[Q];
    // lambda_S(x + c_q, q): index 1 - c of each axis
#pragma unroll
    for (int q = 0; q < Q; q++)
      lp[q] = lin[rb[1 - L::cy[q]][1 - L::cz[q]] + (uint32_t)q * (uint32_t)g.nx + xa[1 - L::cx[q]]];
    const uint32_t row = (uint32_t)y + (uint32_t)g.ny * (uint32_t)z;
    R rho, u[3], Fn[3];
    adj_moments<Q, S, FM>(a, x, row, (int64_t)(row * (uint32_t)g.nx + (uint32_t)x), rho, u, Fn);
    adj_store<Q, S, COLL, FM>(a, (int64_t)(row * Qnx + (uint32_t)x), (int64_t)(row * (uint32_t)g.nx + (uint32_t)x), rho,
                              u, Fn, lp, (R)0);
    return;
  }
  const int64_t nn = (int64_t)b * kAdjThreads + threadIdx.x;
  if (nn >= g.N) return;
  const int x = (int)(nn % g.nx);
  const int64_t r = nn / g.nx;
  const int y = (int)(r % g.ny), z = (int)(r / g.ny);
  for (int p = 0; p < a.np; p++)
    if (in_near_box(a.parts[p], g, x, y, z)) return;
  const int64_t row = (int64_t)y + (int64_t)g.ny * z;
  const int64_t n = row * g.nx + x, own_idx = row * Q * g.nx + x;
  if (is_wall_node(g, x, y, z)) {
    R *lout = static_cast<R *>(a.lam_out);
#pragma unroll
    for (int q = 0; q < Q; q++) store_pop(&lout[own_idx + (int64_t)q * g.nx], (R)0);
    return;
  }
  R rho, u[3], Fn[3], lp[Q];
  adj_moments<Q, S, FM>(a, x, row, n, rho, u, Fn);
  adj_gather<Q, R>(a, x, y, z, own_idx, 0u, lp);
  adj_store<Q, S, COLL, FM>(a, own_idx, n, rho, u, Fn, lp, (R)0);
}

// The owners' Ladd chain of one link slot (x, q), reversed.  Forward:
// S_q(x) = v_J with v_0 = fs = P_q̄(x) and v_i = v_{i-1} + lc rho(x) (u_b^i .
// c_q ls) over the owners i in increasing particle order (om: their bits),
// and the momentum exchange s_i = fs + v_i (+ the deviation shift) into each
// owner's force and torque, whose adjoint is shat_i = -coeff ls [phi_i . c +
// tau_i . ((X - R_i) x c)].  Returns dL/dfs = lam + 2 sum shat and the slot's
// dL/drho(x) term sum_i (lam + sum_{j >= i} shat_j) lc (u_b^i . c ls); when
// particle `want` is an owner, adds its dL/d(B1, B2, node-velocity snapshot,
// R) to acc.  xs: the solid end x - c_q, xc: x.
template <int Q, int S>
__device__ __forceinline__ void slot_chain(const AdjArgs &a, int q, const double (&xs)[3], const double (&xc)[3],
                                          double lam, double fs, double rho, uint32_t om, int want, double &dfs,
                                          double &drho, double (&acc)[kAdjSums]) {
  using L = Lattice<Q>;
  const double cq[3] = {(double)L::cx[q], (double)L::cy[q], (double)L::cz[q]};
  const double w = L::w(q);
  const double lc = 2.0 * w * a.c.invC2s;
  const double shift = S == kF64 ? 0.0 : 2.0 * w * a.c.rho0;
  auto shat = [&](int p) {
    const DevParticle &Pp = a.parts[p];
    const double arm[3] = {xc[0] - Pp.R[0], xc[1] - Pp.R[1], xc[2] - Pp.R[2]};
    const double cr[3] = {arm[1] * cq[2] - arm[2] * cq[1], arm[2] * cq[0] - arm[0] * cq[2],
                          arm[0] * cq[1] - arm[1] * cq[0]};
    const double *pt = a.phitau + p * 6;
    return -a.coeff * a.ls *
           (pt[0] * cq[0] + pt[1] * cq[1] + pt[2] * cq[2] + pt[3] * cr[0] + pt[4] * cr[1] + pt[5] * cr[2]);
  };
  double shat_sum = 0.0;
  for (int p = 0; p < a.np; p++)
    if ((om >> p) & 1u) shat_sum += shat(p);
  dfs = lam + 2.0 * shat_sum;
  drho = 0.0;
  double pre = 0.0, v = fs;
  for (int p = 0; p < a.np; p++) {
    if (!((om >> p) & 1u)) continue;
    const DevParticle &Pp = a.parts[p];
    double ubv[3];
    node_velocity<3>(xs, Pp, LBM_SEMANTICS_GPU, 0.0, ubv);
    const double udc = (ubv[0] * cq[0] + ubv[1] * cq[1] + ubv[2] * cq[2]) * a.ls;
    const double aj = lam + shat_sum - pre;
    drho += aj * lc * udc;
    v = v + lc * rho * udc;
    if (p == want) {
      // u_b of p enters through lc rho (u_b . c ls)
      const double ubar[3] = {aj * lc * rho * cq[0] * a.ls, aj * lc * rho * cq[1] * a.ls,
                              aj * lc * rho * cq[2] * a.ls};
      node_velocity_vjp(xs, Pp, ubar, acc);
      // torque arm: d/dR of -coeff ls tau . ((X - R) x c) s = coeff ls s (c x tau)
      const double s_p = fs + v + shift;
      const double *tp = a.phitau + p * 6 + 3;
      const double cxt[3] = {cq[1] * tp[2] - cq[2] * tp[1], cq[2] * tp[0] - cq[0] * tp[2],
                             cq[0] * tp[1] - cq[1] * tp[0]};
#pragma unroll
      for (int k = 0; k < 3; k++) acc[kAdjR + k] += a.coeff * a.ls * s_p * cxt[k];
    }
    pre += shat(p);
  }
}

// Adds acc (per lane) over the block into tot, in a fixed order: the lanes of
// each wave (DPP), then the waves in order.
__device__ __forceinline__ void adj_block_sum(const double (&acc)[kAdjSums], DD (&red)[kAdjThreads / 64][kAdjSums],
                                              DD *tot) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = __lane_id();
#pragma unroll
  for (int k = 0; k < kAdjSums; k++) {
    const DD s = wave_sum(DD{acc[k], 0.0});
    if (lane == kWaveSumLane) red[wave][k] = s;
  }
  __syncthreads();
  if (tid < kAdjSums) {
    DD s = tot[tid];
    for (int w = 0; w < kAdjThreads / 64; w++) s = dd_merge(s, red[w][tid]);
    tot[tid] = s;
  }
  __syncthreads();
}

// One adjoint tick on the particles' near boxes (block b of particle p
// covers box indices [j 256, (j + 1) 256) of p's near box, j = b -
// near_first[p], x fastest; a node in the boxes of several particles belongs
// to the first).  A lane per node finds the node's link slots and the
// streams links replace, and stages its slots in LDS (node order, each
// node's in increasing q); a lane per slot then runs the owners' Ladd chain
// backwards (slot_chain); the node lanes take the results into LDS rows of
// their own in the order they staged them, then gather, add them and run the
// collision's VJP.  The per-particle sums of the block go to
// partial[b][p][...].
template <int Q, int S, int COLL, int FM>
__device__ __forceinline__ void adjoint_near_block(const AdjArgs &a, int b) {
  using L = Lattice<Q>;
  using T = typename Store<S>::T;
  using R = typename Store<S>::R;
  const BoxGeom &g = a.g;
  const int tid = threadIdx.x, wave = tid >> 6, lane = __lane_id();
  int pb = 0;
  while (pb + 1 < a.np && b >= a.near_first[pb + 1]) pb++;
  int lo[3], hi[3];
  near_box(a.parts[pb], g, lo, hi);
  const int ex = hi[0] - lo[0] + 1, ey = hi[1] - lo[1] + 1, ez = hi[2] - lo[2] + 1;
  const int64_t i0 = (int64_t)(b - a.near_first[pb]) * kAdjThreads;
  auto node_of = [&](int t, int &x, int &y, int &z) {
    const int64_t i = i0 + t;
    x = lo[0] + (int)(i % ex);
    y = lo[1] + (int)((i / ex) % ey);
    z = lo[2] + (int)(i / ((int64_t)ex * ey));
  };
  int x = 0, y = 0, z = 0;
  bool live = i0 + tid < (int64_t)ex * ey * ez;
  if (live) {
    node_of(tid, x, y, z);
    for (int p = 0; p < pb; p++)
      if (in_near_box(a.parts[p], g, x, y, z)) live = false;
  }
  const int64_t row = (int64_t)y + (int64_t)g.ny * z;
  const int64_t n = row * g.nx + x, own_idx = row * Q * g.nx + x;
  const bool fluid = live && !is_wall_node(g, x, y, z);
  const R *lin = static_cast<const R *>(a.lam_in);
  const T *Pk = static_cast<const T *>(a.P);

  R rho = 0, u[3] = {0, 0, 0}, Fn[3] = {0, 0, 0};
  if (fluid) adj_moments<Q, S, FM>(a, x, row, n, rho, u, Fn);

  // link geometry, the forward box blocks' predicates: (x, q) is a link slot
  // of p when x lies in p's link box, r^2 <= d2(x) < (r + 1.5 dx)^2 and
  // d2(x - c_q) < r^2.  own[p]: link slots (x, q) of p; drop: directions q
  // whose stream x -> x + c_q a link of some particle replaces
  uint32_t own[kAdjParticles];
#pragma unroll
  for (int p = 0; p < kAdjParticles; p++) own[p] = 0;
  uint32_t own_any = 0, drop = 0;
  if (fluid) {
    double cxs[3], cys[3], czs[3];  // coordinates at x - 1 .. x + 1 along each axis (periodic)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      cxs[k] = a.coords[0][wrapi(x + k - 1, g.nx)];
      cys[k] = a.coords[1][wrapi(y + k - 1, g.ny)];
      czs[k] = a.coords[2][wrapi(z + k - 1, g.nz)];
    }
    // (X + (-R))_1^2 + ... at the neighbour x + (ox, oy, oz), getSphere's order
    auto d2_at = [&](const DevParticle &Pp, int ox, int oy, int oz) {
#pragma clang fp contract(off)
      const double u = cxs[ox + 1] + (-Pp.Rg[0]), v = cys[oy + 1] + (-Pp.Rg[1]), w = czs[oz + 1] + (-Pp.Rg[2]);
      const double su = u * u, sv = v * v, sw = w * w;
      return (su + sv) + sw;
    };
#pragma unroll
    for (int p = 0; p < kAdjParticles; p++) {
      if (p >= a.np) break;
      const DevParticle &Pp = a.parts[p];
      if (x < Pp.lo[0] - 1 || x > Pp.hi[0] + 1 || y < Pp.lo[1] - 1 || y > Pp.hi[1] + 1 || z < Pp.lo[2] - 1 ||
          z > Pp.hi[2] + 1)
        continue;
      const double r2 = Pp.radius * Pp.radius, ro = Pp.radius + a.shell;
      const double d0 = d2_at(Pp, 0, 0, 0);
      const bool fl0 = in_box(Pp, x, y, z) && d0 >= r2 && d0 < ro * ro;  // may hold link slots
      uint32_t op = 0;
#pragma unroll
      for (int q = 1; q < Q; q++) {
        const int xn = wrapi(x + L::cx[q], g.nx), yn = wrapi(y + L::cy[q], g.ny), zn = wrapi(z + L::cz[q], g.nz);
        const double dn = d2_at(Pp, L::cx[q], L::cy[q], L::cz[q]);
        // (x, q̄) is a link slot: x - c_q̄ = x + c_q solid
        if (fl0 && dn < r2) op |= 1u << L::opp[q];
        // (x + c_q, q) is a link slot: x + c_q fluid within the shell, x solid
        if (d0 < r2 && in_box(Pp, xn, yn, zn) && dn >= r2 && dn < ro * ro) drop |= 1u << q;
      }
      own[p] = op;
      own_any |= op;
    }
  }
  R grho = 0;

  // this node's first slot in the block's list: an exclusive scan of the
  // slot counts over the lanes, then the waves
  __shared__ int wtot[kAdjThreads / 64];
  __shared__ uint32_t s_key[kNearSlots];  // node (bits 0-7), q (8-12), owners (16-23)
  __shared__ R s_a[kNearSlots], s_b[kNearSlots];  // lam and fs in; dL/dfs and the dL/drho term out
  __shared__ R s_rho[kAdjThreads];
  __shared__ DD red[kAdjThreads / 64][kAdjSums];
  __shared__ DD tot[kAdjParticles][kAdjSums];
  __shared__ R s_add[Q][kAdjThreads];  // row tid: dL/dfs of node tid's slots, by the population they add to
  const int cnt = __popc(own_any);
  int incl = cnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int t = __shfl_up(incl, d);
    if (lane >= d) incl += t;
  }
  if (lane == 63) wtot[wave] = incl;
  s_rho[tid] = rho;
  if (tid < a.np * kAdjSums) tot[tid / kAdjSums][tid % kAdjSums] = DD{0.0, 0.0};
  __syncthreads();
  int first = incl - cnt, total = 0;
#pragma unroll
  for (int w = 0; w < kAdjThreads / 64; w++) {
    if (w < wave) first += wtot[w];
    total += wtot[w];
  }

  for (int pass0 = 0; pass0 < total; pass0 += kNearSlots) {
    if (own_any) {
      int s = first - pass0;
#pragma unroll
      for (int q = 1; q < Q; q++) {
        if (!((own_any >> q) & 1u)) continue;
        if (s >= 0 && s < kNearSlots) {
          uint32_t om = 0;
#pragma unroll
          for (int p = 0; p < kAdjParticles; p++) om |= ((own[p] >> q) & 1u) << p;
          s_key[s] = (uint32_t)tid | (uint32_t)q << 8 | om << 16;
          s_a[s] = lin[own_idx + (int64_t)q * g.nx];
          s_b[s] = decode<S, Q>(Pk[own_idx + (int64_t)L::opp[q] * g.nx], L::opp[q]);
        }
        s++;
      }
    }
    __syncthreads();
    const int ns = total - pass0 < kNearSlots ? total - pass0 : kNearSlots;
    // a lane per slot; slot j's node and solid end
    auto slot_at = [&](int j, int &q, uint32_t &om, double (&xs)[3], double (&xc)[3]) {
      const uint32_t key = s_key[j];
      q = (int)((key >> 8) & 31u);
      om = key >> 16;
      int sx, sy, sz;
      node_of((int)(key & 255u), sx, sy, sz);
      xc[0] = a.coords[0][sx];
      xc[1] = a.coords[1][sy];
      xc[2] = a.coords[2][sz];
      xs[0] = a.coords[0][wrapi(sx - L::cx[q], g.nx)];
      xs[1] = a.coords[1][wrapi(sy - L::cy[q], g.ny)];
      xs[2] = a.coords[2][wrapi(sz - L::cz[q], g.nz)];
      return (int)(key & 255u);
    };
    double acc[kAdjSums];
    if (a.np > 1) {
      // each particle's sums, from the slots it owns (inputs still in s_a, s_b)
      for (int pp = 0; pp < a.np; pp++) {
#pragma unroll
        for (int k = 0; k < kAdjSums; k++) acc[k] = 0.0;
        for (int j = tid; j < ns; j += kAdjThreads) {
          int q;
          uint32_t om;
          double xs[3], xc[3];
          const int nd = slot_at(j, q, om, xs, xc);
          if (!((om >> pp) & 1u)) continue;
          double dfs, drho;
          slot_chain<Q, S>(a, q, xs, xc, (double)s_a[j], (double)s_b[j], (double)s_rho[nd], om, pp, dfs, drho, acc);
        }
        adj_block_sum(acc, red, tot[pp]);
      }
    }
#pragma unroll
    for (int k = 0; k < kAdjSums; k++) acc[k] = 0.0;
    for (int j = tid; j < ns; j += kAdjThreads) {
      int q;
      uint32_t om;
      double xs[3], xc[3];
      const int nd = slot_at(j, q, om, xs, xc);
      double dfs, drho;
      slot_chain<Q, S>(a, q, xs, xc, (double)s_a[j], (double)s_b[j], (double)s_rho[nd], om, a.np == 1 ? 0 : -1,
                       dfs, drho, acc);
      s_a[j] = (R)dfs;
      s_b[j] = (R)drho;
    }
    if (a.np == 1) adj_block_sum(acc, red, tot[0]);
    else __syncthreads();
    // the node lanes take their slots' results in the order they staged them
    if (own_any) {
      int s = first - pass0;
#pragma unroll
      for (int q = 1; q < Q; q++) {
        if (!((own_any >> q) & 1u)) continue;
        if (s >= 0 && s < kNearSlots) {
          s_add[L::opp[q]][tid] = s_a[s];
          grho += s_b[s];
        }
        s++;
      }
    }
    __syncthreads();
  }

  if (live) {
    if (!fluid) {
      R *lout = static_cast<R *>(a.lam_out);
#pragma unroll
      for (int q = 0; q < Q; q++) store_pop(&lout[own_idx + (int64_t)q * g.nx], (R)0);
    } else {
      R lp[Q];
      adj_gather<Q, R>(a, x, y, z, own_idx, drop, lp);
#pragma unroll
      for (int q = 1; q < Q; q++)
        if ((own_any >> q) & 1u) lp[L::opp[q]] += s_add[L::opp[q]][tid];
      adj_store<Q, S, COLL, FM>(a, own_idx, n, rho, u, Fn, lp, grho);
    }
  }
  if (tid < a.np * kAdjSums) {
    const DD t = tot[tid / kAdjSums][tid % kAdjSums];
    a.partial[(int64_t)b * a.np * kAdjSums + tid] = t.hi + t.lo;
  }
}

// One adjoint tick (dL/dS_k from dL/dS_{k+1}, with dL/dF +=) is two kernels
// on one stream: the near blocks, then the bulk.  Their register needs
// differ (about 270 against 85): one build of both would leave the bulk at
// two waves per SIMD or, capped at 128 registers, spill the near blocks'
// slot phase (550 B a lane) and make them the launch's critical path.
template <int Q, int S, int COLL, int FM>
__global__ void __launch_bounds__(kAdjThreads) adjoint_kernel(AdjArgs a) {
  static_assert(Q == 19, "the adjoint is built for D3Q19");
  adjoint_bulk_block<Q, S, COLL, FM>(a, blockIdx.x);
}

template <int Q, int S, int COLL, int FM>
__global__ void __launch_bounds__(kAdjThreads) adjoint_near_kernel(AdjArgs a) {
  static_assert(Q == 19, "the adjoint is built for D3Q19");
  adjoint_near_block<Q, S, COLL, FM>(a, blockIdx.x);
}

}  // namespace lbm
