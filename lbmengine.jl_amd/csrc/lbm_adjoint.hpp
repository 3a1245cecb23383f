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
  // the near kernel's blocks: particle p's are [near_first[p], near_first[p + 1]),
  // enough for near_ext[p]^3 nodes of its near box
  int32_t near_first[LBM_MAX_PARTICLES + 1];
  int32_t near_ext[LBM_MAX_PARTICLES];
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
  const double *sums;          // [np][kAdjSums] of tick k's adjoint kernel
  const double *mu_in;
  double *mu;                  // [np][kPState] <- dL/d(state at the start of tick k)
  double *gB;                  // [np][2] += dL/dB1, dL/dB2
  const DevParticle *recs_k;
  int32_t np, sw, direct;      // direct: V_k is one of the averaged speeds
  double gscale;               // 2 (mean - target) / window
};

// Element i of a small register array, and an update of it, without a
// dynamic index: a private array indexed by a value the compiler cannot
// fold is placed in scratch memory, and every access to it then goes through
// the caches to HBM.
// (The read masks the elements' bits: a chain of selects of the elements
// is folded back into one load at a selected address.)
template <int N, class T>
__device__ __forceinline__ T reg_get(const T (&a)[N], int i) {
  using U = std::conditional_t<sizeof(T) == 8, uint64_t, uint32_t>;
  static_assert(sizeof(T) == sizeof(U), "4- or 8-byte elements");
  U v = 0;
#pragma unroll
  for (int k = 0; k < N; k++) v |= __builtin_bit_cast(U, a[k]) & ((U)0 - (U)(i == k));
  return __builtin_bit_cast(T, v);
}
template <int N, class T>
__device__ __forceinline__ void reg_add(T (&a)[N], int i, T d) {
#pragma unroll
  for (int k = 0; k < N; k++) a[k] = i == k ? a[k] + d : a[k];
}
template <int N, class T>
__device__ __forceinline__ void reg_set(T (&a)[N], int i, T v) {
#pragma unroll
  for (int k = 0; k < N; k++) a[k] = i == k ? v : a[k];
}

// Velocity q's components, weight and opposite for a q that varies across
// the lanes (a lane walks its own link slots), from bit masks of the table.
template <int Q>
struct DirMasks {
  static constexpr uint32_t where(const int (&c)[Q], int v) {
    uint32_t m = 0;
    for (int q = 0; q < Q; q++)
      if (c[q] == v) m |= 1u << q;
    return m;
  }
  static constexpr uint32_t cls(int k) {
    uint32_t m = 0;
    for (int q = 0; q < Q; q++)
      if (Lattice<Q>::wclass[q] == k) m |= 1u << q;
    return m;
  }
  static constexpr uint32_t pos[3] = {where(Lattice<Q>::cx, 1), where(Lattice<Q>::cy, 1), where(Lattice<Q>::cz, 1)};
  static constexpr uint32_t neg[3] = {where(Lattice<Q>::cx, -1), where(Lattice<Q>::cy, -1), where(Lattice<Q>::cz, -1)};
  static constexpr uint32_t zero[3] = {where(Lattice<Q>::cx, 0), where(Lattice<Q>::cy, 0), where(Lattice<Q>::cz, 0)};
  static constexpr uint32_t cls1 = cls(1), cls2 = cls(2);
};
template <int Q>
__device__ __forceinline__ int dir_c(int k, int q) {
  using M = DirMasks<Q>;
  return (int)((M::pos[k] >> q) & 1u) - (int)((M::neg[k] >> q) & 1u);
}
template <int Q>
__device__ __forceinline__ double dir_w(int q) {
  using M = DirMasks<Q>;
  using L = Lattice<Q>;
  return ((M::cls1 >> q) & 1u) ? L::wc(1) : (((M::cls2 >> q) & 1u) ? L::wc(2) : L::wc(3));
}
// the index of -c_q: the one velocity whose components are those of c_q negated
template <int Q>
__device__ __forceinline__ int dir_opp(int q) {
  using M = DirMasks<Q>;
  uint32_t m = ~0u;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int c = dir_c<Q>(k, q);
    m &= c > 0 ? M::neg[k] : (c < 0 ? M::pos[k] : M::zero[k]);
  }
  return __builtin_ctz(m);
}

// One adjoint tick: dL/dS_k from dL/dS_{k+1}, with dL/dF +=.  One thread
// per node.  NEAR = false: the bulk, every node outside the particles' near
// boxes (x fastest over the lattice; no particle term reaches these nodes).
// NEAR = true: the nodes of the near boxes (block b of particle p covers
// nodes [j 256, (j + 1) 256) of p's box, j = b - near_first[p]; a node in
// the boxes of several particles belongs to the first), with the particle
// terms and the per-particle sums of the block.
template <int Q, int S, int COLL, int FM, bool NEAR>
__global__ void __launch_bounds__(256) adjoint_kernel(AdjArgs a) {
  using L = Lattice<Q>;
  using T = typename Store<S>::T;
  using R = typename Store<S>::R;
  static_assert(Q == 19, "the adjoint is built for D3Q19");
  const BoxGeom &g = a.g;
  const int tid = threadIdx.x;
  int x = 0, y = 0, z = 0;
  bool live;
  if constexpr (!NEAR) {
    const int64_t nn = (int64_t)blockIdx.x * blockDim.x + tid;
    live = nn < g.N;
    if (live) {
      x = (int)(nn % g.nx);
      const int64_t r = nn / g.nx;
      y = (int)(r % g.ny);
      z = (int)(r / g.ny);
      for (int p = 0; p < a.np; p++)
        if (in_near_box(a.parts[p], g, x, y, z)) live = false;
    }
    if (!live) return;  // (no block-wide step follows in the bulk kernel)
  } else {
    const int b = blockIdx.x;
    int pb = 0;
    while (pb + 1 < a.np && b >= a.near_first[pb + 1]) pb++;
    int lo[3], hi[3];
    near_box(a.parts[pb], g, lo, hi);
    const int ex = hi[0] - lo[0] + 1, ey = hi[1] - lo[1] + 1, ez = hi[2] - lo[2] + 1;
    const int64_t i = (int64_t)(b - a.near_first[pb]) * blockDim.x + tid;
    live = i < (int64_t)ex * ey * ez;
    if (live) {
      x = lo[0] + (int)(i % ex);
      y = lo[1] + (int)((i / ex) % ey);
      z = lo[2] + (int)(i / ((int64_t)ex * ey));
      for (int p = 0; p < pb; p++)
        if (in_near_box(a.parts[p], g, x, y, z)) live = false;
    }
  }
  const int64_t row = (int64_t)y + (int64_t)g.ny * z;
  const int64_t n = row * g.nx + x;
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

  // the coordinates of the node and of its neighbours (x - 1 .. x + 1 along
  // each axis, periodic), loaded together (near kernel)
  double cxs[3] = {0.0, 0.0, 0.0}, cys[3] = {0.0, 0.0, 0.0}, czs[3] = {0.0, 0.0, 0.0};
  if (NEAR && live) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      cxs[k] = a.coords[0][wrapi(x + k - 1, g.nx)];
      cys[k] = a.coords[1][wrapi(y + k - 1, g.ny)];
      czs[k] = a.coords[2][wrapi(z + k - 1, g.nz)];
    }
  }
  // (X + (-R))_1^2 + ... at the neighbour x + (ox, oy, oz), getSphere's order
  auto d2_at = [&](const DevParticle &Pp, int ox, int oy, int oz) {
#pragma clang fp contract(off)
    const double u = cxs[ox + 1] + (-Pp.Rg[0]), v = cys[oy + 1] + (-Pp.Rg[1]), w = czs[oz + 1] + (-Pp.Rg[2]);
    const double su = u * u, sv = v * v, sw = w * w;
    return (su + sv) + sw;
  };

  // link geometry near the particles, the forward box blocks' predicates:
  // (x, q) is a link slot of p when x lies in p's link box, r^2 <= d2(x) <
  // (r + 1.5 dx)^2 and d2(x - c_q) < r^2.  own[p] = link slots (x, q) of p,
  // drop = directions q whose stream x -> x + c_q a link of some particle replaces
  uint32_t own[LBM_MAX_PARTICLES];
  uint32_t own_any = 0, drop = 0;
  for (int p = 0; p < LBM_MAX_PARTICLES; p++) own[p] = 0;
  if (NEAR && live && !wall) {
    for (int p = 0; p < a.np; p++) {
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
        int xn, yn, zn;
        nb(q, xn, yn, zn);
        const double dn = d2_at(Pp, L::cx[q], L::cy[q], L::cz[q]);
        // (x, q̄) is a link slot: x - c_q̄ = x + c_q solid
        if (fl0 && dn < r2) op |= 1u << L::opp[q];
        // (x + c_q, q) is a link slot: x + c_q fluid within the shell, x solid
        if (d0 < r2 && in_box(Pp, xn, yn, zn) && dn >= r2 && dn < ro * ro) drop |= 1u << q;
      }
      reg_set(own, p, op);
      own_any |= op;
    }
  }

  // with a single particle, its per-particle sums (near kernel)
  double acc1[kAdjSums];
#pragma unroll
  for (int k = 0; k < kAdjSums; k++) acc1[k] = 0.0;

  // dL/dP_k at this node
  R lp[Q];
  R grho = 0;
  const double coords_x[3] = {cxs[1], cys[1], czs[1]};
  // dL/dS_{k+1} of this node's own populations, which its link slots read
  R lown[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) lown[q] = 0;
  if (NEAR && own_any)
#pragma unroll
    for (int q = 1; q < Q; q++)
      if ((own_any >> q) & 1) lown[q] = lin[own_idx + (int64_t)q * g.nx];
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
    // particle link slots (x, q'): S_q'(x) = P_q̄'(x) + the owners' Ladd
    // terms; each lane walks its own slots in increasing q' (with a single
    // particle, its per-particle sums are taken in the same walk)
    uint32_t slots = NEAR ? own_any : 0u;
    while (slots) {
      const int qq = __builtin_ctz(slots);
      slots &= slots - 1;
      const int ob = dir_opp<Q>(qq);
      const int c0 = dir_c<Q>(0, qq), c1 = dir_c<Q>(1, qq), c2 = dir_c<Q>(2, qq);
      const double lam = (double)reg_get(lown, qq);
      const double cq[3] = {(double)c0, (double)c1, (double)c2};
      const double xs[3] = {reg_get(cxs, 1 - c0), reg_get(cys, 1 - c1), reg_get(czs, 1 - c2)};
      const double lc = 2.0 * dir_w<Q>(qq) * a.c.invC2s;
      // owners in particle order: u_b . c_q ls and shat_j of each
      double udc_o[LBM_MAX_PARTICLES], sh_o[LBM_MAX_PARTICLES];
      double shat_sum = 0.0;
      for (int p = 0; p < a.np; p++) {
        if (!((reg_get(own, p) >> qq) & 1)) continue;
        const DevParticle &Pp = a.parts[p];
        double ubv[3];
        node_velocity<3>(xs, Pp, LBM_SEMANTICS_GPU, 0.0, ubv);
        const double udc = (ubv[0] * cq[0] + ubv[1] * cq[1] + ubv[2] * cq[2]) * a.ls;
        const double arm[3] = {coords_x[0] - Pp.R[0], coords_x[1] - Pp.R[1], coords_x[2] - Pp.R[2]};
        const double cr[3] = {arm[1] * cq[2] - arm[2] * cq[1], arm[2] * cq[0] - arm[0] * cq[2],
                              arm[0] * cq[1] - arm[1] * cq[0]};
        const double *pt = a.phitau + p * 6;
        const double sh = -a.coeff * a.ls *
                          (pt[0] * cq[0] + pt[1] * cq[1] + pt[2] * cq[2] + pt[3] * cr[0] + pt[4] * cr[1] + pt[5] * cr[2]);
        reg_set(udc_o, p, udc);
        reg_set(sh_o, p, sh);
        shat_sum += sh;
        if (a.np == 1) {
          // the sums below for particle 0 (the only owner): v = fs + its
          // Ladd term, s = fs + v (+ the deviation shift), dL/dv = lam + sh
          const double fs = (double)reg_get(P, ob);
          const double shift = S == kF64 ? 0.0 : 2.0 * dir_w<Q>(qq) * a.c.rho0;
          const double v = fs + lc * (double)rho * udc;
          const double s_p = fs + v + shift;
          const double aj = lam + sh - 0.0;
          const double ubar[3] = {aj * lc * (double)rho * cq[0] * a.ls, aj * lc * (double)rho * cq[1] * a.ls,
                                  aj * lc * (double)rho * cq[2] * a.ls};
          node_velocity_vjp(xs, Pp, ubar, acc1);
          const double *tp = pt + 3;
          const double cxt[3] = {cq[1] * tp[2] - cq[2] * tp[1], cq[2] * tp[0] - cq[0] * tp[2],
                                 cq[0] * tp[1] - cq[1] * tp[0]};
#pragma unroll
          for (int k = 0; k < 3; k++) acc1[kAdjR + k] += a.coeff * a.ls * s_p * cxt[k];
        }
      }
      // dL/dfs = lam + 2 sum shat; dL/dv_j = lam + sum_{i >= j} shat_i
      reg_add(lp, ob, (R)(lam + 2.0 * shat_sum));
      double pre = 0.0;
      for (int p = 0; p < a.np; p++) {
        if (!((reg_get(own, p) >> qq) & 1)) continue;
        const double aj = lam + shat_sum - pre;
        grho += (R)(aj * lc * reg_get(udc_o, p));
        pre += reg_get(sh_o, p);
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

  if constexpr (!NEAR) return;

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
#pragma unroll
    for (int k = 0; k < kAdjSums; k++) acc[k] = 0.0;
    const uint32_t own_p = reg_get(own, p);
    if (a.np == 1) {
#pragma unroll
      for (int k = 0; k < kAdjSums; k++) acc[k] = acc1[k];
    } else if (any_link && own_p != 0) {
      const DevParticle &Pp = a.parts[p];
      uint32_t slots = own_p;
      while (slots) {
        const int qq = __builtin_ctz(slots);
        slots &= slots - 1;
        const int ob = dir_opp<Q>(qq);
        const int c0 = dir_c<Q>(0, qq), c1 = dir_c<Q>(1, qq), c2 = dir_c<Q>(2, qq);
        const double lam = (double)reg_get(lown, qq);
        const double cq[3] = {(double)c0, (double)c1, (double)c2};
        const double xs[3] = {reg_get(cxs, 1 - c0), reg_get(cys, 1 - c1), reg_get(czs, 1 - c2)};
        const double lc = 2.0 * dir_w<Q>(qq) * a.c.invC2s;
        const double shift = S == kF64 ? 0.0 : 2.0 * dir_w<Q>(qq) * a.c.rho0;
        const double fs = (double)reg_get(P, ob);
        // the owner chain up to and after p
        double v = fs, shat_sum = 0.0, pre = 0.0, s_p = 0.0;
        for (int i = 0; i < a.np; i++) {
          if (!((reg_get(own, i) >> qq) & 1)) continue;
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
    if (__any(own_p != 0)) {
#pragma unroll
      for (int k = 0; k < kAdjSums; k++) {
        const DD s = wave_sum(DD{acc[k], 0.0});
        if (lane == kWaveSumLane) red[wave][k] = s;
      }
    } else if (lane == kWaveSumLane) {
      // (a wave without a link slot of p adds zeros)
#pragma unroll
      for (int k = 0; k < kAdjSums; k++) red[wave][k] = DD{0.0, 0.0};
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
