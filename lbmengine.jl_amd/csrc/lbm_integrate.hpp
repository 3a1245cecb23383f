// lbm_integrate.hpp — the particle side of a step: double-double
// accumulation of the momentum exchange, its fixed-order final reduction,
// and moveParticles! (gpu/particles.jl:7-92, cpu/particles.jl:45-100) with
// the link-box update.
#pragma once

#include "lbm_common.hpp"

namespace lbm {

// Double-double accumulation (Knuth's TwoSum) of the momentum-exchange sums:
// they are evaluated essentially exactly and rounded once, so the result does
// not depend on the order in which links, waves and blocks are combined.
struct DD {
  double hi, lo;
};

__device__ __forceinline__ void dd_add(DD &a, double x) {
#pragma clang fp contract(off)
  const double s = a.hi + x;
  const double bb = s - a.hi;
  const double err = (a.hi - (s - bb)) + (x - bb);
  a.hi = s;
  a.lo = a.lo + err;
}

__device__ __forceinline__ DD dd_merge(DD a, DD b) {
#pragma clang fp contract(off)
  DD r{a.hi, a.lo + b.lo};
  dd_add(r, b.hi);
  return r;
}

// One DPP move of a double (two 32-bit halves); lanes without a source read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROW_MASK, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROW_MASK, 0xf, true);
  return __builtin_bit_cast(double, (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32));
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ DD dd_dpp_merge(DD v) {
  return dd_merge(v, DD{dpp_f64<CTRL, ROW_MASK>(v.hi), dpp_f64<CTRL, ROW_MASK>(v.lo)});
}

// Sum over the 64 lanes of a full wave, in a fixed order, valid in lane 63:
// row_shr 1, 2, 4, 8 within each row of 16, then row_bcast 15 and 31 across
// rows (all DPP, no LDS round trips).
__device__ __forceinline__ DD wave_sum(DD v) {
  v = dd_dpp_merge<0x111, 0xf>(v);  // row_shr:1
  v = dd_dpp_merge<0x112, 0xf>(v);  // row_shr:2
  v = dd_dpp_merge<0x114, 0xf>(v);  // row_shr:4
  v = dd_dpp_merge<0x118, 0xf>(v);  // row_shr:8
  v = dd_dpp_merge<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v = dd_dpp_merge<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}
constexpr int kWaveSumLane = 63;

// Most particles one box block works with: the lattice's particles when it
// has at most this many, else those whose owned regions meet its own
// particle's (a link or band node of any other particle lies outside).
constexpr int kLocalParticles = 8;

// What the kernels of a tick need for the particles.  A box kernel reads
// the records of its tick (parts_in) and, when they are pending, first
// applies the move of the previous box kernel from its partials
// (partial_in); it writes its own partials (partial_out) and, from box block
// 0, the records of the next kernel (parts_out, pending).  Records and
// partials are double-buffered by tick parity.
struct StepParticles {
  const DevParticle *parts_in;
  DevParticle *parts_out;
  const double *coords[3];
  const double *partial_in;  // [box block][local particle][6][hi, lo] of the previous box kernel
  double *partial_out;       // the same, of this kernel
  // Lattices with more than kLocalParticles particles (many = 1): the box
  // blocks take the records as moved by the move kernel after each
  // gathering kernel, write their partials under the ids of their local
  // particles, and raise flags[0] when more than kLocalParticles particles
  // reach one block.
  const int32_t *ids_in;     // [box block][local particle]: particle id, or -1
  int32_t *ids_out;
  int32_t *flags;
  int32_t many;
  int32_t reserved_m;
  void *linkval;             // [box block][T (Q - 1)]: link-slot values (compute type)
  const lbm_interaction *inter;  // particle-particle interaction terms, in application order
  double *traj;              // per-tick log: [particle][traj_cap][9], or null
  int64_t traj_cap;
  const double *bsched;      // scheduled magnetic field: [bsched_count][3], entry i of tick bsched_first + i
  int64_t bsched_first, bsched_count;
  int64_t tick0;             // the tick at which particle 0 was added: a move m of it ends tick tick0 + m
  int32_t ninter;
  int32_t np, semantics, dims;
  // box blocks: particle p's are [first_block[p], first_block[p + 1]);
  // nbox = first_block[np] of them lead the grid, padded to nbox_pad (a
  // multiple of the 8 XCDs)
  int32_t first_block[LBM_MAX_PARTICLES + 1];
  int32_t nbox, nbox_pad;
  int32_t coord_stride;  // doubles per axis of a box block's coordinate stage
  double coeff;          // dx^d / dt (gpu semantics) or dx^d (cpu semantics)
  double ls;             // lattice speed dx / dt
  double cutoff;         // 2 dx (cpu-semantics node-velocity range)
  double shell;          // link nodes lie within r + shell of a centre (1.5 dx; 1.75 dx with body diagonals)
  double dt, dx;
  double x0[3];
  int32_t n[3];
};

// Nodes a box kernel owns for a particle: its link box grown by kBoxMargin on
// every side (clamped to the lattice), so that it holds the link nodes of the
// moved particle and of the next tick's move as well.
constexpr int kBoxMargin = 2;

// Link box of a sphere of radius r at Rg (node indices, inclusive): every
// solid node and every fluid node with a solid neighbour lies inside.  A box
// that leaves the lattice raises *err and is clamped into it; the bounds are
// formed and clamped in double precision, so that any centre (huge,
// infinite or NaN, as a diverging run may produce) gives indices inside the
// lattice rather than a saturated or wrapped integer.
__host__ __device__ inline void link_box(const double Rg[3], double r, const double x0[3],
                                         double dx, const int32_t n[3], int dims,
                                         int32_t lo[3], int32_t hi[3], int32_t *err) {
  for (int k = 0; k < 3; k++) {
    if (k >= dims) {
      lo[k] = 0;
      hi[k] = 0;
      continue;
    }
    const double a = (Rg[k] - r - x0[k]) / dx, b = (Rg[k] + r - x0[k]) / dx;
    const double top = (double)(n[k] - 1);
    double l = floor(a) - 2.0, h = ceil(b) + 2.0;
    if (!(l >= 0.0) || !(h <= top)) {  // (NaN fails both)
      *err = 1;
      l = l >= 0.0 ? (l <= top ? l : top) : 0.0;
      h = h <= top ? (h >= 0.0 ? h : 0.0) : top;
      if (!(a == a) || !(b == b)) l = 0.0, h = 0.0;
    }
    lo[k] = (int32_t)l;
    hi[k] = (int32_t)h;
  }
}

// Largest link-box extent along one dimension for radius r.
inline int32_t link_box_extent(double r, double dx) { return (int32_t)ceil(2.0 * r / dx) + 7; }

// (Loops over the dims components run a fixed 3 times with a guard: a small
// array indexed in a loop of runtime trip count is placed in scratch memory,
// and the single thread that moves a particle then waits on a memory round
// trip at every access.)
__host__ __device__ inline bool any_nonzero(const double *v, int n) {
  bool nz = false;
#pragma unroll
  for (int k = 0; k < 3; k++)
    if (k < n && v[k] != 0.0) nz = true;
  return nz;
}

// moveParticles!, gpu/particles.jl:7-92.  The half-kick branch
// (model.tick == 0) is unreachable there because tick! increments the tick
// first (gpu/fluids.jl:217), so only the full kick is restated.
__host__ __device__ inline void move_gpu(DevParticle &p, const double F[3], const double T[3], int dims,
                                double dt, bool &moved, bool &find_nv) {
#pragma clang fp contract(off)
  moved = false;
  find_nv = false;
  if (p.couple_forces) {
    double A[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; k++) if (k < dims) A[k] = F[k] * p.inv_mass;
#pragma unroll
    for (int k = 0; k < 3; k++) if (k < dims) p.V[k] = p.V[k] + dt * A[k];
#pragma unroll
    for (int k = 0; k < 3; k++) if (k < dims) p.R[k] = p.R[k] + dt * p.V[k];
    moved = any_nonzero(p.V, dims);
  }
  if (p.couple_torques) {
    const int nw = dims == 2 ? 1 : 3;
    double Al[3];
#pragma unroll
    for (int k = 0; k < 3; k++) Al[k] = k < nw ? T[k] * p.inv_I : 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) if (k < nw) p.W[k] = p.W[k] + dt * Al[k];
    if (p.kind == LBM_PARTICLE_SQUIRMER) {
      // Rodrigues' rotation of the swimming direction by |omega| dt
      double *e = p.e;
      const double *W = p.W;
      double theta;
      if (dims == 2) {
        theta = fabs(W[0]);
      } else {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 3; k++) s = s + W[k] * W[k];
        theta = sqrt(s);
      }
      if (!(theta == 0.0)) {
        if (dims == 2) {
          const double kk = W[0] / theta;
          theta *= dt;
          const double c = cos(theta), s = sin(theta);
          const double cr[2] = {-kk * e[1], kk * e[0]};
          double vr[2];
          vr[0] = e[0] * c + cr[0] * s;
          vr[1] = e[1] * c + cr[1] * s;
          const double nv = sqrt(vr[0] * vr[0] + vr[1] * vr[1]);
          e[0] = vr[0] / nv;
          e[1] = vr[1] / nv;
        } else {
          const double kk[3] = {W[0] / theta, W[1] / theta, W[2] / theta};
          theta *= dt;
          const double c = cos(theta), s = sin(theta);
          const double cr[3] = {kk[1] * e[2] - kk[2] * e[1], -kk[0] * e[2] + kk[2] * e[0],
                                kk[0] * e[1] - kk[1] * e[0]};
          double vr[3];
#pragma unroll
          for (int k = 0; k < 3; k++) vr[k] = e[k] * c + cr[k] * s;
          double kd = 0.0;
#pragma unroll
          for (int k = 0; k < 3; k++) kd = kd + kk[k] * e[k];
#pragma unroll
          for (int k = 0; k < 3; k++) vr[k] = vr[k] + kk[k] * kd * (1.0 - c);
          double nv = 0.0;
#pragma unroll
          for (int k = 0; k < 3; k++) nv = nv + vr[k] * vr[k];
          nv = sqrt(nv);
#pragma unroll
          for (int k = 0; k < 3; k++) e[k] = vr[k] / nv;
        }
      }
    }
    find_nv = any_nonzero(Al, nw);
  }
  find_nv = moved || find_nv;
}

// moveParticles!, cpu/particles.jl:45-100: v += dP/m, omega += dL/I,
// Euler step only when v != 0, no rotation of the swimming direction.
__device__ inline void move_cpu(DevParticle &p, const double dP[3], const double dL[3], int dims,
                                double dt, bool &moved, bool &find_nv) {
#pragma clang fp contract(off)
  const int nw = dims == 2 ? 1 : 3;
#pragma unroll
  for (int k = 0; k < 3; k++) if (k < dims) p.V[k] = p.V[k] + p.inv_mass * dP[k];
#pragma unroll
  for (int k = 0; k < 3; k++) if (k < nw) p.W[k] = p.W[k] + p.inv_I * dL[k];
  moved = false;
  if (any_nonzero(p.V, dims)) {
#pragma unroll
    for (int k = 0; k < 3; k++) if (k < dims) p.R[k] = p.R[k] + dt * p.V[k];
    moved = true;
  }
  find_nv = any_nonzero(dL, nw) || moved;
}

// ------------------------------------------- particle-particle interactions
// x^n (n >= 1) as Julia's ^(::Float64, ::Integer) evaluates it (Base.Math's
// pow_body: powering by squaring with each product's rounding error carried
// by fma, so the result is within about an ulp of the exact power); n == 3
// is x*x*x, as Julia's literal cube is.
__host__ __device__ inline double julia_pow(double x, int n) {
#pragma clang fp contract(off)
  if (n == 3) return x * x * x;
  double y = 1.0, xnlo = 0.0, ynlo = 0.0;
  while (n > 1) {
    if (n & 1) {
      const double err = fma(y, xnlo, x * ynlo);
      const double hi = x * y;
      ynlo = fma(x, y, -hi) + err;
      y = hi;
    }
    const double err = x * 2.0 * xnlo;
    const double hi = x * x;
    xnlo = fma(x, x, -hi) + err;
    x = hi;
    n >>= 1;
  }
  const double err = fma(y, xnlo, x * ynlo);
  return (isfinite(x) && isfinite(err)) ? fma(x, y, err) : x * y;
}

// norm of a short host vector (LinearAlgebra's generic_norm2 below its BLAS
// cutoff): the squares summed in index order
__host__ __device__ inline double norm_seq(const double *v, int n) {
#pragma clang fp contract(off)
  double s = v[0] * v[0];
#pragma unroll
  for (int k = 1; k < 3; k++)
    if (k < n) s = s + v[k] * v[k];
  return sqrt(s);
}

// The reference's cross (gpu/aux.jl:139-157): 3-vectors as usual; two
// 2-vectors give the scalar z component; a scalar w with a 2-vector V gives
// (-w V_2, w V_1) whichever side the scalar is on.
__host__ __device__ inline void cross3(const double *v, const double *w, double *o) {
#pragma clang fp contract(off)
  o[0] = v[1] * w[2] - v[2] * w[1];
  o[1] = -v[0] * w[2] + v[2] * w[0];
  o[2] = v[0] * w[1] - v[1] * w[0];
}
__host__ __device__ inline double cross2(const double *v, const double *w) {
#pragma clang fp contract(off)
  return v[0] * w[1] - v[1] * w[0];
}
__host__ __device__ inline void cross_sv(double s, const double *V, double *o) {
#pragma clang fp contract(off)
  o[0] = -s * V[1];
  o[1] = s * V[0];
}

// One term of molecularDynamicsTick! (gpu/particles.jl:110-197) from the
// positions R1, R2 (and R3) before the tick's moves.  Pair terms give the
// force on id1 in fa (id2 takes -fa); the polar term gives force321 (on id1)
// in fa and force123 (on id3) in fb (id2 takes -(fa + fb)).  Returns false
// when the term adds nothing (Lennard-Jones beyond its cutoff).
__host__ __device__ inline bool interaction_term(const lbm_interaction &I, const double *R1, const double *R2,
                                                 const double *R3, int d, double fa[3], double fb[3],
                                                 const double *Bt = nullptr) {
#pragma clang fp contract(off)
  for (int k = 0; k < 3; k++) fa[k] = fb[k] = 0.0;
  if (I.kind == LBM_INTERACTION_POLAR) {
    // gpu/particles.jl:136-158
    double d12[3] = {0, 0, 0}, d32[3] = {0, 0, 0}, u12[3] = {0, 0, 0}, u32[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      if (k >= d) continue;
      d12[k] = R1[k] - R2[k];
      d32[k] = R3[k] - R2[k];
    }
    const double n12 = norm_seq(d12, d), n32 = norm_seq(d32, d);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      if (k >= d) continue;
      u12[k] = d12[k] / n12;
      u32[k] = d32[k] / n32;
    }
    double c = u12[0] * u32[0];
#pragma unroll
    for (int k = 1; k < 3; k++)
      if (k < d) c = c + u12[k] * u32[k];
    if (fabs(c) > 1.0) c = c > 0.0 ? 1.0 : -1.0;
    const double angle = acos(c);
    const double s = I.params[1] * (angle - I.params[0]);
    if (d == 3) {
      double td[3], tq[3], c1[3], c3[3];
      cross3(u32, u12, td);
      const double an = norm_seq(td, 3);
      if (an > 0.0)
        for (int k = 0; k < 3; k++) td[k] = td[k] / an;
      for (int k = 0; k < 3; k++) tq[k] = s * td[k];
      cross3(tq, u12, c1);
      cross3(tq, u32, c3);
      for (int k = 0; k < 3; k++) {
        fa[k] = -c1[k] / n12;
        fb[k] = c3[k] / n32;
      }
    } else {
      double td = cross2(u32, u12);
      const double an = fabs(td);
      if (an > 0.0) td = td / an;
      const double tq = s * td;
      double c1[2], c3[2];
      cross_sv(tq, u12, c1);
      cross_sv(tq, u32, c3);
      for (int k = 0; k < 2; k++) {
        fa[k] = -c1[k] / n12;
        fb[k] = c3[k] / n32;
      }
    }
    return true;
  }
  double d21[3] = {0, 0, 0}, u[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 3; k++)
    if (k < d) d21[k] = R2[k] - R1[k];
  const double disp = norm_seq(d21, d);
  if (I.kind == LBM_INTERACTION_LENNARD_JONES && !(disp < I.params[2])) return false;
#pragma unroll
  for (int k = 0; k < 3; k++)
    if (k < d) u[k] = d21[k] / disp;
  double s;
  switch (I.kind) {
    case LBM_INTERACTION_LINEAR:  // gpu/particles.jl:110-121
      s = I.params[1] * (disp - I.params[0]);
      break;
    case LBM_INTERACTION_BISTABLE: {  // gpu/particles.jl:123-134
      const double x = disp - I.params[0];
      s = I.params[1] * julia_pow(x, 3) - I.params[2] * x;
      break;
    }
    case LBM_INTERACTION_LENNARD_JONES: {  // gpu/particles.jl:179-197
      const double sd = I.params[1] / disp;
      s = -4.0 * I.params[0] / disp * (12.0 * julia_pow(sd, 12) - 6.0 * julia_pow(sd, 6));
      break;
    }
    default: {  // LBM_INTERACTION_DIPOLE, gpu/particles.jl:160-177
      // the field of this tick (Bt, when scheduled) or the constant one
      const double *B = Bt ? Bt : I.field;
      if (Bt && B[0] == 0.0 && B[1] == 0.0 && B[2] == 0.0) return false;
      const double pre = -I.params[0] / julia_pow(disp, 4);
      double ccb[3] = {0, 0, 0}, BB, s2;
      if (d == 3) {
        double cb[3];
        cross3(u, B, cb);
        cross3(cb, B, ccb);
        BB = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
        s2 = cb[0] * cb[0] + cb[1] * cb[1] + cb[2] * cb[2];
      } else if (I.params[1] != 0.0) {  // scalar field B = field[0]
        double cb[2];
        cross_sv(B[0], u, cb);
        cross_sv(B[0], cb, ccb);
        BB = B[0] * B[0];
        s2 = cb[0] * cb[0] + cb[1] * cb[1];
      } else {
        const double cb = cross2(u, B);
        cross_sv(cb, B, ccb);
        BB = B[0] * B[0] + B[1] * B[1];
        s2 = cb * cb;
      }
#pragma unroll
      for (int k = 0; k < 3; k++)
        if (k < d) fa[k] = pre * (2.0 * ccb[k] - 2.0 * BB * u[k] + 5.0 * u[k] * s2);
      return true;
    }
  }
#pragma unroll
  for (int k = 0; k < 3; k++)
    if (k < d) fa[k] = s * u[k];
  return true;
}

// forceInput of particle p after every interaction term, in order, from the
// records' positions (all before the tick's moves): F holds the momentum
// exchange on entry.
__device__ inline void add_interaction_forces(const StepParticles &sp, const DevParticle *recs, int p, double F[3]) {
#pragma clang fp contract(off)
  const int d = sp.dims;
  for (int t = 0; t < sp.ninter; t++) {
    const lbm_interaction &I = sp.inter[t];
    const int i1 = I.ids[0], i2 = I.ids[1];
    const bool polar = I.kind == LBM_INTERACTION_POLAR;
    const int i3 = polar ? I.ids[2] : -1;
    if (p != i1 && p != i2 && p != i3) continue;
    double fa[3], fb[3];
    const double *Bt = nullptr;
    if (I.kind == LBM_INTERACTION_DIPOLE && I.params[2] != 0.0) {
      // every particle moves once a tick: the move about to be applied,
      // particle 0's move number moves + 1, ends tick tick0 + moves + 1
      int64_t i = sp.tick0 + recs[0].moves + 1 - sp.bsched_first;
      i = i < 0 ? 0 : (i >= sp.bsched_count ? sp.bsched_count - 1 : i);  // the host keeps runs inside
      Bt = sp.bsched + 3 * i;
    }
    if (!interaction_term(I, recs[i1].R, recs[i2].R, polar ? recs[i3].R : nullptr, d, fa, fb, Bt)) continue;
    if (p == i1)
#pragma unroll
      for (int k = 0; k < 3; k++)
        if (k < d) F[k] = F[k] + fa[k];
    if (polar && p == i3)
#pragma unroll
      for (int k = 0; k < 3; k++)
        if (k < d) F[k] = F[k] + fb[k];
    if (p == i2)
#pragma unroll
      for (int k = 0; k < 3; k++)
        if (k < d) F[k] = F[k] - (polar ? fa[k] + fb[k] : fa[k]);
  }
}

// The move of particle P (index p) from the partials of the previous box
// kernel: tot = forceInput (0..2) and torqueInput (3..5), then
// moveParticles!, the link box and the node-velocity snapshot; with `log`,
// the state after the move also goes to the per-tick log.  Run by one thread.
__device__ inline void move_record(DevParticle &P, int p, const double tot[6], const StepParticles &sp, bool log) {
  const double F[3] = {tot[0], tot[1], sp.dims == 3 ? tot[2] : 0.0};
  const double T[3] = {tot[3], sp.dims == 3 ? tot[4] : 0.0, sp.dims == 3 ? tot[5] : 0.0};
  bool moved, find_nv;
  const double R0[3] = {P.Rg[0], P.Rg[1], P.Rg[2]};
  if (sp.semantics == LBM_SEMANTICS_GPU)
    move_gpu(P, F, T, sp.dims, sp.dt, moved, find_nv);
  else
    move_cpu(P, F, T, sp.dims, sp.dt, moved, find_nv);
  for (int k = 0; k < 3; k++) {
    P.force[k] = F[k];
    P.torque[k] = T[k];
  }
  if (moved) {
    for (int k = 0; k < 3; k++) {
      // the box kernels own a margin of kBoxMargin nodes around the link
      // box for two ticks of motion: half a node per tick at most
      if (!(fabs(P.R[k] - R0[k]) <= 0.5 * sp.dx)) P.err = 1;
      P.Rg[k] = P.R[k];
    }
    int32_t err = 0;
    link_box(P.Rg, P.radius, sp.x0, sp.dx, sp.n, sp.dims, P.lo, P.hi, &err);
    if (err) P.err = 1;
  }
  if (find_nv) {
    for (int k = 0; k < 3; k++) {
      P.nR[k] = P.R[k];
      P.nV[k] = P.V[k];
      P.nW[k] = P.W[k];
      P.ne[k] = P.e[k];
    }
  }
  P.pending = 0;
  P.moves += 1;
  if (log && sp.traj) {
    double *o = sp.traj + ((int64_t)p * sp.traj_cap + (P.moves - 1) % sp.traj_cap) * 9;
    for (int k = 0; k < 3; k++) {
      o[k] = P.R[k];
      o[3 + k] = P.V[k];
      o[6 + k] = P.W[k];
    }
  }
}

// Apply the pending moves of the records in LDS (recs[0..np)) from
// partial_in, over the whole block: each component is reduced over the
// previous kernel's nbox box blocks in a fixed order, so every block that
// does this gets the same records.
// The first partial a lane reduces in apply_pending_moves (component pk =
// wave of box block b = lane), loaded before the records are in: the
// partials' addresses depend on the kernel arguments alone.
__device__ inline DD prefetch_partial(const StepParticles &sp) {
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int pk = tid >> 6, b = __lane_id();
  if (pk >= 6 * sp.np || b >= sp.nbox) return DD{0.0, 0.0};
  const double *src = sp.partial_in + (((int64_t)b * kLocalParticles + pk / 6) * 6 + pk % 6) * 2;
  return DD{src[0], src[1]};
}

// With `log` (one block per lattice and kernel), the moved states also go to
// the per-tick log.
__device__ inline void apply_pending_moves(const StepParticles &sp, DevParticle *recs, const DD *pre = nullptr,
                                           bool log = false) {
  __shared__ double tot[kLocalParticles][6];
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int nthreads = blockDim.x * blockDim.y * blockDim.z;
  // block sizes are whole waves, so every wave reaches wave_sum with 64 lanes
  const int wave = tid >> 6, nwaves = nthreads >> 6, lane = __lane_id();
  for (int pk = wave; pk < 6 * sp.np; pk += nwaves) {
    const int p = pk / 6, k = pk % 6;
    if (!recs[p].pending) continue;  // uniform over the wave
    DD s{0.0, 0.0};
    for (int b = lane; b < sp.nbox; b += 64) {
      if (pre && pk == wave && b == lane) {
        s = dd_merge(s, *pre);
        continue;
      }
      const double *src = sp.partial_in + (((int64_t)b * kLocalParticles + p) * 6 + k) * 2;
      s = dd_merge(s, DD{src[0], src[1]});
    }
    s = wave_sum(s);
    if (lane == kWaveSumLane) {
#pragma clang fp contract(off)
      // forceInput = -coeff * sum(s) * c and torqueInput = -coeff *
      // cross(sum(s (x - R)), c) with c = c_q dx/dt (gpu/fluids.jl:194-203)
      const double v = s.hi + s.lo;
      tot[p][k] = -sp.coeff * (v * sp.ls);
    }
  }
  __syncthreads();
  if (sp.ninter > 0) {
    // molecularDynamicsTick!: every interaction acts on the positions before
    // any particle of the tick moves
    if (tid < sp.np && recs[tid].pending) add_interaction_forces(sp, recs, tid, tot[tid]);
    __syncthreads();
  }
  if (tid < sp.np && recs[tid].pending) move_record(recs[tid], tid, tot[tid], sp, log);
  __syncthreads();
}

// The move of particle p of a lattice with more than kLocalParticles
// particles, after a gathering kernel (one wave): its momentum exchange is
// the sum, over the box blocks in order, of the partials they wrote under
// its id (each lane takes the blocks lane, lane + 64, ...; the waves' sums
// then combine in a fixed order); then the interactions, from the positions
// before any particle of the tick moves, and moveParticles!.  Records from
// parts_in, moved ones to parts_out.
__device__ inline void move_particle_block(const StepParticles &sp, int p) {
  const int lane = __lane_id();
  DD s[6];
#pragma unroll
  for (int k = 0; k < 6; k++) s[k] = DD{0.0, 0.0};
  for (int b = lane; b < sp.nbox; b += 64)
    for (int j = 0; j < kLocalParticles; j++) {
      if (sp.ids_in[(int64_t)b * kLocalParticles + j] != p) continue;
      const double *src = sp.partial_in + ((int64_t)b * kLocalParticles + j) * 12;
#pragma unroll
      for (int k = 0; k < 6; k++) s[k] = dd_merge(s[k], DD{src[2 * k], src[2 * k + 1]});
    }
  __shared__ double tot[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const DD t = wave_sum(s[k]);
    if (lane == kWaveSumLane) {
#pragma clang fp contract(off)
      tot[k] = -sp.coeff * ((t.hi + t.lo) * sp.ls);
    }
  }
  __syncthreads();
  if (lane == 0) {
    double F[6];
    for (int k = 0; k < 6; k++) F[k] = tot[k];
    if (sp.ninter > 0) add_interaction_forces(sp, sp.parts_in, p, F);
    DevParticle P = sp.parts_in[p];
    move_record(P, p, F, sp, true);
    sp.parts_out[p] = P;
  }
}

}  // namespace lbm
