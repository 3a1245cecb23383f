// lbm_common.hpp — lattice tables, device particle record and launch
// parameters shared by the HIP kernels and the host runtime of
// liblbm_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lbm_hip.h"

namespace lbm {

// ---------------------------------------------------------------- lattices
// Velocity sets in the reference's order (gpu/misc.jl:195-272): rest, the
// axis pairs (-,+) per axis, then the comprehensions [[p,q] for p in [-1,1],
// q in [-1,1]] (and [[p,q,r] ...]) with p running fastest.  wclass groups the
// velocities by speed (0 rest, 1 axis, 2 face diagonal, 3 body diagonal);
// wc(k) is the weight of class k.
template <int Q>
struct Lattice;

template <>
struct Lattice<9> {
  static constexpr int D = 2, NC = 3;
  static constexpr int cx[9] = {0, -1, 1, 0, 0, -1, 1, -1, 1};
  static constexpr int cy[9] = {0, 0, 0, -1, 1, -1, -1, 1, 1};
  static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int opp[9] = {0, 2, 1, 4, 3, 8, 7, 6, 5};
  static constexpr int wclass[9] = {0, 1, 1, 1, 1, 2, 2, 2, 2};
  // weights 4/9, 1/9, 1/36; inverse weights exact in FP32/FP64
  __host__ __device__ static constexpr double wc(int k) { return k == 0 ? 4.0 / 9.0 : (k == 1 ? 1.0 / 9.0 : 1.0 / 36.0); }
  __host__ __device__ static constexpr double invwc(int k) { return k == 0 ? 2.25 : (k == 1 ? 9.0 : 36.0); }
  __host__ __device__ static constexpr double w(int i) { return wc(wclass[i]); }
  __host__ __device__ static constexpr double invw(int i) { return invwc(wclass[i]); }
};

template <>
struct Lattice<15> {
  static constexpr int D = 3, NC = 4;  // class 2 unused (no face diagonals)
  static constexpr int cx[15] = {0, -1, 1, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1};
  static constexpr int cy[15] = {0, 0, 0, -1, 1, 0, 0, -1, -1, 1, 1, -1, -1, 1, 1};
  static constexpr int cz[15] = {0, 0, 0, 0, 0, -1, 1, -1, -1, -1, -1, 1, 1, 1, 1};
  static constexpr int opp[15] = {0, 2, 1, 4, 3, 6, 5, 14, 13, 12, 11, 10, 9, 8, 7};
  static constexpr int wclass[15] = {0, 1, 1, 1, 1, 1, 1, 3, 3, 3, 3, 3, 3, 3, 3};
  // weights 2/9, 1/9, 1/72
  __host__ __device__ static constexpr double wc(int k) { return k == 0 ? 2.0 / 9.0 : (k == 1 ? 1.0 / 9.0 : 1.0 / 72.0); }
  __host__ __device__ static constexpr double invwc(int k) { return k == 0 ? 4.5 : (k == 1 ? 9.0 : 72.0); }
  __host__ __device__ static constexpr double w(int i) { return wc(wclass[i]); }
  __host__ __device__ static constexpr double invw(int i) { return invwc(wclass[i]); }
};

template <>
struct Lattice<19> {
  static constexpr int D = 3, NC = 3;
  static constexpr int cx[19] = {0, -1, 1, 0, 0, 0, 0, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1};
  static constexpr int cy[19] = {0, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 0, 0, 0, 0, -1, -1, 1, 1};
  static constexpr int cz[19] = {0, 0, 0, 0, 0, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 0, 0, 0, 0};
  static constexpr int opp[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
  static constexpr int wclass[19] = {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};
  // weights 1/3, 1/18, 1/36
  __host__ __device__ static constexpr double wc(int k) { return k == 0 ? 1.0 / 3.0 : (k == 1 ? 1.0 / 18.0 : 1.0 / 36.0); }
  __host__ __device__ static constexpr double invwc(int k) { return k == 0 ? 3.0 : (k == 1 ? 18.0 : 36.0); }
  __host__ __device__ static constexpr double w(int i) { return wc(wclass[i]); }
  __host__ __device__ static constexpr double invw(int i) { return invwc(wclass[i]); }
};

template <>
struct Lattice<27> {
  static constexpr int D = 3, NC = 4;
  static constexpr int cx[27] = {0, -1, 1, 0, 0, 0, 0, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1,
                                 -1, 1, -1, 1, -1, 1, -1, 1};
  static constexpr int cy[27] = {0, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 0, 0, 0, 0, -1, -1, 1, 1,
                                 -1, -1, 1, 1, -1, -1, 1, 1};
  static constexpr int cz[27] = {0, 0, 0, 0, 0, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 0, 0, 0, 0,
                                 -1, -1, -1, -1, 1, 1, 1, 1};
  static constexpr int opp[27] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15,
                                  26, 25, 24, 23, 22, 21, 20, 19};
  static constexpr int wclass[27] = {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2,
                                     3, 3, 3, 3, 3, 3, 3, 3};
  // weights 8/27, 2/27, 1/54, 1/216
  __host__ __device__ static constexpr double wc(int k) {
    return k == 0 ? 8.0 / 27.0 : (k == 1 ? 2.0 / 27.0 : (k == 2 ? 1.0 / 54.0 : 1.0 / 216.0));
  }
  __host__ __device__ static constexpr double invwc(int k) { return k == 0 ? 3.375 : (k == 1 ? 13.5 : (k == 2 ? 54.0 : 216.0)); }
  __host__ __device__ static constexpr double w(int i) { return wc(wclass[i]); }
  __host__ __device__ static constexpr double invw(int i) { return invwc(wclass[i]); }
};

// FP16 deviation storage: h_i = half((f_i - w_i rho0) * 2^14 / w_i).
constexpr float kF16Scale = 16384.0f;

// Custom 16-bit deviation storage (1 sign, 4 exponent, 11 mantissa bits; the
// FP16C format of Lehmann et al., 2022, with its range placed for deviation
// storage): h_i = (f_i - w_i rho0) / w_i with |h| < 1/8 is kept as bits 26..12
// of the FP32 pattern of h * 2^-108, whose exponent field then runs 0..15:
// 12 significant bits for 2^-18 <= |h| < 2^-3 and FP32's gradual underflow
// (absolute step 2^-29) below.  Larger magnitudes are clamped.  Wall nodes,
// whose stored values nothing reads, hold the clamped -rho0.
constexpr double kF16CDown = 0x1p-108;
constexpr double kF16CUp = 0x1p108;

// ------------------------------------------------------- device particle
// One rigid sphere.  The geometry (solid region) is the sphere at Rg; the
// node velocities are those of the state (nR, nV, nW, ne) at the last
// getNodeVelocity call (gpu/particles.jl:73-91): the reference rebuilds them
// only when the particle moved or its angular velocity changed.
struct DevParticle {
  int32_t kind, couple_forces, couple_torques, err;
  int32_t pending;   // 1: the move from the partials of the last box kernel is still to apply
  int32_t whole;     // 1: the box blocks take its whole owned region, not a band (large radii)
  int64_t moves;     // moves applied since the particle was added (the per-tick log's index)
  double radius, B1, B2, inv_mass, inv_I;
  double R[3], V[3], W[3], e[3];       // state
  double Rg[3];                        // geometry snapshot
  double nR[3], nV[3], nW[3], ne[3];   // node-velocity snapshot
  int32_t lo[3], hi[3];                // link box, node indices, inclusive
  double force[3], torque[3];          // loads consumed by the last move
};

// Walls beyond the layers of walled dimensions (the reference's `walls` and
// `movingWalls` functions, gpu/initMethods.jl:580-620): BoxGeom::walled
// then carries kWallMask, and wmask[n] (node n in Julia order) holds bit 0
// when n is a wall node, bit q (1 <= q < Q) when n is fluid and n - c_q a
// wall node (population q bounces back), and bit 31 when one of those walls
// moves (kMovingWalls; uwall holds their velocities).
constexpr int32_t kWallMask = 8;
constexpr int32_t kMovingWalls = 16;
constexpr uint32_t kMovingLinkBit = 1u << 31;

// Everything the step kernels need about one box of the lattice.
struct BoxGeom {
  int32_t nx, ny, nz;   // nodes per dimension (nz = 1 in 2D)
  int32_t walled;       // bit k: dimension k walled; kWallMask, kMovingWalls
  int64_t N;            // nx*ny*nz
  int32_t off32;        // 1: every population's byte offset in a buffer fits in 32 bits
  int32_t reserved;
  const uint32_t *wmask;  // wall bits per node (with kWallMask), else null
  const void *uwall;      // wall velocities (N, d), compute type (with kMovingWalls), else null
};

struct CollParams {
  double dt, tau, tau_minus;   // reference parameters
  double ls, invC2s, invC4s;   // lattice speed dx/dt and 1/c_s^2, 1/c_s^4
  double rho0;                 // deviation shift
  double F[3];                 // uniform Guo force
  int32_t force_mode;
  int32_t incompressible;      // 1: equilibrium w rho + w rho0 (...) (gpu/fluids.jl:51), rho0 below
  // the FP32 collisions' constants, each rounded once from the above
  // (set_float_params), so that no thread repeats the conversions
  float f_ls, f_invC2s, f_invC4s, f_dt, f_half_dt, f_wp, f_wm, f_rho0;
};

__host__ __device__ inline void set_float_params(CollParams &c) {
  c.f_ls = static_cast<float>(c.ls);
  c.f_invC2s = static_cast<float>(c.invC2s);
  c.f_invC4s = static_cast<float>(c.invC4s);
  c.f_dt = static_cast<float>(c.dt);
  c.f_half_dt = static_cast<float>(0.5 * c.dt);
  c.f_wp = static_cast<float>(c.dt / c.tau);
  c.f_wm = static_cast<float>(c.dt / c.tau_minus);
  c.f_rho0 = static_cast<float>(c.rho0);
}

// Tile of the step kernel: block of bx*by*bz threads, one node each.
struct Tiling {
  int32_t bx, by, bz;   // block shape
  int32_t tx, ty, tz;   // tiles per dimension
  int32_t ntiles;
  int32_t remap;        // 1: xcd_remap the tile order
  uint32_t mtx, mty;    // tile_div multipliers for tx and ty (0: divide)
  int32_t occ;          // 1: the gathering kernels built for half the waves per
                        // SIMD and twice the registers (no spills in the box path)
  uint32_t mbx, mby;    // tile_div multipliers for bx and by over a flat thread
                        // index (the batched kernel's one-dimensional blocks)
};

// Multiplier m with n / d == umulhi(n, m) for every n < nmax, or 0 when
// there is none of 32 bits (then tile_div divides).  m = ceil(2^32 / d)
// gives n m / 2^32 = n / d + n e / (d 2^32) with e = m d - 2^32 < d, whose
// fractional excess stays below 1 / d while n d < 2^32.
inline uint32_t div_magic(uint32_t d, uint64_t nmax) {
  if (d <= 1) return 0;
  const uint64_t m = ((uint64_t(1) << 32) + d - 1) / d;
  if (m > 0xFFFFFFFFull || nmax * d >= (uint64_t(1) << 32)) return 0;
  return (uint32_t)m;
}

// n / d for a tile or thread index: a multiply-high (scalar when n is
// uniform over the block) instead of an integer division
__device__ __forceinline__ uint32_t tile_div(uint32_t n, int32_t d, uint32_t m) {
  return m ? __umulhi(n, m) : (d == 1 ? n : n / (uint32_t)d);
}

__host__ __device__ inline int32_t xcd_remap(int32_t b, int32_t nb) {
  // Workgroups are dispatched round-robin over the 8 XCDs (each with its own
  // L2).  Give XCD k one contiguous range of tiles so neighbouring tiles,
  // which share the cache lines that shifted pushes touch, share an L2.
  const int32_t q = nb >> 3, r = nb & 7;
  const int32_t k = b & 7, slot = b >> 3;
  return k * q + (k < r ? k : r) + slot;
}

}  // namespace lbm
