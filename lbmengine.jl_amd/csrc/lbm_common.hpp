// lbm_common.hpp — lattice tables, device particle record and launch
// parameters shared by the HIP kernels and the host runtime of
// liblbm_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lbm_hip.h"

namespace lbm {

// ---------------------------------------------------------------- lattices
// Velocity sets in the reference's order (gpu/misc.jl:195-210,230-248):
// rest, the axis pairs (-,+) per axis, then the comprehensions
// [[p,q] for p in [-1,1], q in [-1,1]] with p running fastest.
template <int Q>
struct Lattice;

template <>
struct Lattice<9> {
  static constexpr int D = 2;
  static constexpr int cx[9] = {0, -1, 1, 0, 0, -1, 1, -1, 1};
  static constexpr int cy[9] = {0, 0, 0, -1, 1, -1, -1, 1, 1};
  static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int opp[9] = {0, 2, 1, 4, 3, 8, 7, 6, 5};
  // inverse weights (exact in FP32/FP64): 9/4, 9, 36
  static constexpr int wclass[9] = {0, 1, 1, 1, 1, 2, 2, 2, 2};
  __host__ __device__ static constexpr double w(int i) {
    return wclass[i] == 0 ? 4.0 / 9.0 : (wclass[i] == 1 ? 1.0 / 9.0 : 1.0 / 36.0);
  }
  __host__ __device__ static constexpr double invw(int i) {
    return wclass[i] == 0 ? 2.25 : (wclass[i] == 1 ? 9.0 : 36.0);
  }
};

template <>
struct Lattice<19> {
  static constexpr int D = 3;
  static constexpr int cx[19] = {0, -1, 1, 0, 0, 0, 0, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1};
  static constexpr int cy[19] = {0, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 0, 0, 0, 0, -1, -1, 1, 1};
  static constexpr int cz[19] = {0, 0, 0, 0, 0, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 0, 0, 0, 0};
  static constexpr int opp[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
  static constexpr int wclass[19] = {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};
  __host__ __device__ static constexpr double w(int i) {
    return wclass[i] == 0 ? 1.0 / 3.0 : (wclass[i] == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
  }
  __host__ __device__ static constexpr double invw(int i) {
    return wclass[i] == 0 ? 3.0 : (wclass[i] == 1 ? 18.0 : 36.0);
  }
};

// FP16 deviation storage: h_i = half((f_i - w_i rho0) * 2^14 / w_i).
constexpr float kF16Scale = 16384.0f;

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

// Everything the step kernels need about one box of the lattice.
struct BoxGeom {
  int32_t nx, ny, nz;   // nodes per dimension (nz = 1 in 2D)
  int32_t walled;       // bit k: dimension k walled
  int64_t N;            // nx*ny*nz
};

struct CollParams {
  double dt, tau, tau_minus;   // reference parameters
  double ls, invC2s, invC4s;   // lattice speed dx/dt and 1/c_s^2, 1/c_s^4
  double rho0;                 // deviation shift
  double F[3];                 // uniform Guo force
  int32_t force_mode;
};

// Tile of the step kernel: block of bx*by*bz threads, one node each.
struct Tiling {
  int32_t bx, by, bz;   // block shape
  int32_t tx, ty, tz;   // tiles per dimension
  int32_t ntiles;
  int32_t remap;        // 1: xcd_remap the tile order
};

__host__ __device__ inline int32_t xcd_remap(int32_t b, int32_t nb) {
  // Workgroups are dispatched round-robin over the 8 XCDs (each with its own
  // L2).  Give XCD k one contiguous range of tiles so neighbouring tiles,
  // which share the cache lines that shifted pushes touch, share an L2.
  const int32_t q = nb >> 3, r = nb & 7;
  const int32_t k = b & 7, slot = b >> 3;
  return k * q + (k < r ? k : r) + slot;
}

}  // namespace lbm
