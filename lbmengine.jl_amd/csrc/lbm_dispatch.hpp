// lbm_dispatch.hpp — host-callable launchers for the kernels of
// liblbm_hip.so.  lbm_kernels.hip is compiled once per storage kind S
// (FP64, FP32, FP16) and defines the templated launchers for that S.
#pragma once

#include "lbm_step.hpp"

namespace lbm {

struct Launch {
  dim3 grid, block;
  hipStream_t stream;
};

struct IntegrateArgs {
  DevParticle *parts;
  const double *partial;  // [block][np][6]
  int32_t np, nblocks;
  int32_t dims, semantics;
  double dt, dx;
  double coeff;  // dx^d / dt (gpu semantics) or dx^d (cpu semantics)
  double ls;     // lattice speed dx / dt
  double x0[3];
  int32_t n[3];
};

template <int S>
hipError_t launch_step(int Q, int coll, int fm, const StepArgs &a, const Launch &l);
template <int S>
hipError_t launch_link(int Q, int coll, int fm, const LinkArgs &a, const Launch &l);
template <int S>
hipError_t launch_init_uniform(int Q, const AuxArgs &a, const Launch &l);
template <int S>
hipError_t launch_set_populations(int Q, const AuxArgs &a, const Launch &l);
template <int S>
hipError_t launch_get_populations(int Q, const AuxArgs &a, const Launch &l);
template <int S>
hipError_t launch_hydro(int Q, const AuxArgs &a, const Launch &l);

hipError_t launch_link_mask(int Q, const LinkMaskArgs &a, const Launch &l);
hipError_t launch_integrate(const IntegrateArgs &a, hipStream_t stream);

// Link box of a sphere of radius r at Rg (node indices, inclusive): every
// solid node and every fluid node with a solid neighbour lies inside.
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
    int32_t l = (int32_t)floor(a) - 2, h = (int32_t)ceil(b) + 2;
    if (!(a == a) || !(b == b) || l < 0 || h > n[k] - 1) {
      *err = 1;
      l = l < 0 ? 0 : (l > n[k] - 1 ? n[k] - 1 : l);
      h = h > n[k] - 1 ? n[k] - 1 : (h < 0 ? 0 : h);
      if (!(a == a) || !(b == b)) l = 0, h = 0;
    }
    lo[k] = l;
    hi[k] = h;
  }
}

// Largest link-box extent along one dimension for radius r.
inline int32_t link_box_extent(double r, double dx) { return (int32_t)ceil(2.0 * r / dx) + 7; }

}  // namespace lbm
