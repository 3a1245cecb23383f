// Host-side check of the band that splits a gathering kernel's nodes between
// box blocks and tiles (Band, band_row, band_node in lbm_step.hpp), over
// random particles, lattice spacings and dimensions:
//  - the offset tables + band_node list each band node exactly once, and
//    only band nodes (the box blocks' enumeration equals the tiles' test);
//  - every node at distance [r - 1, r + sqrt(3) + 1) from the centre is in
//    the band (this tick's links and the nodes whose rho the next tick's
//    links read, for moves below half a node);
//  - the band never holds more nodes than the host's bound for its radius.
// Prints "ok <cases>" and exits 0, or names the first failure and exits 1.
#include <cmath>
#include <cstdio>
#include <random>
#include <set>
#include <vector>

#include "lbm_step.hpp"

using namespace lbm;

int main(int argc, char **argv) {
  const int cases = argc > 1 ? std::atoi(argv[1]) : 1000;
  std::mt19937_64 rng(1);
  std::uniform_real_distribution<double> U(0, 1);
  int checked = 0;
  for (int trial = 0; trial < cases; trial++) {
    const int dims = trial % 5 == 0 ? 2 : 3;
    const double dx = trial % 3 == 0 ? 0.5 : 1.0;
    const int n = 40 + (int)(U(rng) * 60);
    BoxGeom g{};
    g.nx = n;
    g.ny = n;
    g.nz = dims == 3 ? n : 1;
    g.N = (int64_t)g.nx * g.ny * g.nz;
    StepParticles sp{};
    sp.dims = dims;
    sp.dx = dx;
    for (int k = 0; k < 3; k++) {
      sp.x0[k] = -0.5 * n * dx + 0.25 * U(rng);
      sp.n[k] = k < dims ? n : 1;
    }
    DevParticle P{};
    P.radius = dx * (1.0 + U(rng) * 12.0);
    const double rn = P.radius / dx;
    if (n - 2 * rn - 16 < 1) continue;
    for (int k = 0; k < dims; k++) P.Rg[k] = sp.x0[k] + dx * (rn + 8 + U(rng) * (n - 2 * rn - 16));
    int32_t err = 0;
    link_box(P.Rg, P.radius, sp.x0, dx, sp.n, dims, P.lo, P.hi, &err);
    if (err) continue;
    Band B;
    band_of(P, sp, g, B);
    const int ey = B.hi[1] - B.lo[1] + 1, ez = B.hi[2] - B.lo[2] + 1;
    std::vector<uint16_t> row(ey * ez);
    std::vector<int32_t> pre(ez + 1, 0);
    for (int r = 0; r < ey * ez; r++) {
      int a0, na, b0, nb;
      band_row(B, B.lo[1] + r % ey, B.lo[2] + r / ey, a0, na, b0, nb);
      row[r] = (uint16_t)(na + nb);
    }
    for (int k = 0; k < ez; k++) {
      int s = 0;
      for (int j = 0; j < ey; j++) {
        const int c = row[k * ey + j];
        row[k * ey + j] = (uint16_t)s;
        s += c;
      }
      pre[k + 1] = s;
    }
    for (int k = 1; k <= ez; k++) pre[k] += pre[k - 1];
    std::set<long> listed, tested;
    for (int gi = 0; gi < pre[ez]; gi++) {
      int x, y, z;
      band_node(B, pre.data(), row.data(), ey, ez, gi, x, y, z);
      if (x < B.lo[0] || x > B.hi[0] || !in_band(B, x, y, z)) {
        std::printf("case %d: listed node (%d,%d,%d) is not in the band\n", trial, x, y, z);
        return 1;
      }
      listed.insert(x + 4096L * (y + 4096L * z));
    }
    if ((int)listed.size() != pre[ez]) {
      std::printf("case %d: a band node is listed twice\n", trial);
      return 1;
    }
    for (int z = B.lo[2]; z <= B.hi[2]; z++)
      for (int y = B.lo[1]; y <= B.hi[1]; y++)
        for (int x = B.lo[0]; x <= B.hi[0]; x++) {
          const bool in = in_band(B, x, y, z);
          if (in) tested.insert(x + 4096L * (y + 4096L * z));
          const int c[3] = {x, y, z};
          double d2 = 0.0;
          for (int k = 0; k < dims; k++) {
            const double u = (sp.x0[k] + c[k] * dx - P.Rg[k]) / dx;
            d2 += u * u;
          }
          const double d = std::sqrt(d2);
          if (d >= rn - 1.0 && d < rn + std::sqrt(3.0) + 1.0 && !in) {
            std::printf("case %d: node at distance %.4f (r = %.4f) is outside the band\n", trial, d, rn);
            return 1;
          }
        }
    if (tested != listed) {
      std::printf("case %d: %zu nodes listed, %zu in the band\n", trial, listed.size(), tested.size());
      return 1;
    }
    const double ri = std::max(rn - kBandIn - 1.0, 0.0), ro = rn + kBandOut + 1.0;
    const double v = dims == 3 ? 4.0 / 3.0 * M_PI * (ro * ro * ro - ri * ri * ri) : M_PI * (ro * ro - ri * ri);
    if ((double)pre[ez] > std::ceil(v) + 1) {
      std::printf("case %d: %d band nodes exceed the bound %.0f\n", trial, pre[ez], std::ceil(v) + 1);
      return 1;
    }
    checked++;
  }
  std::printf("ok %d\n", checked);
  return 0;
}
