// Host-side check of link_box (lbm_integrate.hpp) and band_row (lbm_step.hpp)
// at any particle centre, as a diverging run may produce one (huge, infinite
// or NaN): the bounds always lie inside the lattice / the owned row, and for
// the centres of ordinary runs they equal the integer formulas they replace
// (restated below as the expected values), node for node.
// Prints "ok <cases>" and exits 0, or names the first failure and exits 1.
#include <cmath>
#include <cstdio>
#include <limits>
#include <random>

#include "lbm_step.hpp"

using namespace lbm;

// the integer formulas, valid while the centre stays within 2^30 nodes
static void link_box_int(const double Rg[3], double r, const double x0[3], double dx, const int32_t n[3], int dims,
                         int32_t lo[3], int32_t hi[3], int32_t *err) {
  for (int k = 0; k < 3; k++) {
    if (k >= dims) {
      lo[k] = hi[k] = 0;
      continue;
    }
    const double a = (Rg[k] - r - x0[k]) / dx, b = (Rg[k] + r - x0[k]) / dx;
    int32_t l = (int32_t)floor(a) - 2, h = (int32_t)ceil(b) + 2;
    if (!(a == a) || !(b == b) || l < 0 || h > n[k] - 1) {
      *err = 1;
      l = l < 0 ? 0 : (l > n[k] - 1 ? n[k] - 1 : l);
      h = h > n[k] - 1 ? n[k] - 1 : (h < 0 ? 0 : h);
    }
    lo[k] = l;
    hi[k] = h;
  }
}

static void band_row_int(const Band &b, int y, int z, int &a0, int &na, int &b0, int &nb) {
  a0 = b0 = 0;
  na = nb = 0;
  if (y < b.lo[1] || y > b.hi[1] || z < b.lo[2] || z > b.hi[2]) return;
  const double dy = y - b.c[1], dz = z - b.c[2];
  const double rho2 = dy * dy + dz * dz;
  if (rho2 > b.rout2) return;
  const double ao = sqrt(b.rout2 - rho2);
  int lo = (int)ceil(b.c[0] - ao), hi = (int)floor(b.c[0] + ao);
  lo = lo < b.lo[0] ? b.lo[0] : lo;
  hi = hi > b.hi[0] ? b.hi[0] : hi;
  if (rho2 < b.rin2) {
    const double ai = sqrt(b.rin2 - rho2);
    const int l1 = (int)floor(b.c[0] - ai), h0 = (int)ceil(b.c[0] + ai);
    a0 = lo;
    na = (l1 < hi ? l1 : hi) - lo + 1;
    b0 = h0 > lo ? h0 : lo;
    nb = hi - b0 + 1;
  } else {
    a0 = lo;
    na = hi - lo + 1;
  }
  na = na > 0 ? na : 0;
  nb = nb > 0 ? nb : 0;
}

static int fail(const char *what, int trial) {
  std::printf("FAIL %s (case %d)\n", what, trial);
  return 1;
}

int main() {
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(0, 1);
  const double inf = std::numeric_limits<double>::infinity(), nan = std::numeric_limits<double>::quiet_NaN();
  const double wild[] = {1e300, -1e300, inf, -inf, nan, 3e9, -3e9, 2.2e9, -2.2e9, 1e12};
  int cases = 0;
  for (int trial = 0; trial < 20000; trial++) {
    const int dims = trial % 5 == 0 ? 2 : 3;
    const double dx = trial % 3 == 0 ? 0.5 : 1.0;
    const int n = 24 + (int)(U(rng) * 80);
    BoxGeom g{};
    g.nx = n;
    g.ny = n;
    g.nz = dims == 3 ? n : 1;
    StepParticles sp{};
    sp.dims = dims;
    sp.dx = dx;
    int32_t nn[3];
    for (int k = 0; k < 3; k++) {
      sp.x0[k] = -0.5 * n * dx + 0.25 * U(rng);
      nn[k] = sp.n[k] = k < dims ? n : 1;
    }
    DevParticle P{};
    P.radius = dx * (1.0 + U(rng) * 8.0);
    // ordinary centres (some near or past a face), then wild ones
    const bool ordinary = trial % 4 != 0;
    for (int k = 0; k < dims; k++)
      P.Rg[k] = ordinary ? sp.x0[k] + dx * (U(rng) * 1.4 - 0.2) * n : wild[(trial / 4 + 3 * k) % 10];
    int32_t lo[3], hi[3], lo2[3], hi2[3], err = 0, err2 = 0;
    link_box(P.Rg, P.radius, sp.x0, dx, nn, dims, lo, hi, &err);
    for (int k = 0; k < dims; k++)
      if (lo[k] < 0 || hi[k] > nn[k] - 1 || lo[k] > hi[k]) return fail("link box outside the lattice", trial);
    if (ordinary) {
      link_box_int(P.Rg, P.radius, sp.x0, dx, nn, dims, lo2, hi2, &err2);
      for (int k = 0; k < 3; k++)
        if (lo[k] != lo2[k] || hi[k] != hi2[k]) return fail("link box differs from the integer formula", trial);
      if (err != err2) return fail("link box error flag differs", trial);
    } else if (!err) {
      return fail("a wild centre raised no error", trial);
    }
    for (int k = 0; k < 3; k++) {
      P.lo[k] = lo[k];
      P.hi[k] = hi[k];
    }
    Band B;
    band_of(P, sp, g, B);
    for (int z = B.lo[2]; z <= B.hi[2]; z++)
      for (int y = B.lo[1]; y <= B.hi[1]; y++) {
        int a0, na, b0, nb;
        band_row(B, y, z, a0, na, b0, nb);
        if (na < 0 || nb < 0 || na + nb > B.hi[0] - B.lo[0] + 1) return fail("band row count", trial);
        if (na && (a0 < B.lo[0] || a0 + na - 1 > B.hi[0])) return fail("band row run outside the row", trial);
        if (nb && (b0 < B.lo[0] || b0 + nb - 1 > B.hi[0])) return fail("band row run outside the row", trial);
        if (ordinary) {
          int a1, n1, c1, m1;
          band_row_int(B, y, z, a1, n1, c1, m1);
          if (n1 != na || m1 != nb || (na && a1 != a0) || (nb && c1 != b0))
            return fail("band row differs from the integer formula", trial);
        }
      }
    cases++;
  }
  std::printf("ok %d\n", cases);
  return 0;
}
