// lbm_capi.cpp — host runtime of liblbm_hip.so: the extern "C" entry points
// declared in include/lbm_hip.h.  It owns the device state of one lattice
// (ping-pong population buffers, coordinate tables, particle records,
// momentum-exchange partials), launches the step on its own stream, and
// optionally replays the step loop from a captured hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lbm_adjoint.hpp"
#include "lbm_adjoint_move.hpp"
#include "lbm_dispatch.hpp"

using namespace lbm;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return fail(LBM_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(call, what)                  \
  do {                                       \
    hipError_t e_ = (call);                  \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

size_t elem_size(int storage) {
  return storage == LBM_STORAGE_F64 ? 8 : (storage == LBM_STORAGE_F32 ? 4 : 2);  // F16, F16C: 2
}

}  // namespace

struct lbm_handle {
  lbm_config cfg{};
  int dims = 0, Q = 0;
  int32_t n[3] = {1, 1, 1};
  int64_t N = 0;
  double x0[3] = {0, 0, 0};
  std::vector<double> coords_h[3];
  double *coords_d[3] = {nullptr, nullptr, nullptr};
  void *f[2] = {nullptr, nullptr};
  int cur = 0;   // buffer holding the populations (the reference's distributions)
  int pcur = 0;  // particle records of the current tick
  void *rho_side = nullptr;  // rho of the last collision at link-box nodes
  void *linkval = nullptr;   // link-slot values of the box blocks
  size_t linkval_cap = 0;    // bytes
  void *F = nullptr;  // force field (double or float)
  // walls beyond the walled layers (lbm_set_walls, lbm_set_wall_velocity)
  std::vector<uint8_t> wall_h;     // the `walls` region (N), empty if none
  std::vector<double> uwall_h;     // solidNodeVelocity (N, d), empty if none
  uint32_t *wmask_d = nullptr;     // wall bits per node (BoxGeom::wmask)
  unsigned long long *scan_d = nullptr;  // lbm_check_populations' two counters
  void *uwall_d = nullptr;         // wall velocities, compute type
  DevParticle *parts_d = nullptr;
  std::vector<DevParticle> parts_h;  // parameters as added
  double *partial = nullptr;     // 2 parities x [box block][local particle][6][hi, lo]
  size_t partial_cap = 0;        // doubles per parity
  // more than kLocalParticles particles: the ids the partials are written
  // under (2 parities x [box block][local particle]) and the flag word the
  // box blocks raise when too many particles reach one of them
  int32_t *ids = nullptr;
  size_t ids_cap = 0;            // ints per parity
  int32_t *flags_d = nullptr;
  // box blocks leading the step kernel's grid (see StepParticles)
  int32_t first_block[LBM_MAX_PARTICLES + 1] = {0};
  Tiling tiling{};
  hipStream_t stream = nullptr;
  int device = 0;  // the device current at lbm_create (the stream's)
  int64_t tick = 0;
  double time = 0.0;
  bool populated = false;
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev;
  hipEvent_t run_start = nullptr, run_stop = nullptr;
  int64_t timed_steps = 0;
  double step_kernel_ms = 0.0, step_ms = 0.0;
  double adj_forward_ms = 0.0, adj_reverse_ms = 0.0;  // last lbm_adjoint_run
  // particle-particle interaction terms (device copy of inter_h)
  std::vector<lbm_interaction> inter_h;
  lbm_interaction *inter_d = nullptr;
  // per-tick particle log: [LBM_MAX_PARTICLES][traj_cap][9]
  double *traj_d = nullptr;
  int64_t traj_cap = 0;
  int64_t traj_from = 0;  // tick at which the log was turned on
  // scheduled magnetic field (lbm_set_field_schedule): [bsched_count][3]
  double *bsched_d = nullptr;
  int64_t bsched_cap = 0, bsched_first = 0, bsched_count = 0;
  int64_t tick0 = 0;  // tick at which particle 0 was added
  // captured blocks of graph_k ticks, one per starting ping-pong buffer
  // graphs of graph_k fused kernels, per starting population and particle parity
  hipGraphExec_t graph_exec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  int64_t graph_k = 32;
};

namespace {

BoxGeom geom(const lbm_handle *h) {
  BoxGeom g;
  g.nx = h->n[0];
  g.ny = h->n[1];
  g.nz = h->n[2];
  g.walled = h->cfg.walled;
  g.N = h->N;
  g.off32 = (uint64_t)h->N * (uint64_t)h->Q * elem_size(h->cfg.storage) <= 0xFFFFFFFFull;
  g.reserved = 0;
  g.wmask = h->wmask_d;
  g.uwall = h->uwall_d;
  if (h->wmask_d) g.walled |= kWallMask;
  if (h->uwall_d) g.walled |= kMovingWalls;
  return g;
}

CollParams coll(const lbm_handle *h) {
  CollParams c;
  c.dt = h->cfg.dt;
  c.tau = h->cfg.tau;
  c.tau_minus = h->cfg.collision == LBM_COLLISION_TRT ? h->cfg.tau_minus : h->cfg.tau;
  c.ls = h->cfg.dx / h->cfg.dt;
  c.invC2s = 3.0 / (c.ls * c.ls);
  c.invC4s = 9.0 / (c.ls * c.ls * c.ls * c.ls);
  c.rho0 = h->cfg.rho0;
  for (int k = 0; k < 3; k++) c.F[k] = h->cfg.force == LBM_FORCE_UNIFORM ? h->cfg.force_uniform[k] : 0.0;
  c.force_mode = h->cfg.force;
  c.incompressible = (h->cfg.flags & LBM_FLAG_INCOMPRESSIBLE) != 0;
  set_float_params(c);
  return c;
}

Tiling make_tiling(const lbm_handle *h, int bx, int by, int bz) {
  Tiling t;
  t.bx = bx;
  t.by = by;
  t.bz = bz;
  t.tx = (h->n[0] + bx - 1) / bx;
  t.ty = (h->n[1] + by - 1) / by;
  t.tz = (h->n[2] + bz - 1) / bz;
  t.ntiles = t.tx * t.ty * t.tz;
  t.mtx = div_magic((uint32_t)t.tx, (uint64_t)t.ntiles);
  t.mty = div_magic((uint32_t)t.ty, (uint64_t)t.ntiles / (uint64_t)t.tx + 1);
  t.remap = 1;
  t.occ = 0;
  // flat thread index t < bx by bz: t / bx, then (t / bx) / by
  t.mbx = div_magic((uint32_t)bx, (uint64_t)bx * by * bz);
  t.mby = div_magic((uint32_t)by, (uint64_t)by * bz);
  return t;
}

Tiling default_tiling(const lbm_handle *h) {
  // A block spans whole x rows where rows are short, so the x-shifted pushes
  // of a row land in cache lines the same block writes.
  // Blocks are whole waves (the particle reductions rely on it).
  const int bx = h->n[0] <= 256 ? h->n[0] : 256;
  for (int by = std::min(512 / bx, h->n[1]); by >= 1; by--)
    if ((bx * by) % 64 == 0) return make_tiling(h, bx, by, 1);
  return make_tiling(h, 64, 1, 1);
}

// Particle records of ping-pong parity b (double-buffered: the integrator of
// a tick writes the other parity while the step kernel may still read this one).
DevParticle *parts_of(lbm_handle *h, int b) { return h->parts_d + b * LBM_MAX_PARTICLES; }

// Nodes per axis the box blocks may own for a particle: the largest link
// box its radius allows, grown by kBoxMargin on both sides.
int32_t owned_extent(const lbm_handle *h, const DevParticle &p, int k) {
  return std::min(link_box_extent(p.radius, h->cfg.dx) + 2 * kBoxMargin, h->n[k]);
}

// Owned regions wider than this (3D) give their box blocks the whole region
// rather than the band, whose row table would outgrow the blocks' LDS.
constexpr int32_t kBandMaxExtent = 96;

// Most nodes a particle's box blocks take: its band holds at most as many
// nodes as the shell one node wider on either side has volume (each node's
// unit cell lies inside it), and never more than its owned region.
int64_t box_nodes_bound(const lbm_handle *h, const DevParticle &p) {
  int64_t cube = 1;
  for (int k = 0; k < h->dims; k++) cube *= owned_extent(h, p, k);
  if (p.whole) return cube;
  const double r = p.radius / h->cfg.dx;
  const double ri = std::max(r - kBandIn - 1.0, 0.0), ro = r + kBandOut + 1.0;
  const double v = h->dims == 3 ? 4.0 / 3.0 * M_PI * (ro * ro * ro - ri * ri * ri) : M_PI * (ro * ro - ri * ri);
  return std::min(cube, (int64_t)std::ceil(v) + 1);
}

// Box blocks for blocks of `threads` threads: enough per particle to cover
// the most nodes it may take; partials for every box block.
int plan_boxes(lbm_handle *h, int64_t threads) {
  int32_t blocks = 0;
  for (size_t p = 0; p < h->parts_h.size(); p++) {
    h->first_block[p] = blocks;
    const int64_t vol = box_nodes_bound(h, h->parts_h[p]);
    blocks += (int32_t)((vol + threads - 1) / threads);
  }
  for (size_t p = h->parts_h.size(); p <= LBM_MAX_PARTICLES; p++) h->first_block[p] = blocks;
  const size_t rsize = h->cfg.storage == LBM_STORAGE_F64 ? 8 : 4;
  if (!h->parts_h.empty() && !h->rho_side) HIP_TRY(hipMalloc(&h->rho_side, (size_t)h->N * rsize), "hipMalloc rho");
  const size_t lv = (size_t)std::max(blocks, 1) * threads * (h->Q - 1) * rsize;
  if (lv > h->linkval_cap) {
    HIP_TRY(hipStreamSynchronize(h->stream), "sync");
    if (h->linkval) hipFree(h->linkval);
    h->linkval = nullptr;
    h->linkval_cap = 0;
    HIP_TRY(hipMalloc(&h->linkval, lv), "hipMalloc link values");
    h->linkval_cap = lv;
  }
  if (h->parts_h.size() > (size_t)kLocalParticles) {
    const size_t nid = (size_t)std::max(blocks, 1) * kLocalParticles;
    if (nid > h->ids_cap) {
      HIP_TRY(hipStreamSynchronize(h->stream), "sync");
      if (h->ids) hipFree(h->ids);
      h->ids = nullptr;
      h->ids_cap = 0;
      HIP_TRY(hipMalloc(&h->ids, 2 * nid * sizeof(int32_t)), "hipMalloc partial ids");
      h->ids_cap = nid;
    }
    if (!h->flags_d) {
      HIP_TRY(hipMalloc(&h->flags_d, 256), "hipMalloc flags");
      HIP_TRY(hipMemset(h->flags_d, 0, 256), "clear flags");
    }
  }
  // (between runs no move is pending, so the partials may be reallocated)
  const size_t need = (size_t)std::max(blocks, 1) * kLocalParticles * 6 * 2;
  if (need > h->partial_cap) {
    HIP_TRY(hipStreamSynchronize(h->stream), "sync");
    if (h->partial) hipFree(h->partial);
    h->partial = nullptr;
    h->partial_cap = 0;
    HIP_TRY(hipMalloc(&h->partial, 2 * need * sizeof(double)), "hipMalloc partials");
    h->partial_cap = need;
  }
  return LBM_OK;
}

int64_t tile_threads(const Tiling &t) { return (int64_t)t.bx * t.by * t.bz; }

// Doubles per axis of a box block's coordinate stage: the largest owned
// extent plus one node on each side.
int32_t coord_stride(const lbm_handle *h) {
  int32_t most = 1;
  for (const DevParticle &p : h->parts_h)
    for (int k = 0; k < h->dims; k++)
      most = std::max(most, owned_extent(h, p, k) + 2);
  return most;
}

StepArgs step_args(lbm_handle *h, const void *fin, void *fout, int pin) {
  StepArgs a;
  std::memset(&a, 0, sizeof a);  // (batches compare argument blocks bytewise)
  a.fin = fin;
  a.fout = fout;
  a.Ffield = h->F;
  a.rho_side = h->rho_side;
  a.g = geom(h);
  a.c = coll(h);
  a.t = h->tiling;
  StepParticles &sp = a.sp;
  sp.parts_in = parts_of(h, pin);
  sp.parts_out = parts_of(h, pin ^ 1);
  for (int k = 0; k < 3; k++) sp.coords[k] = h->coords_d[k];
  sp.partial_in = h->partial + (size_t)pin * h->partial_cap;
  sp.partial_out = h->partial + (size_t)(pin ^ 1) * h->partial_cap;
  sp.many = h->parts_h.size() > (size_t)kLocalParticles;
  if (sp.many) {
    sp.ids_in = h->ids + (size_t)pin * h->ids_cap;
    sp.ids_out = h->ids + (size_t)(pin ^ 1) * h->ids_cap;
    sp.flags = h->flags_d;
  }
  sp.linkval = h->linkval;
  sp.inter = h->inter_d;
  sp.ninter = (int32_t)h->inter_h.size();
  sp.traj = h->traj_d;
  sp.traj_cap = h->traj_cap;
  sp.bsched = h->bsched_d;
  sp.bsched_first = h->bsched_first;
  sp.bsched_count = h->bsched_count;
  sp.tick0 = h->tick0;
  sp.np = (int32_t)h->parts_h.size();
  sp.semantics = h->cfg.semantics;
  sp.dims = h->dims;
  for (int p = 0; p <= LBM_MAX_PARTICLES; p++) sp.first_block[p] = h->first_block[p];
  sp.nbox = h->first_block[LBM_MAX_PARTICLES];
  sp.nbox_pad = (sp.nbox + 7) / 8 * 8;
  sp.coord_stride = coord_stride(h);
  double coeff = 1.0;
  for (int k = 0; k < h->dims; k++) coeff *= h->cfg.dx;
  if (h->cfg.semantics == LBM_SEMANTICS_GPU) coeff = coeff / h->cfg.dt;
  sp.coeff = coeff;
  sp.ls = h->cfg.dx / h->cfg.dt;
  sp.cutoff = 2.0 * h->cfg.dx;
  // a link's fluid end lies within |c| of the sphere: sqrt(2) dx on D2Q9 and
  // D3Q19, sqrt(3) dx on the sets with body diagonals
  sp.shell = (h->Q == 15 || h->Q == 27 ? 1.75 : 1.5) * h->cfg.dx;
  sp.dt = h->cfg.dt;
  sp.dx = h->cfg.dx;
  for (int k = 0; k < 3; k++) {
    sp.x0[k] = h->x0[k];
    sp.n[k] = h->n[k];
  }
  return a;
}

// Dynamic LDS of a box block: the coordinate stage, 2t link slots (f*,
// value, key), t rho and node words, the plane offsets and the row offsets
// of the largest band region.
size_t box_shmem(const lbm_handle *h, const StepArgs &a) {
  const size_t rsize = h->cfg.storage == LBM_STORAGE_F64 ? 8 : 4;
  const size_t t = (size_t)tile_threads(a.t);
  const size_t cs = (size_t)a.sp.coord_stride;
  size_t rows = 0;
  for (const DevParticle &p : h->parts_h)
    if (!p.whole) {
      size_t r = 1;
      for (int k = 1; k < h->dims; k++) r *= (size_t)owned_extent(h, p, k);
      rows = std::max(rows, r);
    }
  return 3 * cs * 8 + 5 * t * rsize + 2 * t * 4 + t * 4 + (cs + 1) * 4 + rows * 2;
}

// Launch shape of a step kernel of the given mode: in the gathering modes the
// box blocks come first, then the tiles; the box blocks' link lists take
// dynamic LDS.
Launch step_launch(lbm_handle *h, const StepArgs &a, int mode) {
  const int nbox = mode == kCollide ? 0 : a.sp.nbox_pad;
  Launch l{dim3((unsigned)(nbox + h->tiling.ntiles)), dim3(h->tiling.bx, h->tiling.by, h->tiling.bz), h->stream};
  if (nbox) l.shmem = box_shmem(h, a);
  return l;
}

hipError_t launch_args(lbm_handle *h, int mode, const StepArgs &a) {
  Launch l = step_launch(h, a, mode);
  switch (h->cfg.storage) {
    case LBM_STORAGE_F64: return launch_step<kF64>(h->Q, h->cfg.collision, h->cfg.force, mode, a, l);
    case LBM_STORAGE_F32: return launch_step<kF32>(h->Q, h->cfg.collision, h->cfg.force, mode, a, l);
    case LBM_STORAGE_F16C: return launch_step<kF16C>(h->Q, h->cfg.collision, h->cfg.force, mode, a, l);
    default: return launch_step<kF16>(h->Q, h->cfg.collision, h->cfg.force, mode, a, l);
  }
}

hipError_t launch_mode(lbm_handle *h, int mode, const void *fin, void *fout, int pin) {
  return launch_args(h, mode, step_args(h, fin, fout, pin));
}

// The move kernel after a gathering kernel that read records `pin` (more
// than kLocalParticles particles): it reads that kernel's partials and ids
// (parity pin ^ 1) and records pin, and writes the moved records to pin ^ 1.
hipError_t launch_moves(lbm_handle *h, int pin) {
  StepArgs m = step_args(h, nullptr, nullptr, pin ^ 1);
  m.sp.parts_in = parts_of(h, pin);
  m.sp.parts_out = parts_of(h, pin ^ 1);
  return launch_move(m.sp, h->stream);
}

bool many_particles(const lbm_handle *h) { return h->parts_h.size() > (size_t)kLocalParticles; }

// One fused kernel (gather tick k, collide tick k + 1) from buffer `in`
// with particle records `pin` (then, with many particles, the move kernel);
// the timing events bracket the step kernel.
hipError_t fused_once(lbm_handle *h, int in, int pin, hipEvent_t e0, hipEvent_t e1) {
  if (e0) hipEventRecord(e0, h->stream);
  hipError_t e = launch_mode(h, kFused, h->f[in], h->f[in ^ 1], pin);
  if (e != hipSuccess) return e;
  if (e1) hipEventRecord(e1, h->stream);
  if (many_particles(h)) e = launch_moves(h, pin);
  return e;
}

void drop_graph(lbm_handle *h) {
  for (int b = 0; b < 2; b++)
    for (int c = 0; c < 2; c++) {
      if (h->graph_exec[b][c]) hipGraphExecDestroy(h->graph_exec[b][c]);
      h->graph_exec[b][c] = nullptr;
    }
}

// Capture k fused kernels starting from buffer `start` and particle records
// `pstart` into an executable graph.  The kernels take every pointer by value
// and read the particle state from device memory, so the graph replays
// correctly as long as the handle's buffers, particles and block shape stay
// as they are (drop_graph otherwise).
int capture_graph(lbm_handle *h, int start, int pstart, int64_t k, hipGraphExec_t *out) {
  HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal), "begin capture");
  hipError_t e = hipSuccess;
  for (int64_t i = 0; i < k && e == hipSuccess; i++)
    e = fused_once(h, start ^ (int)(i & 1), pstart ^ (int)(i & 1), nullptr, nullptr);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(h->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) hipGraphDestroy(g);
    return hip_fail(e != hipSuccess ? e : e2, "capture step graph");
  }
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail(e, "instantiate step graph");
  return LBM_OK;
}

int aux_launch(lbm_handle *h, int which, AuxArgs &a) {
  a.g = geom(h);
  a.c = coll(h);
  a.Ffield = h->F;
  const int threads = 256;
  Launch l{dim3((unsigned)((h->N + threads - 1) / threads)), dim3(threads), h->stream};
  hipError_t e = hipErrorInvalidValue;
#define LBM_AUX(SK)                                                     \
  switch (which) {                                                      \
    case 0: e = launch_init_uniform<SK>(h->Q, a, l); break;             \
    case 1: e = launch_set_populations<SK>(h->Q, a, l); break;          \
    case 2: e = launch_get_populations<SK>(h->Q, a, l); break;          \
    case 3: e = launch_hydro<SK>(h->Q, a, l); break;                    \
    case 4: e = launch_stress<SK>(h->Q, a, l); break;                   \
  }
  switch (h->cfg.storage) {
    case LBM_STORAGE_F64: LBM_AUX(kF64); break;
    case LBM_STORAGE_F32: LBM_AUX(kF32); break;
    case LBM_STORAGE_F16C: LBM_AUX(kF16C); break;
    default: LBM_AUX(kF16); break;
  }
#undef LBM_AUX
  if (e != hipSuccess) return hip_fail(e, "aux kernel launch");
  return LBM_OK;
}

// Whether node n is a wall node: a layer of a walled dimension, a node of
// the `walls` region or a moving wall node (non-zero solidNodeVelocity).
bool host_wall(const lbm_handle *h, int64_t n) {
  const int64_t id[3] = {n % h->n[0], (n / h->n[0]) % h->n[1], n / ((int64_t)h->n[0] * h->n[1])};
  for (int k = 0; k < h->dims; k++)
    if (((h->cfg.walled >> k) & 1) && (id[k] == 0 || id[k] == h->n[k] - 1)) return true;
  if (!h->wall_h.empty() && h->wall_h[n]) return true;
  if (!h->uwall_h.empty())
    for (int k = 0; k < h->dims; k++)
      if (h->uwall_h[(size_t)k * h->N + n] != 0.0) return true;
  return false;
}

// BoxGeom::wmask of the handle's wall region: bounceBackPrep's links
// (gpu/aux.jl:187-198: fluid node x and wall node x - c_q) and the moving
// walls among them.
template <int Q>
void wall_mask_of(const lbm_handle *h, std::vector<uint32_t> &m) {
  using L = Lattice<Q>;
  const int64_t N = h->N;
  std::vector<uint8_t> wall(N), moving(N, 0);
  for (int64_t n = 0; n < N; n++) {
    wall[n] = host_wall(h, n);
    if (!h->uwall_h.empty())
      for (int k = 0; k < h->dims; k++)
        if (h->uwall_h[(size_t)k * N + n] != 0.0) moving[n] = 1;
  }
  m.assign(N, 0u);
  const int nx = h->n[0], ny = h->n[1], nz = h->n[2];
  for (int64_t n = 0; n < N; n++) {
    if (wall[n]) {
      m[n] = 1u;
      continue;
    }
    const int x = n % nx, y = (n / nx) % ny, z = n / ((int64_t)nx * ny);
    uint32_t b = 0;
    for (int q = 1; q < Q; q++) {
      const int xs = (x - L::cx[q] + nx) % nx, ys = (y - L::cy[q] + ny) % ny, zs = (z - L::cz[q] + nz) % nz;
      const int64_t s = xs + (int64_t)nx * (ys + (int64_t)ny * zs);
      if (wall[s]) {
        b |= 1u << q;
        if (moving[s]) b |= kMovingLinkBit;
      }
    }
    m[n] = b;
  }
}

// Rebuild the device wall bits and wall velocities from wall_h / uwall_h.
int upload_walls(lbm_handle *h) {
  std::vector<uint32_t> m;
  switch (h->Q) {
    case 9: wall_mask_of<9>(h, m); break;
    case 15: wall_mask_of<15>(h, m); break;
    case 19: wall_mask_of<19>(h, m); break;
    default: wall_mask_of<27>(h, m); break;
  }
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  if (!h->wmask_d) HIP_TRY(hipMalloc(&h->wmask_d, (size_t)h->N * sizeof(uint32_t)), "hipMalloc wall mask");
  HIP_TRY(hipMemcpy(h->wmask_d, m.data(), m.size() * sizeof(uint32_t), hipMemcpyHostToDevice), "copy wall mask");
  if (!h->uwall_h.empty()) {
    const size_t cnt = (size_t)h->N * h->dims;
    if (h->cfg.storage == LBM_STORAGE_F64) {
      if (!h->uwall_d) HIP_TRY(hipMalloc(&h->uwall_d, cnt * sizeof(double)), "hipMalloc wall velocity");
      HIP_TRY(hipMemcpy(h->uwall_d, h->uwall_h.data(), cnt * sizeof(double), hipMemcpyHostToDevice), "copy wall velocity");
    } else {
      std::vector<float> uf(h->uwall_h.begin(), h->uwall_h.end());
      if (!h->uwall_d) HIP_TRY(hipMalloc(&h->uwall_d, cnt * sizeof(float)), "hipMalloc wall velocity");
      HIP_TRY(hipMemcpy(h->uwall_d, uf.data(), cnt * sizeof(float), hipMemcpyHostToDevice), "copy wall velocity");
    }
    const size_t rsize = h->cfg.storage == LBM_STORAGE_F64 ? 8 : 4;
    if (!h->rho_side) HIP_TRY(hipMalloc(&h->rho_side, (size_t)h->N * rsize), "hipMalloc rho");
  }
  drop_graph(h);
  return LBM_OK;
}

int check_particles(lbm_handle *h) {
  if (h->parts_h.empty()) return LBM_OK;
  if (h->flags_d) {
    int32_t fl = 0;
    HIP_TRY(hipMemcpy(&fl, h->flags_d, sizeof fl, hipMemcpyDeviceToHost), "copy flags");
    if (fl) {
      HIP_TRY(hipMemset(h->flags_d, 0, sizeof fl), "clear flags");
      return fail(LBM_ERANGE, "more than " + std::to_string(kLocalParticles) +
                                  " particles came within reach of one another (their owned regions, the link "
                                  "boxes grown by two nodes, overlapped); the populations are not reliable");
    }
  }
  std::vector<DevParticle> tmp(h->parts_h.size());
  HIP_TRY(hipMemcpy(tmp.data(), parts_of(h, h->pcur), tmp.size() * sizeof(DevParticle), hipMemcpyDeviceToHost),
          "copy particles");
  for (size_t p = 0; p < tmp.size(); p++)
    if (tmp[p].err)
      return fail(LBM_ERANGE, "particle " + std::to_string(p) +
                                  " came within two nodes of the lattice edge; the unwrapped sphere "
                                  "of the reference is only supported away from the faces");
  return LBM_OK;
}

}  // namespace

extern "C" {

const char *lbm_last_error(void) { return g_err.c_str(); }

const char *lbm_version(void) { return "lbm_hip 0.1 (gfx950)"; }

int lbm_device_count(int32_t *count) {
  if (!count) return fail(LBM_EINVAL, "null argument");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return LBM_OK;
}

int lbm_destroy(lbm_handle *h) {
  if (!h) return LBM_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  drop_graph(h);
  for (int k = 0; k < 3; k++)
    if (h->coords_d[k]) hipFree(h->coords_d[k]);
  for (int b = 0; b < 2; b++)
    if (h->f[b]) hipFree(h->f[b]);
  if (h->F) hipFree(h->F);
  if (h->wmask_d) hipFree(h->wmask_d);
  if (h->scan_d) hipFree(h->scan_d);
  if (h->uwall_d) hipFree(h->uwall_d);
  if (h->parts_d) hipFree(h->parts_d);
  if (h->partial) hipFree(h->partial);
  if (h->rho_side) hipFree(h->rho_side);
  if (h->linkval) hipFree(h->linkval);
  if (h->inter_d) hipFree(h->inter_d);
  if (h->traj_d) hipFree(h->traj_d);
  if (h->bsched_d) hipFree(h->bsched_d);
  if (h->ids) hipFree(h->ids);
  if (h->flags_d) hipFree(h->flags_d);
  for (auto e : h->ev) hipEventDestroy(e);
  if (h->run_start) hipEventDestroy(h->run_start);
  if (h->run_stop) hipEventDestroy(h->run_stop);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return LBM_OK;
}

int lbm_create(const lbm_config *cfg, lbm_handle **out) {
  if (!cfg || !out) return fail(LBM_EINVAL, "null argument");
  *out = nullptr;
  if (!((cfg->dims == 2 && cfg->q == 9) || (cfg->dims == 3 && (cfg->q == 15 || cfg->q == 19 || cfg->q == 27))))
    return fail(LBM_EINVAL, "lattice must be D2Q9 (dims 2) or D3Q15, D3Q19, D3Q27 (dims 3)");
  if (cfg->flags & ~LBM_FLAG_INCOMPRESSIBLE) return fail(LBM_EINVAL, "unknown configuration flags");
  if ((cfg->flags & LBM_FLAG_INCOMPRESSIBLE) && !(cfg->rho0 > 0))
    return fail(LBM_EINVAL, "the incompressible equilibrium needs the initial density rho0 > 0");
  if (cfg->collision != LBM_COLLISION_BGK && cfg->collision != LBM_COLLISION_TRT)
    return fail(LBM_EINVAL, "collision must be BGK or TRT");
  if (cfg->storage < LBM_STORAGE_F64 || cfg->storage > LBM_STORAGE_F16C)
    return fail(LBM_EINVAL, "storage must be F64, F32, F16 or F16C");
  if (cfg->semantics != LBM_SEMANTICS_GPU && cfg->semantics != LBM_SEMANTICS_CPU)
    return fail(LBM_EINVAL, "semantics must be GPU or CPU");
  if (cfg->force < LBM_FORCE_NONE || cfg->force > LBM_FORCE_FIELD)
    return fail(LBM_EINVAL, "force mode must be NONE, UNIFORM or FIELD");
  if (!(cfg->dt > 0) || !(cfg->dx > 0) || !(cfg->tau > 0))
    return fail(LBM_EINVAL, "dt, dx and tau must be positive");
  if (cfg->collision == LBM_COLLISION_TRT && !(cfg->tau_minus > 0))
    return fail(LBM_EINVAL, "tau_minus must be positive for TRT");
  if (cfg->walled & ~((1 << cfg->dims) - 1)) return fail(LBM_EINVAL, "walled names a dimension the lattice lacks");
  int64_t N = 1;
  for (int k = 0; k < cfg->dims; k++) {
    if (cfg->n[k] < 1 || cfg->n[k] > (1 << 30)) return fail(LBM_EINVAL, "bad lattice size");
    if (((cfg->walled >> k) & 1) && cfg->n[k] < 3) return fail(LBM_EINVAL, "a walled dimension needs >= 3 nodes");
    if (!cfg->coords[k]) return fail(LBM_EINVAL, "missing coordinate table");
    N *= cfg->n[k];
  }
  if (N * (int64_t)cfg->q >= ((int64_t)1 << 31) || N >= ((int64_t)1 << 31))
    return fail(LBM_EINVAL, "lattice too large for one box (N*Q must stay below 2^31)");
  lbm_handle *h = new (std::nothrow) lbm_handle;
  if (!h) return fail(LBM_ENOMEM, "out of host memory");
  h->cfg = *cfg;
  h->dims = cfg->dims;
  h->Q = cfg->q;
  for (int k = 0; k < 3; k++) h->n[k] = k < cfg->dims ? (int32_t)cfg->n[k] : 1;
  h->N = N;
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return hip_fail(e, "hipStreamCreate");
  }
  auto bail = [&](int code) {
    lbm_destroy(h);
    return code;
  };
  for (int k = 0; k < 3; k++) {
    if (k < cfg->dims)
      h->coords_h[k].assign(cfg->coords[k], cfg->coords[k] + cfg->n[k]);
    else
      h->coords_h[k].assign(1, 0.0);
    h->x0[k] = h->coords_h[k][0];
    if (hipMalloc(&h->coords_d[k], h->coords_h[k].size() * sizeof(double)) != hipSuccess)
      return bail(fail(LBM_ENOMEM, "hipMalloc coordinates"));
    if (hipMemcpy(h->coords_d[k], h->coords_h[k].data(), h->coords_h[k].size() * sizeof(double),
                  hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(LBM_EDEVICE, "copy coordinates"));
  }
  h->cfg.coords[0] = h->cfg.coords[1] = h->cfg.coords[2] = nullptr;
  if ((uint64_t)N * elem_size(cfg->storage) >= (1ull << 31))
    return bail(fail(LBM_EINVAL, "lattice too large: one population plane (nodes x element size) must stay "
                                 "below 2 GiB"));
  const size_t bytes = (size_t)N * cfg->q * elem_size(cfg->storage);
  for (int b = 0; b < 2; b++)
    if (hipMalloc(&h->f[b], bytes) != hipSuccess) return bail(fail(LBM_ENOMEM, "hipMalloc populations"));
  if (cfg->force == LBM_FORCE_FIELD) {
    const size_t fb = (size_t)N * cfg->dims * (cfg->storage == LBM_STORAGE_F64 ? 8 : 4);
    if (hipMalloc(&h->F, fb) != hipSuccess) return bail(fail(LBM_ENOMEM, "hipMalloc force field"));
    if (hipMemset(h->F, 0, fb) != hipSuccess) return bail(fail(LBM_EDEVICE, "clear force field"));
  }
  if (hipMalloc(&h->parts_d, sizeof(DevParticle) * LBM_MAX_PARTICLES * 2) != hipSuccess)
    return bail(fail(LBM_ENOMEM, "hipMalloc particles"));
  h->tiling = default_tiling(h);
  if (hipEventCreate(&h->run_start) != hipSuccess || hipEventCreate(&h->run_stop) != hipSuccess)
    return bail(fail(LBM_EDEVICE, "hipEventCreate"));
  *out = h;
  return LBM_OK;
}

int lbm_init_uniform(lbm_handle *h, double rho, const double *u) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  AuxArgs a{};
  a.f = h->f[h->cur];
  a.rho_init = rho;
  for (int k = 0; k < 3; k++) a.u_init[k] = (u && k < h->dims) ? u[k] : 0.0;
  int rc = aux_launch(h, 0, a);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream), "init");
  h->tick = 0;
  h->time = 0.0;
  h->populated = true;
  return LBM_OK;
}

int lbm_set_force_field(lbm_handle *h, const double *force) {
  if (!h || !force) return fail(LBM_EINVAL, "null argument");
  if (h->cfg.force != LBM_FORCE_FIELD) return fail(LBM_ESTATE, "the model was created without a force field");
  const int64_t N = h->N;
  const int d = h->dims;
  std::vector<double> F(force, force + (size_t)N * d);
  for (int64_t n = 0; n < N; n++)
    if (host_wall(h, n))
      for (int k = 0; k < d; k++) F[(size_t)k * N + n] = 0.0;
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  if (h->cfg.storage == LBM_STORAGE_F64) {
    HIP_TRY(hipMemcpy(h->F, F.data(), F.size() * sizeof(double), hipMemcpyHostToDevice), "copy force");
  } else {
    std::vector<float> Ff(F.begin(), F.end());
    HIP_TRY(hipMemcpy(h->F, Ff.data(), Ff.size() * sizeof(float), hipMemcpyHostToDevice), "copy force");
  }
  return LBM_OK;
}

int lbm_set_walls(lbm_handle *h, const uint8_t *wall) {
  if (!h || !wall) return fail(LBM_EINVAL, "null argument");
  if (h->populated || !h->parts_h.empty())
    return fail(LBM_ESTATE, "walls are set before the populations and the particles");
  h->wall_h.assign(wall, wall + h->N);
  return upload_walls(h);
}

int lbm_set_wall_velocity(lbm_handle *h, const double *u) {
  if (!h || !u) return fail(LBM_EINVAL, "null argument");
  if (h->populated || !h->parts_h.empty())
    return fail(LBM_ESTATE, "walls are set before the populations and the particles");
  h->uwall_h.assign(u, u + (size_t)h->N * h->dims);
  return upload_walls(h);
}

int lbm_set_populations(lbm_handle *h, const double *f) {
  if (!h || !f) return fail(LBM_EINVAL, "null argument");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  const size_t bytes = (size_t)h->N * h->Q * sizeof(double);
  double *tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, bytes), "hipMalloc staging");
  hipError_t e = hipMemcpy(tmp, f, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(tmp);
    return hip_fail(e, "copy populations");
  }
  AuxArgs a{};
  a.f = h->f[h->cur];
  a.f64 = tmp;
  int rc = aux_launch(h, 1, a);
  e = hipStreamSynchronize(h->stream);
  hipFree(tmp);
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "set populations");
  h->populated = true;
  return LBM_OK;
}

int lbm_get_populations(lbm_handle *h, double *f) {
  if (!h || !f) return fail(LBM_EINVAL, "null argument");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  const size_t bytes = (size_t)h->N * h->Q * sizeof(double);
  double *tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, bytes), "hipMalloc staging");
  AuxArgs a{};
  a.f = h->f[h->cur];
  a.out64 = tmp;
  int rc = aux_launch(h, 2, a);
  hipError_t e = hipStreamSynchronize(h->stream);
  if (!rc && e == hipSuccess) e = hipMemcpy(f, tmp, bytes, hipMemcpyDeviceToHost);
  hipFree(tmp);
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "get populations");
  return LBM_OK;
}

int lbm_get_hydro(lbm_handle *h, double *rho, double *u) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  const size_t N = (size_t)h->N;
  double *tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, N * (1 + h->dims) * sizeof(double)), "hipMalloc staging");
  AuxArgs a{};
  a.f = h->f[h->cur];
  a.out64 = tmp;
  a.uout = tmp + N;
  int rc = aux_launch(h, 3, a);
  hipError_t e = hipStreamSynchronize(h->stream);
  if (!rc && e == hipSuccess && rho) e = hipMemcpy(rho, tmp, N * sizeof(double), hipMemcpyDeviceToHost);
  if (!rc && e == hipSuccess && u) e = hipMemcpy(u, tmp + N, N * h->dims * sizeof(double), hipMemcpyDeviceToHost);
  hipFree(tmp);
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "get hydro");
  return LBM_OK;
}

int lbm_check_populations(lbm_handle *h, int64_t *nonfinite, int64_t *saturated) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  if (!h->populated) return fail(LBM_ESTATE, "populations not initialised");
  if (!h->scan_d) HIP_TRY(hipMalloc(&h->scan_d, 2 * sizeof(unsigned long long)), "hipMalloc scan counters");
  HIP_TRY(hipMemsetAsync(h->scan_d, 0, 2 * sizeof(unsigned long long), h->stream), "clear scan counters");
  AuxArgs a{};
  a.f = h->f[h->cur];
  a.g = geom(h);
  a.c = coll(h);
  const int threads = 256;
  const int64_t want = (h->N * h->Q + threads - 1) / threads;
  Launch l{dim3((unsigned)std::min<int64_t>(want, 4096)), dim3(threads), h->stream};
  hipError_t e = hipErrorInvalidValue;
  switch (h->cfg.storage) {
    case LBM_STORAGE_F64: e = launch_scan<kF64>(h->Q, a, h->scan_d, l); break;
    case LBM_STORAGE_F32: e = launch_scan<kF32>(h->Q, a, h->scan_d, l); break;
    case LBM_STORAGE_F16C: e = launch_scan<kF16C>(h->Q, a, h->scan_d, l); break;
    default: e = launch_scan<kF16>(h->Q, a, h->scan_d, l); break;
  }
  if (e != hipSuccess) return hip_fail(e, "scan launch");
  unsigned long long c[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(c, h->scan_d, sizeof c, hipMemcpyDeviceToHost, h->stream), "copy scan counters");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  if (nonfinite) *nonfinite = (int64_t)c[0];
  if (saturated) *saturated = (int64_t)c[1];
  return LBM_OK;
}

int lbm_viscous_stress(lbm_handle *h, double *sigma) {
  if (!h || !sigma) return fail(LBM_EINVAL, "null argument");
  if (!h->populated) return fail(LBM_ESTATE, "populations not initialised");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  const size_t bytes = (size_t)h->N * h->dims * h->dims * sizeof(double);
  double *tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, bytes), "hipMalloc staging");
  AuxArgs a{};
  a.f = h->f[h->cur];
  a.out64 = tmp;
  int rc = aux_launch(h, 4, a);
  hipError_t e = hipStreamSynchronize(h->stream);
  if (!rc && e == hipSuccess) e = hipMemcpy(sigma, tmp, bytes, hipMemcpyDeviceToHost);
  hipFree(tmp);
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "viscous stress");
  return LBM_OK;
}

int lbm_add_particle(lbm_handle *h, const lbm_particle *p, int32_t *id) {
  if (!h || !p) return fail(LBM_EINVAL, "null argument");
  if ((int)h->parts_h.size() >= LBM_MAX_PARTICLES) return fail(LBM_EINVAL, "too many particles");
  if (p->kind != LBM_PARTICLE_BEAD && p->kind != LBM_PARTICLE_SQUIRMER)
    return fail(LBM_EINVAL, "particle kind must be BEAD or SQUIRMER");
  if (!(p->radius >= h->cfg.dx))
    return fail(LBM_EINVAL, "particle radius must be at least one lattice spacing");
  for (int k = 0; k < h->dims; k++)
    if (!(2.0 * p->radius + 6.0 * h->cfg.dx < h->n[k] * h->cfg.dx))
      return fail(LBM_EINVAL, "particle diameter must be smaller than the box");
  DevParticle d{};
  d.kind = p->kind;
  d.couple_forces = p->couple_forces != 0;
  d.couple_torques = p->couple_torques != 0;
  d.radius = p->radius;
  d.B1 = p->B1;
  d.B2 = p->B2;
  d.inv_mass = p->inverse_mass;
  d.inv_I = p->inverse_moment_of_inertia;
  for (int k = 0; k < 3; k++) {
    const bool in = k < h->dims;
    d.R[k] = in ? p->position[k] : 0.0;
    d.V[k] = in ? p->velocity[k] : 0.0;
    d.W[k] = (h->dims == 3 || k == 0) ? p->angular_velocity[k] : 0.0;
    d.e[k] = in ? p->swimming_direction[k] : 0.0;
    d.Rg[k] = d.R[k];
    d.nR[k] = d.R[k];
    d.nV[k] = d.V[k];
    d.nW[k] = d.W[k];
    d.ne[k] = d.e[k];
  }
  int32_t widest = 1;
  for (int k = 0; k < h->dims; k++) widest = std::max(widest, owned_extent(h, d, k));
  if (widest > 1023) return fail(LBM_EINVAL, "particle radius too large (owned region over 1023 nodes)");
  d.whole = h->dims == 3 && widest > kBandMaxExtent;
  int32_t err = 0;
  link_box(d.Rg, d.radius, h->x0, h->cfg.dx, h->n, h->dims, d.lo, d.hi, &err);
  if (err)
    return fail(LBM_ERANGE, "particle must lie at least two nodes inside every face of the lattice");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  const int idx = (int)h->parts_h.size();
  HIP_TRY(hipMemcpy(parts_of(h, h->pcur) + idx, &d, sizeof d, hipMemcpyHostToDevice), "copy particle");
  h->parts_h.push_back(d);
  int rc = plan_boxes(h, tile_threads(h->tiling));
  if (rc) {
    h->parts_h.pop_back();
    return rc;
  }
  if (idx == 0) h->tick0 = h->tick;
  drop_graph(h);
  if (id) *id = idx;
  return LBM_OK;
}

int lbm_get_particle(lbm_handle *h, int32_t id, lbm_particle *out) {
  if (!h || !out) return fail(LBM_EINVAL, "null argument");
  if (id < 0 || id >= (int)h->parts_h.size()) return fail(LBM_EINVAL, "bad particle id");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  DevParticle d;
  HIP_TRY(hipMemcpy(&d, parts_of(h, h->pcur) + id, sizeof d, hipMemcpyDeviceToHost), "copy particle");
  std::memset(out, 0, sizeof *out);
  out->kind = d.kind;
  out->couple_forces = d.couple_forces;
  out->couple_torques = d.couple_torques;
  out->radius = d.radius;
  for (int k = 0; k < 3; k++) {
    out->position[k] = d.R[k];
    out->velocity[k] = d.V[k];
    out->angular_velocity[k] = d.W[k];
    out->swimming_direction[k] = d.e[k];
  }
  out->B1 = d.B1;
  out->B2 = d.B2;
  out->inverse_mass = d.inv_mass;
  out->inverse_moment_of_inertia = d.inv_I;
  return LBM_OK;
}

int lbm_get_particle_load(lbm_handle *h, int32_t id, double force[3], double torque[3]) {
  if (!h || !force || !torque) return fail(LBM_EINVAL, "null argument");
  if (id < 0 || id >= (int)h->parts_h.size()) return fail(LBM_EINVAL, "bad particle id");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  DevParticle d;
  HIP_TRY(hipMemcpy(&d, parts_of(h, h->pcur) + id, sizeof d, hipMemcpyDeviceToHost), "copy particle");
  for (int k = 0; k < 3; k++) {
    force[k] = d.force[k];
    torque[k] = d.torque[k];
  }
  return LBM_OK;
}

int lbm_set_interactions(lbm_handle *h, const lbm_interaction *terms, int32_t n) {
  if (!h || (n > 0 && !terms)) return fail(LBM_EINVAL, "null argument");
  if (n < 0 || n > LBM_MAX_INTERACTIONS) return fail(LBM_EINVAL, "interaction count out of range");
  if (n > 0 && h->cfg.semantics != LBM_SEMANTICS_GPU)
    return fail(LBM_EINVAL, "particle interactions exist only in the GPU semantics");
  const int np = (int)h->parts_h.size();
  for (int t = 0; t < n; t++) {
    const lbm_interaction &I = terms[t];
    if (I.kind < LBM_INTERACTION_LINEAR || I.kind > LBM_INTERACTION_LENNARD_JONES)
      return fail(LBM_EINVAL, "unknown interaction kind");
    const int nid = I.kind == LBM_INTERACTION_POLAR ? 3 : 2;
    for (int k = 0; k < nid; k++)
      if (I.ids[k] < 0 || I.ids[k] >= np) return fail(LBM_EINVAL, "interaction names a particle that does not exist");
    if (I.kind == LBM_INTERACTION_DIPOLE && I.params[1] != 0.0 && h->dims != 2)
      return fail(LBM_EINVAL, "a scalar magnetic field is a 2D field");
  }
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  if (!h->inter_d)
    HIP_TRY(hipMalloc(&h->inter_d, sizeof(lbm_interaction) * LBM_MAX_INTERACTIONS), "allocate interactions");
  if (n > 0)
    HIP_TRY(hipMemcpy(h->inter_d, terms, sizeof(lbm_interaction) * n, hipMemcpyHostToDevice), "copy interactions");
  h->inter_h.assign(terms, terms + n);
  drop_graph(h);
  return LBM_OK;
}

int lbm_set_field_schedule(lbm_handle *h, const double *B, int64_t first_tick, int64_t count) {
  if (!h || (count > 0 && !B)) return fail(LBM_EINVAL, "null argument");
  if (count < 0 || count > (int64_t(1) << 26)) return fail(LBM_EINVAL, "schedule length out of range");
  for (int64_t i = 0; i < 3 * count; i++)
    if (!std::isfinite(B[i])) return fail(LBM_EINVAL, "the scheduled field must be finite");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  if (count > h->bsched_cap) {
    if (h->bsched_d) HIP_TRY(hipFree(h->bsched_d), "free schedule");
    h->bsched_d = nullptr;
    h->bsched_cap = 0;
    HIP_TRY(hipMalloc(&h->bsched_d, sizeof(double) * 3 * count), "allocate schedule");
    h->bsched_cap = count;
  }
  if (count > 0) HIP_TRY(hipMemcpy(h->bsched_d, B, sizeof(double) * 3 * count, hipMemcpyHostToDevice), "copy schedule");
  h->bsched_first = first_tick;
  h->bsched_count = count;
  drop_graph(h);
  return LBM_OK;
}

int lbm_set_trajectory_log(lbm_handle *h, int64_t capacity) {
  if (!h) return fail(LBM_EINVAL, "null argument");
  if (capacity < 0 || capacity > (int64_t(1) << 24)) return fail(LBM_EINVAL, "log capacity out of range");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  if (h->traj_d) HIP_TRY(hipFree(h->traj_d), "free log");
  h->traj_d = nullptr;
  h->traj_cap = 0;
  if (capacity > 0) {
    HIP_TRY(hipMalloc(&h->traj_d, sizeof(double) * 9 * LBM_MAX_PARTICLES * capacity), "allocate log");
    h->traj_cap = capacity;
  }
  h->traj_from = h->tick;
  drop_graph(h);
  return LBM_OK;
}

int lbm_get_trajectory(lbm_handle *h, int32_t id, int64_t first, int64_t count, double *out) {
  if (!h || (count > 0 && !out)) return fail(LBM_EINVAL, "null argument");
  if (!h->traj_d) return fail(LBM_ESTATE, "the trajectory log is off");
  if (id < 0 || id >= (int)h->parts_h.size()) return fail(LBM_EINVAL, "bad particle id");
  if (count < 0) return fail(LBM_EINVAL, "negative count");
  if (count == 0) return LBM_OK;
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  DevParticle d;
  HIP_TRY(hipMemcpy(&d, parts_of(h, h->pcur) + id, sizeof d, hipMemcpyDeviceToHost), "copy particle");
  // every tick moves every particle once: tick t's state is move
  // d.moves - (tick - t)
  const int64_t last = first + count - 1;
  const int64_t m0 = d.moves - (h->tick - first);
  if (first <= h->traj_from || m0 < 1 || last > h->tick || h->tick - first >= h->traj_cap)
    return fail(LBM_EINVAL, "ticks outside what the trajectory log holds");
  const double *ring = h->traj_d + (size_t)id * h->traj_cap * 9;
  int64_t done = 0;
  while (done < count) {
    const int64_t slot = (m0 - 1 + done) % h->traj_cap;
    const int64_t run = std::min(count - done, h->traj_cap - slot);
    HIP_TRY(hipMemcpy(out + done * 9, ring + slot * 9, sizeof(double) * 9 * run, hipMemcpyDeviceToHost),
            "copy log");
    done += run;
  }
  return LBM_OK;
}

namespace {

// A double as Julia prints it: the shortest decimal that reads back to the
// same value, plain for 1e-4 <= |x| < 1e6 (with ".0" on whole numbers) and
// d.ddde<exp> outside.
int julia_double(double x, char *out) {
  if (std::isnan(x)) return std::sprintf(out, "NaN");
  if (std::isinf(x)) return std::sprintf(out, x < 0 ? "-Inf" : "Inf");
  char *o = out;
  if (std::signbit(x)) *o++ = '-';
  x = std::fabs(x);
  if (x == 0.0) {
    std::memcpy(o, "0.0", 3);
    return (int)(o - out) + 3;
  }
  char sci[40];
  const auto r = std::to_chars(sci, sci + sizeof sci, x, std::chars_format::scientific);
  *r.ptr = '\0';
  char digits[32];
  int nd = 0;
  const char *c = sci;
  for (; *c && *c != 'e'; c++)
    if (*c != '.') digits[nd++] = *c;
  const int e10 = std::atoi(c + 1);
  const int pt = e10 + 1;  // digits before the decimal point
  if (pt > -4 && pt <= 6) {
    if (pt <= 0) {
      *o++ = '0';
      *o++ = '.';
      for (int i = 0; i < -pt; i++) *o++ = '0';
      for (int i = 0; i < nd; i++) *o++ = digits[i];
    } else if (pt < nd) {
      for (int i = 0; i < pt; i++) *o++ = digits[i];
      *o++ = '.';
      for (int i = pt; i < nd; i++) *o++ = digits[i];
    } else {
      for (int i = 0; i < nd; i++) *o++ = digits[i];
      for (int i = nd; i < pt; i++) *o++ = '0';
      *o++ = '.';
      *o++ = '0';
    }
  } else {
    *o++ = digits[0];
    *o++ = '.';
    if (nd == 1) *o++ = '0';
    for (int i = 1; i < nd; i++) *o++ = digits[i];
    o += std::sprintf(o, "e%d", e10);
  }
  return (int)(o - out);
}

}  // namespace

namespace {

// One row per node in column-major order: tick, time, the 1-based indices,
// the coordinates, then the given columns (one FP64 value per node each),
// as the reference's DataFrame + CSV.write produce them
// (gpu/aux.jl:271-317,364-400).
int write_node_csv(const char *path, int32_t dims, const int64_t n[3], int64_t tick, double time,
                   const double *const coords[3], const std::vector<std::string> &names,
                   const std::vector<const double *> &cols) {
  int64_t N = 1;
  for (int k = 0; k < dims; k++) N *= n[k];
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return fail(LBM_EINVAL, std::string("cannot open ") + path);
  static const char *ax[3] = {"x", "y", "z"};
  std::string line = "tick,time";
  for (int k = 0; k < dims; k++) line += std::string(",id_") + ax[k];
  for (int k = 0; k < dims; k++) line += std::string(",coordinate_") + ax[k];
  for (const std::string &c : names) line += "," + c;
  line += "\n";
  bool ok = std::fwrite(line.data(), 1, line.size(), fp) == line.size();
  // tick and time open every row
  char head[96];
  int hl = std::sprintf(head, "%lld,", (long long)tick);
  hl += julia_double(time, head + hl);
  std::vector<char> buf;
  buf.reserve(1 << 21);
  char tmp[64];
  for (int64_t x = 0; x < N && ok; x++) {
    buf.insert(buf.end(), head, head + hl);
    int64_t idx[3], r = x;
    for (int k = 0; k < dims; k++) {
      idx[k] = r % n[k];
      r /= n[k];
    }
    for (int k = 0; k < dims; k++) {
      const int l = std::sprintf(tmp, ",%lld", (long long)(idx[k] + 1));
      buf.insert(buf.end(), tmp, tmp + l);
    }
    auto put = [&](double v) {
      tmp[0] = ',';
      const int l = julia_double(v, tmp + 1);
      buf.insert(buf.end(), tmp, tmp + l + 1);
    };
    for (int k = 0; k < dims; k++) put(coords[k][idx[k]]);
    for (const double *c : cols) put(c[x]);
    buf.push_back('\n');
    if (buf.size() > (1u << 20)) {
      ok = std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
      buf.clear();
    }
  }
  if (ok) ok = std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
  if (std::fclose(fp) != 0 || !ok) return fail(LBM_EINVAL, std::string("write failed: ") + path);
  return LBM_OK;
}

int check_csv_args(const char *path, int32_t dims, const int64_t n[3], const double *const coords[3],
                   int64_t *N) {
  if (!path || !n || !coords) return fail(LBM_EINVAL, "null argument");
  if (dims != 2 && dims != 3) return fail(LBM_EINVAL, "dims must be 2 or 3");
  *N = 1;
  for (int k = 0; k < dims; k++) {
    if (n[k] < 1 || !coords[k]) return fail(LBM_EINVAL, "bad lattice size or coordinates");
    *N *= n[k];
  }
  return LBM_OK;
}

}  // namespace

int lbm_write_fluid_csv(const char *path, int32_t dims, const int64_t n[3], int64_t tick, double time,
                        const double *const coords[3], const double *rho, const double *u, int32_t q,
                        const double *f) {
  int64_t N;
  if (int rc = check_csv_args(path, dims, n, coords, &N)) return rc;
  if (!rho || !u || !f) return fail(LBM_EINVAL, "null argument");
  if (q < 1) return fail(LBM_EINVAL, "bad population count");
  static const char *ax[3] = {"x", "y", "z"};
  std::vector<std::string> names{"massDensity"};
  std::vector<const double *> cols{rho};
  for (int k = 0; k < dims; k++) {
    names.push_back(std::string("fluidVelocity_") + ax[k]);
    cols.push_back(u + (int64_t)k * N);
  }
  for (int i = 0; i < q; i++) {
    names.push_back("f" + std::to_string(i + 1));
    cols.push_back(f + (int64_t)i * N);
  }
  return write_node_csv(path, dims, n, tick, time, coords, names, cols);
}

int lbm_write_tensor_csv(const char *path, int32_t dims, const int64_t n[3], int64_t tick, double time,
                         const double *const coords[3], const double *t) {
  int64_t N;
  if (int rc = check_csv_args(path, dims, n, coords, &N)) return rc;
  if (!t) return fail(LBM_EINVAL, "null argument");
  static const char *ax[3] = {"x", "y", "z"};
  std::vector<std::string> names;
  std::vector<const double *> cols;
  for (int b = 0; b < dims; b++)
    for (int a = 0; a < dims; a++) {
      names.push_back(std::string("component_") + ax[a] + ax[b]);
      cols.push_back(t + (int64_t)(a + dims * b) * N);
    }
  return write_node_csv(path, dims, n, tick, time, coords, names, cols);
}

int lbm_get_links(lbm_handle *h, int32_t id, int64_t *node, int32_t *dir, int64_t cap, int64_t *count) {
  if (!h || !count) return fail(LBM_EINVAL, "null argument");
  if (id < 0 || id >= (int)h->parts_h.size()) return fail(LBM_EINVAL, "bad particle id");
  if (cap > 0 && (!node || !dir)) return fail(LBM_EINVAL, "null output buffer");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  DevParticle d;
  HIP_TRY(hipMemcpy(&d, parts_of(h, h->pcur) + id, sizeof d, hipMemcpyDeviceToHost), "copy particle");
  LinkMaskArgs a;
  a.g = geom(h);
  a.part = parts_of(h, h->pcur) + id;
  for (int k = 0; k < 3; k++) {
    a.coords[k] = h->coords_d[k];
    a.lo[k] = d.lo[k];
    a.ext[k] = d.hi[k] - d.lo[k] + 1;
  }
  const int64_t tot = (int64_t)a.ext[0] * a.ext[1] * a.ext[2];
  uint32_t *mask = nullptr;
  HIP_TRY(hipMalloc(&mask, tot * sizeof(uint32_t)), "hipMalloc link mask");
  a.mask = mask;
  Launch l{dim3((unsigned)((tot + 255) / 256)), dim3(256), h->stream};
  hipError_t e = launch_link_mask(h->Q, a, l);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  std::vector<uint32_t> m(tot);
  if (e == hipSuccess) e = hipMemcpy(m.data(), mask, tot * sizeof(uint32_t), hipMemcpyDeviceToHost);
  hipFree(mask);
  if (e != hipSuccess) return hip_fail(e, "link mask");
  // findall over (N..., Q): node index fastest, direction slowest; the box
  // does not wrap, so box order (x fastest) is node order.
  int64_t c = 0;
  for (int j = 0; j < h->Q; j++)
    for (int z = 0; z < a.ext[2]; z++)
      for (int y = 0; y < a.ext[1]; y++)
        for (int x = 0; x < a.ext[0]; x++) {
          const int64_t t = x + (int64_t)a.ext[0] * (y + (int64_t)a.ext[1] * z);
          if ((m[t] >> j) & 1) {
            if (c < cap) {
              node[c] = (a.lo[0] + x) + (int64_t)h->n[0] * ((a.lo[1] + y) + (int64_t)h->n[1] * (a.lo[2] + z));
              dir[c] = j;
            }
            c++;
          }
        }
  *count = c;
  return LBM_OK;
}

// A run of n ticks is one collide-only kernel, n - 1 fused kernels (replayed
// from captured graphs of graph_k where possible) and one gather-only kernel,
// so the populations between runs are the reference's distributions.  With
// `direct`, every kernel is launched directly (no graph is captured) and no
// timing event is recorded, so the run allocates nothing.
}  // extern "C"

namespace {

// A run of nsteps ticks reads the scheduled field of ticks tick + 1 ..
// tick + nsteps when a dipole term has the schedule flag: refuse it when the
// schedule does not hold them all.
int check_schedule(const lbm_handle *h, int64_t nsteps) {
  bool scheduled = false;
  for (const lbm_interaction &I : h->inter_h)
    scheduled = scheduled || (I.kind == LBM_INTERACTION_DIPOLE && I.params[2] != 0.0);
  if (!scheduled || h->parts_h.empty()) return LBM_OK;
  const int64_t a = h->tick + 1, b = h->tick + nsteps;
  if (h->bsched_count == 0 || a < h->bsched_first || b > h->bsched_first + h->bsched_count - 1)
    return fail(LBM_ESTATE, "the magnetic-field schedule does not cover ticks " + std::to_string(a) + " .. " +
                                std::to_string(b) + " (lbm_set_field_schedule)");
  return LBM_OK;
}

int run_ticks(lbm_handle *h, int64_t nsteps, bool direct) {
  if (const int rc = check_schedule(h, nsteps)) return rc;
  const int64_t nfused = nsteps - 1;
  const bool timing = h->timing && !direct;
  if (timing && nfused > 0) {
    if ((int64_t)h->ev.size() < 2 * nfused) {
      const size_t old = h->ev.size();
      h->ev.resize(2 * nfused);
      for (size_t i = old; i < h->ev.size(); i++)
        HIP_TRY(hipEventCreate(&h->ev[i]), "hipEventCreate");
    }
  }
  if (!direct) HIP_TRY(hipEventRecord(h->run_start, h->stream), "event record");
  int a = h->cur;
  hipError_t e = launch_mode(h, kCollide, h->f[a], h->f[a ^ 1], h->pcur);
  if (e != hipSuccess) return hip_fail(e, "step launch");
  a ^= 1;
  int64_t s = 0;
  if (!timing && !direct && h->graph_k > 0) {
    // replay captured blocks of graph_k fused kernels (graph_k is even, so a
    // block leaves both parities where it found them)
    while (nfused - s >= h->graph_k) {
      hipGraphExec_t &g = h->graph_exec[a][h->pcur];
      if (!g) {
        int rc = capture_graph(h, a, h->pcur, h->graph_k, &g);
        if (rc) return rc;
      }
      HIP_TRY(hipGraphLaunch(g, h->stream), "graph launch");
      s += h->graph_k;
    }
  }
  for (int64_t i = 0; s < nfused; s++, i++) {
    hipEvent_t e0 = timing ? h->ev[2 * i] : nullptr;
    hipEvent_t e1 = timing ? h->ev[2 * i + 1] : nullptr;
    e = fused_once(h, a, h->pcur, e0, e1);
    if (e != hipSuccess) return hip_fail(e, "step launch");
    a ^= 1;
    h->pcur ^= 1;
  }
  e = launch_mode(h, kGather, h->f[a], h->f[a ^ 1], h->pcur);
  if (e != hipSuccess) return hip_fail(e, "step launch");
  if (many_particles(h)) {
    e = launch_moves(h, h->pcur);
    if (e != hipSuccess) return hip_fail(e, "move launch");
  }
  a ^= 1;
  h->pcur ^= 1;
  h->cur = a;
  if (!h->parts_h.empty() && !many_particles(h)) {
    // the last tick's move
    StepArgs fa = step_args(h, nullptr, nullptr, h->pcur);
    fa.sp.parts_out = parts_of(h, h->pcur);  // in place
    e = launch_finalize(fa.sp, h->stream);
    if (e != hipSuccess) return hip_fail(e, "finalize launch");
  }
  h->tick += nsteps;
  for (int64_t k = 0; k < nsteps; k++) h->time += h->cfg.dt;
  if (direct) return LBM_OK;
  HIP_TRY(hipEventRecord(h->run_stop, h->stream), "event record");
  if (timing) {
    HIP_TRY(hipEventSynchronize(h->run_stop), "sync");
    double sum = 0.0;
    for (int64_t k = 0; k < nfused; k++) {
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, h->ev[2 * k], h->ev[2 * k + 1]), "event time");
      sum += ms;
    }
    float tot = 0.0f;
    HIP_TRY(hipEventElapsedTime(&tot, h->run_start, h->run_stop), "event time");
    h->step_kernel_ms = nfused > 0 ? sum / (double)nfused : 0.0;
    h->step_ms = tot / (double)nsteps;
  }
  return LBM_OK;
}

// ------------------------------------------------- allocation-free run
// lbm_run_noalloc's scratch: the momentum-exchange partials (two parities)
// and the link-slot values of the box blocks, in one device array of fixed
// size.  Host code uses the per-device address from hipGetSymbolAddress,
// never the array's name.
constexpr size_t kStaticScratchBytes = size_t(8) << 20;
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices] = {};

int static_scratch(int dev, void **out);

}  // namespace

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

// Address of this device's copy of `scratch` (device `dev` must be current),
// looked up once per device.
int static_scratch(int dev, void **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0 || dev >= kMaxScratchDevices) return fail(LBM_EINVAL, "device index beyond the scratch table");
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_run(lbm_handle *h, int64_t nsteps) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  if (nsteps < 0) return fail(LBM_EINVAL, "nsteps must be non-negative");
  if (!h->populated) return fail(LBM_ESTATE, "populations not initialised (lbm_init_uniform or lbm_set_populations)");
  if (nsteps == 0) return LBM_OK;
  return run_ticks(h, nsteps, false);
}

int lbm_run_noalloc(lbm_handle *h, int64_t nsteps) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  if (h->parts_h.size() > (size_t)kLocalParticles)
    return fail(LBM_EINVAL, "lbm_run_noalloc takes at most " + std::to_string(kLocalParticles) +
                                " particles; use lbm_run");
  if (nsteps < 0) return fail(LBM_EINVAL, "nsteps must be non-negative");
  if (!h->populated) return fail(LBM_ESTATE, "populations not initialised (lbm_init_uniform or lbm_set_populations)");
  if (nsteps == 0) return LBM_OK;
  // the partials (both parities) and link-slot values the box blocks of this
  // block shape use, as plan_boxes sizes them
  const int64_t blocks = std::max<int64_t>(h->first_block[LBM_MAX_PARTICLES], 1);
  const size_t rsize = h->cfg.storage == LBM_STORAGE_F64 ? 8 : 4;
  const size_t part_doubles = (size_t)blocks * LBM_MAX_PARTICLES * 6 * 2;
  const size_t part_bytes = (2 * part_doubles * sizeof(double) + 255) / 256 * 256;
  const size_t lv_bytes = (size_t)blocks * (size_t)tile_threads(h->tiling) * (h->Q - 1) * rsize;
  if (part_bytes + lv_bytes > kStaticScratchBytes)
    return fail(LBM_EINVAL, "lbm_run_noalloc: the particles' scratch (" + std::to_string(part_bytes + lv_bytes) +
                                " bytes for this block shape and these particles) exceeds the fixed " +
                                std::to_string(kStaticScratchBytes) + " bytes; use lbm_run");
  int prev = 0;
  HIP_TRY(hipGetDevice(&prev), "hipGetDevice");
  if (prev != h->device) HIP_TRY(hipSetDevice(h->device), "hipSetDevice");
  void *base = nullptr;
  int rc = static_scratch(h->device, &base);
  if (rc == LBM_OK) {
    // the run's kernels take the static scratch in place of the handle's
    // buffers (they take every pointer by value at launch)
    double *const own_partial = h->partial;
    const size_t own_cap = h->partial_cap;
    void *const own_linkval = h->linkval;
    h->partial = static_cast<double *>(base);
    h->partial_cap = part_doubles;
    h->linkval = static_cast<unsigned char *>(base) + part_bytes;
    rc = run_ticks(h, nsteps, true);
    h->partial = own_partial;
    h->partial_cap = own_cap;
    h->linkval = own_linkval;
  }
  if (prev != h->device) hipSetDevice(prev);
  return rc;
}

int lbm_sync(lbm_handle *h) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  return check_particles(h);
}

int lbm_get_tick(lbm_handle *h, int64_t *tick, double *time) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  if (tick) *tick = h->tick;
  if (time) *time = h->time;
  return LBM_OK;
}

// Snapshot layout: StateHeader, the populations of the current buffer in the
// storage's own encoding and the row-blocked layout, then the particles'
// device records of the current parity.
struct StateHeader {
  char magic[8];  // "LBMSTAT1"
  int32_t Q, storage, np, record_bytes;
  int64_t N, pop_bytes, tick, tick0;
  double time;
};
constexpr char kStateMagic[8] = {'L', 'B', 'M', 'S', 'T', 'A', 'T', '1'};

static int64_t state_bytes(const lbm_handle *h) {
  return (int64_t)sizeof(StateHeader) + (int64_t)h->N * h->Q * (int64_t)elem_size(h->cfg.storage) +
         (int64_t)h->parts_h.size() * (int64_t)sizeof(DevParticle);
}

int lbm_state_size(lbm_handle *h, int64_t *bytes) {
  if (!h || !bytes) return fail(LBM_EINVAL, "null argument");
  *bytes = state_bytes(h);
  return LBM_OK;
}

int lbm_save_state(lbm_handle *h, void *buf, int64_t bytes) {
  if (!h || !buf) return fail(LBM_EINVAL, "null argument");
  if (bytes < state_bytes(h)) return fail(LBM_EINVAL, "state buffer too small (see lbm_state_size)");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  StateHeader hd{};
  std::memcpy(hd.magic, kStateMagic, sizeof hd.magic);
  hd.Q = h->Q;
  hd.storage = h->cfg.storage;
  hd.np = (int32_t)h->parts_h.size();
  hd.record_bytes = (int32_t)sizeof(DevParticle);
  hd.N = h->N;
  hd.pop_bytes = (int64_t)h->N * h->Q * (int64_t)elem_size(h->cfg.storage);
  hd.tick = h->tick;
  hd.tick0 = h->tick0;
  hd.time = h->time;
  char *out = static_cast<char *>(buf);
  std::memcpy(out, &hd, sizeof hd);
  HIP_TRY(hipMemcpy(out + sizeof hd, h->f[h->cur], (size_t)hd.pop_bytes, hipMemcpyDeviceToHost), "copy populations");
  if (hd.np)
    HIP_TRY(hipMemcpy(out + sizeof hd + hd.pop_bytes, parts_of(h, h->pcur), (size_t)hd.np * sizeof(DevParticle),
                      hipMemcpyDeviceToHost),
            "copy particles");
  return LBM_OK;
}

int lbm_load_state(lbm_handle *h, const void *buf, int64_t bytes) {
  if (!h || !buf) return fail(LBM_EINVAL, "null argument");
  if (bytes < (int64_t)sizeof(StateHeader)) return fail(LBM_EINVAL, "not a state snapshot");
  StateHeader hd;
  const char *in = static_cast<const char *>(buf);
  std::memcpy(&hd, in, sizeof hd);
  if (std::memcmp(hd.magic, kStateMagic, sizeof hd.magic) != 0) return fail(LBM_EINVAL, "not a state snapshot");
  const int64_t pop_bytes = (int64_t)h->N * h->Q * (int64_t)elem_size(h->cfg.storage);
  if (hd.Q != h->Q || hd.storage != h->cfg.storage || hd.N != h->N || hd.pop_bytes != pop_bytes ||
      hd.record_bytes != (int32_t)sizeof(DevParticle) || hd.np != (int32_t)h->parts_h.size())
    return fail(LBM_EINVAL, "the snapshot was taken of a lattice or particle set other than this handle's");
  if (bytes < state_bytes(h)) return fail(LBM_EINVAL, "state snapshot truncated");
  std::vector<DevParticle> recs((size_t)hd.np);
  if (hd.np) std::memcpy(recs.data(), in + sizeof hd + pop_bytes, (size_t)hd.np * sizeof(DevParticle));
  // the box plan follows the handle's particles (their radii): the
  // snapshot's must be the same particles, with no move left pending
  for (size_t p = 0; p < recs.size(); p++) {
    const DevParticle &a = recs[p], &b = h->parts_h[p];
    if (a.kind != b.kind || a.radius != b.radius || a.whole != b.whole || a.pending)
      return fail(LBM_EINVAL, "particle " + std::to_string(p) + " of the snapshot is not this handle's");
  }
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  HIP_TRY(hipMemcpy(h->f[h->cur], in + sizeof hd, (size_t)pop_bytes, hipMemcpyHostToDevice), "copy populations");
  if (hd.np)
    HIP_TRY(hipMemcpy(parts_of(h, h->pcur), recs.data(), recs.size() * sizeof(DevParticle), hipMemcpyHostToDevice),
            "copy particles");
  h->parts_h = recs;
  h->tick = hd.tick;
  h->tick0 = hd.tick0;
  h->time = hd.time;
  h->populated = true;
  h->traj_from = h->tick;  // the log holds no tick of the restored past
  return LBM_OK;
}

int lbm_set_block(lbm_handle *h, int32_t bx, int32_t by, int32_t bz) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  Tiling t;
  if (bx == 0 && by == 0 && bz == 0) {
    t = default_tiling(h);
  } else {
    const int64_t nt = (int64_t)bx * by * bz;
    if (bx < 1 || by < 1 || bz < 1 || nt < 64 || nt > 512 || nt % 64)
      return fail(LBM_EINVAL, "block shape must have 64..512 threads, a multiple of 64");
    t = make_tiling(h, bx, by, bz);
  }
  t.occ = h->tiling.occ;  // (the occupancy variant stays)
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  int rc = plan_boxes(h, tile_threads(t));
  if (rc) return rc;
  h->tiling = t;
  drop_graph(h);
  return LBM_OK;
}

int lbm_last_run_ms(lbm_handle *h, double *ms) {
  if (!h || !ms) return fail(LBM_EINVAL, "null argument");
  *ms = h->step_kernel_ms;
  return LBM_OK;
}

// Extra entry points used by the host front end and the benchmark.

// Turn per-step kernel timing (HIP events around every step kernel) on/off.
int lbm_set_timing(lbm_handle *h, int on) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  h->timing = on != 0;
  return LBM_OK;
}

// Average device time per step kernel and per whole step of the last timed run.
int lbm_last_run_times(lbm_handle *h, double *kernel_ms, double *step_ms) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  if (kernel_ms) *kernel_ms = h->step_kernel_ms;
  if (step_ms) *step_ms = h->step_ms;
  return LBM_OK;
}

// Ticks per captured graph (even, >= 2), or 0 to launch every tick directly.
int lbm_set_graph(lbm_handle *h, int32_t ticks_per_graph) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  if (ticks_per_graph < 0 || ticks_per_graph == 1 || (ticks_per_graph & 1))
    return fail(LBM_EINVAL, "ticks per graph must be 0 or an even number");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  drop_graph(h);
  h->graph_k = ticks_per_graph;
  return LBM_OK;
}

// Autotuner: time the fused step kernel of this lattice over candidate block
// shapes and tile orders (reading the current populations, writing a scratch
// buffer, so the simulation state is untouched; the trial kernels also
// rewrite the particle records of the other parity and the box nodes' rho,
// which every run rewrites before reading) and keep the fastest.
int lbm_autotune(lbm_handle *h, int32_t reps, int32_t best[3], double *best_ms) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  if (!h->populated) return fail(LBM_ESTATE, "populations not initialised");
  if (reps < 1) reps = 1;
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  std::vector<int32_t> bxs = {16, 32, 64, 96, 128, 192, 256, 512};
  if (h->n[0] <= 512) bxs.push_back(h->n[0]);
  std::sort(bxs.begin(), bxs.end());
  bxs.erase(std::unique(bxs.begin(), bxs.end()), bxs.end());
  struct Cand {
    int32_t bx, by, bz, remap, occ;
  };
  std::vector<Cand> cands;
  // both occupancy builds: the low one runs the box path without spills,
  // and the tiles too where the collision needs more registers than the
  // default cap (FP64 with a force: 157 against 128)
  const int nocc = 2;
  for (int32_t bx : bxs)
    for (int32_t by : {1, 2, 4, 8, 16})
      for (int32_t bz : {1, 2, 4}) {
        const int64_t t = (int64_t)bx * by * bz;
        if (t < 64 || t > 512 || t % 64 || bx > h->n[0] || by > h->n[1] || bz > h->n[2]) continue;
        const Tiling tl = make_tiling(h, bx, by, bz);
        const double waste = (double)tl.ntiles * t / (double)h->N - 1.0;
        if (waste > 0.25) continue;
        for (int occ = 0; occ < nocc; occ++) {
          cands.push_back({bx, by, bz, 1, occ});
          cands.push_back({bx, by, bz, 0, occ});
        }
      }
  // One-row tiles wider than the lattice: the lanes past the row's end
  // return at once (whole waves of them without issuing a load), and the
  // box blocks keep the block's thread count (96^3 FP32 with a force field
  // and a squirmer: 192 x 1 x 1 runs a tick in 29.8 us, the best shape that
  // fits the rows 32.6 us)
  for (int32_t bx = (h->n[0] + 63) / 64 * 64, k = 0; k < 2 && bx <= 512; bx += 64, k++) {
    if (bx == h->n[0] || h->n[0] < 16) continue;
    for (int occ = 0; occ < nocc; occ++) {
      cands.push_back({bx, 1, 1, 1, occ});
      cands.push_back({bx, 1, 1, 0, occ});
    }
  }
  if (cands.empty()) return fail(LBM_EINVAL, "no block shape fits this lattice");
  // The trial kernels step a copy of the populations back and forth between
  // two scratch buffers, as a run does: stepping the same input again and
  // again would leave it in the last-level cache (the streamed stores do not
  // evict it) and time a rate no run reaches.
  const size_t bytes = (size_t)h->N * h->Q * elem_size(h->cfg.storage);
  char *scratch = nullptr;
  HIP_TRY(hipMalloc(&scratch, 2 * bytes), "hipMalloc autotune scratch");
  // A run's first kernel collides without gathering and records the rho
  // that the first gathering kernel's links read (rho_side); the trials
  // start the same way, from a copy of the populations in the second
  // scratch buffer into the first, so that they read no rho that no kernel
  // has written (a handle that never ran holds none).
  hipError_t ce = hipMemcpyAsync(scratch + bytes, h->f[h->cur], bytes, hipMemcpyDeviceToDevice, h->stream);
  if (ce == hipSuccess) ce = launch_mode(h, kCollide, scratch + bytes, scratch, h->pcur);
  if (ce != hipSuccess) {
    hipStreamSynchronize(h->stream);
    hipFree(scratch);
    return hip_fail(ce, "autotune copy");
  }
  // The trials move the particles as a run does: each trial kernel applies
  // the move the one before it left pending, so the box blocks' whole chain
  // is timed.  Their records alternate parity from the handle's own, which
  // are kept aside and put back before every candidate and at the end; the
  // trials write nothing to the per-tick log.  (With more than
  // kLocalParticles particles the moves belong to the move kernel, and every
  // trial reads the handle's records.)
  const size_t rec_bytes = h->parts_h.size() * sizeof(DevParticle);
  const bool moving = rec_bytes > 0 && !many_particles(h);
  DevParticle *recs_kept = nullptr;
  if (moving) {
    hipError_t ce = hipMalloc(&recs_kept, rec_bytes);
    if (ce == hipSuccess)
      ce = hipMemcpyAsync(recs_kept, parts_of(h, h->pcur), rec_bytes, hipMemcpyDeviceToDevice, h->stream);
    if (ce != hipSuccess) {
      if (recs_kept) hipFree(recs_kept);
      hipFree(scratch);
      return hip_fail(ce, "autotune record copy");
    }
  }
  auto put_back = [&]() {
    hipError_t ce = hipSuccess;
    if (moving) ce = hipMemcpyAsync(parts_of(h, h->pcur), recs_kept, rec_bytes, hipMemcpyDeviceToDevice, h->stream);
    return ce;
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const Tiling saved = h->tiling;
  double best_t = 1e300;
  Cand best_c = cands[0];
  hipError_t e = hipSuccess;
  auto release = [&]() {
    hipStreamSynchronize(h->stream);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (recs_kept) hipFree(recs_kept);
    hipFree(scratch);
  };
  // One candidate's time per fused kernel over `nrep` trial kernels (after
  // two warm-up ones); rc is set when the box plan cannot be made.
  int rc = LBM_OK;
  auto timed = [&](const Cand &c, int nrep) -> double {
    Tiling t = make_tiling(h, c.bx, c.by, c.bz);
    t.remap = c.remap;
    t.occ = c.occ;
    if ((rc = plan_boxes(h, tile_threads(t)))) return 1e300;
    h->tiling = t;
    // fused kernels between the two scratch buffers (an even count of
    // warm-up steps, so the timed ones start from the copy's buffer), the
    // first from the handle's records
    e = put_back();
    int side = 0;
    auto trial = [&]() {
      char *from = scratch + side * bytes;
      StepArgs a = step_args(h, from, scratch + (side ^ 1) * bytes, moving ? h->pcur ^ side : h->pcur);
      a.sp.traj = nullptr;
      side ^= 1;
      return launch_args(h, kFused, a);
    };
    for (int r = 0; r < 2 && e == hipSuccess; r++) e = trial();  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
    for (int r = 0; r < nrep && e == hipSuccess; r++) e = trial();
    if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    return e == hipSuccess ? (double)ms / nrep : 1e300;
  };
  // A first pass over every candidate, then the best few again, three times
  // in turn with more kernels each, ranked by their median: single passes
  // of a few kernels vary by a few per cent from one to the next, as much as
  // the best shapes differ.
  std::vector<std::pair<double, size_t>> first;
  for (size_t i = 0; i < cands.size() && e == hipSuccess && rc == LBM_OK; i++) first.push_back({timed(cands[i], reps), i});
  if (e == hipSuccess && rc == LBM_OK) {
    std::stable_sort(first.begin(), first.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    const size_t top = std::min<size_t>(first.size(), 6);
    const int nrep = std::max(reps, 30);
    std::vector<std::vector<double>> again(top);
    for (int round = 0; round < 3 && e == hipSuccess && rc == LBM_OK; round++)
      for (size_t k = 0; k < top && e == hipSuccess && rc == LBM_OK; k++)
        again[k].push_back(timed(cands[first[k].second], nrep));
    for (size_t k = 0; k < top && e == hipSuccess && rc == LBM_OK; k++) {
      std::sort(again[k].begin(), again[k].end());
      const double med = again[k][again[k].size() / 2];
      if (med < best_t) {
        best_t = med;
        best_c = cands[first[k].second];
      }
    }
  }
  if (rc) {
    h->tiling = saved;
    plan_boxes(h, tile_threads(saved));
    put_back();
    release();
    return rc;
  }
  if (e == hipSuccess) e = put_back();
  else put_back();
  release();
  if (e != hipSuccess) {
    h->tiling = saved;
    plan_boxes(h, tile_threads(saved));
    return hip_fail(e, "autotune");
  }
  h->tiling = make_tiling(h, best_c.bx, best_c.by, best_c.bz);
  h->tiling.remap = best_c.remap;
  h->tiling.occ = best_c.occ;
  if (const int rc = plan_boxes(h, tile_threads(h->tiling))) return rc;
  drop_graph(h);
  if (best) {
    best[0] = best_c.bx;
    best[1] = best_c.by;
    best[2] = best_c.bz;
  }
  if (best_ms) *best_ms = best_t;
  return LBM_OK;
}

// Device-to-device copy of `bytes` bytes (multiple of 16), `reps` times on a
// private stream; *ms receives the average time per copy (HIP events).
int lbm_measure_copy(int64_t bytes, int32_t reps, double *ms) {
  if (!ms || bytes < 16 || bytes % 16 || reps < 1) return fail(LBM_EINVAL, "bad copy size or repetitions");
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, bytes);
  if (e == hipSuccess) e = hipMalloc(&b, bytes);
  if (e == hipSuccess) e = hipMemset(a, 1, bytes);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_copy(a, b, bytes, s);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  // back and forth, as the step's populations go
  for (int r = 0; r < reps && e == hipSuccess; r++) e = (r & 1) ? launch_copy(b, a, bytes, s) : launch_copy(a, b, bytes, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (s) hipStreamDestroy(s);
  if (a) hipFree(a);
  if (b) hipFree(b);
  if (e != hipSuccess) return hip_fail(e, "copy measurement");
  *ms = t / reps;
  return LBM_OK;
}

int lbm_measure_pull(int32_t storage, int32_t nx, int32_t ny, int32_t nz, int32_t reps, double *ms) {
  if (!ms || reps < 1 || nx < 1 || ny < 1 || nz < 1) return fail(LBM_EINVAL, "bad lattice or repetitions");
  if (storage < LBM_STORAGE_F64 || storage > LBM_STORAGE_F16C) return fail(LBM_EINVAL, "unknown storage");
  const int eb = (int)elem_size(storage);
  const int64_t bytes = (int64_t)nx * ny * nz * 19 * eb;
  if ((int64_t)ny * nz * ((nx + 255) / 256) > INT32_MAX) return fail(LBM_EINVAL, "lattice too large");
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, bytes);
  if (e == hipSuccess) e = hipMalloc(&b, bytes);
  if (e == hipSuccess) e = hipMemset(a, 0, bytes);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_pull(eb, a, b, nx, ny, nz, s);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  // back and forth, as the step's populations go
  for (int r = 0; r < reps && e == hipSuccess; r++)
    e = (r & 1) ? launch_pull(eb, b, a, nx, ny, nz, s) : launch_pull(eb, a, b, nx, ny, nz, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (s) hipStreamDestroy(s);
  if (a) hipFree(a);
  if (b) hipFree(b);
  if (e != hipSuccess) return hip_fail(e, "pull measurement");
  *ms = t / reps;
  return LBM_OK;
}

// Current block shape of the step kernel.
int lbm_get_block(lbm_handle *h, int32_t *bx, int32_t *by, int32_t *bz) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  if (bx) *bx = h->tiling.bx;
  if (by) *by = h->tiling.by;
  if (bz) *bz = h->tiling.bz;
  return LBM_OK;
}

// Order of the step kernel's tiles over the grid: 1 gives each of the 8
// XCDs one contiguous range of tiles, 0 the plain order.
int lbm_set_tile_order(lbm_handle *h, int32_t xcd_remap) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  h->tiling.remap = xcd_remap != 0;
  drop_graph(h);
  return LBM_OK;
}

int lbm_get_tile_order(lbm_handle *h, int32_t *xcd_remap) {
  if (!h || !xcd_remap) return fail(LBM_EINVAL, "null argument");
  *xcd_remap = h->tiling.remap;
  return LBM_OK;
}

int lbm_set_occupancy(lbm_handle *h, int32_t low) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  h->tiling.occ = low != 0;
  drop_graph(h);
  return LBM_OK;
}

int lbm_get_occupancy(lbm_handle *h, int32_t *low) {
  if (!h || !low) return fail(LBM_EINVAL, "null argument");
  *low = h->tiling.occ;
  return LBM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batches
// A batch advances independent lattices together: every step kernel of a
// run covers all of them in one launch (batch_step_kernel), the launches of
// a run are the one-lattice run's (collide, fused..., gather, finalize), and
// the fused ones replay from captured graphs.  Each kernel reads lattice
// i's arguments from a device table, so the graphs stay valid while the
// lattices swap buffers.

// Sections of a batch's argument table, n entries each.
enum BatchSection { kSecCollide = 0, kSecStep0 = 1, kSecStep1 = 2, kSecFinal0 = 3, kSecFinal1 = 4, kSections = 5 };

struct lbm_batch {
  std::vector<lbm_handle *> boxes;
  int Q = 0, storage = 0, coll = 0, force = 0;
  Tiling shape{};  // block shape of every box (bx, by, bz) unless row tiles, tile order, occupancy
  int32_t rows = 0;  // > 0: every box takes row tiles (row_tile) of this many threads
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  // argument tables, one per slot (box 0's population buffer at run start)
  StepArgs *args_d[2] = {nullptr, nullptr};
  std::vector<StepArgs> args_h[2];  // what the device tables hold
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int64_t graph_k = 32;
  bool timing = false;
  std::vector<hipEvent_t> ev;
  hipEvent_t run_start = nullptr, run_stop = nullptr;
  double step_kernel_ms = 0.0, step_ms = 0.0;
};

namespace {

int n_boxes(const lbm_batch *b) { return (int)b->boxes.size(); }

void drop_batch_graphs(lbm_batch *b) {
  for (int s = 0; s < 2; s++) {
    if (b->graph[s]) hipGraphExecDestroy(b->graph[s]);
    b->graph[s] = nullptr;
  }
}

// Tile of `threads` threads for box h whose x extent spans whole rows: bx
// the least divisor of `threads` of at least 16 that holds a row (the whole
// block when none does), by the largest divisor of the rest that fits the
// box's y extent, bz what remains (tiles overhanging the box leave threads
// idle).  A row of the box then lies in lines its own tile reads and
// writes, as in the one-lattice kernel's default tiles, whatever the other
// boxes' sizes.
void row_tile(const lbm_handle *h, int32_t threads, int32_t &bx, int32_t &by, int32_t &bz) {
  bx = threads;
  for (int32_t d = 16; d <= threads; d++)
    if (threads % d == 0 && d >= h->n[0]) {
      bx = d;
      break;
    }
  const int32_t rest = threads / bx;
  by = 1;
  for (int32_t d = 1; d <= rest; d++)
    if (rest % d == 0 && d <= h->n[1]) by = d;
  bz = rest / by;
}

// Box h's tile shape in batch b, with the batch's tile order and occupancy.
Tiling lattice_shape(const lbm_batch *b, const lbm_handle *h) {
  Tiling t = b->shape;
  if (b->rows > 0) row_tile(h, b->rows, t.bx, t.by, t.bz);
  return t;
}

// Give box h the block shape `shape` (and box-block plan for it).
int adopt_shape(lbm_handle *h, const Tiling &shape) {
  if (h->tiling.bx == shape.bx && h->tiling.by == shape.by && h->tiling.bz == shape.bz &&
      h->tiling.remap == shape.remap)
    return LBM_OK;
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  Tiling t = make_tiling(h, shape.bx, shape.by, shape.bz);
  t.remap = shape.remap;
  if (const int rc = plan_boxes(h, tile_threads(t))) return rc;
  h->tiling = t;
  drop_graph(h);
  return LBM_OK;
}

// Box i's arguments for every section, from its current buffers and records:
// collide reads f[cur]; fused (and gather) kernel k of the run reads
// f[cur ^ 1 ^ (k & 1)] with records pcur ^ (k & 1), as lbm_run does; the
// finalize after a run whose fused count has parity ph works in place on
// records pcur ^ ph ^ 1.
void box_sections(lbm_handle *h, StepArgs out[kSections]) {
  const int c = h->cur, p = h->pcur;
  out[kSecCollide] = step_args(h, h->f[c], h->f[c ^ 1], p);
  for (int ph = 0; ph < 2; ph++) {
    out[kSecStep0 + ph] = step_args(h, h->f[c ^ 1 ^ ph], h->f[c ^ ph], p ^ ph);
    StepArgs fa = step_args(h, nullptr, nullptr, p ^ ph ^ 1);
    fa.sp.parts_out = parts_of(h, p ^ ph ^ 1);
    out[kSecFinal0 + ph] = fa;
  }
}

// Refresh slot s's device table from the boxes' current state (a copy only
// when something changed, after the batch's earlier work is done).
int upload_table(lbm_batch *b, int s) {
  const int n = n_boxes(b);
  std::vector<StepArgs> t((size_t)n * kSections);
  for (int i = 0; i < n; i++) {
    StepArgs sec[kSections];
    box_sections(b->boxes[i], sec);
    for (int k = 0; k < kSections; k++) t[(size_t)k * n + i] = sec[k];
  }
  if (b->args_h[s].size() == t.size() && std::memcmp(b->args_h[s].data(), t.data(), t.size() * sizeof(StepArgs)) == 0)
    return LBM_OK;
  HIP_TRY(hipStreamSynchronize(b->stream), "sync");
  if (!b->args_h[s].empty()) drop_batch_graphs(b);  // (block counts may have changed)
  b->args_h[s] = std::move(t);
  // stream-ordered; the host copy stays untouched until the stream is idle
  HIP_TRY(hipMemcpyAsync(b->args_d[s], b->args_h[s].data(), b->args_h[s].size() * sizeof(StepArgs),
                         hipMemcpyHostToDevice, b->stream),
          "copy batch table");
  return LBM_OK;
}

// Lattices by which a box segment leads its tiles in a batched grid: a
// sixteenth of the batch (64 boxes of 48-96 nodes a side with one squirmer
// each, row tiles of 192 threads: 0.7505 ms per fused kernel leading by 32
// lattices, 0.7429 by 16, 0.7364 by 10, 0.7327 by 8, 0.7288 by 5, 0.7272
// by 4, 0.7436 with each box segment right before its tiles; with one
// shared 32 x 1 x 2 shape the best lead had been 8).
int batch_lead(int n) { return n / 16; }

// Grid of a batched kernel over the boxes' arguments `host` (one section):
// the box blocks of lattice i + lead, then the tiles of lattice i, for i =
// -lead .. n - 1 (segments that do not exist are left out).  Every segment
// starts at a multiple of 8, so each box's tiles keep the XCD order the
// one-lattice kernel gives them.
Launch batch_launch(lbm_batch *b, const StepArgs *host, const StepArgs *dev, int mode, BatchArgs &ba) {
  const int n = n_boxes(b);
  const int lead = batch_lead(n);
  ba.boxes = dev;
  ba.n = 0;
  ba.start[0] = 0;
  size_t shmem = 0;
  auto push = [&](int lat, int tiles, int blocks) {
    if (blocks == 0) return;
    ba.seg[ba.n] = (uint16_t)(lat << 1 | tiles);
    ba.start[ba.n + 1] = ba.start[ba.n] + (blocks + 7) / 8 * 8;
    ba.n++;
  };
  for (int i = -lead; i < n; i++) {
    const int j = i + lead;
    if (j < n) {
      const int nbox = mode == kCollide ? 0 : host[j].sp.nbox_pad;
      if (nbox) shmem = std::max(shmem, box_shmem(b->boxes[j], host[j]));
      push(j, 0, nbox);
    }
    if (i >= 0) push(i, 1, host[i].t.ntiles);
  }
  const int threads = b->rows > 0 ? b->rows : b->shape.bx * b->shape.by * b->shape.bz;
  Launch l{dim3((unsigned)ba.start[ba.n]), dim3((unsigned)threads), b->stream};
  l.shmem = shmem;
  return l;
}

hipError_t launch_batch_mode(lbm_batch *b, int mode, const StepArgs *host, const StepArgs *dev) {
  BatchArgs ba;
  Launch l = batch_launch(b, host, dev, mode, ba);
  switch (b->storage) {
    case LBM_STORAGE_F64: return launch_batch_step<kF64>(b->Q, b->coll, b->force, mode, b->shape.occ, ba, l);
    case LBM_STORAGE_F32: return launch_batch_step<kF32>(b->Q, b->coll, b->force, mode, b->shape.occ, ba, l);
    case LBM_STORAGE_F16C: return launch_batch_step<kF16C>(b->Q, b->coll, b->force, mode, b->shape.occ, ba, l);
    default: return launch_batch_step<kF16>(b->Q, b->coll, b->force, mode, b->shape.occ, ba, l);
  }
}

// Section k of slot s: host copy and device address.
const StepArgs *sec_host(const lbm_batch *b, int s, int k) { return b->args_h[s].data() + (size_t)k * n_boxes(b); }
const StepArgs *sec_dev(const lbm_batch *b, int s, int k) { return b->args_d[s] + (size_t)k * n_boxes(b); }

hipError_t batch_fused(lbm_batch *b, int s, int ph) {
  return launch_batch_mode(b, kFused, sec_host(b, s, kSecStep0 + ph), sec_dev(b, s, kSecStep0 + ph));
}

int capture_batch_graph(lbm_batch *b, int s, hipGraphExec_t *out) {
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin capture");
  hipError_t e = hipSuccess;
  for (int64_t i = 0; i < b->graph_k && e == hipSuccess; i++) e = batch_fused(b, s, (int)(i & 1));
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) hipGraphDestroy(g);
    return hip_fail(e != hipSuccess ? e : e2, "capture batch graph");
  }
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail(e, "instantiate batch graph");
  return LBM_OK;
}

bool batch_has_particles(const lbm_batch *b) {
  for (const lbm_handle *h : b->boxes)
    if (!h->parts_h.empty()) return true;
  return false;
}

// Every box populated and at the batch's block shape, and its own stream idle.
int prepare_boxes(lbm_batch *b) {
  for (lbm_handle *h : b->boxes) {
    if (!h->populated) return fail(LBM_ESTATE, "a box of the batch has no populations yet");
    if (h->parts_h.size() > (size_t)kLocalParticles)
      return fail(LBM_EINVAL, "a box of a batch takes at most " + std::to_string(kLocalParticles) +
                                  " particles; run a box with more alone");
    if (const int rc = adopt_shape(h, lattice_shape(b, h))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  }
  return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_batch_destroy(lbm_batch *b) {
  if (!b) return LBM_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  drop_batch_graphs(b);
  for (int s = 0; s < 2; s++)
    if (b->args_d[s]) hipFree(b->args_d[s]);
  for (auto e : b->ev) hipEventDestroy(e);
  if (b->run_start) hipEventDestroy(b->run_start);
  if (b->run_stop) hipEventDestroy(b->run_stop);
  if (b->done) hipEventDestroy(b->done);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return LBM_OK;
}

int lbm_batch_create(lbm_handle *const *boxes, int32_t n, lbm_batch **out) {
  if (!boxes || !out) return fail(LBM_EINVAL, "null argument");
  *out = nullptr;
  if (n < 1 || n > kMaxBatch) return fail(LBM_EINVAL, "a batch holds 1.." + std::to_string(kMaxBatch) + " boxes");
  for (int i = 0; i < n; i++) {
    if (!boxes[i]) return fail(LBM_EINVAL, "null box handle");
    for (int j = 0; j < i; j++)
      if (boxes[j] == boxes[i]) return fail(LBM_EINVAL, "a box appears twice in the batch");
    const lbm_config &c0 = boxes[0]->cfg, &c = boxes[i]->cfg;
    if (c.q != c0.q || c.storage != c0.storage || c.collision != c0.collision || c.force != c0.force)
      return fail(LBM_EINVAL, "the boxes of a batch share lattice, storage, collision and force kinds");
  if (boxes[0]->Q != 9 && boxes[0]->Q != 19) return fail(LBM_EINVAL, "batches run D2Q9 and D3Q19 lattices");
  }
  lbm_batch *b = new (std::nothrow) lbm_batch;
  if (!b) return fail(LBM_ENOMEM, "out of host memory");
  b->boxes.assign(boxes, boxes + n);
  b->Q = boxes[0]->Q;
  b->storage = boxes[0]->cfg.storage;
  b->coll = boxes[0]->cfg.collision;
  b->force = boxes[0]->cfg.force;
  auto bail = [&](int code) {
    lbm_batch_destroy(b);
    return code;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess ||
      hipEventCreate(&b->run_start) != hipSuccess || hipEventCreate(&b->run_stop) != hipSuccess)
    return bail(fail(LBM_EDEVICE, "create batch stream and events"));
  for (int s = 0; s < 2; s++)
    if (hipMalloc(&b->args_d[s], (size_t)n * kSections * sizeof(StepArgs)) != hipSuccess)
      return bail(fail(LBM_ENOMEM, "hipMalloc batch tables"));
  // default shape: box 0's current one, if every box admits it
  Tiling t = boxes[0]->tiling;
  for (int i = 0; i < n; i++)
    if (t.bx > boxes[i]->n[0] || t.by > boxes[i]->n[1] || t.bz > boxes[i]->n[2]) t = Tiling{64, 1, 1, 0, 0, 0, 0, 1};
  b->shape = t;
  for (lbm_handle *h : b->boxes)
    if (const int rc = adopt_shape(h, lattice_shape(b, h))) return bail(rc);
  *out = b;
  return LBM_OK;
}

int lbm_batch_run(lbm_batch *b, int64_t nsteps) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  if (nsteps < 0) return fail(LBM_EINVAL, "nsteps must be non-negative");
  if (nsteps == 0) return LBM_OK;
  for (const lbm_handle *h : b->boxes)
    if (const int rc = check_schedule(h, nsteps)) return rc;
  if (const int rc = prepare_boxes(b)) return rc;
  const int s = b->boxes[0]->cur;
  if (const int rc = upload_table(b, s)) return rc;
  const int64_t nfused = nsteps - 1;
  if (b->timing && nfused > 0 && (int64_t)b->ev.size() < 2 * nfused) {
    const size_t old = b->ev.size();
    b->ev.resize(2 * nfused);
    for (size_t i = old; i < b->ev.size(); i++) HIP_TRY(hipEventCreate(&b->ev[i]), "hipEventCreate");
  }
  HIP_TRY(hipEventRecord(b->run_start, b->stream), "event record");
  hipError_t e = launch_batch_mode(b, kCollide, sec_host(b, s, kSecCollide), sec_dev(b, s, kSecCollide));
  if (e != hipSuccess) return hip_fail(e, "batch step launch");
  int64_t k = 0;
  if (!b->timing && b->graph_k > 0) {
    while (nfused - k >= b->graph_k) {
      if (!b->graph[s])
        if (const int rc = capture_batch_graph(b, s, &b->graph[s])) return rc;
      HIP_TRY(hipGraphLaunch(b->graph[s], b->stream), "graph launch");
      k += b->graph_k;
    }
  }
  for (int64_t i = 0; k < nfused; k++, i++) {
    if (b->timing) hipEventRecord(b->ev[2 * i], b->stream);
    e = batch_fused(b, s, (int)(k & 1));
    if (e != hipSuccess) return hip_fail(e, "batch step launch");
    if (b->timing) hipEventRecord(b->ev[2 * i + 1], b->stream);
  }
  const int ph = (int)(nfused & 1);
  e = launch_batch_mode(b, kGather, sec_host(b, s, kSecStep0 + ph), sec_dev(b, s, kSecStep0 + ph));
  if (e != hipSuccess) return hip_fail(e, "batch step launch");
  if (batch_has_particles(b)) {
    e = launch_finalize_batch(sec_dev(b, s, kSecFinal0 + ph), n_boxes(b), b->stream);
    if (e != hipSuccess) return hip_fail(e, "finalize launch");
  }
  HIP_TRY(hipEventRecord(b->run_stop, b->stream), "event record");
  HIP_TRY(hipEventRecord(b->done, b->stream), "event record");
  for (lbm_handle *h : b->boxes) {
    // the boxes' own later work waits for the batch
    HIP_TRY(hipStreamWaitEvent(h->stream, b->done, 0), "stream wait");
    // as after lbm_run: nsteps + 1 buffer swaps, nsteps record swaps
    h->cur ^= (int)((nsteps + 1) & 1);
    h->pcur ^= (int)(nsteps & 1);
    h->tick += nsteps;
    for (int64_t i = 0; i < nsteps; i++) h->time += h->cfg.dt;
  }
  if (b->timing) {
    HIP_TRY(hipEventSynchronize(b->run_stop), "sync");
    double sum = 0.0;
    for (int64_t i = 0; i < nfused; i++) {
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, b->ev[2 * i], b->ev[2 * i + 1]), "event time");
      sum += ms;
    }
    float tot = 0.0f;
    HIP_TRY(hipEventElapsedTime(&tot, b->run_start, b->run_stop), "event time");
    b->step_kernel_ms = nfused > 0 ? sum / (double)nfused : 0.0;
    b->step_ms = tot / (double)nsteps;
  }
  return LBM_OK;
}

int lbm_batch_sync(lbm_batch *b) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  HIP_TRY(hipStreamSynchronize(b->stream), "sync");
  for (lbm_handle *h : b->boxes) {
    HIP_TRY(hipStreamSynchronize(h->stream), "sync");
    if (const int rc = check_particles(h)) return rc;
  }
  return LBM_OK;
}

int lbm_batch_set_block(lbm_batch *b, int32_t bx, int32_t by, int32_t bz) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  const int64_t nt = (int64_t)bx * by * bz;
  if (bx < 1 || by < 1 || bz < 1 || nt < 64 || nt > 512 || nt % 64)
    return fail(LBM_EINVAL, "block shape must have 64..512 threads, a multiple of 64");
  HIP_TRY(hipStreamSynchronize(b->stream), "sync");
  Tiling t{};
  t.bx = bx;
  t.by = by;
  t.bz = bz;
  t.remap = 1;
  t.occ = b->shape.occ;  // (the occupancy variant stays)
  b->shape = t;
  b->rows = 0;
  drop_batch_graphs(b);
  for (lbm_handle *h : b->boxes)
    if (const int rc = adopt_shape(h, lattice_shape(b, h))) return rc;
  return LBM_OK;
}

int lbm_batch_set_row_tiles(lbm_batch *b, int32_t threads) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  if (threads < 64 || threads > 512 || threads % 64)
    return fail(LBM_EINVAL, "row tiles take 64..512 threads, a multiple of 64");
  HIP_TRY(hipStreamSynchronize(b->stream), "sync");
  b->rows = threads;
  drop_batch_graphs(b);
  for (lbm_handle *h : b->boxes)
    if (const int rc = adopt_shape(h, lattice_shape(b, h))) return rc;
  return LBM_OK;
}

int lbm_batch_get_row_tiles(lbm_batch *b, int32_t *threads) {
  if (!b || !threads) return fail(LBM_EINVAL, "null argument");
  *threads = b->rows;
  return LBM_OK;
}

int lbm_batch_set_tile_order(lbm_batch *b, int32_t xcd_remap) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  HIP_TRY(hipStreamSynchronize(b->stream), "sync");
  b->shape.remap = xcd_remap != 0;
  drop_batch_graphs(b);
  for (lbm_handle *h : b->boxes)
    if (const int rc = adopt_shape(h, lattice_shape(b, h))) return rc;
  return LBM_OK;
}

int lbm_batch_get_tile_order(lbm_batch *b, int32_t *xcd_remap) {
  if (!b || !xcd_remap) return fail(LBM_EINVAL, "null argument");
  *xcd_remap = b->shape.remap;
  return LBM_OK;
}

int lbm_batch_set_occupancy(lbm_batch *b, int32_t low) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  HIP_TRY(hipStreamSynchronize(b->stream), "sync");
  b->shape.occ = low != 0;
  drop_batch_graphs(b);
  return LBM_OK;
}

int lbm_batch_get_occupancy(lbm_batch *b, int32_t *low) {
  if (!b || !low) return fail(LBM_EINVAL, "null argument");
  *low = b->shape.occ;
  return LBM_OK;
}

int lbm_batch_get_block(lbm_batch *b, int32_t *bx, int32_t *by, int32_t *bz) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  if (bx) *bx = b->shape.bx;
  if (by) *by = b->shape.by;
  if (bz) *bz = b->shape.bz;
  return LBM_OK;
}

int lbm_batch_set_graph(lbm_batch *b, int32_t ticks_per_graph) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  if (ticks_per_graph < 0 || (ticks_per_graph & 1)) return fail(LBM_EINVAL, "ticks per graph must be 0 or an even number");
  HIP_TRY(hipStreamSynchronize(b->stream), "sync");
  drop_batch_graphs(b);
  b->graph_k = ticks_per_graph;
  return LBM_OK;
}

int lbm_batch_set_timing(lbm_batch *b, int on) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  b->timing = on != 0;
  return LBM_OK;
}

int lbm_batch_last_run_times(lbm_batch *b, double *kernel_ms, double *step_ms) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  if (kernel_ms) *kernel_ms = b->step_kernel_ms;
  if (step_ms) *step_ms = b->step_ms;
  return LBM_OK;
}

// Autotuner of a batch: times the batched fused kernel (stepping a copy of
// every box's populations back and forth between two scratch buffers, as a
// run does; the records and partials it also writes are the ones every run
// rewrites before reading) over the block shapes that fit every box and both
// tile orders, and keeps the fastest.
int lbm_batch_autotune(lbm_batch *b, int32_t reps, int32_t best[3], double *best_ms) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  if (reps < 1) reps = 1;
  if (const int rc = prepare_boxes(b)) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream), "sync");
  const int n = n_boxes(b);
  int32_t mx[3] = {1, 1, 1};
  int64_t nodes = 0;
  size_t scratch_bytes = 0;
  std::vector<size_t> off(n);
  std::vector<int32_t> bxs = {16, 32, 64, 96, 128, 192, 256, 512};
  for (int i = 0; i < n; i++) {
    const lbm_handle *h = b->boxes[i];
    for (int k = 0; k < 3; k++) mx[k] = std::max(mx[k], h->n[k]);
    if (h->n[0] <= 512) bxs.push_back(h->n[0]);
    nodes += h->N;
    off[i] = scratch_bytes;
    scratch_bytes += ((size_t)h->N * h->Q * elem_size(b->storage) + 255) / 256 * 256;
  }
  std::sort(bxs.begin(), bxs.end());
  bxs.erase(std::unique(bxs.begin(), bxs.end()), bxs.end());
  // shapes whose idle threads (tiles overhanging the boxes) stay within 25%
  // of the nodes, or within 15% of the least any shape achieves
  // one shape for every box (rows = 0), or row tiles of `rows` threads
  struct Shape {
    int32_t bx, by, bz, rows;
    double waste;
  };
  std::vector<Shape> shapes;
  double least = 1e300;
  auto add_shape = [&](int32_t bx, int32_t by, int32_t bz, int32_t rows) {
    int64_t covered = 0;
    for (const lbm_handle *h : b->boxes) {
      int32_t x = bx, y = by, z = bz;
      if (rows) row_tile(h, rows, x, y, z);
      covered += (int64_t)make_tiling(h, x, y, z).ntiles * x * y * z;
    }
    const double waste = (double)covered / (double)nodes - 1.0;
    shapes.push_back({bx, by, bz, rows, waste});
    least = std::min(least, waste);
  };
  for (int32_t bx : bxs)
    for (int32_t by : {1, 2, 4, 8, 16})
      for (int32_t bz : {1, 2, 4}) {
        const int64_t t = (int64_t)bx * by * bz;
        if (t < 64 || t > 512 || t % 64 || bx > mx[0] || by > mx[1] || bz > mx[2]) continue;
        add_shape(bx, by, bz, 0);
      }
  for (int32_t rows = 64; rows <= 512; rows += 64) add_shape(0, 0, 0, rows);
  struct Cand {
    Tiling t;
    int32_t rows;
  };
  std::vector<Cand> cands;
  for (const Shape &sh : shapes) {
    if (sh.waste > std::max(0.25, least + 0.15)) continue;
    for (int occ = 0; occ < 2; occ++)
      for (int remap : {1, 0}) {
        Cand c{};
        c.t.bx = sh.rows ? b->shape.bx : sh.bx;
        c.t.by = sh.rows ? b->shape.by : sh.by;
        c.t.bz = sh.rows ? b->shape.bz : sh.bz;
        c.t.remap = remap;
        c.t.occ = occ;
        c.rows = sh.rows;
        cands.push_back(c);
      }
  }
  if (cands.empty()) return fail(LBM_EINVAL, "no block shape fits the boxes");
  char *scratch = nullptr;
  StepArgs *trial_d = nullptr;  // two tables: buffer 0 -> 1, then 1 -> 0
  HIP_TRY(hipMalloc(&scratch, 2 * scratch_bytes), "hipMalloc autotune scratch");
  if (hipMalloc(&trial_d, 2 * (size_t)n * sizeof(StepArgs)) != hipSuccess) {
    hipFree(scratch);
    return fail(LBM_ENOMEM, "hipMalloc autotune table");
  }
  // As a run does, the trials start with a collide-only kernel, from a copy
  // of each box's populations in the second scratch buffer into the first,
  // which records the rho the first gathering kernel's links read.
  {
    hipError_t ce = hipSuccess;
    std::vector<StepArgs> col((size_t)n);
    for (int i = 0; i < n && ce == hipSuccess; i++) {
      lbm_handle *h = b->boxes[i];
      char *s0 = scratch + off[i], *s1 = s0 + scratch_bytes;
      ce = hipMemcpyAsync(s1, h->f[h->cur], (size_t)h->N * h->Q * elem_size(b->storage), hipMemcpyDeviceToDevice,
                          b->stream);
      col[i] = step_args(h, s1, s0, h->pcur);
    }
    if (ce == hipSuccess)
      ce = hipMemcpyAsync(trial_d, col.data(), (size_t)n * sizeof(StepArgs), hipMemcpyHostToDevice, b->stream);
    if (ce == hipSuccess) ce = launch_batch_mode(b, kCollide, col.data(), trial_d);
    if (ce == hipSuccess) ce = hipStreamSynchronize(b->stream);
    if (ce != hipSuccess) {
      hipStreamSynchronize(b->stream);
      hipFree(trial_d);
      hipFree(scratch);
      return hip_fail(ce, "batch autotune copy");
    }
  }
  // The trials move the particles as runs do (records alternating parity
  // from each box's own, kept aside and put back before every candidate and
  // at the end; nothing goes to the per-tick logs), so every box's whole
  // chain of box steps, the pending move included, is timed.
  std::vector<size_t> rec_off(n + 1, 0);
  for (int i = 0; i < n; i++) rec_off[i + 1] = rec_off[i] + b->boxes[i]->parts_h.size();
  DevParticle *recs_kept = nullptr;
  if (rec_off[n] > 0) {
    hipError_t ce = hipMalloc(&recs_kept, rec_off[n] * sizeof(DevParticle));
    for (int i = 0; i < n && ce == hipSuccess; i++) {
      lbm_handle *h = b->boxes[i];
      const size_t nb = h->parts_h.size() * sizeof(DevParticle);
      if (nb) ce = hipMemcpyAsync(recs_kept + rec_off[i], parts_of(h, h->pcur), nb, hipMemcpyDeviceToDevice, b->stream);
    }
    if (ce != hipSuccess) {
      if (recs_kept) hipFree(recs_kept);
      hipFree(trial_d);
      hipFree(scratch);
      return hip_fail(ce, "batch autotune record copy");
    }
  }
  auto put_back = [&]() {
    hipError_t ce = hipSuccess;
    for (int i = 0; i < n && ce == hipSuccess; i++) {
      lbm_handle *h = b->boxes[i];
      const size_t nb = h->parts_h.size() * sizeof(DevParticle);
      if (nb) ce = hipMemcpyAsync(parts_of(h, h->pcur), recs_kept + rec_off[i], nb, hipMemcpyDeviceToDevice, b->stream);
    }
    return ce;
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const Cand saved{b->shape, b->rows};
  double best_t = 1e300;
  Cand best_c = cands[0];
  int rc = LBM_OK;
  hipError_t e = hipSuccess;
  std::vector<StepArgs> trial(2 * (size_t)n);
  // One candidate's time per batched fused kernel over `nrep` trial kernels
  // (after two warm-up ones).
  auto timed = [&](const Cand &c, int nrep) -> double {
    b->shape = c.t;
    b->rows = c.rows;
    for (lbm_handle *h : b->boxes)
      if ((rc = adopt_shape(h, lattice_shape(b, h)))) return 1e300;
    for (int i = 0; i < n; i++) {
      lbm_handle *h = b->boxes[i];
      char *s0 = scratch + off[i], *s1 = s0 + scratch_bytes;
      trial[i] = step_args(h, s0, s1, h->pcur);
      trial[n + i] = step_args(h, s1, s0, h->pcur ^ 1);
      trial[i].sp.traj = nullptr;
      trial[n + i].sp.traj = nullptr;
    }
    e = hipMemcpyAsync(trial_d, trial.data(), 2 * (size_t)n * sizeof(StepArgs), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = put_back();
    // (an even count of warm-up steps, so the timed ones start from buffer 0)
    int side = 0;
    auto step = [&]() {
      const hipError_t le = launch_batch_mode(b, kFused, trial.data() + side * n, trial_d + side * n);
      side ^= 1;
      return le;
    };
    for (int r = 0; r < 2 && e == hipSuccess; r++) e = step();
    if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
    for (int r = 0; r < nrep && e == hipSuccess; r++) e = step();
    if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    return e == hipSuccess ? (double)ms / nrep : 1e300;
  };
  // every candidate once, then the best few three times in turn with more
  // kernels each, ranked by their median (single passes vary by a few per
  // cent, as much as the best shapes differ)
  std::vector<std::pair<double, size_t>> first;
  for (size_t i = 0; i < cands.size() && e == hipSuccess && rc == LBM_OK; i++) first.push_back({timed(cands[i], reps), i});
  if (e == hipSuccess && rc == LBM_OK) {
    std::stable_sort(first.begin(), first.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    const size_t top = std::min<size_t>(first.size(), 4);
    const int nrep = std::max(reps, 10);
    std::vector<std::vector<double>> again(top);
    for (int round = 0; round < 3 && e == hipSuccess && rc == LBM_OK; round++)
      for (size_t k = 0; k < top && e == hipSuccess && rc == LBM_OK; k++)
        again[k].push_back(timed(cands[first[k].second], nrep));
    for (size_t k = 0; k < top && e == hipSuccess && rc == LBM_OK; k++) {
      std::sort(again[k].begin(), again[k].end());
      const double med = again[k][again[k].size() / 2];
      if (med < best_t) {
        best_t = med;
        best_c = cands[first[k].second];
      }
    }
  }
  if (const hipError_t pe = put_back(); e == hipSuccess) e = pe;
  hipStreamSynchronize(b->stream);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (recs_kept) hipFree(recs_kept);
  hipFree(trial_d);
  hipFree(scratch);
  if (!rc && e != hipSuccess) rc = hip_fail(e, "batch autotune");
  const Cand &keep = rc ? saved : best_c;
  b->shape = keep.t;
  b->rows = keep.rows;
  drop_batch_graphs(b);
  for (lbm_handle *h : b->boxes) {
    const int rc2 = adopt_shape(h, lattice_shape(b, h));
    if (!rc) rc = rc2;
  }
  if (rc) return rc;
  if (best) {
    // (row tiles: the thread count, then zeros)
    best[0] = best_c.rows ? best_c.rows : best_c.t.bx;
    best[1] = best_c.rows ? 0 : best_c.t.by;
    best[2] = best_c.rows ? 0 : best_c.t.bz;
  }
  if (best_ms) *best_ms = best_t;
  return LBM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ adjoint
namespace {

// Device buffers of one adjoint run, freed on every exit.
struct AdjBuffers {
  std::vector<void *> ptrs;
  ~AdjBuffers() {
    for (void *p : ptrs) hipFree(p);
  }
  template <class Tp>
  hipError_t alloc(Tp **out, size_t bytes) {
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
    if (e == hipSuccess) ptrs.push_back(p);
    *out = static_cast<Tp *>(p);
    return e;
  }
};

// Stream-ordered copies on the handle's stream, waited for.
hipError_t copy_sync(lbm_handle *h, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  return e;
}

}  // namespace

extern "C" {

int lbm_adjoint_run(lbm_handle *h, const lbm_adjoint_config *cfg, double *loss, double *dB1, double *dB2,
                    double *dF) {
  if (!h || !cfg) return fail(LBM_EINVAL, "null argument");
  if (h->dims != 3 || h->Q != 19) return fail(LBM_EINVAL, "the adjoint is built for D3Q19");
  if (h->wmask_d || (h->cfg.flags & LBM_FLAG_INCOMPRESSIBLE))
    return fail(LBM_EINVAL, "the adjoint covers walled layers and the compressible equilibrium only");
  if (h->cfg.semantics != LBM_SEMANTICS_GPU) return fail(LBM_EINVAL, "the adjoint follows the GPU semantics");
  if (!h->populated) return fail(LBM_ESTATE, "populations not initialised");
  if (!h->inter_h.empty()) return fail(LBM_EINVAL, "the adjoint does not differentiate particle interactions");
  if (h->parts_h.size() > (size_t)kAdjParticles)
    return fail(LBM_EINVAL, "the adjoint takes at most " + std::to_string(kAdjParticles) + " particles");
  const int np = (int)h->parts_h.size();
  const int64_t T = cfg->ticks, K = cfg->checkpoint_every, Wn = cfg->window;
  if (T < 1 || K < 1 || Wn < 1 || Wn > T) return fail(LBM_EINVAL, "need ticks >= 1, checkpoint_every >= 1, 1 <= window <= ticks");
  if (cfg->particle < 0 || cfg->particle >= np) return fail(LBM_EINVAL, "bad swimmer id");
  if (dF && h->cfg.force != LBM_FORCE_FIELD) return fail(LBM_EINVAL, "dF needs a model with a force field");
  const int sw = cfg->particle;
  const size_t N = (size_t)h->N;
  const size_t pop_bytes = N * h->Q * elem_size(h->cfg.storage);
  const size_t rsize = h->cfg.storage == LBM_STORAGE_F64 ? 8 : 4;
  const int threads = kAdjThreads;
  const int64_t nblocks = ((int64_t)N + threads - 1) / threads;  // the bulk kernel's
  // the near kernel's blocks: particle p's near box (link box and one node)
  // spans at most 2 r / dx + 9 nodes an axis
  int32_t near_first[kAdjParticles + 1] = {0};
  int32_t near_ext[kAdjParticles] = {0};
  for (int p = 0; p < np; p++) {
    const int64_t e = (int64_t)std::ceil(2.0 * h->parts_h[p].radius / h->cfg.dx) + 10;
    near_ext[p] = (int32_t)e;
    near_first[p + 1] = near_first[p] + (int32_t)((e * e * e + threads - 1) / threads);
  }
  for (int p = np; p < kAdjParticles; p++) near_first[p + 1] = near_first[p];
  const int64_t nblocks_near = near_first[np];
  const int64_t nck = (T + K - 1) / K, seg_len = std::min(K, T);
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");

  AdjBuffers buf;
  char *ck = nullptr, *seg = nullptr, *hyd = nullptr, *fin = nullptr;
  const size_t hyd_bytes = N * 4 * rsize;  // one tick's moments (rho, u)
  void *lam[2] = {nullptr, nullptr};
  double *gF = nullptr, *partial = nullptr;
  // reduced storage with a force field: each segment's dL/dF summed in FP32
  // (half the bytes of an FP64 update per node and tick), folded into gF
  // after the segment
  float *gF32 = nullptr;
  const bool fold = h->cfg.storage != LBM_STORAGE_F64 && h->cfg.force == LBM_FORCE_FIELD;
  DevParticle *recs_seg = nullptr;  // the records of every tick of a segment (two sets)
  DevParticle *ck_recs_d = nullptr;  // the checkpoints' records
  if (buf.alloc(&ck, (size_t)nck * pop_bytes) != hipSuccess ||
      buf.alloc(&seg, 2 * (size_t)seg_len * pop_bytes) != hipSuccess ||
      buf.alloc(&hyd, 2 * (size_t)seg_len * hyd_bytes) != hipSuccess ||
      buf.alloc(&fin, pop_bytes) != hipSuccess || buf.alloc(&lam[0], N * h->Q * rsize) != hipSuccess ||
      buf.alloc(&lam[1], N * h->Q * rsize) != hipSuccess || buf.alloc(&gF, N * 3 * sizeof(double)) != hipSuccess ||
      buf.alloc(&partial, (size_t)std::max<int64_t>(nblocks_near, 1) * np * kAdjSums * sizeof(double)) != hipSuccess ||
      buf.alloc(&recs_seg, 2 * (size_t)(seg_len + 1) * np * sizeof(DevParticle)) != hipSuccess ||
      buf.alloc(&ck_recs_d, (size_t)nck * np * sizeof(DevParticle)) != hipSuccess)
    return fail(LBM_ENOMEM, "hipMalloc adjoint buffers (checkpoints, segment states, adjoint fields)");
  std::vector<DevParticle> fin_recs(np);
  auto save_recs = [&](std::vector<DevParticle> &dst) {
    return copy_sync(h, dst.data(), parts_of(h, h->pcur), np * sizeof(DevParticle), hipMemcpyDeviceToHost);
  };
  auto load_recs = [&](const std::vector<DevParticle> &src) {
    return copy_sync(h, parts_of(h, h->pcur), src.data(), np * sizeof(DevParticle), hipMemcpyHostToDevice);
  };
  auto run = [&](int64_t n) -> int {
    if (const int rc = lbm_run(h, n)) return rc;
    return LBM_OK;
  };
  const auto t_start = std::chrono::steady_clock::now();

  // ---- forward: checkpoints, and the swimmer's speed over the window
  const int64_t ws = T - Wn;  // V_t for t in (ws, T]
  std::vector<double> speeds;
  std::vector<std::array<double, 3>> vel;
  // the window's velocities come from a per-tick log of the moves (the
  // handle's own log, if any, is set aside meanwhile)
  double *ring = nullptr;
  if (buf.alloc(&ring, sizeof(double) * 9 * LBM_MAX_PARTICLES * (size_t)Wn) != hipSuccess)
    return fail(LBM_ENOMEM, "hipMalloc adjoint speed log");
  // (the handle's own log was not written meanwhile: on restore it holds
  // only ticks after the one the handle has reached)
  struct LogSwap {
    lbm_handle *h;
    double *d;
    int64_t cap, from;
    ~LogSwap() {
      h->traj_d = d;
      h->traj_cap = cap;
      h->traj_from = std::max(from, h->tick);
      drop_graph(h);
    }
  } swap{h, h->traj_d, h->traj_cap, h->traj_from};
  h->traj_d = ring;
  h->traj_cap = Wn;
  h->traj_from = h->tick;
  drop_graph(h);
  const int64_t tick0 = h->tick;
  for (int64_t t = 0; t < T;) {
    if (t % K == 0) {
      HIP_TRY(hipMemcpyAsync(ck + (size_t)(t / K) * pop_bytes, h->f[h->cur], pop_bytes, hipMemcpyDeviceToDevice, h->stream),
              "save checkpoint");
      HIP_TRY(hipMemcpyAsync(ck_recs_d + (size_t)(t / K) * np, parts_of(h, h->pcur), np * sizeof(DevParticle),
                             hipMemcpyDeviceToDevice, h->stream),
              "keep checkpoint records");
    }
    const int64_t n = std::min((t / K + 1) * K, T) - t;
    if (const int rc = run(n)) return rc;
    t += n;
  }
  {
    std::vector<double> lg((size_t)Wn * 9);
    if (const int rc = lbm_get_trajectory(h, sw, tick0 + ws + 1, Wn, lg.data())) return rc;
    for (int64_t i = 0; i < Wn; i++) {
      std::array<double, 3> v{lg[i * 9 + 3], lg[i * 9 + 4], lg[i * 9 + 5]};
      vel.push_back(v);
      speeds.push_back(std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    }
  }
  HIP_TRY(hipStreamSynchronize(h->stream), "sync");
  if (const int rc = check_particles(h)) return rc;
  HIP_TRY(hipMemcpyAsync(fin, h->f[h->cur], pop_bytes, hipMemcpyDeviceToDevice, h->stream), "save final state");
  HIP_TRY(save_recs(fin_recs), "save final records");
  const int64_t fin_tick = h->tick;
  const double fin_time = h->time;
  double mean = 0.0;
  for (double s : speeds) mean += s;
  mean /= (double)Wn;
  double Fsq = 0.0;
  std::vector<double> Fh;
  if (h->cfg.force == LBM_FORCE_FIELD) {
    Fh.resize(N * 3);
    if (h->cfg.storage == LBM_STORAGE_F64) {
      HIP_TRY(copy_sync(h, Fh.data(), h->F, N * 3 * sizeof(double), hipMemcpyDeviceToHost), "read force field");
    } else {
      std::vector<float> Ff(N * 3);
      HIP_TRY(copy_sync(h, Ff.data(), h->F, N * 3 * sizeof(float), hipMemcpyDeviceToHost), "read force field");
      for (size_t i = 0; i < Ff.size(); i++) Fh[i] = Ff[i];
    }
    for (double v : Fh) Fsq += v * v;
  }
  const double L = (mean - cfg->target_speed) * (mean - cfg->target_speed) + cfg->force_penalty * Fsq;
  const auto t_fwd = std::chrono::steady_clock::now();

  // ---- reverse
  const double gscale = 2.0 * (mean - cfg->target_speed) / (double)Wn;
  auto direct = [&](const DevParticle &p, double (&mu)[kPState]) {  // dL/dV_t of the speed term
    const double s = std::sqrt(p.V[0] * p.V[0] + p.V[1] * p.V[1] + p.V[2] * p.V[2]);
    if (s > 0.0)
      for (int k = 0; k < 3; k++) mu[kStV + k] += gscale * p.V[k] / s;
  };
  std::vector<std::array<double, kPState>> mu(np);
  for (auto &m : mu) m.fill(0.0);
  {
    double m[kPState];
    for (int i = 0; i < kPState; i++) m[i] = 0.0;
    direct(fin_recs[sw], m);
    for (int i = 0; i < kPState; i++) mu[sw][i] = m[i];
  }
  // the particles' adjoint state and the B1, B2 gradients live on the device:
  // the reverse sweep never waits for the host
  double *mu_d = nullptr, *muin_d = nullptr, *pt_d = nullptr, *gB_d = nullptr;
  if (buf.alloc(&mu_d, (size_t)np * kPState * sizeof(double)) != hipSuccess ||
      buf.alloc(&muin_d, (size_t)np * kPState * sizeof(double)) != hipSuccess ||
      buf.alloc(&pt_d, (size_t)np * 6 * sizeof(double)) != hipSuccess ||
      buf.alloc(&gB_d, (size_t)np * 2 * sizeof(double)) != hipSuccess)
    return fail(LBM_ENOMEM, "hipMalloc adjoint particle state");
  {
    std::vector<double> mu0((size_t)np * kPState);
    for (int p = 0; p < np; p++)
      for (int i = 0; i < kPState; i++) mu0[(size_t)p * kPState + i] = mu[p][i];
    HIP_TRY(copy_sync(h, mu_d, mu0.data(), mu0.size() * sizeof(double), hipMemcpyHostToDevice), "upload adjoint state");
  }
  int a = 0;
  HIP_TRY(hipMemsetAsync(lam[a], 0, N * h->Q * rsize, h->stream), "clear adjoint");
  HIP_TRY(hipMemsetAsync(gF, 0, N * 3 * sizeof(double), h->stream), "clear force gradient");
  if (fold) {
    if (buf.alloc(&gF32, N * 3 * sizeof(float)) != hipSuccess) return fail(LBM_ENOMEM, "hipMalloc segment force gradient");
    HIP_TRY(hipMemsetAsync(gF32, 0, N * 3 * sizeof(float), h->stream), "clear segment force gradient");
  }
  HIP_TRY(hipMemsetAsync(gB_d, 0, (size_t)np * 2 * sizeof(double), h->stream), "clear B gradients");
  AdjArgs aa;
  std::memset(&aa, 0, sizeof aa);
  aa.gF = fold ? static_cast<void *>(gF32) : static_cast<void *>(gF);
  aa.Ffield = h->F;
  aa.g = geom(h);
  aa.c = coll(h);
  for (int k = 0; k < 3; k++) aa.coords[k] = h->coords_d[k];
  aa.partial = partial;
  aa.np = np;
  {
    double coeff = h->cfg.dx * h->cfg.dx * h->cfg.dx / h->cfg.dt;
    aa.coeff = coeff;
  }
  aa.ls = h->cfg.dx / h->cfg.dt;
  aa.shell = 1.5 * h->cfg.dx;
  aa.phitau = pt_d;
  for (int p = 0; p <= kAdjParticles; p++) aa.near_first[p] = near_first[p];
  for (int p = 0; p < kAdjParticles; p++) aa.near_ext[p] = near_ext[p];
  aa.fast = aa.g.walled == 0 && (uint64_t)N * h->Q < (uint64_t(1) << 31);
  aa.mnx = div_magic((uint32_t)h->n[0], (uint64_t)N);
  aa.mny = div_magic((uint32_t)h->n[1], (uint64_t)N / (uint64_t)h->n[0] + 1);
  AdjMoveArgs ma{};
  ma.mu = mu_d;
  ma.mu_in = muin_d;
  ma.phitau = pt_d;
  ma.dt = h->cfg.dt;
  AdjFinishArgs fa2{};
  fa2.mu_in = muin_d;
  fa2.mu = mu_d;
  fa2.gB = gB_d;
  fa2.np = np;
  fa2.sw = sw;
  fa2.gscale = gscale;
  // Segment c's recomputation (on a second stream, into slot set c % 2)
  // overlaps the reverse sweep of segment c + 1 (on the handle's stream):
  // the two share only the read-only checkpoints.  The recomputation uses
  // the lattice's own working state (populations' spare buffer, records,
  // partials, rho), which the reverse sweep does not touch.
  // Two streams.  The handle's (s_bulk) takes each reverse tick's bulk
  // kernel and, one kernel per tick behind it, the recomputation of the
  // next segment; the other (s_near) takes each tick's near kernel and
  // tail.  Tick k's near kernel waits for tick k + 1's bulk and tick k's
  // bulk for tick k + 1's near kernel (the adjoint populations and dL/dF
  // they read and update; events of the tick's parity).
  hipStream_t s_bulk = h->stream, s_near = nullptr;
  hipEvent_t ev_rec[2] = {}, ev_b[2] = {}, ev_n[2] = {}, ev_fold = nullptr, ev_segdone = nullptr;
  constexpr int kEvents = 8;
  struct Streams {
    lbm_handle *h;
    hipStream_t main, *extra;
    hipEvent_t *evs;
    ~Streams() {
      h->stream = main;
      if (*extra) {
        hipStreamSynchronize(*extra);
        hipStreamDestroy(*extra);
      }
      for (int i = 0; i < kEvents; i++)
        if (evs[i]) hipEventDestroy(evs[i]);
    }
  };
  hipEvent_t evs[kEvents] = {};
  Streams guard{h, s_bulk, &s_near, evs};
  HIP_TRY(hipStreamCreate(&s_near), "create near-kernel stream");
  for (int i = 0; i < kEvents; i++) HIP_TRY(hipEventCreateWithFlags(&evs[i], hipEventDisableTiming), "create event");
  ev_rec[0] = evs[0];
  ev_rec[1] = evs[1];
  ev_b[0] = evs[2];
  ev_b[1] = evs[3];
  ev_n[0] = evs[4];
  ev_n[1] = evs[5];
  ev_fold = evs[6];
  ev_segdone = evs[7];
  // (every event starts recorded after the set-up above on the handle's stream)
  for (hipEvent_t ev : {ev_b[0], ev_b[1], ev_n[0], ev_n[1], ev_fold, ev_segdone})
    HIP_TRY(hipEventRecord(ev, s_bulk), "event");
  HIP_TRY(hipStreamWaitEvent(s_near, ev_fold, 0), "wait for the set-up");
  // segment c's tick j (0 <= j < its length): P_k = C(S_k) and the
  // moments of S_k in slot j of the segment's set, the records at its start
  // in recs_of(c, j) (and after the segment's last move in recs_of(c, len))
  auto Pk = [&](int64_t c, int64_t j) -> void * {
    return seg + ((size_t)(c & 1) * seg_len + (size_t)j) * pop_bytes;
  };
  auto Hk = [&](int64_t c, int64_t j) -> void * { return hyd + ((size_t)(c & 1) * seg_len + (size_t)j) * hyd_bytes; };
  auto recs_of = [&](int64_t c, int64_t j) { return recs_seg + ((size_t)(c & 1) * (seg_len + 1) + (size_t)j) * np; };
  // The recomputation of segment c as a run of its ticks, issued piece by
  // piece on s_bulk: piece 0 restores the checkpoint's records and collides
  // (keeping P_{t0} and the moments), pieces 1 .. len - 1 are fused kernels
  // (keeping each P_k and its moments), piece len the last tick's gather,
  // piece len + 1 its move, after which ev_rec of the segment's set is
  // recorded.  Each gathering kernel writes the records it streamed with
  // straight into the segment's record slots.
  struct RecState {
    int64_t c = -1, len = 0, m = 0;
    int pp = 0;  // the parity of the partials the next gathering kernel reads
  } rs;
  auto keep = [&](int mode, const void *fin, void *fout, void *moments, const DevParticle *pin, DevParticle *pout) {
    StepArgs sa = step_args(h, fin, fout, rs.pp);
    sa.hydro = moments;
    sa.sp.parts_in = pin;
    sa.sp.parts_out = pout;
    const Launch lk = step_launch(h, sa, base_mode(mode));
    switch (h->cfg.storage) {
      case LBM_STORAGE_F64: return launch_step_keep<kF64>(h->cfg.collision, h->cfg.force, mode, sa, lk);
      case LBM_STORAGE_F32: return launch_step_keep<kF32>(h->cfg.collision, h->cfg.force, mode, sa, lk);
      case LBM_STORAGE_F16C: return launch_step_keep<kF16C>(h->cfg.collision, h->cfg.force, mode, sa, lk);
      default: return launch_step_keep<kF16>(h->cfg.collision, h->cfg.force, mode, sa, lk);
    }
  };
  auto rec_piece = [&]() -> int {
    if (rs.c < 0) return LBM_OK;
    const int64_t c = rs.c, len = rs.len, m = rs.m;
    // gathering kernel m (1..len) streams tick t0 + m - 1 with the records
    // recs_of(c, m - 1), which it writes (the checkpoint's, for m = 1)
    auto rin = [&](int64_t mm) -> const DevParticle * { return mm == 1 ? parts_of(h, h->pcur) : recs_of(c, mm - 2); };
    hipError_t e = hipSuccess;
    if (m == 0) {
      e = hipMemcpyAsync(parts_of(h, h->pcur), ck_recs_d + (size_t)c * np, np * sizeof(DevParticle),
                         hipMemcpyDeviceToDevice, s_bulk);
      // (the collide-only kernel has no box blocks and writes no records)
      if (e == hipSuccess)
        e = keep(kCollideKeep, ck + (size_t)c * pop_bytes, Pk(c, 0), Hk(c, 0), parts_of(h, h->pcur),
                 parts_of(h, h->pcur ^ 1));
    } else if (m < len) {
      e = keep(kFusedKeep, Pk(c, m - 1), Pk(c, m), Hk(c, m), rin(m), recs_of(c, m - 1));
      rs.pp ^= 1;
    } else if (m == len) {
      StepArgs sa = step_args(h, Pk(c, len - 1), h->f[h->cur ^ 1], rs.pp);
      sa.sp.parts_in = rin(len);
      sa.sp.parts_out = recs_of(c, len - 1);
      const Launch lg = step_launch(h, sa, kGather);
      switch (h->cfg.storage) {
        case LBM_STORAGE_F64: e = launch_step<kF64>(h->Q, h->cfg.collision, h->cfg.force, kGather, sa, lg); break;
        case LBM_STORAGE_F32: e = launch_step<kF32>(h->Q, h->cfg.collision, h->cfg.force, kGather, sa, lg); break;
        case LBM_STORAGE_F16C: e = launch_step<kF16C>(h->Q, h->cfg.collision, h->cfg.force, kGather, sa, lg); break;
        default: e = launch_step<kF16>(h->Q, h->cfg.collision, h->cfg.force, kGather, sa, lg); break;
      }
      rs.pp ^= 1;
    } else {
      StepArgs fa = step_args(h, nullptr, nullptr, rs.pp);
      fa.sp.parts_in = recs_of(c, len - 1);
      fa.sp.parts_out = recs_of(c, len);
      e = launch_finalize(fa.sp, s_bulk);
      if (e == hipSuccess) e = hipEventRecord(ev_rec[c & 1], s_bulk);
      rs.c = -1;
    }
    if (e != hipSuccess) return hip_fail(e, "segment recomputation launch");
    rs.m++;
    return LBM_OK;
  };
  auto rec_start = [&](int64_t c) {
    rs.c = c;
    rs.len = std::min(c * K + K, T) - c * K;
    rs.m = 0;
    rs.pp = h->pcur;
  };
  auto rec_finish = [&]() -> int {
    while (rs.c >= 0)
      if (const int rc = rec_piece()) return rc;
    return LBM_OK;
  };
  const Launch l{dim3((unsigned)nblocks), dim3(threads), s_bulk};
  const Launch ln{dim3((unsigned)std::max<int64_t>(nblocks_near, 1)), dim3(threads), s_near};
  AdjTailArgs ta{};
  ta.partial = partial;
  ta.nblocks = (int32_t)nblocks_near;
  ta.mv = ma;
  rec_start(nck - 1);
  if (const int rc = rec_finish()) return rc;
  for (int64_t c = nck - 1; c >= 0; c--) {
    const int64_t t0 = c * K, t1 = std::min(t0 + K, T);
    // segment c's slots (written on s_bulk) and dL/dF folded
    HIP_TRY(hipStreamWaitEvent(s_near, ev_rec[c & 1], 0), "wait for the recomputation");
    HIP_TRY(hipStreamWaitEvent(s_near, ev_fold, 0), "wait for the fold");
    if (c > 0) {
      // the next recomputation writes the other set, which the near stream
      // read until segment c + 1's last tail
      HIP_TRY(hipStreamWaitEvent(s_bulk, ev_segdone, 0), "wait for the near stream");
      rec_start(c - 1);
    }
    for (int64_t k = t1 - 1; k >= t0; k--) {
      // tick k + 1's kernels, recorded in the events of its parity
      hipError_t e = hipStreamWaitEvent(s_near, ev_b[(k + 1) & 1], 0);
      if (e == hipSuccess) e = hipStreamWaitEvent(s_bulk, ev_n[(k + 1) & 1], 0);
      aa.parts = recs_of(c, k - t0);
      aa.P = Pk(c, k - t0);
      aa.H = Hk(c, k - t0);
      aa.lam_in = lam[a];
      aa.lam_out = lam[a ^ 1];
      // near stream: the move's reverse (dL/dF, dL/dT of tick k: on its own
      // at the segment's last tick, else in the previous tail), the near
      // kernel, then the tail: the particles' adjoint state at the start of
      // tick k and the reverse of tick k - 1's move when that is in the segment
      if (e == hipSuccess && k == t1 - 1) {
        ma.recs_k = recs_of(c, k - t0);
        ma.recs_k1 = recs_of(c, k - t0 + 1);
        e = launch_adjoint_move(ma, np, s_near);
      }
      if (e == hipSuccess) e = launch_adjoint(h->cfg.storage, h->cfg.collision, h->cfg.force, true, aa, ln);
      if (e == hipSuccess) e = hipEventRecord(ev_n[k & 1], s_near);
      ta.fin = fa2;
      ta.fin.recs_k = recs_of(c, k - t0);
      ta.fin.direct = k > ws;  // V_k is one of the averaged speeds
      ta.do_move = k > t0;
      if (k > t0) {
        ta.mv.recs_k = recs_of(c, k - 1 - t0);
        ta.mv.recs_k1 = recs_of(c, k - t0);
      }
      if (e == hipSuccess) e = launch_adjoint_tail(ta, s_near);
      // bulk stream: the bulk kernel, then a piece of the next recomputation
      if (e == hipSuccess) e = launch_adjoint(h->cfg.storage, h->cfg.collision, h->cfg.force, false, aa, l);
      if (e == hipSuccess) e = hipEventRecord(ev_b[k & 1], s_bulk);
      if (e != hipSuccess) return hip_fail(e, "adjoint launch");
      if (const int rc = rec_piece()) return rc;
      a ^= 1;
    }
    if (const int rc = rec_finish()) return rc;
    HIP_TRY(hipEventRecord(ev_segdone, s_near), "event");
    // the segment's dL/dF, once its last near kernel has added to it
    HIP_TRY(hipStreamWaitEvent(s_bulk, ev_n[t0 & 1], 0), "wait for the near stream");
    if (fold) HIP_TRY(launch_adjoint_fold(gF32, gF, (int64_t)N * 3, s_bulk), "fold force gradient");
    HIP_TRY(hipEventRecord(ev_fold, s_bulk), "event");
  }
  // the last tail (the B gradients and the particles' adjoint state)
  HIP_TRY(hipStreamWaitEvent(s_bulk, ev_segdone, 0), "wait for the near stream");
  std::vector<double> gB((size_t)np * 2);
  HIP_TRY(copy_sync(h, gB.data(), gB_d, gB.size() * sizeof(double), hipMemcpyDeviceToHost), "read B gradients");
  std::vector<double> gB1(np), gB2(np);
  for (int p = 0; p < np; p++) {
    gB1[p] = gB[(size_t)p * 2];
    gB2[p] = gB[(size_t)p * 2 + 1];
  }
  // the state the forward run reached
  HIP_TRY(hipMemcpyAsync(h->f[h->cur], fin, pop_bytes, hipMemcpyDeviceToDevice, h->stream), "restore final state");
  HIP_TRY(load_recs(fin_recs), "restore final records");
  h->tick = fin_tick;
  h->time = fin_time;
  if (dF) {
    HIP_TRY(copy_sync(h, dF, gF, N * 3 * sizeof(double), hipMemcpyDeviceToHost), "read force gradient");
    for (size_t i = 0; i < N * 3; i++) dF[i] += 2.0 * cfg->force_penalty * Fh[i];
  }
  if (loss) *loss = L;
  for (int p = 0; p < np; p++) {
    if (dB1) dB1[p] = gB1[p];
    if (dB2) dB2[p] = gB2[p];
  }
  const auto t_end = std::chrono::steady_clock::now();
  h->adj_forward_ms = std::chrono::duration<double, std::milli>(t_fwd - t_start).count();
  h->adj_reverse_ms = std::chrono::duration<double, std::milli>(t_end - t_fwd).count();
  return LBM_OK;
}

int lbm_adjoint_last_times(lbm_handle *h, double *forward_ms, double *reverse_ms) {
  if (!h) return fail(LBM_EINVAL, "null handle");
  if (forward_ms) *forward_ms = h->adj_forward_ms;
  if (reverse_ms) *reverse_ms = h->adj_reverse_ms;
  return LBM_OK;
}

}  // extern "C"
