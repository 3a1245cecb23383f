// lbm_particles.hip — the kernels that move the particles between step
// kernels (the last pending move of a run; every move on lattices with more
// than kLocalParticles particles), the link-bit kernel used to list a
// particle's boundary links (lbm_get_links), the plain copy kernel that
// measures the HBM ceiling, and the pull kernel that measures the ceiling of
// the step's access pattern.
#include "lbm_dispatch.hpp"

namespace lbm {

// Plain device copy of n16 16-byte words, one per thread: the measured HBM
// ceiling, and the kernel of known traffic (n16*16 bytes read and written)
// that scales the byte counters of the profiler.  Its stores are
// non-temporal, like the step kernels' population stores.
__global__ void __launch_bounds__(256) copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                   int64_t n16) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n16)
    __builtin_nontemporal_store(reinterpret_cast<const u32x4 *>(src)[i], reinterpret_cast<u32x4 *>(dst) + i);
}

hipError_t launch_copy(const void *src, void *dst, int64_t bytes, hipStream_t stream) {
  const int64_t n16 = bytes / 16;
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((n16 + 255) / 256)), dim3(256), 0, stream,
                     static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n16);
  return hipGetLastError();
}

// The D3Q19 step's data movement alone, with no collision: every node of a
// periodic nx x ny x nz lattice in the row-blocked layout pulls f_q(x - c_q)
// into registers (one node per lane, its 19 loads at once) and stores them
// at x with non-temporal stores, as the step kernels' tiles do.  The ceiling
// the step's access pattern allows: its bytes over this kernel's time.
// Block: one x-segment of `blockDim.x` nodes; grid: segments x rows.
template <typename T>
__global__ void __launch_bounds__(256) pull_kernel(const T *__restrict__ f, T *__restrict__ g, int nx, int ny, int nz,
                                                   int nseg) {
  using L = Lattice<19>;
  const int seg = (int)(blockIdx.x % (unsigned)nseg), row = (int)(blockIdx.x / (unsigned)nseg);
  const int x = seg * blockDim.x + threadIdx.x;
  if (x >= nx) return;
  const int y = row % ny, z = row / ny;
  const int xs[3] = {x == nx - 1 ? 0 : x + 1, x, x == 0 ? nx - 1 : x - 1};  // x - c for c = -1, 0, 1
  const int ys[3] = {y == ny - 1 ? 0 : y + 1, y, y == 0 ? ny - 1 : y - 1};
  const int zs[3] = {z == nz - 1 ? 0 : z + 1, z, z == 0 ? nz - 1 : z - 1};
  T v[19];
#pragma unroll
  for (int q = 0; q < 19; q++) {
    const int64_t src_row = ys[L::cy[q] + 1] + (int64_t)ny * zs[L::cz[q] + 1];
    v[q] = f[(src_row * 19 + q) * nx + xs[L::cx[q] + 1]];
  }
#pragma unroll
  for (int q = 0; q < 19; q++) __builtin_nontemporal_store(v[q], g + ((int64_t)row * 19 + q) * nx + x);
}

hipError_t launch_pull(int elem_bytes, const void *src, void *dst, int nx, int ny, int nz, hipStream_t stream) {
  const int threads = nx >= 256 ? 256 : (nx + 63) / 64 * 64;
  const int nseg = (nx + threads - 1) / threads;
  const dim3 grid((unsigned)((int64_t)nseg * ny * nz)), block((unsigned)threads);
  switch (elem_bytes) {
    case 2:
      hipLaunchKernelGGL(pull_kernel<uint16_t>, grid, block, 0, stream, static_cast<const uint16_t *>(src),
                         static_cast<uint16_t *>(dst), nx, ny, nz, nseg);
      break;
    case 4:
      hipLaunchKernelGGL(pull_kernel<uint32_t>, grid, block, 0, stream, static_cast<const uint32_t *>(src),
                         static_cast<uint32_t *>(dst), nx, ny, nz, nseg);
      break;
    case 8:
      hipLaunchKernelGGL(pull_kernel<uint64_t>, grid, block, 0, stream, static_cast<const uint64_t *>(src),
                         static_cast<uint64_t *>(dst), nx, ny, nz, nseg);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Applies the move that the last box kernel of a run left pending.
__device__ __forceinline__ void finalize_block(const StepParticles &sp) {
  __shared__ DevParticle recs[kLocalParticles];
  constexpr int kWords = sizeof(DevParticle) / 8;
  const int tid = threadIdx.x;
  for (int w = tid; w < sp.np * kWords; w += blockDim.x)
    reinterpret_cast<uint64_t *>(recs)[w] = reinterpret_cast<const uint64_t *>(sp.parts_in)[w];
  __syncthreads();
  apply_pending_moves(sp, recs, nullptr, true);
  for (int w = tid; w < sp.np * kWords; w += blockDim.x)
    reinterpret_cast<uint64_t *>(sp.parts_out)[w] = reinterpret_cast<const uint64_t *>(recs)[w];
}

__global__ void __launch_bounds__(256) finalize_kernel(StepParticles sp) { finalize_block(sp); }

// Lattices with more than kLocalParticles particles: every particle's move
// after a gathering kernel, one wave per particle.
__global__ void __launch_bounds__(64) move_kernel(StepParticles sp) { move_particle_block(sp, blockIdx.x); }

hipError_t launch_move(const StepParticles &sp, hipStream_t stream) {
  hipLaunchKernelGGL(move_kernel, dim3((unsigned)sp.np), dim3(64), 0, stream, sp);
  return hipGetLastError();
}

// One block per lattice of a batch; boxes[i].sp names lattice i's records
// and the partials of its last box kernel.
__global__ void __launch_bounds__(256) finalize_batch_kernel(const StepArgs *boxes) {
  const StepParticles &sp = boxes[blockIdx.x].sp;
  if (sp.np > 0) finalize_block(sp);
}

hipError_t launch_finalize(const StepParticles &sp, hipStream_t stream) {
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(256), 0, stream, sp);
  return hipGetLastError();
}

hipError_t launch_finalize_batch(const StepArgs *boxes, int n, hipStream_t stream) {
  hipLaunchKernelGGL(finalize_batch_kernel, dim3((unsigned)n), dim3(256), 0, stream, boxes);
  return hipGetLastError();
}

hipError_t launch_link_mask(int Q, const LinkMaskArgs &a, const Launch &l) {
  switch (Q) {
    case 9: hipLaunchKernelGGL(link_mask_kernel<9>, l.grid, l.block, l.shmem, l.stream, a); break;
    case 15: hipLaunchKernelGGL(link_mask_kernel<15>, l.grid, l.block, l.shmem, l.stream, a); break;
    case 19: hipLaunchKernelGGL(link_mask_kernel<19>, l.grid, l.block, l.shmem, l.stream, a); break;
    case 27: hipLaunchKernelGGL(link_mask_kernel<27>, l.grid, l.block, l.shmem, l.stream, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace lbm
