// lbm_particles.hip — the kernels that move the particles between step
// kernels (the last pending move of a run; every move on lattices with more
// than kLocalParticles particles), the link-bit kernel used to list a
// particle's boundary links (lbm_get_links), and the plain copy kernel that
// measures the HBM ceiling.
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
