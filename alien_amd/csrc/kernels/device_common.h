// Device helpers shared by the gfx950 kernels: launch shapes, wavefront and
// workgroup reductions, and the per-workgroup partials of the grid reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace alien_hip {

constexpr int kWave = 64;    // CDNA wavefront
constexpr int kBlock = 256;    // 4 waves per workgroup (streaming kernels)
constexpr int kRowBlock = 256; // 4 waves per workgroup (fused row kernels: one bundle per wave)
constexpr int kEPT = 4;        // elements per thread of the streaming kernels
constexpr int kMaxSlots = 4; // up to 4 fused reductions per launch

// Streaming kernels: one workgroup per kBlock*kEPT elements, each lane issues
// kEPT independent loads per array before using any (memory-level
// parallelism), no grid-stride loop.
inline long ept_blocks(long n) {
  long g = (n + (long)kBlock * kEPT - 1) / ((long)kBlock * kEPT);
  return g < 1 ? 1 : g;
}
// Fused row kernels on the RB-CSR layout: one wavefront per 64-row bundle,
// kRowBlock/64 bundles per workgroup.
inline long bundle_blocks(long n) {
  long nb = (n + 63) / 64;
  long g = (nb + (kRowBlock / kWave) - 1) / (kRowBlock / kWave);
  return g < 1 ? 1 : g;
}
// Shape of the simple element-wise kernels (grid-stride).
inline int grid_for(long n, int per_thread = 1) {
  long g = (n + (long)kBlock * per_thread - 1) / ((long)kBlock * per_thread);
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return (int)g;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  return v;
}

// Sum over an NT-thread workgroup (lds holds NT/64 doubles); the result is
// valid in thread 0. Waves are combined in a fixed order.
template <int NT = kBlock>
__device__ inline double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NT / kWave; ++q) r += lds[q];
  }
  return r;
}

// Workspace of the deterministic reductions. A reducing kernel stores one
// partial per workgroup (store_partials); a second, single-workgroup kernel
// (launch_finalize, reduce.hip) sums them in a fixed order and consumes the
// totals. The kernel boundary between the two makes the partials visible on
// every XCD, so no grid-wide fence or atomic is needed, and the result depends
// only on the launch shape (a function of n): deterministic.
struct RedWork {
  double* partial; // kMaxSlots * cap
  int cap;         // workgroups per launch the buffer can hold
};

template <int NSLOT>
__device__ inline void store_partials(const double (&vals)[NSLOT], RedWork w) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) w.partial[(long)s * w.cap + blockIdx.x] = vals[s];
  }
}

} // namespace alien_hip
