// Device helpers shared by the gfx950 kernels: launch shapes, wavefront and
// workgroup reductions, and the one-pass deterministic grid reduction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace alien_hip {

constexpr int kWave = 64;    // CDNA wavefront
constexpr int kBlock = 256;  // 4 waves per workgroup
constexpr int kEPT = 4;      // elements per thread of the streaming kernels
constexpr int kGroup = 256;  // workgroups per first-level reduction group
constexpr int kMaxSlots = 4; // up to 4 fused reductions per launch

// Streaming kernels: one workgroup per kBlock*kEPT elements, each lane issues
// kEPT independent loads per array before using any (memory-level
// parallelism), no grid-stride loop.
inline long ept_blocks(long n) {
  long g = (n + (long)kBlock * kEPT - 1) / ((long)kBlock * kEPT);
  return g < 1 ? 1 : g;
}
// Row kernels on the RB-CSR layout: one wavefront per 64-row bundle.
inline long bundle_blocks(long n) {
  long nb = (n + 63) / 64;
  long g = (nb + (kBlock / kWave) - 1) / (kBlock / kWave);
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

// Sum over a 256-thread workgroup; the result is valid in thread 0.
__device__ inline double block_sum(double v, double* lds4) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  if (lane == 0) lds4[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
  return r;
}

// Workspace of the one-pass reductions. Level 1: one partial per workgroup;
// the last workgroup of each group of kGroup workgroups sums its group's
// partials (level 2); the last group to finish sums the group partials and
// hands the totals to the caller. Every sum runs in a fixed order, so the
// result depends only on the launch shape (a function of n): deterministic.
// Tickets are reset by the workgroups that consume them, so no memset is
// ever needed between launches (graph-capture friendly).
struct RedWork {
  double* partial;   // kMaxSlots * cap
  double* gpartial;  // kMaxSlots * gcap
  unsigned* tickets; // 1 + gcap
  int cap;           // workgroups per launch the buffers can hold
  int gcap;          // = ceil(cap / kGroup)
};

template <int NSLOT, typename Fin>
__device__ inline void grid_reduce_finish(const double (&vals)[NSLOT], RedWork w, Fin fin) {
  __shared__ bool is_last;
  __shared__ double lds[4 * NSLOT];
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = tid / kWave;
  const int nblk = gridDim.x;
  const int grp = blockIdx.x / kGroup;
  const int ngroups = (nblk + kGroup - 1) / kGroup;
  const int gsize = min(kGroup, nblk - grp * kGroup);
  if (tid == 0) {
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) w.partial[(long)s * w.cap + blockIdx.x] = vals[s];
    __threadfence();
    const unsigned t = atomicAdd(&w.tickets[1 + grp], 1u);
    is_last = (t == (unsigned)gsize - 1);
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  // level 2: this group's partials, one per lane
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    double acc = (tid < gsize) ? w.partial[(long)s * w.cap + (long)grp * kGroup + tid] : 0.0;
    acc = wave_sum(acc);
    if (lane == 0) lds[s * 4 + wid] = acc;
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int s = 0; s < NSLOT; ++s)
      w.gpartial[(long)s * w.gcap + grp] = (lds[s * 4 + 0] + lds[s * 4 + 1]) + (lds[s * 4 + 2] + lds[s * 4 + 3]);
    w.tickets[1 + grp] = 0u;
    __threadfence();
    const unsigned t = atomicAdd(&w.tickets[0], 1u);
    is_last = (t == (unsigned)ngroups - 1);
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  // level 3: the group partials, in group order
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    double acc = 0.0;
    for (int q = tid; q < ngroups; q += kBlock) acc += w.gpartial[(long)s * w.gcap + q];
    acc = wave_sum(acc);
    if (lane == 0) lds[s * 4 + wid] = acc;
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int s = 0; s < NSLOT; ++s)
      fin(s, (lds[s * 4 + 0] + lds[s * 4 + 1]) + (lds[s * 4 + 2] + lds[s * 4 + 3]));
    w.tickets[0] = 0u;
  }
}

} // namespace alien_hip
