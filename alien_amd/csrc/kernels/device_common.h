// Device helpers shared by the gfx950 kernels: launch shapes, wavefront and
// workgroup reductions, and the one-pass deterministic reduction epilogue.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace alien_hip {

constexpr int kWave = 64;       // CDNA wavefront
constexpr int kBlock = 256;     // 4 waves per workgroup
constexpr int kRedBlocks = 2048; // largest reduction grid: 8 workgroups per CU on 256 CUs

// Fixed launch shape for grid-stride vector kernels: enough workgroups to fill
// the 256 CUs several times over, never more than the work needs.
inline int grid_for(long n, int per_thread = 1) {
  long g = (n + (long)kBlock * per_thread - 1) / ((long)kBlock * per_thread);
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return (int)g;
}

// Reduction grid: a function of n only, so the combination order (and the
// rounding) is the same on every call.
inline int red_grid(long n) {
  long g = (n + kBlock * 4 - 1) / (kBlock * 4);
  if (g < 1) g = 1;
  if (g > kRedBlocks) g = kRedBlocks;
  return (int)g;
}

// Grid of the fused row kernels (one wavefront per 64-row bundle, grid-stride
// over bundles): a function of n only, at most kRedBlocks workgroups.
inline int bundle_grid(long n) {
  long nb = (n + 63) / 64;
  long g = (nb + 3) / 4;
  if (g < 1) g = 1;
  if (g > kRedBlocks) g = kRedBlocks;
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

// Workspace of the one-pass reductions: per-workgroup partials and a ticket
// counter that the last workgroup resets, so no memset is ever needed between
// launches (graph-capture friendly).
struct RedWork {
  double* partial;    // kRedBlocks * kMaxSlots doubles
  unsigned* ticket;   // 1 counter
};
constexpr int kMaxSlots = 4; // up to 4 fused dot products per launch

// Epilogue: thread 0 of each workgroup holds `nslot` partial sums in vals[].
// The last workgroup to finish combines all partials in workgroup order and
// hands the totals to `fin(slot, total)` (called by thread 0 only).
template <int NSLOT, typename Fin>
__device__ inline void grid_reduce_finish(const double (&vals)[NSLOT], RedWork w, Fin fin) {
  __shared__ bool is_last;
  __shared__ double lds[4 * NSLOT];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) w.partial[s * kRedBlocks + blockIdx.x] = vals[s];
    __threadfence();
    unsigned t = atomicAdd(w.ticket, 1u);
    is_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kBlock)
      acc += w.partial[s * kRedBlocks + b];
    acc = wave_sum(acc);
    if (lane == 0) lds[s * 4 + wid] = acc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < NSLOT; ++s)
      fin(s, (lds[s * 4 + 0] + lds[s * 4 + 1]) + (lds[s * 4 + 2] + lds[s * 4 + 3]));
    *w.ticket = 0u;
  }
}

} // namespace alien_hip
