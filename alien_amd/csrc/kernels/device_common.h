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
// Row kernels on the row-bundled layouts, one wavefront per bundle of 64 rows
// or, when the grid is capped, a walk over several bundles per wavefront.
// Capped grids hold a multiple of the XCD count; workgroups are dealt to the
// XCDs round-robin, so the bundles are split into one contiguous share per XCD
// and the waves of an XCD stride through its share together: the waves running
// at any moment read neighbouring bundles and their x-gathers hit that XCD's
// L2, and per-wave start-up and the per-workgroup reduction are paid once per
// walk. One-bundle-per-wave grids are mapped in chunks per XCD (bundle_walk). Measured at 256^3
// (7-point): on RB-CSR / RB-DIA the FP32 row kernels run fastest capped at 4096 workgroups, the
// FP64 ones uncapped; on the uniform RB-DIA (no per-bundle descriptor loads) see uni_blocks. The
// environment variables ALIEN_HIP_ROW_BLOCKS_F64 / _F32 / _UNI / _UNI_WIDE, read when a context is
// created, override the caps.
constexpr int kXcds = 8;
constexpr int kRowWaves = kRowBlock / kWave;
constexpr long kUncapped = 1L << 30;
extern long g_row_blocks_f64, g_row_blocks_f32, g_row_blocks_uni, g_row_blocks_uni_wide;
inline long round_xcds(long g) {
  g = (g + kXcds - 1) / kXcds * kXcds;
  return g < kXcds ? kXcds : g;
}
// workgroups of a row kernel over n rows with elements of `elem` bytes
// Uncapped grids are rounded to whole chunks of kXcdChunk workgroups per XCD
// (bundle_walk's chunked XCD mapping needs a bijection).
#ifndef ALIEN_HIP_XCD_CHUNK
#define ALIEN_HIP_XCD_CHUNK 64
#endif
constexpr long kXcdChunk = ALIEN_HIP_XCD_CHUNK;
inline long round_chunks(long g) {
  const long q = kXcds * kXcdChunk;
  return (g + q - 1) / q * q;
}
inline long bundle_blocks(long n, int elem) {
  const long nb = (n + kWave - 1) / kWave;
  const long g = (nb + kRowWaves - 1) / kRowWaves;
  const long cap = elem == 8 ? g_row_blocks_f64 : g_row_blocks_f32;
  if (g > cap) return round_xcds(cap);
  return round_chunks(g);
}
// workgroups of a row kernel over a uniform RB-DIA (RBMat::uni) of k slots per bundle, both
// precisions. Measured at 256^3 inside the solves (caps 8192 .. 65536 and uncapped): the 7-slot
// operator's q = A p step runs best at 32768 workgroups (FP32 100 us, FP64 216 us), the 4-slot
// factors' steps at 16384 (67 and 69 us; 72 and 81 us at 32768).
inline long uni_blocks(long n, int k) {
  const long nb = (n + kWave - 1) / kWave;
  const long g = (nb + kRowWaves - 1) / kRowWaves;
  const long cap = k >= 6 ? g_row_blocks_uni_wide : g_row_blocks_uni;
  if (g > cap) return round_xcds(cap);
  return round_chunks(g);
}
// the largest grid any row kernel uses (reduction workspace)
inline long bundle_blocks_max(long n) {
  const long nb = (n + kWave - 1) / kWave;
  return round_chunks((nb + kRowWaves - 1) / kRowWaves);
}
#if defined(__HIPCC__)
// Bundles of this wave: b = first, first + stride, ... < end.
struct BundleWalk {
  int first, end, stride;
};
// Walk over nb units (bundles, or pairs of bundles) of a row kernel's grid. `chunk`: the XCD
// chunk of the one-unit-per-wave mapping below, used when the grid is a multiple of 8 chunks
// (kXcdChunk otherwise: every grid is rounded to that).
__device__ __forceinline__ BundleWalk bundle_walk_units(int nb, int chunk = (int)kXcdChunk) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
  BundleWalk w;
  if ((long)gridDim.x * kRowWaves >= nb) {
    // one bundle per wave; workgroup p (dealt to XCD p % 8) takes logical
    // workgroup (i / C) 8C + (p % 8) C + i % C, i = p / 8: each XCD gets chunks of
    // C consecutive workgroups (the y-neighbour rows of a bundle mostly fall in
    // its own chunk, so in its XCD's L2), while the workgroups in flight still
    // cover one contiguous window of 8C. C = 64 measured best at 256^3 (8..128 swept).
    const int C = (gridDim.x % (unsigned)(kXcds * chunk) == 0) ? chunk : (int)kXcdChunk;
    const int p = blockIdx.x, i = p / kXcds;
    const int logical = (i / C) * (kXcds * C) + (p % kXcds) * C + i % C;
    w.first = logical * kRowWaves + wave;
    w.end = min(w.first + 1, nb);
    w.stride = 1;
    return w;
  }
  const int per_xcd = gridDim.x / kXcds; // a capped grid is a multiple of kXcds
  const int xcd = blockIdx.x % kXcds;
  const int share = (nb + kXcds - 1) / kXcds;
  const int start = xcd * share;
  w.first = start + (int)(blockIdx.x / kXcds) * kRowWaves + wave;
  w.end = min(start + share, nb);
  w.stride = per_xcd * kRowWaves;
  return w;
}
// Bundles of a row kernel over n rows: b = first, first + stride, ... < end.
__device__ __forceinline__ BundleWalk bundle_walk(int n) { return bundle_walk_units((n + kWave - 1) / kWave); }
#endif
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
