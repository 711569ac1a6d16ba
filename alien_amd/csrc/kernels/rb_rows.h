// Row sums over the row-bundled layouts (one wavefront per bundle of 64 rows,
// lane l = row 64 b + l; entry j of the bundle's rows sits at slot
// (bo[b] + j) * 64 + l).
//
//  * RB-CSR (column mode): each slot holds its column (-1 = padding).
//  * BEllPack (slice mode): Alien's sliced ELL read in place, below.
//  * RB-DIA (offset mode): slot j of bundle b is one "diagonal" of the bundle,
//    the j-th smallest distinct column offset c - r among its rows' entries
//    (doff[bo[b] + j]); bit l of dmask[bo[b] + j] says whether row 64 b + l has
//    that entry. No column index is streamed: the gather address is r + d.
//
// In both modes each row's entries are visited in ascending column order and
// absent entries are skipped, so the sum is the sequential CSR sum, bit for bit.
// A batch of slots is loaded before any is used, so a wave keeps that many
// value loads and gathers in flight even for short rows.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.h"
#include "krylov_state.h"

namespace alien_hip {

// stored value -> arithmetic type (bf16 bits are the upper half of an FP32 value)
template <typename T>
__device__ __forceinline__ T widen(T v) {
  return v;
}
template <typename T>
__device__ __forceinline__ T widen(bf16_bits v) {
  return __uint_as_float((unsigned)v << 16);
}

// U slots per batch. Slot indices past the bundle width are clamped to its
// last slot, so every load is in bounds and unconditional (no scalar waits
// between them); the clamped duplicates are masked out of the sum.
template <int U, typename T, typename V, typename XS>
__device__ __forceinline__ T rb_sum_u(const RBMat<T, V>& A, int b, int s0, int w, XS xs) {
  const int lane = threadIdx.x & (kWave - 1);
  const long base = (long)s0 * kWave + lane;
  T acc = 0;
  if (A.doff) {
    const int row = b * kWave + lane;
    const unsigned long long bit = 1ull << lane;
    for (int j = 0; j < w; j += U) {
      int d[U];
      unsigned long long m[U];
      T v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int sl = min(j + u, w - 1);
        d[u] = A.doff[s0 + sl];
        m[u] = A.dmask[s0 + sl];
        v[u] = widen<T>(__builtin_nontemporal_load(A.bv + base + (long)sl * kWave));
      }
      bool on[U];
      T xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        on[u] = (j + u < w) && (m[u] & bit) != 0;
        xv[u] = on[u] ? xs(row + d[u]) : T(0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (on[u]) acc = acc + v[u] * xv[u];
    }
  } else {
    for (int j = 0; j < w; j += U) {
      int c[U];
      T v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int sl = min(j + u, w - 1);
        c[u] = __builtin_nontemporal_load(A.bc + base + (long)sl * kWave);
        v[u] = widen<T>(__builtin_nontemporal_load(A.bv + base + (long)sl * kWave));
      }
      bool on[U];
      T xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        on[u] = (j + u < w) && c[u] >= 0;
        xv[u] = on[u] ? xs(c[u]) : T(0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (on[u]) acc = acc + v[u] * xv[u];
    }
  }
  return acc;
}

// Pair mode of RB-DIA (RBMat::pair): slot j of the bundle is the column pair
// (base + d, base + d + 1), base = row & ~1, d = doff[s0 + j]; the row's two
// values of the pair are taken in column order, as the CSR sum takes them.
template <int U, typename T, typename V, typename XS>
__device__ __forceinline__ T rb_sum_pair_u(const RBMat<T, V>& A, int b, int s0, int w, XS xs) {
  const int lane = threadIdx.x & (kWave - 1);
  const int base = (b * kWave + lane) & ~1;
  const unsigned long long bit = 1ull << lane;
  T acc = 0;
  for (int j = 0; j < w; j += U) {
    int d[U];
    unsigned long long m[U];
    T v0[U], v1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int sl = min(j + u, w - 1);
      d[u] = A.doff[s0 + sl];
      m[u] = A.dmask[s0 + sl];
      const long e = (2L * (s0 + sl)) * kWave + lane;
      v0[u] = widen<T>(__builtin_nontemporal_load(A.bv + e));
      v1[u] = widen<T>(__builtin_nontemporal_load(A.bv + e + kWave));
    }
    bool on[U];
    T x0[U], x1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      on[u] = (j + u < w) && (m[u] & bit) != 0;
      x0[u] = on[u] ? xs(base + d[u]) : T(0);
      x1[u] = on[u] ? xs(base + d[u] + 1) : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (on[u]) {
        acc = acc + v0[u] * x0[u];
        acc = acc + v1[u] * x1[u];
      }
  }
  return acc;
}

// BEllPack mode (RBMat::ell = E bundles per slice): the 64 rows of bundle b are rows
// 64 (b % E) .. + 63 of slice s = b / E, whose entry j sits at
// (bro[s] E + j E + b % E) 64 + lane: the wave reads 64 consecutive values and columns
// per slot, slots 64 E elements apart. Every slot of the slice's width is summed,
// padding (column 0, value 0) included, in slot order: MatrixInternal::mult's
// arithmetic (SYCLBEllPackMatrix.cc:647-685), +0 * x[0] for each padding slot.
template <int U, typename T, typename V, typename XS>
__device__ __forceinline__ T ell_sum_u(const RBMat<T, V>& A, long base, long stride, int w, XS xs) {
  T acc = 0;
  for (int j = 0; j < w; j += U) {
    int c[U];
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long e = base + (long)min(j + u, w - 1) * stride;
      c[u] = __builtin_nontemporal_load(A.bc + e);
      v[u] = widen<T>(__builtin_nontemporal_load(A.bv + e));
    }
    T xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = xs(c[u]);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (j + u < w) acc = acc + v[u] * xv[u];
  }
  return acc;
}
template <typename T, typename V, typename XS>
__device__ __forceinline__ T ell_sum(const RBMat<T, V>& A, int b, XS xs) {
  const int E = A.ell;
  const int s = b / E;
  const int w = A.bo[s + 1] - A.bo[s];
  if (w <= 0) return T(0);
  const long stride = (long)E * kWave;
  const long base = ((long)A.bo[s] * E + (b - s * E)) * kWave + (threadIdx.x & (kWave - 1));
  return w <= 4 ? ell_sum_u<4, T>(A, base, stride, w, xs) : ell_sum_u<7, T>(A, base, stride, w, xs);
}

// Uniform RB-DIA (RBMat::uni = K): the bundle's K slots sit at b K .. b K + K - 1 with the
// column offsets A.uoff (kernel arguments, no per-bundle descriptor loads): one independent value
// load and one gather per slot, so a bundle costs one memory round trip. Absent entries are stored
// as zeros and their gathers clamped into [0, nrows): each present entry is added in ascending
// column order as in the CSR sum, the absent ones add +0 (the same sum up to the sign of a zero).
template <int K, typename T, typename V, typename XS>
__device__ __forceinline__ T uni_sum_k(const RBMat<T, V>& A, int b, XS xs) {
  const int lane = threadIdx.x & (kWave - 1);
  const long row = (long)b * kWave + lane;
  const long base = (long)b * K * kWave + lane;
  T v[K], xv[K];
#pragma unroll
  for (int u = 0; u < K; ++u) v[u] = widen<T>(__builtin_nontemporal_load(A.bv + base + (long)u * kWave));
#pragma unroll
  for (int u = 0; u < K; ++u) {
    const long c = row + A.uoff[u];
    xv[u] = xs((int)(c < 0 ? 0 : c >= A.nrows ? A.nrows - 1 : c));
  }
  T acc = 0;
#pragma unroll
  for (int u = 0; u < K; ++u) acc = acc + v[u] * xv[u];
  return acc;
}
template <typename T, typename V, typename XS>
__device__ __forceinline__ T uni_sum(const RBMat<T, V>& A, int b, XS xs) {
  switch (A.uni) {
  case 4: return uni_sum_k<4, T>(A, b, xs);
  case 7: return uni_sum_k<7, T>(A, b, xs);
  case 1: return uni_sum_k<1, T>(A, b, xs);
  case 2: return uni_sum_k<2, T>(A, b, xs);
  case 3: return uni_sum_k<3, T>(A, b, xs);
  case 5: return uni_sum_k<5, T>(A, b, xs);
  case 6: return uni_sum_k<6, T>(A, b, xs);
  default: return uni_sum_k<8, T>(A, b, xs);
  }
}

// Uniform operators of at most this many slots are walked two bundles at a time by the CG step
// kernel (k_cg_q); the factor kernels always pair.
#ifndef ALIEN_HIP_UNI_PAIR_MAX_SLOTS
#define ALIEN_HIP_UNI_PAIR_MAX_SLOTS 8
#endif
constexpr int kUniPairMaxSlots = ALIEN_HIP_UNI_PAIR_MAX_SLOTS;
// XCD chunk of the CG step's pair walk (one pair per wave on its 32,768 workgroups at 256^3): with
// 1,024-workgroup chunks each XCD's share is a contiguous run of z-planes and p is fetched into its
// L2 1.03 times instead of 3 (measured at 256^3: FETCH_SIZE 270,151 against 328,006 KiB; the time
// is the same, the Infinity Cache served the refetches)
#ifndef ALIEN_HIP_UNI_Q_CHUNK
#define ALIEN_HIP_UNI_Q_CHUNK 1024
#endif
constexpr int kUniQChunk = ALIEN_HIP_UNI_Q_CHUNK;

// Two bundles of a uniform RB-DIA at once (b0 and b1, b1 may repeat b0): both bundles' value
// loads and gathers are issued before any product, twice the memory-level parallelism of a wave
// walking one bundle at a time. Each row's sum is uni_sum_k's.
template <int K, typename T, typename V, typename XS>
__device__ __forceinline__ void uni_sum2_k(const RBMat<T, V>& A, int b0, int b1, XS xs, T& t0, T& t1) {
  const int lane = threadIdx.x & (kWave - 1);
  const long r0 = (long)b0 * kWave + lane, r1 = (long)b1 * kWave + lane;
  const long base0 = (long)b0 * K * kWave + lane, base1 = (long)b1 * K * kWave + lane;
  T v0[K], v1[K], x0[K], x1[K];
#pragma unroll
  for (int u = 0; u < K; ++u) {
    v0[u] = widen<T>(__builtin_nontemporal_load(A.bv + base0 + (long)u * kWave));
    v1[u] = widen<T>(__builtin_nontemporal_load(A.bv + base1 + (long)u * kWave));
  }
  const long last = A.nrows - 1;
#pragma unroll
  for (int u = 0; u < K; ++u) {
    const long c0 = r0 + A.uoff[u], c1 = r1 + A.uoff[u];
    x0[u] = xs((int)(c0 < 0 ? 0 : c0 > last ? last : c0));
    x1[u] = xs((int)(c1 < 0 ? 0 : c1 > last ? last : c1));
  }
  T a0 = 0, a1 = 0;
#pragma unroll
  for (int u = 0; u < K; ++u) {
    a0 = a0 + v0[u] * x0[u];
    a1 = a1 + v1[u] * x1[u];
  }
  t0 = a0;
  t1 = a1;
}
template <typename T, typename V, typename XS>
__device__ __forceinline__ void uni_sum2(const RBMat<T, V>& A, int b0, int b1, XS xs, T& t0, T& t1) {
  switch (A.uni) {
  case 4: uni_sum2_k<4, T>(A, b0, b1, xs, t0, t1); return;
  case 7: uni_sum2_k<7, T>(A, b0, b1, xs, t0, t1); return;
  case 1: uni_sum2_k<1, T>(A, b0, b1, xs, t0, t1); return;
  case 2: uni_sum2_k<2, T>(A, b0, b1, xs, t0, t1); return;
  case 3: uni_sum2_k<3, T>(A, b0, b1, xs, t0, t1); return;
  case 5: uni_sum2_k<5, T>(A, b0, b1, xs, t0, t1); return;
  case 6: uni_sum2_k<6, T>(A, b0, b1, xs, t0, t1); return;
  default: uni_sum2_k<8, T>(A, b0, b1, xs, t0, t1); return;
  }
}

// Row sum of this lane's row of bundle b over the vector given by xs(column).
// Short bundles (factors of M = G G^T: 4 slots at a 7-point stencil) use a
// batch of 4, longer ones a batch of 7.
// rb_sum_single: operands never in pair mode (the learned factors), so the
// kernel carries no pair-mode code.
template <typename T, typename V, typename XS>
__device__ __forceinline__ T rb_sum_single(const RBMat<T, V>& A, int b, XS xs) {
  const int s0 = A.bo[b];
  const int w = A.bo[b + 1] - s0;
  if (w <= 0) return T(0);
  if (w <= 4) return rb_sum_u<4>(A, b, s0, w, xs);
  return rb_sum_u<7>(A, b, s0, w, xs);
}
// LAY selects the layout class at compile time: kernels are instantiated once per class (the
// launchers dispatch with by_layout), so the row-bundled kernels carry no BEllPack or uniform
// code and compile as before.
template <typename T, int LAY = kRowBundled, typename V, typename XS>
__device__ __forceinline__ T rb_sum(const RBMat<T, V>& A, int b, XS xs) {
  if constexpr (LAY == kRowEll) {
    return ell_sum<T>(A, b, xs);
  } else if constexpr (LAY == kRowUniform) {
    return uni_sum<T>(A, b, xs);
  } else {
    if (A.pair) {
      const int s0 = A.bo[b];
      const int w = A.bo[b + 1] - s0;
      if (w <= 0) return T(0);
      return w <= 4 ? rb_sum_pair_u<4>(A, b, s0, w, xs) : rb_sum_pair_u<7>(A, b, s0, w, xs);
    }
    return rb_sum_single<T>(A, b, xs);
  }
}

// Row sum of the operands that are never in pair or BEllPack mode (the learned factors):
// uniform or row-bundled.
template <typename T, int LAY, typename V, typename XS>
__device__ __forceinline__ T rb_factor_sum(const RBMat<T, V>& A, int b, XS xs) {
  if constexpr (LAY == kRowUniform) return uni_sum<T>(A, b, xs);
  else return rb_sum_single<T>(A, b, xs);
}

// Workgroups of a row kernel over operand A: BEllPack (slice mode) runs one bundle per wave
// (measured at 256^3: FP32 157 us uncapped against 192 us at the 4096 cap), the row-bundled
// layouts follow bundle_blocks' caps.
template <typename T, typename V>
inline long rb_blocks(long n, const RBMat<T, V>& A) {
  return A.ell ? bundle_blocks_max(n) : A.uni ? uni_blocks(n, A.uni) : bundle_blocks(n, (int)sizeof(T));
}

// f(std::integral_constant<int, L>{}) with L the layout class of A (host side, at launch)
template <typename T, typename V, typename F>
inline void by_layout(const RBMat<T, V>& A, F f) {
  if (A.ell) f(std::integral_constant<int, kRowEll>{});
  else if (A.uni) f(std::integral_constant<int, kRowUniform>{});
  else f(std::integral_constant<int, kRowBundled>{});
}

} // namespace alien_hip
