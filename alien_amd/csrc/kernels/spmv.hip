// SpMV kernels for gfx950.
//
// Every kernel keeps the reference's per-row summation order (the entries of
// a row in ascending column order, accumulated left to right, separate
// multiply and add), so results are bit-exact with
// SimpleCSRMatrixMultT::_seqMult (simple_csr/algebra/SimpleCSRMatrixMultT.h:172-191);
// what changes between layouts is only how the value and column streams are
// laid out in HBM:
//
//  * CSR: one lane per row over Alien's own arrays (no conversion).
//  * RB-CSR (row-bundled CSR): 64 rows per bundle = one wavefront, entry j of
//    the bundle's rows stored contiguously, so each of the width loads of a
//    wavefront is one 512-byte (FP64) / 256-byte (FP32, int32) coalesced
//    access. Padding slots carry column -1 and are skipped.
//  * BEllPack: Alien's sliced-ELL layout (slice height 1024, padding column 0
//    value 0), computed exactly as MatrixInternal::mult
//    (sycl/data/SYCLBEllPackMatrix.cc:623-688), padding products included.
//
// The matrix streams are read once per SpMV and loaded non-temporally so that
// L2 / Infinity Cache capacity stays with the gathered x vector.
#include "../alien_hip_internal.h"
#include "../alien/hip/HIPContext.h"
#include "device_common.h"
#include "launch.h"
#include "rb_rows.h"

namespace alien_hip {

template <typename T>
__device__ __forceinline__ T stream_load(const T* p) {
  return __builtin_nontemporal_load(p);
}

// ---- CSR, one lane per row ------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_spmv_csr(int n, const int* __restrict__ ro,
                                                      const int* __restrict__ cols,
                                                      const T* __restrict__ vals,
                                                      const T* __restrict__ x, T* __restrict__ y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  T acc = 0;
  const int e = ro[i + 1];
  for (int k = ro[i]; k < e; ++k) acc = acc + vals[k] * x[cols[k]];
  y[i] = acc;
}

// y = A x and out = y . x
template <typename T>
__global__ void __launch_bounds__(kBlock) k_spmv_dot_csr(int n, const int* __restrict__ ro,
                                                          const int* __restrict__ cols,
                                                          const T* __restrict__ vals,
                                                          const T* __restrict__ x, T* __restrict__ y,
                                                          RedWork w) {
  __shared__ double lds4[4];
  double s = 0.0;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    T acc = 0;
    const int e = ro[i + 1];
    for (int k = ro[i]; k < e; ++k) acc = acc + vals[k] * x[cols[k]];
    y[i] = acc;
    s = (double)acc * (double)x[i];
  }
  double v[1] = {block_sum(s, lds4)};
  store_partials<1>(v, w);
}

// ---- RB-CSR / RB-DIA --------------------------------------------------------------
// One wavefront per bundle (rb_rows.h); two entry points so that profiles name
// the layout.
template <typename T, int LAY = kRowBundled>
__device__ __forceinline__ void spmv_rb_body(int n, const RBMat<T>& A, const T* __restrict__ x,
                                             T* __restrict__ y) {
  const int lane = threadIdx.x & (kWave - 1);
  const BundleWalk bw = bundle_walk(n);
  for (int b = bw.first; b < bw.end; b += bw.stride) {
    const T acc = rb_sum<T, LAY>(A, b, [&](int c) { return x[c]; });
    const int row = b * kBundle + lane;
    if (row < n) y[row] = acc;
  }
}
template <typename T>
__global__ void __launch_bounds__(kRowBlock) k_spmv_rbcsr(int n, RBMat<T> A, const T* __restrict__ x,
                                                           T* __restrict__ y) {
  spmv_rb_body<T>(n, A, x, y);
}
template <typename T>
__global__ void __launch_bounds__(kRowBlock) k_spmv_rbdia(int n, RBMat<T> A, const T* __restrict__ x,
                                                           T* __restrict__ y) {
  spmv_rb_body<T>(n, A, x, y);
}
// uniform RB-DIA (RBMat::uni)
template <typename T>
__global__ void __launch_bounds__(kRowBlock) k_spmv_udia(int n, RBMat<T> A, const T* __restrict__ x,
                                                          T* __restrict__ y) {
  spmv_rb_body<T, kRowUniform>(n, A, x, y);
}

// Device conversion CSR -> RB-DIA: lane = row; walks the row's entries and the
// bundle's ascending offsets together, zero where the row has no entry.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_rbdia_fill(int n, const int* __restrict__ ro,
                                                        const int* __restrict__ cols,
                                                        const double* __restrict__ vals,
                                                        const int* __restrict__ bo, const int* __restrict__ doff,
                                                        T* __restrict__ bv) {
  const int nb = (n + kBundle - 1) / kBundle;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int b = t / kBundle, lane = t % kBundle;
  if (b >= nb) return;
  const int row = t;
  const int s0 = bo[b], w = bo[b + 1] - s0;
  const long base = (long)s0 * kBundle + lane;
  int k = 0, e = 0;
  if (row < n) {
    k = ro[row];
    e = ro[row + 1];
  }
  for (int j = 0; j < w; ++j) {
    T v = T(0);
    if (k < e && cols[k] - row == doff[s0 + j]) v = (T)vals[k++];
    bv[base + (long)j * kBundle] = v;
  }
}

// Device RB-DIA structure, one wavefront per bundle (lane = row): the 64 rows'
// ascending offsets col - row are merged by a wave-wide minimum per step, so
// the bundle's distinct offsets come out ascending, each with the ballot of the
// rows that hold it as its lane mask -- the structure rbdia_structure builds on
// the host. COUNT pass: width[b] = number of distinct offsets, or -1 when it
// exceeds max_w or a row's columns are not strictly ascending. WRITE pass:
// offsets and masks at the bundle offsets bo of a successful COUNT pass.
// PAIR: the structure of RBMat's pair mode; every row must hold column pairs
// (c, c + 1) with c even, and a slot is a pair at offset c - (row & ~1).
__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}

template <bool WRITE, bool PAIR>
__global__ void __launch_bounds__(kBlock) k_rbdia_struct(int n, const int* __restrict__ ro,
                                                          const int* __restrict__ cols, int max_w,
                                                          const int* __restrict__ bo, int* __restrict__ width,
                                                          int* __restrict__ doff,
                                                          unsigned long long* __restrict__ dmask) {
  constexpr int kEnd = 0x7fffffff;
  const int nb = (n + kBundle - 1) / kBundle;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int b = t / kBundle, lane = t % kBundle;
  if (b >= nb) return;
  const int row = t;
  int k = 0, e = 0;
  if (row < n) {
    k = ro[row];
    e = ro[row + 1];
  }
  constexpr int step = PAIR ? 2 : 1;
  const int base = PAIR ? (row & ~1) : row;
  // a pair starts at an even column and is followed by the next column
  auto bad_pair = [&](int kk) { return PAIR && kk < e && (kk + 1 >= e || (cols[kk] & 1) || cols[kk + 1] != cols[kk] + 1); };
  int cur = k < e ? cols[k] - base : kEnd;
  bool bad = bad_pair(k);
  const int s0 = WRITE ? bo[b] : 0;
  const int cap = WRITE ? bo[b + 1] - s0 : max_w + 1;
  int w = 0;
  while (w < cap) {
    const int m = wave_min_i32(cur);
    if (m == kEnd) break;
    const bool hit = cur == m;
    const unsigned long long mask = __ballot(hit);
    if (WRITE && lane == 0) {
      doff[s0 + w] = m;
      dmask[s0 + w] = mask;
    }
    ++w;
    if (hit) {
      k += step;
      const int nx = k < e ? cols[k] - base : kEnd;
      bad |= nx <= m || bad_pair(k);
      cur = nx;
    }
  }
  if (!WRITE) {
    const bool any_bad = __ballot(bad) != 0ull;
    if (lane == 0) width[b] = (any_bad || w > max_w) ? -1 : w;
  }
}

// Device conversion CSR -> RB-DIA pair mode: lane = row; a slot whose pair the
// row holds takes its two values (planes 2 s, 2 s + 1), zeros elsewhere.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_rbdia2_fill(int n, const int* __restrict__ ro,
                                                         const int* __restrict__ cols,
                                                         const double* __restrict__ vals,
                                                         const int* __restrict__ bo, const int* __restrict__ doff,
                                                         T* __restrict__ bv) {
  const int nb = (n + kBundle - 1) / kBundle;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int b = t / kBundle, lane = t % kBundle;
  if (b >= nb) return;
  const int row = t;
  const int base = row & ~1;
  const int s0 = bo[b], w = bo[b + 1] - s0;
  int k = 0, e = 0;
  if (row < n) {
    k = ro[row];
    e = ro[row + 1];
  }
  for (int j = 0; j < w; ++j) {
    T v0 = T(0), v1 = T(0);
    if (k + 1 < e && cols[k] - base == doff[s0 + j]) {
      v0 = (T)vals[k];
      v1 = (T)vals[k + 1];
      k += 2;
    }
    const long s = (long)s0 + j;
    bv[(2 * s) * kWave + lane] = v0;
    bv[(2 * s + 1) * kWave + lane] = v1;
  }
}

// Device conversion CSR -> RB-CSR (one lane per row writes its slots).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_rbcsr_fill(int n, const int* __restrict__ ro,
                                                        const int* __restrict__ cols,
                                                        const double* __restrict__ vals,
                                                        const int* __restrict__ bo, int* __restrict__ bc,
                                                        T* __restrict__ bv) {
  const int nb = (n + kBundle - 1) / kBundle;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int b = t / kBundle, lane = t % kBundle;
  if (b >= nb) return;
  const int row = t;
  const int s0 = bo[b], w = bo[b + 1] - s0;
  const long base = (long)s0 * kBundle + lane;
  int len = 0, r0 = 0;
  if (row < n) {
    r0 = ro[row];
    len = ro[row + 1] - r0;
  }
  for (int j = 0; j < w; ++j) {
    const long e = base + (long)j * kBundle;
    if (j < len) {
      bc[e] = cols[r0 + j];
      bv[e] = (T)vals[r0 + j];
    } else {
      bc[e] = -1;
      bv[e] = T(0);
    }
  }
}

// ---- BEllPack (Alien's sliced ELL) ------------------------------------------------------
// y = A x read from the BEllPack arrays in place: one wavefront per 64 rows of a slice
// (rb_rows.h, slice mode), batched coalesced value / column loads, the row's slots
// summed in the reference's order (padding included).
template <typename T>
__global__ void __launch_bounds__(kRowBlock) k_spmv_bellpack(int n, RBMat<T> A, const T* __restrict__ x,
                                                              T* __restrict__ y) {
  spmv_rb_body<T, kRowEll>(n, A, x, y);
}

// Device conversion CSR -> BEllPack (the scatter of SYCLBEllPackMatrix.cc:127-170 and
// 422-460, masks of :235-315): one lane per row slot of the padded slices (rows past n
// included), so every storage entry is written: entry j of row r at
// bro[s] slice + j slice + r % slice, column 0 / value 0 / masks 0 past the row's end.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bellpack_fill(int n, int ns, int slice, const int* __restrict__ ro,
                                                           const int* __restrict__ cols,
                                                           const double* __restrict__ vals,
                                                           const int* __restrict__ dcol,
                                                           const int* __restrict__ bro, int* __restrict__ bc,
                                                           T* __restrict__ bv, uint8_t* __restrict__ lmask,
                                                           uint8_t* __restrict__ umask) {
  const long i = blockIdx.x * (long)kBlock + threadIdx.x;
  if (i >= (long)ns * slice) return;
  const int s = (int)(i / slice), l = (int)(i - (long)s * slice);
  const int w = bro[s + 1] - bro[s];
  const long base = (long)bro[s] * slice + l;
  const int r0 = i < n ? ro[i] : 0;
  const int len = i < n ? ro[i + 1] - r0 : 0;
  const int dk = (i < n && (lmask || umask)) ? dcol[i] : 0;
  for (int j = 0; j < w; ++j) {
    const long e = base + (long)j * slice;
    const bool on = j < len;
    if (bc) bc[e] = on ? cols[r0 + j] : 0;
    if (bv) bv[e] = on ? (T)vals[r0 + j] : T(0);
    if (lmask) lmask[e] = (on && r0 + j < dk) ? 1 : 0;
    if (umask) umask[e] = (on && r0 + j > dk) ? 1 : 0;
  }
}

// the reference's one-work-item-per-row loop, for slice heights that are not a multiple
// of the wavefront
template <typename T>
__global__ void __launch_bounds__(kBlock) k_spmv_bellpack_rows(int n, int slice, const int* __restrict__ bro,
                                                           const int* __restrict__ bc,
                                                           const T* __restrict__ bv,
                                                           const T* __restrict__ x, T* __restrict__ y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int s = i / slice, l = i - s * slice;
  const int w = bro[s + 1] - bro[s];
  const long base = (long)bro[s] * slice + l;
  T acc = 0;
  for (int j = 0; j < w; ++j) {
    const long e = base + (long)j * slice;
    acc = acc + stream_load(bv + e) * x[stream_load(bc + e)];
  }
  y[i] = acc;
}

// y_i += alpha * sum_j mask * a * x  (SYCLBEllPackMatrix.cc:796-835 / 872-908):
// the product is formed as ((alpha * mask) * a) * x, as there.
__global__ void __launch_bounds__(kBlock) k_addmult_bellpack(int n, int slice, double alpha,
                                                              const int* __restrict__ bro,
                                                              const int* __restrict__ bc,
                                                              const double* __restrict__ bv,
                                                              const uint8_t* __restrict__ mask,
                                                              const double* __restrict__ x,
                                                              double* __restrict__ y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int s = i / slice, l = i - s * slice;
  const int w = bro[s + 1] - bro[s];
  const long base = (long)bro[s] * slice + l;
  double acc = y[i];
  for (int j = 0; j < w; ++j) {
    const long e = base + (long)j * slice;
    acc = acc + ((alpha * (double)mask[e]) * bv[e]) * x[bc[e]];
  }
  y[i] = acc;
}

// CSR addLMult / addUMult: tmpy += (alpha * a) * x over k < dcol / k > dcol
__global__ void __launch_bounds__(kBlock) k_addmult_csr(int n, int upper, double alpha,
                                                         const int* __restrict__ ro,
                                                         const int* __restrict__ cols,
                                                         const int* __restrict__ dcol,
                                                         const double* __restrict__ vals,
                                                         const double* __restrict__ x,
                                                         double* __restrict__ y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int b = upper ? dcol[i] + 1 : ro[i];
  const int e = upper ? ro[i + 1] : dcol[i];
  double acc = y[i];
  for (int k = b; k < e; ++k) acc = acc + (alpha * vals[k]) * x[cols[k]];
  y[i] = acc;
}

// ---- 2x2 block CSR ---------------------------------------------------------------------
template <typename T>
struct Vec4;
template <>
struct Vec4<double> { typedef double4 type; typedef double2 half; };
template <>
struct Vec4<float> { typedef float4 type; typedef float2 half; };

template <typename T>
__global__ void __launch_bounds__(kBlock) k_spmv_bsr2(int nb, const int* __restrict__ ro,
                                                       const int* __restrict__ cols,
                                                       const T* __restrict__ vals,
                                                       const T* __restrict__ x, T* __restrict__ y) {
  typedef typename Vec4<T>::type V4;
  typedef typename Vec4<T>::half V2;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nb) return;
  T y0 = 0, y1 = 0;
  const int e = ro[i + 1];
  for (int k = ro[i]; k < e; ++k) {
    const V4 m = reinterpret_cast<const V4*>(vals)[k];
    const V2 xv = reinterpret_cast<const V2*>(x)[cols[k]];
    // row ieq = 0 then 1, columns iu = 0 then 1: m[ieq + 2 iu]
    y0 = y0 + m.x * xv.x;
    y0 = y0 + m.z * xv.y;
    y1 = y1 + m.y * xv.x;
    y1 = y1 + m.w * xv.y;
  }
  V2 out;
  out.x = y0;
  out.y = y1;
  reinterpret_cast<V2*>(y)[i] = out;
}

// ---- row-bundled 2x2 block CSR ("RB-BSR2") -------------------------------------------------
// Bundles of 64 block rows (one wavefront, lane = block row); slot j of bundle
// b holds block j of each lane's row: column at bc[s 64 + lane] (-1 = padding)
// and the four values in planes, value q (column-major m[ieq + 2 iu]) at
// bv[(4 s + q) 64 + lane], s = bo[b] + j. Every load of a slot is one
// contiguous 64-lane access; per lane the blocks and the four products are
// taken in the order of k_spmv_bsr2, so the result is the same bit for bit.
template <typename T>
__global__ void __launch_bounds__(kRowBlock) k_spmv_rbbsr2(int nbr, const int* __restrict__ bo,
                                                            const int* __restrict__ bc,
                                                            const T* __restrict__ bv,
                                                            const T* __restrict__ x, T* __restrict__ y) {
  typedef typename Vec4<T>::half V2;
  constexpr int U = 4;
  const int lane = threadIdx.x & (kWave - 1);
  const BundleWalk bw = bundle_walk(nbr);
  for (int b = bw.first; b < bw.end; b += bw.stride) {
    const int s0 = bo[b], w = bo[b + 1] - s0;
    T y0 = 0, y1 = 0;
    for (int j = 0; j < w; j += U) {
      int c[U];
      T m[U][4];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long sl = (long)s0 + min(j + u, w - 1);
        c[u] = __builtin_nontemporal_load(bc + sl * kWave + lane);
#pragma unroll
        for (int q = 0; q < 4; ++q) m[u][q] = __builtin_nontemporal_load(bv + (4 * sl + q) * kWave + lane);
      }
      bool on[U];
      V2 xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        on[u] = (j + u < w) && c[u] >= 0;
        xv[u] = reinterpret_cast<const V2*>(x)[on[u] ? c[u] : 0];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (on[u]) {
          y0 = y0 + m[u][0] * xv[u].x;
          y0 = y0 + m[u][2] * xv[u].y;
          y1 = y1 + m[u][1] * xv[u].x;
          y1 = y1 + m[u][3] * xv[u].y;
        }
    }
    const int row = b * kWave + lane;
    if (row < nbr) {
      V2 out;
      out.x = y0;
      out.y = y1;
      reinterpret_cast<V2*>(y)[row] = out;
    }
  }
}

// Device conversion 2x2 block CSR -> RB-BSR2 (one lane per block row writes its slots).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_rbbsr2_fill(int nbr, const int* __restrict__ ro,
                                                         const int* __restrict__ cols,
                                                         const double* __restrict__ vals,
                                                         const int* __restrict__ bo, int* __restrict__ bc,
                                                         T* __restrict__ bv) {
  const int nb = (nbr + kBundle - 1) / kBundle;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int b = t / kBundle, lane = t % kBundle;
  if (b >= nb) return;
  const int row = t;
  const int s0 = bo[b], w = bo[b + 1] - s0;
  int k0 = 0, len = 0;
  if (row < nbr) {
    k0 = ro[row];
    len = ro[row + 1] - k0;
  }
  for (int j = 0; j < w; ++j) {
    const long sl = (long)s0 + j;
    const bool on = j < len;
    bc[sl * kWave + lane] = on ? cols[k0 + j] : -1;
#pragma unroll
    for (int q = 0; q < 4; ++q) bv[(4 * sl + q) * kWave + lane] = on ? (T)vals[4L * (k0 + j) + q] : T(0);
  }
}

// ---- variable-block CSR -------------------------------------------------------------------
// _seqMultVariableBlock (SimpleCSRMatrixMultT.h:439-476), one lane per block row:
// output row kr of block row i sums, block by block, the row-major block row kr
// against x's block, in the reference's order.
__global__ void __launch_bounds__(kBlock) k_spmv_vblock(int nbr, const int* __restrict__ ro,
                                                         const int* __restrict__ cols,
                                                         const int* __restrict__ bcols,
                                                         const int* __restrict__ bsize,
                                                         const int* __restrict__ boff,
                                                         const double* __restrict__ vals,
                                                         const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nbr) return;
  const int bsr = bsize[i], k0 = ro[i], k1 = ro[i + 1];
  for (int kr = 0; kr < bsr; ++kr) {
    double acc = 0.;
    for (int j = k0; j < k1; ++j) {
      const int col = cols[j];
      const int bsc = bsize[col];
      const double* m = vals + (long)bcols[j] + (long)kr * bsc;
      const double* xb = x + boff[col];
      for (int kc = 0; kc < bsc; ++kc) acc = acc + m[kc] * xb[kc];
    }
    y[boff[i] + kr] = acc;
  }
}

// dcol[i] = first k of row i with cols[k] >= i (CSRStructInfo::computeUpperDiagOffset,
// simple_csr/CSRStructInfo.h:291-309), one lane per row
__global__ void __launch_bounds__(kBlock) k_dcol(int n, const int* __restrict__ ro, const int* __restrict__ cols,
                                                  int* __restrict__ dcol) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  int k = ro[i];
  const int e = ro[i + 1];
  while (k < e && cols[k] < i) ++k;
  dcol[i] = k;
}

// ---- inverse diagonal -------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_invdiag_csr(int n, const int* __restrict__ dcol,
                                                         const double* __restrict__ vals,
                                                         double* __restrict__ y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) y[i] = 1. / vals[dcol[i]];
}

__global__ void __launch_bounds__(kBlock) k_invdiag_bellpack(int n, int slice, const int* __restrict__ bro,
                                                              const int* __restrict__ bc,
                                                              const double* __restrict__ bv,
                                                              double* __restrict__ y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int s = i / slice, l = i - s * slice;
  const int w = bro[s + 1] - bro[s];
  const long base = (long)bro[s] * slice + l;
  for (int j = 0; j < w; ++j) {
    const long e = base + (long)j * slice;
    if (bc[e] == i && bv[e] != 0) y[i] = 1. / bv[e];
  }
}

// ---- launchers ---------------------------------------------------------------------------
static inline dim3 rows_grid(long n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

template <typename T>
void launch_spmv_csr(hipStream_t st, int n, const int* ro, const int* cols, const T* vals, const T* x, T* y) {
  if (n > 0) hipLaunchKernelGGL(k_spmv_csr<T>, rows_grid(n), dim3(kBlock), 0, st, n, ro, cols, vals, x, y);
}
template <typename T>
void launch_spmv_dot_csr(hipStream_t st, RedWork w, int n, const int* ro, const int* cols, const T* vals,
                         const T* x, T* y, double* out) {
  if (n <= 0) {
    launch_finalize(st, w, 0, 1, kFinStore, out, nullptr);
    return;
  }
  hipLaunchKernelGGL(k_spmv_dot_csr<T>, rows_grid(n), dim3(kBlock), 0, st, n, ro, cols, vals, x, y, w);
  launch_finalize(st, w, rows_grid(n).x, 1, kFinStore, out, nullptr);
}
template <typename T>
void launch_spmv_rb(hipStream_t st, int n, RBMat<T> A, const T* x, T* y) {
  if (n <= 0) return;
  const dim3 g((unsigned)rb_blocks(n, A));
  if (A.ell) hipLaunchKernelGGL(k_spmv_bellpack<T>, g, dim3(kRowBlock), 0, st, n, A, x, y);
  else if (A.uni) hipLaunchKernelGGL(k_spmv_udia<T>, g, dim3(kRowBlock), 0, st, n, A, x, y);
  else if (A.doff) hipLaunchKernelGGL(k_spmv_rbdia<T>, g, dim3(kRowBlock), 0, st, n, A, x, y);
  else hipLaunchKernelGGL(k_spmv_rbcsr<T>, g, dim3(kRowBlock), 0, st, n, A, x, y);
}
template <typename T>
void launch_spmv_rbcsr(hipStream_t st, int n, const int* bo, const int* bc, const T* bv, const T* x, T* y) {
  launch_spmv_rb<T>(st, n, RBMat<T>{bo, bc, bv}, x, y);
}
template <typename T>
void launch_rbdia_fill(hipStream_t st, int n, const int* ro, const int* cols, const double* vals, const int* bo,
                       const int* doff, T* bv) {
  if (n > 0) hipLaunchKernelGGL(k_rbdia_fill<T>, rows_grid(n), dim3(kBlock), 0, st, n, ro, cols, vals, bo, doff, bv);
}
long rbdia_count_device(hipStream_t st, int n, const int* ro, const int* cols, int max_w, std::vector<int>& bo,
                        bool pair) {
  const int nb = (n + kBundle - 1) / kBundle;
  bo.assign(nb + 1, 0);
  if (nb == 0) return 0;
  DeviceArray<int> width(nb);
  if (pair)
    hipLaunchKernelGGL((k_rbdia_struct<false, true>), rows_grid(n), dim3(kBlock), 0, st, n, ro, cols, max_w,
                       (const int*)nullptr, width.data(), (int*)nullptr, (unsigned long long*)nullptr);
  else
    hipLaunchKernelGGL((k_rbdia_struct<false, false>), rows_grid(n), dim3(kBlock), 0, st, n, ro, cols, max_w,
                       (const int*)nullptr, width.data(), (int*)nullptr, (unsigned long long*)nullptr);
  check_launch("rbdia_struct count");
  ALIEN_HIP_CHECK(hipMemcpyAsync(bo.data(), width.data(), nb * sizeof(int), hipMemcpyDeviceToHost, st));
  ALIEN_HIP_CHECK(hipStreamSynchronize(st));
  long acc = 0;
  for (int b = 0; b < nb; ++b) {
    const int w = bo[b];
    if (w < 0) return -1;
    bo[b] = (int)acc;
    acc += w;
    if (acc * kBundle > 0x7fffffffL) return -1;
  }
  bo[nb] = (int)acc;
  return acc;
}

void launch_rbdia_struct(hipStream_t st, int n, const int* ro, const int* cols, const int* bo, int* doff,
                         unsigned long long* dmask, bool pair) {
  if (n <= 0) return;
  if (pair)
    hipLaunchKernelGGL((k_rbdia_struct<true, true>), rows_grid(n), dim3(kBlock), 0, st, n, ro, cols, 0, bo,
                       (int*)nullptr, doff, dmask);
  else
    hipLaunchKernelGGL((k_rbdia_struct<true, false>), rows_grid(n), dim3(kBlock), 0, st, n, ro, cols, 0, bo,
                       (int*)nullptr, doff, dmask);
}
// Uniform RB-DIA structure (RBMat::uni): every bundle gets the K slots of the offsets U (the union of
// the matrix's diagonal offsets, ascending), bo[b] = b K, with each slot's lane mask as an ordinary
// RB-DIA (rows lacking that offset have no bit, and their values are filled as zeros). One wave per
// bundle, lane = row: the row's ascending offsets are merged with U.
struct UniOffsets {
  int k;
  int off[kUniMaxSlots];
};
__global__ void __launch_bounds__(kBlock) k_rbdia_uniform_struct(int n, const int* __restrict__ ro,
                                                                 const int* __restrict__ cols, UniOffsets U,
                                                                 int* __restrict__ bo, int* __restrict__ doff,
                                                                 unsigned long long* __restrict__ dmask) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int nb = (n + kBundle - 1) / kBundle;
  const int b = t / kBundle, lane = t % kBundle;
  if (b >= nb) return; // the whole wave: a wave is one bundle
  const int row = t;
  int k = 0, e = 0;
  if (row < n) {
    k = ro[row];
    e = ro[row + 1];
  }
  for (int j = 0; j < U.k; ++j) {
    const bool on = k < e && cols[k] - row == U.off[j];
    if (on) ++k;
    const unsigned long long m = __ballot(on);
    if (lane == 0) {
      doff[(long)b * U.k + j] = U.off[j];
      dmask[(long)b * U.k + j] = m;
    }
  }
  if (lane == 0) {
    bo[b] = b * U.k;
    if (b == nb - 1) bo[nb] = nb * U.k;
  }
}
void launch_rbdia_uniform_struct(hipStream_t st, int n, const int* ro, const int* cols, int k, const int* off,
                                 int* bo, int* doff, unsigned long long* dmask) {
  if (n <= 0) return;
  if (k < 1 || k > kUniMaxSlots) throw Error(ALIEN_HIP_EINVAL, "uniform RB-DIA: 1 to 8 offsets");
  UniOffsets U{};
  U.k = k;
  for (int j = 0; j < k; ++j) U.off[j] = off[j];
  hipLaunchKernelGGL(k_rbdia_uniform_struct, rows_grid(n), dim3(kBlock), 0, st, n, ro, cols, U, bo, doff, dmask);
}

template <typename T>
void launch_rbdia2_fill(hipStream_t st, int n, const int* ro, const int* cols, const double* vals, const int* bo,
                        const int* doff, T* bv) {
  if (n > 0) hipLaunchKernelGGL(k_rbdia2_fill<T>, rows_grid(n), dim3(kBlock), 0, st, n, ro, cols, vals, bo, doff, bv);
}

long rbdia_structure_device(hipStream_t st, int n, const int* ro, const int* cols, int max_w, std::vector<int>& bo,
                            DeviceArray<int>& d_bo, DeviceArray<int>& doff, DeviceArray<unsigned long long>& dmask) {
  const long slots = rbdia_count_device(st, n, ro, cols, max_w, bo, false);
  if (slots < 0) return -1;
  d_bo.upload(bo.data(), bo.size(), st);
  doff.resize(slots);
  dmask.resize(slots);
  launch_rbdia_struct(st, n, ro, cols, d_bo.data(), doff.data(), dmask.data(), false);
  check_launch("rbdia_struct write");
  return slots;
}

template <typename T>
void launch_rbcsr_fill(hipStream_t st, int n, const int* ro, const int* cols, const double* vals,
                       const int* bo, int* bc, T* bv) {
  if (n > 0) hipLaunchKernelGGL(k_rbcsr_fill<T>, rows_grid(n), dim3(kBlock), 0, st, n, ro, cols, vals, bo, bc, bv);
}
template <typename T>
void launch_spmv_bellpack(hipStream_t st, int n, int slice, const int* bro, const int* bc, const T* bv,
                          const T* x, T* y) {
  if (n <= 0) return;
  if (slice % kWave == 0) {
    RBMat<T> A{bro, bc, bv};
    A.ell = slice / kWave;
    const dim3 g((unsigned)rb_blocks(n, A));
    hipLaunchKernelGGL(k_spmv_bellpack<T>, g, dim3(kRowBlock), 0, st, n, A, x, y);
  } else {
    hipLaunchKernelGGL(k_spmv_bellpack_rows<T>, rows_grid(n), dim3(kBlock), 0, st, n, slice, bro, bc, bv, x, y);
  }
}
template <typename T>
void launch_bellpack_fill(hipStream_t st, int n, int slice, const int* ro, const int* cols, const double* vals,
                          const int* dcol, const int* bro, int* bc, T* bv, uint8_t* lmask, uint8_t* umask) {
  const int ns = (n + slice - 1) / slice;
  const long lanes = (long)ns * slice;
  if (lanes > 0)
    hipLaunchKernelGGL(k_bellpack_fill<T>, dim3((unsigned)((lanes + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, n,
                       ns, slice, ro, cols, vals, dcol, bro, bc, bv, lmask, umask);
}
void launch_addmult_bellpack(hipStream_t st, int n, int slice, double alpha, const int* bro, const int* bc,
                             const double* bv, const uint8_t* mask, const double* x, double* y) {
  if (n > 0)
    hipLaunchKernelGGL(k_addmult_bellpack, rows_grid(n), dim3(kBlock), 0, st, n, slice, alpha, bro, bc, bv,
                       mask, x, y);
}
void launch_addmult_csr(hipStream_t st, int n, int upper, double alpha, const int* ro, const int* cols,
                        const int* dcol, const double* vals, const double* x, double* y) {
  if (n > 0)
    hipLaunchKernelGGL(k_addmult_csr, rows_grid(n), dim3(kBlock), 0, st, n, upper, alpha, ro, cols, dcol,
                       vals, x, y);
}
template <typename T>
void launch_spmv_bsr2(hipStream_t st, int nb, const int* ro, const int* cols, const T* vals, const T* x, T* y) {
  if (nb > 0) hipLaunchKernelGGL(k_spmv_bsr2<T>, rows_grid(nb), dim3(kBlock), 0, st, nb, ro, cols, vals, x, y);
}
template <typename T>
void launch_spmv_rbbsr2(hipStream_t st, int nbr, const int* bo, const int* bc, const T* bv, const T* x, T* y) {
  if (nbr > 0)
    hipLaunchKernelGGL(k_spmv_rbbsr2<T>, dim3((unsigned)bundle_blocks(nbr, (int)sizeof(T))), dim3(kRowBlock), 0, st,
                       nbr, bo, bc, bv, x, y);
}
template <typename T>
void launch_rbbsr2_fill(hipStream_t st, int nbr, const int* ro, const int* cols, const double* vals, const int* bo,
                        int* bc, T* bv) {
  if (nbr > 0)
    hipLaunchKernelGGL(k_rbbsr2_fill<T>, rows_grid(nbr), dim3(kBlock), 0, st, nbr, ro, cols, vals, bo, bc, bv);
}
void launch_spmv_vblock(hipStream_t st, int nbr, const int* ro, const int* cols, const int* bcols, const int* bsize,
                        const int* boff, const double* vals, const double* x, double* y) {
  if (nbr > 0)
    hipLaunchKernelGGL(k_spmv_vblock, rows_grid(nbr), dim3(kBlock), 0, st, nbr, ro, cols, bcols, bsize, boff, vals,
                       x, y);
}
void launch_dcol(hipStream_t st, int n, const int* ro, const int* cols, int* dcol) {
  if (n > 0) hipLaunchKernelGGL(k_dcol, rows_grid(n), dim3(kBlock), 0, st, n, ro, cols, dcol);
}
void launch_invdiag_csr(hipStream_t st, int n, const int* dcol, const double* vals, double* y) {
  if (n > 0) hipLaunchKernelGGL(k_invdiag_csr, rows_grid(n), dim3(kBlock), 0, st, n, dcol, vals, y);
}
void launch_invdiag_bellpack(hipStream_t st, int n, int slice, const int* bro, const int* bc, const double* bv,
                             double* y) {
  if (n > 0)
    hipLaunchKernelGGL(k_invdiag_bellpack, rows_grid(n), dim3(kBlock), 0, st, n, slice, bro, bc, bv, y);
}

#define ALIEN_HIP_INST_SPMV(T)                                                                          \
  template void launch_spmv_csr<T>(hipStream_t, int, const int*, const int*, const T*, const T*, T*);   \
  template void launch_spmv_dot_csr<T>(hipStream_t, RedWork, int, const int*, const int*, const T*,     \
                                       const T*, T*, double*);                                          \
  template void launch_spmv_rbcsr<T>(hipStream_t, int, const int*, const int*, const T*, const T*, T*); \
  template void launch_spmv_rb<T>(hipStream_t, int, RBMat<T>, const T*, T*);                           \
  template void launch_rbdia_fill<T>(hipStream_t, int, const int*, const int*, const double*, const int*,\
                                     const int*, T*);                                                    \
  template void launch_rbdia2_fill<T>(hipStream_t, int, const int*, const int*, const double*,           \
                                      const int*, const int*, T*);                                       \
  template void launch_rbcsr_fill<T>(hipStream_t, int, const int*, const int*, const double*,           \
                                     const int*, int*, T*);                                             \
  template void launch_spmv_bellpack<T>(hipStream_t, int, int, const int*, const int*, const T*,        \
                                        const T*, T*);                                                  \
  template void launch_bellpack_fill<T>(hipStream_t, int, int, const int*, const int*, const double*,    \
                                        const int*, const int*, int*, T*, uint8_t*, uint8_t*);          \
  template void launch_spmv_bsr2<T>(hipStream_t, int, const int*, const int*, const T*, const T*, T*);     \
  template void launch_spmv_rbbsr2<T>(hipStream_t, int, const int*, const int*, const T*, const T*, T*);   \
  template void launch_rbbsr2_fill<T>(hipStream_t, int, const int*, const int*, const double*, const int*, \
                                      int*, T*);
ALIEN_HIP_INST_SPMV(double)
ALIEN_HIP_INST_SPMV(float)

} // namespace alien_hip
