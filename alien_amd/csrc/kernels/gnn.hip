// Learned preconditioner M = G G^T: graph network forward, backward, losses and
// optimiser on gfx950 (no counterpart in the reference; see the oracle's
// statement of the model in oracle/gnn_oracle.py for the mathematics).
//
// Model, for an SPD matrix A with a symmetric pattern (d_i = a_ii):
//   e_ij = -a_ij / sqrt(d_i d_j),  f_i = [1, s_i, 1 - s_i, max_j e_ij],  s_i = sum_j e_ij
//   h0 = relu(f W_in + b_in)                                      (N x 32)
//   h_l = h_{l-1} + relu([h_{l-1}, E h_{l-1}] W_l + b_l)   l = 1..3
//   Ghat_ii = exp(h3_i . w_d + c_d),  Ghat_ij = e_ij (h3_i . w_a + h3_j . w_b + c_e)  (j < i)
//   G = D^-1/2 Ghat,  M = G G^T
// Dense layers run on the matrix cores: v_mfma_f32_32x32x8_bf16 with bf16
// weights and bf16-rounded inputs, FP32 accumulation. Every reduction over
// rows (losses, weight gradients) uses per-workgroup partials summed in a fixed
// order: the results are deterministic.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <vector>

#include "../alien_hip_internal.h"
#include "device_common.h"
#include "gnn.h"
#include "launch.h"

namespace alien_hip {
namespace gnn {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// float -> bfloat16 bits, round to nearest even (NaN stays NaN, quieted). Branch-free: both
// results are formed and one selected (v_cndmask), where an early return for NaN made every
// conversion a divergent branch and left the dense layers' operand conversion instruction-bound.
__device__ __forceinline__ short to_bf16(float x) {
  const unsigned u = __float_as_uint(x);
  const unsigned rounded = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
  const unsigned quiet = (u >> 16) | 0x40u;
  return (short)((u & 0x7fffffffu) > 0x7f800000u ? quiet : rounded);
}
__device__ __forceinline__ float bf16_round(float x) {
  return __uint_as_float(((unsigned)(unsigned short)to_bf16(x)) << 16);
}

// ---- graph preparation ----------------------------------------------------------------
__global__ void k_prep_dis(int n, const int* __restrict__ dcol, const double* __restrict__ vals,
                           float* __restrict__ dis) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dis[i] = (float)(1.0 / sqrt(vals[dcol[i]]));
}
// e per CSR entry (0 on the diagonal) and the node features
__global__ void k_prep_feat(int n, const int* __restrict__ ro, const int* __restrict__ cols,
                            const double* __restrict__ vals, const float* __restrict__ dis, float* __restrict__ e,
                            float* __restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f, mx = 0.f, up = 0.f, lo = 0.f;
  bool any = false;
  for (int k = ro[i]; k < ro[i + 1]; ++k) {
    const int j = cols[k];
    float v = 0.f;
    if (j != i) {
      v = (float)(-vals[k] * (double)dis[i] * (double)dis[j]);
      s += v;
      mx = any ? fmaxf(mx, v) : v;
      any = true;
      if (j > i) up += v * v;
      else lo += v * v;
    }
    e[k] = v;
  }
  float* fi = f + (long)i * kFeat;
  fi[0] = 1.f;
  fi[1] = s;
  fi[2] = 1.f - s;
  fi[3] = mx;
  fi[4] = up;
  fi[5] = lo;
}

// ---- dense layer on MFMA --------------------------------------------------------------------
// Z = [X0 | X1] W + b ; H = Hres + relu(Z)   (Hres may be NULL: H = relu(Z))
// X0 holds input columns [0, c0), X1 columns [c0, kin); kin <= KIN (multiple of 8).
// One wavefront computes 32 rows x 32 outputs: KIN/8 mfma 32x32x8 steps.
// Lane l holds A[row = l%32][k = 4(l/32)..+3], B[k][col = l%32], and the result
// D[8 (r/4) + 4 (l/32) + r%4][l%32] in accumulator register r.
//
// Message-passing layers (KIN = 64: X0 = h, X1 = m, both n x 32, Hres = h): each wave first copies
// its 32 rows of h and m into LDS with 16-byte loads (a contiguous 4 KB block each), then reads
// the MFMA operands and the residual term from there: h is read from memory once, not twice, and
// the row-strided 4-byte operand loads become coalesced 16-byte ones. Same values, same MFMA
// sequence: the results are those of the direct form below.
constexpr int kDenseLdsPitch = 2 * kHidden + 1; // one pad float: the 32 rows of a column hit 32 banks
template <int KIN>
__global__ void __launch_bounds__(256) k_dense(int n, const float* __restrict__ X0, int ld0, int c0,
                                               const float* __restrict__ X1, int ld1, int kin,
                                               const float* __restrict__ W, const float* __restrict__ b,
                                               const float* __restrict__ Hres, float* __restrict__ Z,
                                               float* __restrict__ Hout) {
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * 256 + threadIdx.x) >> 6;
  const int r0 = wave * 32;
  const int kb = lane >> 5;
  const int col = lane & 31;
  if constexpr (KIN == 2 * kHidden) {
    __shared__ float xs[4][32][kDenseLdsPitch];
    const bool staged = ld0 == kHidden && ld1 == kHidden && c0 == kHidden && kin == KIN && (!Hres || Hres == X0);
    if (staged) { // uniform over the grid
      float(*x)[kDenseLdsPitch] = xs[threadIdx.x >> 6];
      // the lane's weights for every step, loaded beside the activations (one round trip for both,
      // where loading them inside the step loop cost a round trip per step)
      float wv[KIN / 8][4];
#pragma unroll
      for (int ks = 0; ks < KIN / 8; ++ks)
#pragma unroll
        for (int v = 0; v < 4; ++v) wv[ks][v] = W[(ks * 8 + kb * 4 + v) * kHidden + col];
      const float bias = b[col];
#pragma unroll
      for (int part = 0; part < 2; ++part) {
        const float* src = part ? X1 : X0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int f = j * 64 + lane; // float4 index in the wave's 32 x 32 block
          const int rr = f >> 3, c4 = (f & 7) * 4;
          float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
          if (r0 + rr < n) v = *reinterpret_cast<const float4*>(src + (long)(r0 + rr) * kHidden + c4);
          float* d = &x[rr][part * kHidden + c4];
          d[0] = v.x;
          d[1] = v.y;
          d[2] = v.z;
          d[3] = v.w;
        }
      }
      __syncthreads(); // every wave reaches it (no early return above)
      if (r0 >= n) return;
      f32x16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < KIN / 8; ++ks) {
        bf16x4 a, bb;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          a[v] = to_bf16(x[lane & 31][ks * 8 + kb * 4 + v]);
          bb[v] = to_bf16(wv[ks][v]);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, bb, acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int rl = 8 * (r >> 2) + 4 * kb + (r & 3);
        const int row = r0 + rl;
        if (row < n) {
          const float z = acc[r] + bias;
          const long o = (long)row * kHidden + col;
          if (Z) Z[o] = z;
          Hout[o] = (Hres ? x[rl][col] : 0.f) + fmaxf(z, 0.f);
        }
      }
      return;
    }
  }
  if (r0 >= n) return;
  const int arow = r0 + (lane & 31);
  f32x16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < KIN / 8; ++ks) {
    bf16x4 a, bb;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = ks * 8 + kb * 4 + v;
      float xa = 0.f, wb = 0.f;
      if (k < kin) {
        if (arow < n) xa = (k < c0) ? X0[(long)arow * ld0 + k] : X1[(long)arow * ld1 + (k - c0)];
        wb = W[k * kHidden + col];
      }
      a[v] = to_bf16(xa);
      bb[v] = to_bf16(wb);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, bb, acc, 0, 0, 0);
  }
  const float bias = b[col];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = r0 + 8 * (r >> 2) + 4 * kb + (r & 3);
    if (row < n) {
      const float z = acc[r] + bias;
      const long o = (long)row * kHidden + col;
      if (Z) Z[o] = z;
      Hout[o] = (Hres ? Hres[o] : 0.f) + fmaxf(z, 0.f);
    }
  }
}

// m_i = sum_{j != i} e_ij h_j   (32 lanes per row, one feature per lane)
// 8 lanes per row, 4 features per lane (one 16-byte load of the neighbour's row
// chunk), 32 rows per workgroup; workgroups are remapped so that each XCD takes
// chunks of consecutive rows (its L2 then holds the x- and y-neighbours' rows).
// Per feature the sum runs over the row's entries in order, as one lane per feature.
constexpr int kAggRows = 32;
constexpr int kAggChunk = 64;
__global__ void __launch_bounds__(256) k_aggregate(int n, const int* __restrict__ ro, const int* __restrict__ cols,
                                                   const float* __restrict__ e, const float* __restrict__ h,
                                                   float* __restrict__ m) {
  const int p = blockIdx.x, q = p / kXcds;
  const int logical = (q / kAggChunk) * (kXcds * kAggChunk) + (p % kXcds) * kAggChunk + q % kAggChunk;
  const int i = logical * kAggRows + (threadIdx.x >> 3);
  const int c4 = threadIdx.x & 7;
  if (i >= n) return;
  float4 acc = {0.f, 0.f, 0.f, 0.f};
  const float4* __restrict__ h4 = reinterpret_cast<const float4*>(h);
  for (int k = ro[i]; k < ro[i + 1]; ++k) {
    const float w = e[k];
    if (w != 0.f) {
      const float4 v = h4[(long)cols[k] * (kHidden / 4) + c4];
      acc.x += w * v.x;
      acc.y += w * v.y;
      acc.z += w * v.z;
      acc.w += w * v.w;
    }
  }
  reinterpret_cast<float4*>(m)[(long)i * (kHidden / 4) + c4] = acc;
}
inline dim3 aggregate_grid(int n) {
  const long g = (n + kAggRows - 1) / kAggRows;
  const long q = (long)kXcds * kAggChunk;
  return dim3((unsigned)((g + q - 1) / q * q));
}

// ---- output head ------------------------------------------------------------------------------
__device__ __forceinline__ float dot32(const float* __restrict__ a, const float* __restrict__ w) {
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < kHidden; ++q) s += a[q] * w[q];
  return s;
}

// ghat per CSR entry: diagonal exp(h_i.w_d + c_d), strictly lower e_ij (h_i.w_a + h_j.w_b + c_e),
// upper entries 0 (G is lower triangular)
__global__ void k_head(int n, const int* __restrict__ ro, const int* __restrict__ cols, const int* __restrict__ dcol,
                       const float* __restrict__ e, const float* __restrict__ h, const float* __restrict__ P,
                       float* __restrict__ ghat) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* hi = h + (long)i * kHidden;
  const float ai = dot32(hi, P + kOffWa);
  ghat[dcol[i]] = expf(dot32(hi, P + kOffWd) + P[kOffCd]);
  for (int k = ro[i]; k < dcol[i]; ++k) {
    const float* hj = h + (long)cols[k] * kHidden;
    ghat[k] = e[k] * (ai + dot32(hj, P + kOffWb) + P[kOffCe]);
  }
  for (int k = dcol[i] + 1; k < ro[i + 1]; ++k) ghat[k] = 0.f;
}

// ---- losses ------------------------------------------------------------------------------------
// Kaporin term. For each lower entry (i,j) of Ghat: AG_ij = sum_k Ahat_ik Ghat_kj,
// Ahat_ik = 1 (k = i) or -e_ik; Ghat_kj read at the position of column j in row k.
__device__ __forceinline__ int find_col(const int* __restrict__ cols, int b, int e, int j) {
  for (int k = b; k < e; ++k)
    if (cols[k] == j) return k;
  return -1;
}
// The gradient uses the symmetric part, ((A + A^T) Ghat)_ij / 2, read through the
// transpose positions (A^T)_ic = A_ci = -e at tpos (equal to AG for symmetric A).
__global__ void __launch_bounds__(256) k_kcond(int n, const int* __restrict__ ro, const int* __restrict__ cols,
                                               const int* __restrict__ dcol, const int* __restrict__ tpos,
                                               const float* __restrict__ e, const float* __restrict__ ghat,
                                               float* __restrict__ ag, RedWork w) {
  __shared__ double lds[4];
  double tr = 0.0, ld = 0.0;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    for (int p = ro[i]; p <= dcol[i]; ++p) { // lower entries and the diagonal of row i
      const int j = cols[p];
      float s = 0.f, st = 0.f;
      for (int k2 = ro[i]; k2 < ro[i + 1]; ++k2) {
        const int c = cols[k2];
        if (c < j) continue;                       // Ghat_cj = 0 above the diagonal
        const int q = find_col(cols, ro[c], dcol[c] + 1, j); // (c, j) in c's lower part
        if (q < 0) continue;
        const float a = (c == i) ? 1.f : -e[k2];
        const float at = (c == i) ? 1.f : -e[tpos[k2]];
        s += a * ghat[q];
        st += at * ghat[q];
      }
      ag[p] = 0.5f * (s + st);
      tr += (double)ghat[p] * (double)s;
    }
    ld = log((double)ghat[dcol[i]]);
  }
  double v[2];
  v[0] = block_sum(tr, lds);
  __syncthreads();
  v[1] = block_sum(ld, lds);
  store_partials<2>(v, w);
}

// probe term, K probes stored row-major (N x K)
template <int K>
__global__ void k_ahat_mult(int n, const int* __restrict__ ro, const int* __restrict__ cols, const float* __restrict__ e,
                            const float* __restrict__ V, float* __restrict__ W) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float acc[K];
#pragma unroll
  for (int c = 0; c < K; ++c) acc[c] = 0.f;
  for (int k = ro[i]; k < ro[i + 1]; ++k) {
    const int j = cols[k];
    const float a = (j == i) ? 1.f : -e[k];
#pragma unroll
    for (int c = 0; c < K; ++c) acc[c] += a * V[(long)j * K + c];
  }
#pragma unroll
  for (int c = 0; c < K; ++c) W[(long)i * K + c] = acc[c];
}
// Y = Ghat X (lower rows) or Y = Ghat^T X (upper rows through the transpose positions)
template <int K, bool TRANS>
__global__ void k_ghat_mult(int n, const int* __restrict__ ro, const int* __restrict__ cols, const int* __restrict__ dcol,
                            const int* __restrict__ tpos, const float* __restrict__ ghat, const float* __restrict__ X,
                            float* __restrict__ Y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float acc[K];
#pragma unroll
  for (int c = 0; c < K; ++c) acc[c] = 0.f;
  const int b = TRANS ? dcol[i] : ro[i];
  const int en = TRANS ? ro[i + 1] : dcol[i] + 1;
  for (int k = b; k < en; ++k) {
    const int j = cols[k];
    const float g = TRANS ? ghat[tpos[k]] : ghat[k];
#pragma unroll
    for (int c = 0; c < K; ++c) acc[c] += g * X[(long)j * K + c];
  }
#pragma unroll
  for (int c = 0; c < K; ++c) Y[(long)i * K + c] = acc[c];
}
// R = V - Y, dY = -2 R / (N K), loss partial sum R^2
template <int K>
__global__ void __launch_bounds__(256) k_probe_resid(int n, const float* __restrict__ V, const float* __restrict__ Y,
                                                     float* __restrict__ dY, float scale, RedWork w) {
  __shared__ double lds[4];
  double s = 0.0;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const float r = V[(long)i * K + c] - Y[(long)i * K + c];
      dY[(long)i * K + c] = -2.f * r * scale;
      s += (double)r * (double)r;
    }
  }
  double v[1] = {block_sum(s, lds)};
  store_partials<1>(v, w);
}

// ---- dL/dGhat -> dL/d(head inputs) --------------------------------------------------------------
// g_ij (per lower CSR entry) = kw (2 AG_ij / tr - (2/N) [i==j] / Ghat_ii) + beta (sum_c dY_ic U_jc + W_ic Q_jc)
template <int K>
__global__ void k_grad_ghat(int n, const int* __restrict__ ro, const int* __restrict__ cols, const int* __restrict__ dcol,
                            const float* __restrict__ ghat, const float* __restrict__ ag, const double* __restrict__ sums,
                            float kw, float beta, const float* __restrict__ dY, const float* __restrict__ U,
                            const float* __restrict__ Wp, const float* __restrict__ Q, float* __restrict__ gg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float tr = kw != 0.f ? (float)sums[0] : 1.f;
  for (int k = ro[i]; k <= dcol[i]; ++k) {
    const int j = cols[k];
    float g = 0.f;
    if (kw != 0.f) {
      g = 2.f * ag[k] / tr;
      if (k == dcol[i]) g -= 2.f / ((float)n * ghat[k]);
      g *= kw;
    }
    if (beta != 0.f) {
      float p = 0.f;
#pragma unroll
      for (int c = 0; c < K; ++c) p += dY[(long)i * K + c] * U[(long)j * K + c] + Wp[(long)i * K + c] * Q[(long)j * K + c];
      g += beta * p;
    }
    gg[k] = g;
  }
}

// ---- k-step CG loss (N x 4 column blocks, one right-hand side per column) ----
struct Coef4 {
  float c[4];
};
// per-column dot products X_c . Y_c into 4 partial slots
__global__ void __launch_bounds__(256) k_dot4(int n, const float* __restrict__ X, const float* __restrict__ Y,
                                              RedWork w) {
  __shared__ double lds[4];
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] = (double)X[(long)i * 4 + c] * (double)Y[(long)i * 4 + c];
  }
  double v[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    v[c] = block_sum(a[c], lds);
    __syncthreads();
  }
  store_partials<4>(v, w);
}
// Y_c = a_c X_c + b_c Z_c (Z may be NULL; Y may alias X or Z)
__global__ void k_lin4(int n, Coef4 a, const float* X, Coef4 b, const float* Z, float* Y) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n * 4) return;
  const int c = (int)(t & 3);
  float v = a.c[c] * X[t];
  if (Z) v += b.c[c] * Z[t];
  Y[t] = v;
}
// dL/dGhat_il += sum_c U_ic V_lc (+ W_ic X_lc) over the lower pattern (l <= i)
__global__ void k_edge_acc4(int n, const int* __restrict__ ro, const int* __restrict__ cols,
                            const int* __restrict__ dcol, const float* __restrict__ U, const float* __restrict__ V,
                            const float* __restrict__ W, const float* __restrict__ X, float* __restrict__ gg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = ro[i]; k <= dcol[i]; ++k) {
    const int l = cols[k];
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      s += U[(long)i * 4 + c] * V[(long)l * 4 + c];
      if (W) s += W[(long)i * 4 + c] * X[(long)l * 4 + c];
    }
    gg[k] += s;
  }
}

// head backward: do_i = g_ii ghat_ii, dc_ij = g_ij e_ij;
// dh_i = do_i w_d + sum_{j<i} dc_ij w_a + sum_{m>i} dc_mi w_b ; per-row parameter-gradient terms
// for (w_d, c_d, w_a, w_b, c_e) go to `prow` (N x kHeadGrads) and are summed by k_colsum.
__global__ void k_head_bwd(int n, const int* __restrict__ ro, const int* __restrict__ cols, const int* __restrict__ dcol,
                           const int* __restrict__ tpos, const float* __restrict__ e, const float* __restrict__ ghat,
                           const float* __restrict__ gg, const float* __restrict__ h, const float* __restrict__ P,
                           float* __restrict__ dh, float* __restrict__ prow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* hi = h + (long)i * kHidden;
  const float dO = gg[dcol[i]] * ghat[dcol[i]];
  float sa = 0.f, sb = 0.f;      // sum of dc over row i's lower edges / over edges (m, i), m > i
  float wb_acc[kHidden];
#pragma unroll
  for (int q = 0; q < kHidden; ++q) wb_acc[q] = 0.f;
  for (int k = ro[i]; k < dcol[i]; ++k) sa += gg[k] * e[k];
  for (int k = dcol[i] + 1; k < ro[i + 1]; ++k) {
    const int t = tpos[k]; // position of (m, i) in row m, m = cols[k] > i
    const float dc = gg[t] * e[t];
    sb += dc;
    const float* hm = h + (long)cols[k] * kHidden;
#pragma unroll
    for (int q = 0; q < kHidden; ++q) wb_acc[q] += dc * hi[q];
    (void)hm;
  }
  float* pr = prow + (long)i * kHeadGrads;
#pragma unroll
  for (int q = 0; q < kHidden; ++q) {
    dh[(long)i * kHidden + q] = dO * P[kOffWd + q] + sa * P[kOffWa + q] + sb * P[kOffWb + q];
    pr[q] = dO * hi[q];               // w_d
    pr[kHidden + 1 + q] = sa * hi[q]; // w_a: sum over lower edges (i,j) of dc_ij h_i
    pr[2 * kHidden + 1 + q] = wb_acc[q]; // w_b: sum over edges (m,i) of dc_mi h_i
  }
  pr[kHidden] = dO;                    // c_d
  pr[3 * kHidden + 1] = sa;            // c_e (each lower edge counted once, in its row)
}

// dense layer backward: dz = dh * (z > 0) ; du = dz W^T (W bf16-rounded) ;
// dh_prev = dh + du[:, :32] (the E du[:, 32:] part is added by k_aggregate_add) ; dmsg = du[:, 32:]
template <int KIN>
__global__ void k_dense_bwd(int n, const float* __restrict__ dh, const float* __restrict__ Z, const float* __restrict__ W,
                            int kin, float* __restrict__ dz, float* __restrict__ dh_prev, float* __restrict__ dmsg) {
  __shared__ float Ws[KIN * kHidden];
  for (int t = threadIdx.x; t < KIN * kHidden; t += blockDim.x)
    Ws[t] = (t / kHidden < kin) ? bf16_round(W[t]) : 0.f;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float d[kHidden];
#pragma unroll
  for (int q = 0; q < kHidden; ++q) {
    const long o = (long)i * kHidden + q;
    d[q] = Z[o] > 0.f ? dh[o] : 0.f;
    dz[o] = d[q];
  }
  if (!dh_prev) return;
  for (int k = 0; k < kin; ++k) {
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < kHidden; ++q) s += d[q] * Ws[k * kHidden + q];
    if (k < kHidden) dh_prev[(long)i * kHidden + k] = dh[(long)i * kHidden + k] + s;
    else dmsg[(long)i * kHidden + (k - kHidden)] = s;
  }
}

// The same for the message-passing layers (kin = 64) on MFMA: du = dz W^T is a 32-row x 64 x 32
// product per wave, 2 output tiles of 16 v_mfma_f32_32x32x2_f32 steps (FP32 operands: dz and the
// bf16-rounded weights, as in the scalar form; only the summation order of the 32 terms differs).
// Each wave stages its 32 rows of dh and dz in LDS (16-byte loads), the workgroup the weights.
__global__ void __launch_bounds__(256) k_dense_bwd_mfma(int n, const float* __restrict__ dh,
                                                        const float* __restrict__ Z, const float* __restrict__ W,
                                                        float* __restrict__ dz, float* __restrict__ dh_prev,
                                                        float* __restrict__ dmsg) {
  __shared__ float ds[4][32][kHidden + 1];
  __shared__ float hs[4][32][kHidden + 1];
  __shared__ float ws[2 * kHidden][kHidden + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r0 = ((blockIdx.x * 256 + threadIdx.x) >> 6) * 32;
  for (int t = threadIdx.x; t < 2 * kHidden * kHidden; t += 256) ws[t / kHidden][t % kHidden] = bf16_round(W[t]);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int f = j * 64 + lane; // float4 index in the wave's 32 x 32 block
    const int rr = f >> 3, c4 = (f & 7) * 4;
    float4 h = make_float4(0.f, 0.f, 0.f, 0.f), z = make_float4(0.f, 0.f, 0.f, 0.f);
    const long o = (long)(r0 + rr) * kHidden + c4;
    if (r0 + rr < n) {
      h = *reinterpret_cast<const float4*>(dh + o);
      z = *reinterpret_cast<const float4*>(Z + o);
    }
    const float4 d = make_float4(z.x > 0.f ? h.x : 0.f, z.y > 0.f ? h.y : 0.f, z.z > 0.f ? h.z : 0.f,
                                 z.w > 0.f ? h.w : 0.f);
    if (r0 + rr < n) *reinterpret_cast<float4*>(dz + o) = d;
    hs[w][rr][c4] = h.x;
    hs[w][rr][c4 + 1] = h.y;
    hs[w][rr][c4 + 2] = h.z;
    hs[w][rr][c4 + 3] = h.w;
    ds[w][rr][c4] = d.x;
    ds[w][rr][c4 + 1] = d.y;
    ds[w][rr][c4 + 2] = d.z;
    ds[w][rr][c4 + 3] = d.w;
  }
  __syncthreads(); // every wave reaches it
  if (r0 >= n) return;
  const int half = lane >> 5, col = lane & 31;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    f32x16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int qs = 0; qs < kHidden / 2; ++qs) {
      const int q = 2 * qs + half;
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ds[w][col][q], ws[t * kHidden + col][q], acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int rl = 8 * (r >> 2) + 4 * half + (r & 3);
      const int row = r0 + rl;
      if (row < n) {
        const long o = (long)row * kHidden + col;
        if (t == 0) dh_prev[o] = hs[w][rl][col] + acc[r];
        else dmsg[o] = acc[r];
      }
    }
  }
}

// dh[i] += sum_{j != i} e_ji dmsg[j]   (E^T dmsg; e_ji read at the transpose position)
__global__ void __launch_bounds__(256) k_aggregate_add(int n, const int* __restrict__ ro, const int* __restrict__ cols,
                                                       const int* __restrict__ tpos, const float* __restrict__ e,
                                                       const float* __restrict__ dmsg, float* __restrict__ dh) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  const int i = (int)(t >> 5);
  const int c = (int)(t & 31);
  if (i >= n) return;
  float acc = 0.f;
  for (int k = ro[i]; k < ro[i + 1]; ++k) {
    const float w = e[tpos[k]];
    if (w != 0.f) acc += w * dmsg[(long)cols[k] * kHidden + c];
  }
  dh[(long)i * kHidden + c] += acc;
}

// weight-gradient partials: per workgroup of 256 rows, sum_r bf16(X_r) (x) dz_r  (kin x 32) and sum_r dz_r (32)
template <int KIN>
__global__ void __launch_bounds__(256) k_wgrad_partials(int n, const float* __restrict__ X0, int ld0, int c0,
                                                        const float* __restrict__ X1, int ld1, int kin,
                                                        const float* __restrict__ dz, double* __restrict__ part,
                                                        int cap) {
  constexpr int NOUT = KIN * kHidden + kHidden;
  constexpr int CH = 32; // rows staged per chunk
  __shared__ float xs[CH][KIN];
  __shared__ float ds[CH][kHidden];
  const int r0 = blockIdx.x * 256;
  double acc[(NOUT + 255) / 256];
#pragma unroll
  for (int q = 0; q < (NOUT + 255) / 256; ++q) acc[q] = 0.0;
  for (int c = 0; c < 256; c += CH) {
    for (int t = threadIdx.x; t < CH * KIN; t += 256) {
      const int rr = t / KIN, k = t % KIN, row = r0 + c + rr;
      float v = 0.f;
      if (row < n && k < kin) v = bf16_round((k < c0) ? X0[(long)row * ld0 + k] : X1[(long)row * ld1 + (k - c0)]);
      xs[rr][k] = v;
    }
    for (int t = threadIdx.x; t < CH * kHidden; t += 256) {
      const int rr = t / kHidden, q = t % kHidden, row = r0 + c + rr;
      ds[rr][q] = (row < n) ? dz[(long)row * kHidden + q] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < (NOUT + 255) / 256; ++q) {
      const int o = threadIdx.x + q * 256;
      if (o < NOUT) {
        float s = 0.f;
        if (o < KIN * kHidden) {
          const int k = o / kHidden, j = o % kHidden;
          for (int rr = 0; rr < CH; ++rr) s += xs[rr][k] * ds[rr][j];
        } else {
          const int j = o - KIN * kHidden;
          for (int rr = 0; rr < CH; ++rr) s += ds[rr][j];
        }
        acc[q] += (double)s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < (NOUT + 255) / 256; ++q) {
    const int o = threadIdx.x + q * 256;
    if (o < NOUT) part[(long)o * cap + blockIdx.x] = acc[q];
  }
}

// out[o] (+)= sum_b part[o * cap + b] in block order (one lane per output)
__global__ void k_colsum(int nout, int nblk, const double* __restrict__ part, int cap, float* __restrict__ out,
                         int accumulate) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nout) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[(long)o * cap + b];
  out[o] = accumulate ? out[o] + (float)s : (float)s;
}
// per-row head gradient terms (N x kHeadGrads) -> partials per 256 rows
__global__ void __launch_bounds__(256) k_rowsum_partials(int n, int ncol, const float* __restrict__ prow,
                                                         double* __restrict__ part, int cap) {
  const int r0 = blockIdx.x * 256;
  for (int o = threadIdx.x; o < ncol; o += 256) {
    double s = 0.0;
    for (int r = r0; r < min(n, r0 + 256); ++r) s += (double)prow[(long)r * ncol + o];
    part[(long)o * cap + blockIdx.x] = s;
  }
}

// Adam (Kingma & Ba 2015) on the FP32 master weights
__global__ void k_adam(int np, float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Transpose positions: tpos[k] = position of (j, i) in row j for the entry
// k = (i, j); a missing mirror entry (pattern not symmetric) raises *bad.
__global__ void k_tpos(int n, const int* __restrict__ ro, const int* __restrict__ cols, const int* __restrict__ dcol,
                       const double* __restrict__ vals, int* __restrict__ tpos, int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int d = dcol[i];
  if (d >= ro[i + 1] || cols[d] != i) atomicOr(bad, 2);
  else if (!(vals[d] > 0)) atomicOr(bad, 4);
  for (int k = ro[i]; k < ro[i + 1]; ++k) {
    const int j = cols[k];
    int q = -1;
    for (int t = ro[j]; t < ro[j + 1]; ++t)
      if (cols[t] == i) {
        q = t;
        break;
      }
    tpos[k] = q;
    if (q < 0) atomicOr(bad, 1);
  }
}

// a factor value in its storage type: FP64, FP32, or bfloat16 (FP32 rounded to
// nearest even, then its upper half rounded to nearest even)
template <typename T>
__device__ __forceinline__ T factor_store(double g) {
  return (T)g;
}
template <>
__device__ __forceinline__ bf16_bits factor_store<bf16_bits>(double g) {
  return (bf16_bits)to_bf16((float)g);
}

// Row-bundled values of G (TRANS = false: rows = lower + diagonal of A,
// G_ij = dis_i Ghat_ij) or of G^T (TRANS: rows = diagonal + upper, (G^T)_ij =
// G_ji = dis_j Ghat_ji, read at the transpose position). Lane = row; the row's
// entries are walked together with the bundle's slots: slot j of an RB-CSR
// bundle is the row's j-th entry, an RB-DIA slot the entry at offset doff.
template <typename T, bool TRANS>
__global__ void k_fill_factor(int n, const int* __restrict__ ro, const int* __restrict__ cols,
                              const int* __restrict__ dcol, const int* __restrict__ tpos, const int* __restrict__ bo,
                              const int* __restrict__ doff, const float* __restrict__ ghat,
                              const float* __restrict__ dis, T* __restrict__ bv) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nb = (n + kBundle - 1) / kBundle;
  const int b = t / kBundle, lane = t % kBundle;
  if (b >= nb) return;
  const int row = t;
  const int s0 = bo[b], w = bo[b + 1] - s0;
  int k = 0, e = 0;
  if (row < n) {
    k = TRANS ? dcol[row] : ro[row];
    e = TRANS ? ro[row + 1] : dcol[row] + 1;
  }
  for (int j = 0; j < w; ++j) {
    T v = T(0);
    if (k < e && (!doff || cols[k] - row == doff[s0 + j])) {
      const double g = TRANS ? (double)dis[cols[k]] * (double)ghat[tpos[k]] : (double)dis[row] * (double)ghat[k];
      v = factor_store<T>(g);
      ++k;
    }
    bv[((long)s0 + j) * kBundle + lane] = v;
  }
}

} // namespace gnn
} // namespace alien_hip

// =====================================================================================
// Host engine: device buffers of one system and the launch sequences
// =====================================================================================
#include "../alien/hip/HIPContext.h"

namespace alien_hip {
namespace gnn {

namespace {
inline dim3 g1(long n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }
} // namespace

Engine::Engine(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, const double* vals) : m_ctx(ctx), m_n(n) {
  const int nnz = ro[n];
  h_ro.assign(ro, ro + n + 1);
  h_cols.assign(cols, cols + nnz);
  h_dcol.resize(n);
  host_dcol(n, ro, cols, h_dcol.data());
  hipStream_t s = m_ctx->stream;
  d_ro.upload(h_ro.data(), h_ro.size(), s);
  d_cols.upload(h_cols.data(), h_cols.size(), s);
  d_dcol.upload(h_dcol.data(), h_dcol.size(), s);
  d_vals.upload(vals, nnz, s);
  SystemView v;
  v.n = n;
  v.ro = h_ro.data();
  v.cols = h_cols.data();
  v.dcol = h_dcol.data();
  v.vals = vals;
  v.d_ro = d_ro.data();
  v.d_cols = d_cols.data();
  v.d_dcol = d_dcol.data();
  v.d_vals = d_vals.data();
  init(v);
}

Engine::Engine(alien_hip_ctx* ctx, const SystemView& v) : m_ctx(ctx), m_n(v.n) { init(v); }

// Checks of the system (stored positive diagonal, symmetric pattern), the
// transpose positions and the per-system buffers; host values are read here only.
void Engine::init(const SystemView& v) {
  const int n = v.n;
  m_nnz = v.ro[n];
  m_v = v;
  m_v.vals = nullptr;
  hipStream_t s = m_ctx->stream;
  d_tpos.resize(m_nnz);
  {
    // transpose positions, and the checks of the stored diagonal and the pattern's symmetry
    DeviceArray<int> bad(1);
    ALIEN_HIP_CHECK(hipMemsetAsync(bad.data(), 0, sizeof(int), s));
    hipLaunchKernelGGL(k_tpos, g1(n), dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, m_v.d_dcol, m_v.d_vals, d_tpos.data(),
                       bad.data());
    check_launch("gnn tpos");
    int hbad = 0;
    ALIEN_HIP_CHECK(hipMemcpyAsync(&hbad, bad.data(), sizeof(int), hipMemcpyDeviceToHost, s));
    ALIEN_HIP_CHECK(hipStreamSynchronize(s));
    if (hbad & 2) throw Error(ALIEN_HIP_EINVAL, "learned preconditioner: missing diagonal");
    if (hbad & 4) throw Error(ALIEN_HIP_EINVAL, "learned preconditioner: non-positive diagonal");
    if (hbad & 1) throw Error(ALIEN_HIP_EINVAL, "learned preconditioner: the pattern must be symmetric");
  }
  d_dis.resize(n);
  d_e.resize(m_nnz);
  d_f.resize((size_t)n * kFeat);
  d_ghat.resize(m_nnz);
}

// Activation buffers (n x 32 each): every layer's input, pre-activation and
// message for the backward pass, or two inputs and one message for inference.
void Engine::ensure_activations(bool all_layers) {
  const size_t sz = (size_t)m_n * kHidden;
  for (int l = 0; l < 4; ++l) {
    const bool need_h = all_layers || l < 2, need_z = all_layers, need_m = all_layers || l == 1;
    if (need_h && d_h[l].empty()) d_h[l].resize(sz);
    if (need_z && d_z[l].empty()) d_z[l].resize(sz);
    if (need_m && d_m[l].empty()) d_m[l].resize(sz);
  }
}

// scaled couplings and node features from A's values
void Engine::prep_features() {
  hipStream_t s = m_ctx->stream;
  const int n = m_n;
  hipLaunchKernelGGL(k_prep_dis, g1(n), dim3(256), 0, s, n, m_v.d_dcol, m_v.d_vals, d_dis.data());
  hipLaunchKernelGGL(k_prep_feat, g1(n), dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, m_v.d_vals, d_dis.data(),
                     d_e.data(), d_f.data());
}

void Engine::forward(const float* P) {
  ensure_activations(true);
  hipStream_t s = m_ctx->stream;
  const int n = m_n;
  prep_features();
  const dim3 gd((unsigned)(((n + 31) / 32 + 3) / 4)); // 32 rows per wavefront, 4 wavefronts per workgroup
  hipLaunchKernelGGL(k_dense<8>, gd, dim3(256), 0, s, n, d_f.data(), kFeat, kFeat, (const float*)nullptr, 0, kFeat,
                     P + kOffWin, P + kOffBin, (const float*)nullptr, d_z[0].data(), d_h[0].data());
  const int offs[3] = {kOffW1, kOffW2, kOffW3};
  for (int l = 1; l <= 3; ++l) {
    hipLaunchKernelGGL(k_aggregate, aggregate_grid(n), dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, d_e.data(),
                       d_h[l - 1].data(), d_m[l].data());
    hipLaunchKernelGGL(k_dense<64>, gd, dim3(256), 0, s, n, d_h[l - 1].data(), kHidden, kHidden, d_m[l].data(), kHidden,
                       2 * kHidden, P + offs[l - 1], P + offs[l - 1] + 2 * kHidden * kHidden, d_h[l - 1].data(),
                       d_z[l].data(), d_h[l].data());
  }
  hipLaunchKernelGGL(k_head, g1(n), dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, m_v.d_dcol, d_e.data(),
                     d_h[3].data(), P, d_ghat.data());
  check_launch("gnn forward");
}

// The same arithmetic as forward(), layer l reading buffer (l - 1) % 2 and
// writing l % 2, pre-activations not stored.
void Engine::infer(const float* P) {
  ensure_activations(false);
  hipStream_t s = m_ctx->stream;
  const int n = m_n;
  prep_features();
  const dim3 gd((unsigned)(((n + 31) / 32 + 3) / 4));
  hipLaunchKernelGGL(k_dense<8>, gd, dim3(256), 0, s, n, d_f.data(), kFeat, kFeat, (const float*)nullptr, 0, kFeat,
                     P + kOffWin, P + kOffBin, (const float*)nullptr, (float*)nullptr, d_h[0].data());
  const int offs[3] = {kOffW1, kOffW2, kOffW3};
  for (int l = 1; l <= 3; ++l) {
    const float* hin = d_h[(l - 1) & 1].data();
    float* hout = d_h[l & 1].data();
    hipLaunchKernelGGL(k_aggregate, aggregate_grid(n), dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, d_e.data(), hin,
                       d_m[1].data());
    hipLaunchKernelGGL(k_dense<64>, gd, dim3(256), 0, s, n, hin, kHidden, kHidden, d_m[1].data(), kHidden,
                       2 * kHidden, P + offs[l - 1], P + offs[l - 1] + 2 * kHidden * kHidden, hin, (float*)nullptr,
                       hout);
  }
  hipLaunchKernelGGL(k_head, g1(n), dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, m_v.d_dcol, d_e.data(),
                     d_h[1].data(), P, d_ghat.data());
  check_launch("gnn inference");
}

double Engine::loss_grad(const float* P, const float* V, int K, float beta, int cg_steps, float* grad,
                         double* parts) {
  if (K != 4) throw Error(ALIEN_HIP_EINVAL, "learned preconditioner: 4 probe vectors are supported");
  if (cg_steps < 0 || cg_steps > kMaxCgSteps) throw Error(ALIEN_HIP_EINVAL, "learned preconditioner: bad cg_steps");
  hipStream_t s = m_ctx->stream;
  const int n = m_n;
  forward(P);
  // ---- Kaporin term (cg_steps == 0)
  m_ctx->reserve_for(n);
  DeviceArray<float> ag(m_nnz), gg(m_nnz);
  DeviceArray<double> sums(4);
  const long nb = (n + 255) / 256;
  RedWork w = m_ctx->red();
  const bool kap = cg_steps == 0;
  if (kap) {
    hipLaunchKernelGGL(k_kcond, dim3((unsigned)nb), dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, m_v.d_dcol,
                       d_tpos.data(), d_e.data(), d_ghat.data(), ag.data(), w);
    launch_finalize(s, w, nb, 2, kFinStore, sums.data(), nullptr);
  }
  // ---- probe term
  constexpr int KP = 4;
  DeviceArray<float> W_((size_t)n * KP), U((size_t)n * KP), Y((size_t)n * KP), dY((size_t)n * KP), Q((size_t)n * KP);
  hipLaunchKernelGGL((k_ahat_mult<KP>), g1(n), dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, d_e.data(), V, W_.data());
  hipLaunchKernelGGL((k_ghat_mult<KP, true>), g1(n), dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, m_v.d_dcol,
                     d_tpos.data(), d_ghat.data(), W_.data(), U.data());
  hipLaunchKernelGGL((k_ghat_mult<KP, false>), g1(n), dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, m_v.d_dcol,
                     d_tpos.data(), d_ghat.data(), U.data(), Y.data());
  hipLaunchKernelGGL((k_probe_resid<KP>), dim3((unsigned)nb), dim3(256), 0, s, n, V, Y.data(), dY.data(),
                     (float)(1.0 / ((double)n * KP)), w);
  launch_finalize(s, w, nb, 1, kFinStore, sums.data() + 2, nullptr);
  hipLaunchKernelGGL((k_ghat_mult<KP, true>), g1(n), dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, m_v.d_dcol,
                     d_tpos.data(), d_ghat.data(), dY.data(), Q.data());
  hipLaunchKernelGGL((k_grad_ghat<KP>), g1(n), dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, m_v.d_dcol,
                     d_ghat.data(), ag.data(), sums.data(), kap ? 1.f : 0.f, beta, dY.data(), U.data(), W_.data(),
                     Q.data(), gg.data());
  // ---- CG term (cg_steps > 0): k PCG steps on Ahat with Ghat Ghat^T, reverse pass into gg
  double lcg = 0.0;
  if (!kap) lcg = cg_term(V, cg_steps, gg.data());
  // ---- head backward
  DeviceArray<float> dh[2], dz((size_t)n * kHidden), dmsg((size_t)n * kHidden), prow((size_t)n * kHeadGrads);
  dh[0].resize((size_t)n * kHidden);
  dh[1].resize((size_t)n * kHidden);
  hipLaunchKernelGGL(k_head_bwd, g1(n), dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, m_v.d_dcol, d_tpos.data(),
                     d_e.data(), d_ghat.data(), gg.data(), d_h[3].data(), P, dh[0].data(), prow.data());
  const int cap = (int)nb;
  DeviceArray<double> part((size_t)(2 * kHidden * kHidden + kHidden) * cap);
  hipLaunchKernelGGL(k_rowsum_partials, dim3((unsigned)nb), dim3(256), 0, s, n, kHeadGrads, prow.data(), part.data(), cap);
  hipLaunchKernelGGL(k_colsum, g1(kHeadGrads), dim3(256), 0, s, kHeadGrads, (int)nb, part.data(), cap, grad + kOffWd, 0);
  // ---- message-passing layers, last to first
  const int offs[3] = {kOffW1, kOffW2, kOffW3};
  int cur = 0;
  for (int l = 3; l >= 1; --l) {
    const float* Wl = P + offs[l - 1];
    hipLaunchKernelGGL(k_dense_bwd_mfma, dim3((unsigned)(((n + 31) / 32 + 3) / 4)), dim3(256), 0, s, n, dh[cur].data(),
                       d_z[l].data(), Wl, dz.data(), dh[cur ^ 1].data(), dmsg.data());
    hipLaunchKernelGGL(k_aggregate_add, g1((long)n * 32), dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols,
                       d_tpos.data(), d_e.data(), dmsg.data(), dh[cur ^ 1].data());
    hipLaunchKernelGGL(k_wgrad_partials<64>, dim3((unsigned)nb), dim3(256), 0, s, n, d_h[l - 1].data(), kHidden, kHidden,
                       d_m[l].data(), kHidden, 2 * kHidden, dz.data(), part.data(), cap);
    hipLaunchKernelGGL(k_colsum, g1(kLayerSize), dim3(256), 0, s, kLayerSize, (int)nb, part.data(), cap,
                       grad + offs[l - 1], 0);
    cur ^= 1;
  }
  // ---- input layer
  hipLaunchKernelGGL(k_dense_bwd<kFeat>, g1(n), dim3(256), 0, s, n, dh[cur].data(), d_z[0].data(), P + kOffWin, kFeat,
                     dz.data(), (float*)nullptr, (float*)nullptr);
  hipLaunchKernelGGL(k_wgrad_partials<kFeat>, dim3((unsigned)nb), dim3(256), 0, s, n, d_f.data(), kFeat, kFeat,
                     (const float*)nullptr, 0, kFeat, dz.data(), part.data(), cap);
  hipLaunchKernelGGL(k_colsum, g1(kFeat * kHidden + kHidden), dim3(256), 0, s, kFeat * kHidden + kHidden, (int)nb,
                     part.data(), cap, grad + kOffWin, 0);
  check_launch("gnn backward");
  double hs[4];
  ALIEN_HIP_CHECK(hipMemcpyAsync(hs, sums.data(), 3 * sizeof(double), hipMemcpyDeviceToHost, s));
  ALIEN_HIP_CHECK(hipStreamSynchronize(s));
  const double lk = kap ? std::log(hs[0] / n) - 2.0 * hs[1] / n : lcg;
  const double lp = hs[2] / ((double)n * KP);
  if (parts) {
    parts[0] = lk;
    parts[1] = lp;
  }
  return lk + beta * lp;
}

// Residual reduction after k PCG steps on Ahat with M = Ghat Ghat^T, x0 = 0, the
// 4 columns of B as right-hand sides: L = mean_c log(|r_k|^2 / |r_0|^2), rho =
// |Ghat^T r|^2. Forward keeps r_j, t_j = Ghat^T r_j, p_j, q_j = Ahat p_j; the
// reverse pass (the adjoint of each step, last to first) accumulates dL/dGhat
// on the lower pattern into gg. Step scalars live on the host (one dot4
// readback each), vectors in FP32, dots in FP64.
double Engine::cg_term(const float* B, int k, float* gg) {
  hipStream_t s = m_ctx->stream;
  const int n = m_n;
  const size_t nv = (size_t)n * 4;
  const dim3 gv = g1((long)nv), gr = g1(n);
  RedWork w = m_ctx->red();
  const long nb = (n + 255) / 256;
  DeviceArray<double> dsum(4);
  auto dot4 = [&](const float* X, const float* Y) {
    std::array<double, 4> h{};
    hipLaunchKernelGGL(k_dot4, dim3((unsigned)nb), dim3(256), 0, s, n, X, Y, w);
    launch_finalize(s, w, nb, 4, kFinStore, dsum.data(), nullptr);
    ALIEN_HIP_CHECK(hipMemcpyAsync(h.data(), dsum.data(), sizeof(double) * 4, hipMemcpyDeviceToHost, s));
    ALIEN_HIP_CHECK(hipStreamSynchronize(s));
    return h;
  };
  auto lin = [&](std::array<double, 4> a, const float* X, std::array<double, 4> b, const float* Z, float* Y) {
    Coef4 ca, cb;
    for (int c = 0; c < 4; ++c) {
      ca.c[c] = (float)a[c];
      cb.c[c] = (float)b[c];
    }
    hipLaunchKernelGGL(k_lin4, gv, dim3(256), 0, s, n, ca, X, cb, Z, Y);
  };
  auto amul = [&](const float* X, float* Y) {
    hipLaunchKernelGGL((k_ahat_mult<4>), gr, dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, d_e.data(), X, Y);
  };
  auto gmul = [&](const float* X, float* Y) { // Y = Ghat X
    hipLaunchKernelGGL((k_ghat_mult<4, false>), gr, dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, m_v.d_dcol,
                       d_tpos.data(), d_ghat.data(), X, Y);
  };
  auto gtmul = [&](const float* X, float* Y) { // Y = Ghat^T X
    hipLaunchKernelGGL((k_ghat_mult<4, true>), gr, dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, m_v.d_dcol,
                       d_tpos.data(), d_ghat.data(), X, Y);
  };
  auto acc = [&](const float* U, const float* V, const float* W, const float* X) {
    hipLaunchKernelGGL(k_edge_acc4, gr, dim3(256), 0, s, n, m_v.d_ro, m_v.d_cols, m_v.d_dcol, U, V, W, X,
                       gg);
  };
  typedef std::array<double, 4> S4;
  const S4 one{1, 1, 1, 1}, zero{0, 0, 0, 0};
  std::vector<DeviceArray<float>> R(k + 1), T(k), Pv(k), Q(k);
  for (auto& a : R) a.resize(nv);
  for (auto& a : T) a.resize(nv);
  for (auto& a : Pv) a.resize(nv);
  for (auto& a : Q) a.resize(nv);
  DeviceArray<float> z(nv), rbar(nv), pbar(nv), pbar2(nv), qbar(nv), tbar(nv), tmp(nv);
  std::vector<S4> rho(k), sig(k), alpha(k), beta(k);
  // forward
  ALIEN_HIP_CHECK(hipMemcpyAsync(R[0].data(), B, sizeof(float) * nv, hipMemcpyDeviceToDevice, s));
  gtmul(R[0].data(), T[0].data());
  gmul(T[0].data(), Pv[0].data());
  rho[0] = dot4(T[0].data(), T[0].data());
  const S4 rr0 = dot4(R[0].data(), R[0].data());
  for (int j = 0; j < k; ++j) {
    amul(Pv[j].data(), Q[j].data());
    sig[j] = dot4(Pv[j].data(), Q[j].data());
    S4 na;
    for (int c = 0; c < 4; ++c) {
      alpha[j][c] = rho[j][c] / sig[j][c];
      na[c] = -alpha[j][c];
    }
    lin(one, R[j].data(), na, Q[j].data(), R[j + 1].data());
    if (j < k - 1) {
      gtmul(R[j + 1].data(), T[j + 1].data());
      gmul(T[j + 1].data(), z.data());
      rho[j + 1] = dot4(T[j + 1].data(), T[j + 1].data());
      for (int c = 0; c < 4; ++c) beta[j][c] = rho[j + 1][c] / rho[j][c];
      lin(one, z.data(), beta[j], Pv[j].data(), Pv[j + 1].data());
    }
  }
  const S4 rrk = dot4(R[k].data(), R[k].data());
  double L = 0.0;
  S4 c0;
  for (int c = 0; c < 4; ++c) {
    L += std::log(rrk[c] / rr0[c]) / 4.0;
    c0[c] = 2.0 / (4.0 * rrk[c]);
  }
  if (!std::isfinite(L)) return L;
  // reverse
  lin(c0, R[k].data(), zero, nullptr, rbar.data());
  S4 rhobar_next = zero;
  for (int j = k - 1; j >= 0; --j) {
    S4 rhobar_j = zero;
    if (j < k - 1) {
      // p_{j+1} = z_{j+1} + beta_j p_j ; zbar = pbar_{j+1} (held in pbar)
      const S4 betabar = dot4(pbar.data(), Pv[j].data());
      S4 rhobar_jp1, two_rho;
      for (int c = 0; c < 4; ++c) {
        rhobar_jp1[c] = rhobar_next[c] + betabar[c] / rho[j][c];
        rhobar_j[c] -= betabar[c] * rho[j + 1][c] / (rho[j][c] * rho[j][c]);
        two_rho[c] = 2.0 * rhobar_jp1[c];
      }
      // z = Ghat t: dGhat += zbar t^T ; tbar = 2 rhobar t + Ghat^T zbar
      gtmul(pbar.data(), tmp.data());
      lin(two_rho, T[j + 1].data(), one, tmp.data(), tbar.data());
      // t = Ghat^T r: dGhat += r tbar^T ; rbar += Ghat tbar
      acc(pbar.data(), T[j + 1].data(), R[j + 1].data(), tbar.data());
      gmul(tbar.data(), tmp.data());
      lin(one, rbar.data(), one, tmp.data(), rbar.data());
      lin(beta[j], pbar.data(), zero, nullptr, pbar2.data()); // pbar_j = beta_j pbar_{j+1}
    } else {
      ALIEN_HIP_CHECK(hipMemsetAsync(pbar2.data(), 0, sizeof(float) * nv, s)); // p_k is not formed
    }
    // r_{j+1} = r_j - alpha q_j
    const S4 qr = dot4(Q[j].data(), rbar.data());
    S4 na, sigbar;
    for (int c = 0; c < 4; ++c) {
      const double alphabar = -qr[c];
      rhobar_j[c] += alphabar / sig[j][c];
      sigbar[c] = -alphabar * rho[j][c] / (sig[j][c] * sig[j][c]);
      na[c] = -alpha[j][c];
    }
    lin(na, rbar.data(), sigbar, Pv[j].data(), qbar.data()); // qbar = -alpha rbar + sigbar p
    lin(one, pbar2.data(), sigbar, Q[j].data(), pbar2.data()); // pbar += sigbar q
    amul(qbar.data(), tmp.data());
    lin(one, pbar2.data(), one, tmp.data(), pbar.data()); // pbar_j = ... + Ahat qbar
    rhobar_next = rhobar_j;
  }
  // p_0 = Ghat t_0, rho_0 = t_0.t_0, t_0 = Ghat^T r_0
  S4 two_rho;
  for (int c = 0; c < 4; ++c) two_rho[c] = 2.0 * rhobar_next[c];
  gtmul(pbar.data(), tmp.data());
  lin(two_rho, T[0].data(), one, tmp.data(), tbar.data());
  acc(pbar.data(), T[0].data(), R[0].data(), tbar.data());
  check_launch("gnn cg loss");
  ALIEN_HIP_CHECK(hipStreamSynchronize(s));
  return L;
}

// Row-bundled layouts of G (rows = lower + diagonal of A) and of G^T (rows =
// diagonal + upper of A; the pattern is symmetric): RB-DIA when it stores at
// most kRbdiaMaxGrowth times the RB-CSR slots, RB-CSR otherwise. When A itself
// is RB-DIA, the factors' offsets and lane masks are A's with offset <= 0 (G)
// and >= 0 (G^T): a split of A's structure, no pass over the entries.
void Engine::build_factor_layout() {
  if (!m_gbo.empty()) return;
  const int n = m_n;
  const int nb = (n + kBundle - 1) / kBundle;
  hipStream_t s = m_ctx->stream;
  const bool split = m_v.rb_bo && m_v.d_rb_doff && m_v.d_rb_dmask;
  std::vector<int> adoff;
  std::vector<unsigned long long> admask;
  if (split) {
    const size_t aslots = (size_t)m_v.rb_bo[nb];
    adoff.resize(aslots);
    admask.resize(aslots);
    ALIEN_HIP_CHECK(hipMemcpyAsync(adoff.data(), m_v.d_rb_doff, aslots * sizeof(int), hipMemcpyDeviceToHost, s));
    ALIEN_HIP_CHECK(hipMemcpyAsync(admask.data(), m_v.d_rb_dmask, aslots * sizeof(unsigned long long),
                                   hipMemcpyDeviceToHost, s));
    ALIEN_HIP_CHECK(hipStreamSynchronize(s));
  }
  // RB-CSR widths of both factors in one pass: row r of G has dcol - ro + 1 entries, of G^T ro' - dcol
  std::vector<int> cbos[2] = {std::vector<int>(nb + 1, 0), std::vector<int>(nb + 1, 0)};
  long totals[2] = {0, 0};
  {
    long* acc = totals;
    for (int b = 0; b < nb; ++b) {
      cbos[0][b] = (int)acc[0];
      cbos[1][b] = (int)acc[1];
      int w0 = 0, w1 = 0;
      for (int r = b * kBundle; r < std::min(n, (b + 1) * kBundle); ++r) {
        w0 = std::max(w0, m_v.dcol[r] - m_v.ro[r] + 1);
        w1 = std::max(w1, m_v.ro[r + 1] - m_v.dcol[r]);
      }
      acc[0] += w0;
      acc[1] += w1;
    }
    cbos[0][nb] = (int)acc[0];
    cbos[1][nb] = (int)acc[1];
  }
  std::vector<int> kb, ke; // entry ranges per row, built only for a host structure pass
  for (int t = 0; t < 2; ++t) {
    const std::vector<int>& cbo = cbos[t];
    const long acc = totals[t];
    auto ranges = [&]() {
      kb.resize(n);
      ke.resize(n);
      for (int r = 0; r < n; ++r) {
        kb[r] = t == 0 ? m_v.ro[r] : m_v.dcol[r];
        ke[r] = t == 0 ? m_v.dcol[r] + 1 : m_v.ro[r + 1];
      }
    };
    if (acc * kBundle > 0x7fffffffL) throw Error(ALIEN_HIP_EINVAL, "learned preconditioner: factor too large");
    std::vector<int> dbo, doff;
    std::vector<uint64_t> dmask;
    long dslots;
    if (split) {
      dbo.assign(nb + 1, 0);
      doff.reserve(adoff.size());
      dmask.reserve(adoff.size());
      for (int b = 0; b < nb; ++b) {
        dbo[b] = (int)doff.size();
        for (int j = m_v.rb_bo[b]; j < m_v.rb_bo[b + 1]; ++j) {
          if (t == 0 ? adoff[j] > 0 : adoff[j] < 0) continue;
          doff.push_back(adoff[j]);
          dmask.push_back(admask[j]);
        }
      }
      dbo[nb] = (int)doff.size();
      dslots = dbo[nb];
    } else {
      ranges();
      dslots = rbdia_structure(n, m_v.ro, m_v.cols, kb.data(), ke.data(), kRbdiaMaxWidth, dbo, doff, dmask);
    }
    const bool dia = dslots >= 0 && dslots <= kRbdiaMaxGrowth * cbo[nb];
    std::vector<int>& bo = t == 0 ? m_gbo : m_gtbo;
    bo = dia ? dbo : cbo;
    const size_t slots = (size_t)bo[nb] * kBundle;
    Factor& F = t == 0 ? m_g : m_gt;
    F.dia = dia;
    F.uni = 0;
    if (dia && nb > 0) {
      // uniform when every bundle holds the same offsets (the split of a uniform A)
      const int k = bo[1] - bo[0];
      bool same = k >= 1 && k <= kUniMaxSlots;
      for (int b = 0; same && b < nb; ++b) {
        same = bo[b] == b * k && bo[b + 1] == (b + 1) * k;
        for (int j = 0; same && j < k; ++j) same = doff[(size_t)b * k + j] == doff[j];
      }
      if (same) {
        F.uni = k;
        F.nrows = n;
        for (int j = 0; j < k; ++j) F.uoff[j] = doff[j];
      }
    }
    F.bo.upload(bo.data(), bo.size(), s);
    if (dia) {
      F.doff.upload(doff.data(), doff.size(), s);
      F.dmask.upload(reinterpret_cast<const unsigned long long*>(dmask.data()), dmask.size(), s);
    } else {
      ranges();
      std::vector<int> bc(slots, -1);
      for (int r = 0; r < n; ++r) {
        const int b = r / kBundle, lane = r % kBundle;
        for (int k = kb[r]; k < ke[r]; ++k) bc[((size_t)bo[b] + (k - kb[r])) * kBundle + lane] = m_v.cols[k];
      }
      F.bc.upload(bc.data(), slots, s);
    }
    F.slots = slots;
  }
  ALIEN_HIP_CHECK(hipStreamSynchronize(s));
}

template <typename T>
void Engine::fill_factor(T* g_vals, T* gt_vals) {
  build_factor_layout();
  hipStream_t s = m_ctx->stream;
  const int nrows64 = (m_n + kBundle - 1) / kBundle * kBundle;
  hipLaunchKernelGGL((k_fill_factor<T, false>), g1(nrows64), dim3(256), 0, s, m_n, m_v.d_ro, m_v.d_cols,
                     m_v.d_dcol, d_tpos.data(), m_g.bo.data(), m_g.dia ? m_g.doff.data() : nullptr,
                     d_ghat.data(), d_dis.data(), g_vals);
  hipLaunchKernelGGL((k_fill_factor<T, true>), g1(nrows64), dim3(256), 0, s, m_n, m_v.d_ro, m_v.d_cols,
                     m_v.d_dcol, d_tpos.data(), m_gt.bo.data(), m_gt.dia ? m_gt.doff.data() : nullptr,
                     d_ghat.data(), d_dis.data(), gt_vals);
  check_launch("gnn fill factor");
}
void Engine::release_network_buffers() {
  ALIEN_HIP_CHECK(hipStreamSynchronize(m_ctx->stream));
  for (int l = 0; l < 4; ++l) {
    d_h[l].reset();
    d_z[l].reset();
    d_m[l].reset();
  }
  d_e.reset();
  d_f.reset();
  d_ghat.reset();
  d_dis.reset();
  d_tpos.reset();
}

template void Engine::fill_factor<float>(float*, float*);
template void Engine::fill_factor<bf16_bits>(bf16_bits*, bf16_bits*);
template void Engine::fill_factor<double>(double*, double*);

void launch_adam(hipStream_t s, int np, float* P, const float* g, float* m, float* v, float lr, int step) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(k_adam, g1(np), dim3(256), 0, s, np, P, g, m, v, lr, b1, b2, 1e-8f, c1, c2);
}

} // namespace gnn
} // namespace alien_hip
