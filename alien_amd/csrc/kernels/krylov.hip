// Fused Krylov kernels: a preconditioned CG iteration in a few launches, with
// every step coefficient kept on the device.
//
// The reference (CG.h:111-145 with ChebyshevPreconditioner::_algo2,
// ChebyshevPreconditioner.h:392-434, or DiagPreconditioner.h:61-67) issues
// per iteration one SpMV, three dot products read back by the host and a
// string of BLAS-1 launches, plus the preconditioner's own. Here:
//
//   cheb_step x m  row kernel: t = (A y)_i, r = x_i - t, z = invD_i r,
//                  p_i = b p_i + z, y'_i = y_i + a p_i  (+ rho1 = r.z on the last)
//   cg_p           p = z + (rho1/rho) p (p = z on the first step), streaming
//   cg_q           q = A p, pq = p.q
//   cg_xr          x += alpha p, r -= alpha q, rr = r.r, the stop test, ++iter;
//                  with the Diag preconditioner also z = invD r and rho1 = r.z
//                  for the next iteration
//
// Each element-wise step rounds exactly like the reference's separate BLAS-1
// calls (same operands, same order, no FMA), so apart from the summation
// order of the dot products the iteration is the reference's.
//
// Each reduction ends with launch_finalize (reduce.hip), which also applies
// the stop test of Iteration::stop (Iteration.h:83-90) after cg_xr and counts
// the iteration; once `done` is set every later kernel returns at entry, so a
// hipGraph holding several iterations can be replayed without overshooting
// and the iteration count stays exact.
//
// Row kernels process RB-CSR / RB-DIA bundles, one wavefront per 64-row bundle;
// streaming kernels handle kEPT elements per lane. Launch shapes depend on n
// only, so the reductions are deterministic.
#include <hip/hip_ext.h>

#include "../alien_hip_internal.h"
#include "device_common.h"
#include "krylov_state.h"
#include "launch.h"
#include "rb_rows.h"

#include <type_traits>

namespace alien_hip {

#define EPT_INDEX(k) (blockIdx.x * (long)(kBlock * kEPT) + (k) * kBlock + threadIdx.x)

// Row kernels walk their bundles with bundle_walk (device_common.h).

// r = b - A x  (copy(b,r); mult(A,x,p); axpy(-1,p,r)); out = r.r
template <typename T, int LAY = kRowBundled>
__global__ void __launch_bounds__(kRowBlock) k_residual(int n, RBMat<T> A, const T* __restrict__ x,
                                                      const T* __restrict__ rhs, T* __restrict__ r,
                                                      RedWork w) {
  __shared__ double lds4[kRowBlock / kWave];
  double s = 0.0;
  const BundleWalk bw = bundle_walk(n);
  for (int b = bw.first; b < bw.end; b += bw.stride) {
    const int row = b * kBundle + (threadIdx.x & (kWave - 1));
    const T br = rhs[min(row, n - 1)];
    const T t = rb_sum<T, LAY>(A, b, [&](int c) { return x[c]; });
    if (row < n) {
      const T v = br + T(-1) * t;
      r[row] = v;
      s += (double)v * (double)v;
    }
  }
  double v[1] = {block_sum<kRowBlock>(s, lds4)};
  store_partials<1>(v, w);
}

// One Chebyshev step (ChebyshevPreconditioner::_algo2). FIRST: y_in = x,
// p = z, y' = x + a z. Otherwise p = b p + z, y' = y + a p. LAST also
// accumulates rho1 = x . y' into the CG state.
template <typename T, bool FIRST, bool LAST, int LAY = kRowBundled>
__global__ void __launch_bounds__(kRowBlock) k_cheb_step(int n, RBMat<T> A, const T* __restrict__ xin,
                                                       const T* __restrict__ yin, T* __restrict__ yout,
                                                       T* __restrict__ p, const T* __restrict__ invd, T a, T bcoef,
                                                       const int* __restrict__ done, int rho, RedWork w) {
  if (*done) return;
  __shared__ double lds4[kRowBlock / kWave];
  double s = 0.0;
  const T* src = FIRST ? xin : yin;
  const BundleWalk bw = bundle_walk(n);
  for (int b = bw.first; b < bw.end; b += bw.stride) {
    const T t = rb_sum<T, LAY>(A, b, [&](int c) { return src[c]; });
    const int row = b * kBundle + (threadIdx.x & (kWave - 1));
    if (row < n) {
      const T xi = xin[row];
      const T rr = xi + T(-1) * t;
      const T z = invd[row] * rr;
      T pn, yo;
      if (FIRST) {
        pn = z;
        yo = xi + a * pn;
      } else {
        pn = bcoef * p[row];
        pn = pn + T(1) * z;
        yo = yin[row] + a * pn;
      }
      p[row] = pn;
      yout[row] = yo;
      if (LAST && rho) s += (double)xi * (double)yo;
    }
  }
  if (LAST && rho) {
    double v[1] = {block_sum<kRowBlock>(s, lds4)};
    store_partials<1>(v, w);
  }
}

// One step of the Neumann polynomial apply (NeumannPolyPreconditioner.h:85-175,
// the algebra calls of its solve): yout = (yin + (-f) (A yin)) + x, the last
// step scaled by f; with rdot the last step also stores the partials of
// rdot . yout. Returns at entry once *done.
template <typename T, bool LAST, int LAY = kRowBundled>
__global__ void __launch_bounds__(kRowBlock) k_neumann_step(int n, RBMat<T> A, const T* __restrict__ x,
                                                          const T* __restrict__ yin, T* __restrict__ yout, T f,
                                                          const int* __restrict__ done, const T* __restrict__ rdot,
                                                          RedWork w) {
  if (done && *done) return;
  __shared__ double lds4[kRowBlock / kWave];
  double s = 0.0;
  const BundleWalk bw = bundle_walk(n);
  for (int b = bw.first; b < bw.end; b += bw.stride) {
    const T t = rb_sum<T, LAY>(A, b, [&](int c) { return yin[c]; });
    const int row = b * kBundle + (threadIdx.x & (kWave - 1));
    if (row < n) {
      T v = yin[row] + (-f) * t;
      v = v + T(1) * x[row];
      if (LAST) v = f * v;
      yout[row] = v;
      if (LAST && rdot) s += (double)rdot[row] * (double)v;
    }
  }
  if (LAST && rdot) {
    double v[1] = {block_sum<kRowBlock>(s, lds4)};
    store_partials<1>(v, w);
  }
}

// Jacobi apply for the Diag preconditioner: z = invD .* r, rho1 = r . z
template <typename T>
__global__ void __launch_bounds__(kBlock) k_diag_apply(int n, const T* __restrict__ r, const T* __restrict__ invd,
                                                        T* __restrict__ z, CGState* st, RedWork w) {
  if (st->done) return;
  __shared__ double lds4[4];
  T rv[kEPT], dv[kEPT];
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    rv[k] = (i < n) ? r[i] : T(0);
    dv[k] = (i < n) ? invd[i] : T(0);
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    const T v = dv[k] * rv[k];
    if (i < n) {
      z[i] = v;
      s += (double)rv[k] * (double)v;
    }
  }
  double v[1] = {block_sum(s, lds4)};
  store_partials<1>(v, w);
}

// p = z + (rho1/rho) p  (axpy(alpha,p,z); copy(z,p)), or p = z on the first step.
// Streaming, in place: each element is read and written by the same lane.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_p(int n, const T* __restrict__ z, T* __restrict__ p, CGState* st,
                                                  int first) {
  if (st->done) return;
  const T beta = first ? T(0) : (T)(st->rho1 / st->rho);
  T zv[kEPT], pv[kEPT];
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    zv[k] = (i < n) ? z[i] : T(0);
    pv[k] = (i < n && !first) ? p[i] : T(0);
  }
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i < n) p[i] = first ? zv[k] : zv[k] + beta * pv[k];
  }
}

// q = A p ; pq = p . q (a zero pq raises the breakdown flag in the finalize)
template <typename T, int LAY = kRowBundled>
__global__ void __launch_bounds__(kRowBlock) k_cg_q(int n, RBMat<T> A, const T* __restrict__ p, T* __restrict__ q,
                                                     CGState* st, RedWork w) {
  if (st->done) return;
  __shared__ double lds4[kRowBlock / kWave];
  double s = 0.0;
  bool paired = false;
  if constexpr (LAY == kRowUniform) {
    if (A.uni <= kUniPairMaxSlots) {
      // bundles in pairs (uni_sum2)
      paired = true;
      const int nb = (n + kWave - 1) / kWave;
      const BundleWalk bw = bundle_walk_units((nb + 1) / 2, kUniQChunk);
      for (int u = bw.first; u < bw.end; u += bw.stride) {
        const int b0 = 2 * u, b1 = min(2 * u + 1, nb - 1);
        const int row0 = b0 * kBundle + (threadIdx.x & (kWave - 1)), row1 = row0 + kBundle;
        const T p0 = p[min(row0, n - 1)], p1 = p[min(row1, n - 1)];
        T t0, t1;
        uni_sum2<T>(A, b0, b1, [&](int c) { return p[c]; }, t0, t1);
        if (row0 < n) {
          q[row0] = t0;
          s += (double)p0 * (double)t0;
        }
        if (row1 < n) {
          q[row1] = t1;
          s += (double)p1 * (double)t1;
        }
      }
    }
  }
  if (!paired) {
    const BundleWalk bw = bundle_walk(n);
    for (int b = bw.first; b < bw.end; b += bw.stride) {
      const int row = b * kBundle + (threadIdx.x & (kWave - 1));
      const T prow = p[min(row, n - 1)]; // unconditional (clamped): issued with the row's loads
      const T t = rb_sum<T, LAY>(A, b, [&](int c) { return p[c]; });
      if (row < n) {
        q[row] = t;
        s += (double)prow * (double)t;
      }
    }
  }
  double v[1] = {block_sum<kRowBlock>(s, lds4)};
  store_partials<1>(v, w);
}

// y = A x with rho1 = w . y into the CG state (second SpMV of a factored
// preconditioner z = G (G^T r): x = G^T r, w = r)
template <typename T, int LAY = kRowBundled>
__global__ void __launch_bounds__(kRowBlock) k_rb_spmv_dot(int n, RBMat<T> A, const T* __restrict__ x,
                                                            const T* __restrict__ wv, T* __restrict__ y, CGState* st,
                                                            RedWork w) {
  if (st->done) return;
  __shared__ double lds4[kRowBlock / kWave];
  double s = 0.0;
  const BundleWalk bw = bundle_walk(n);
  for (int b = bw.first; b < bw.end; b += bw.stride) {
    const T t = rb_factor_sum<T, LAY>(A, b, [&](int c) { return x[c]; });
    const int row = b * kBundle + (threadIdx.x & (kWave - 1));
    if (row < n) {
      y[row] = t;
      s += (double)wv[row] * (double)t;
    }
  }
  double v[1] = {block_sum<kRowBlock>(s, lds4)};
  store_partials<1>(v, w);
}

// y = A x (RB-CSR), returning early once the CG state is done
template <typename T, int LAY = kRowBundled>
__global__ void __launch_bounds__(kRowBlock) k_rb_spmv_guarded(int n, RBMat<T> A, const T* __restrict__ x,
                                                                T* __restrict__ y, const CGState* st) {
  if (st->done) return;
  const BundleWalk bw = bundle_walk(n);
  for (int b = bw.first; b < bw.end; b += bw.stride) {
    const T t = rb_factor_sum<T, LAY>(A, b, [&](int c) { return x[c]; });
    const int row = b * kBundle + (threadIdx.x & (kWave - 1));
    if (row < n) y[row] = t;
  }
}

// Factored preconditioner M = G G^T, first half: t = G^T r and rho1 = t.t
// (= r.M r, the r.z of CG.h:124 with z = G t, without forming z), plus r.r:
// the stop test of the previous update is taken here (its finalize runs
// kFinXRDiag), so the x/r update itself needs no reduction.
template <typename T, typename V, int LAY = kRowBundled>
__global__ void __launch_bounds__(kRowBlock) k_factor_t(int n, RBMat<T, V> GT, const T* __restrict__ r,
                                                         T* __restrict__ t, CGState* st, RedWork w) {
  if (st->done) return;
  __shared__ double lds4[kRowBlock / kWave];
  double srr = 0.0, stt = 0.0;
  if constexpr (LAY == kRowUniform) {
    // bundles in pairs (uni_sum2)
    const int nb = (n + kWave - 1) / kWave;
    const BundleWalk bw = bundle_walk_units((nb + 1) / 2);
    for (int u = bw.first; u < bw.end; u += bw.stride) {
      const int b0 = 2 * u, b1 = min(2 * u + 1, nb - 1);
      const int row0 = b0 * kBundle + (threadIdx.x & (kWave - 1)), row1 = row0 + kBundle;
      const T r0 = r[min(row0, n - 1)], r1 = r[min(row1, n - 1)];
      T v0, v1;
      uni_sum2<T>(GT, b0, b1, [&](int c) { return r[c]; }, v0, v1);
      if (row0 < n) {
        t[row0] = v0;
        srr += (double)r0 * (double)r0;
        stt += (double)v0 * (double)v0;
      }
      if (row1 < n) {
        t[row1] = v1;
        srr += (double)r1 * (double)r1;
        stt += (double)v1 * (double)v1;
      }
    }
  } else {
    const BundleWalk bw = bundle_walk(n);
    for (int b = bw.first; b < bw.end; b += bw.stride) {
      const int row = b * kBundle + (threadIdx.x & (kWave - 1));
      const T rv = r[min(row, n - 1)];
      const T v = rb_factor_sum<T, LAY>(GT, b, [&](int c) { return r[c]; });
      if (row < n) {
        t[row] = v;
        srr += (double)rv * (double)rv;
        stt += (double)v * (double)v;
      }
    }
  }
  double v[2];
  v[0] = block_sum<kRowBlock>(srr, lds4);
  __syncthreads();
  v[1] = block_sum<kRowBlock>(stt, lds4);
  store_partials<2>(v, w);
}

// x += alpha p ; r += (-alpha) q without a reduction (the learned-M iteration
// takes its stop test in the next k_factor_t)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_update(int n, const T* __restrict__ p, const T* __restrict__ q,
                                                       T* __restrict__ x, T* __restrict__ r, const CGState* st) {
  if (st->done) return;
  const T a = (T)(st->rho1 / st->pq);
  T pv[kEPT], qv[kEPT], xv[kEPT], rv[kEPT];
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    const bool ok = i < n;
    pv[k] = ok ? p[i] : T(0);
    qv[k] = ok ? q[i] : T(0);
    xv[k] = ok ? x[i] : T(0);
    rv[k] = ok ? r[i] : T(0);
  }
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i < n) {
      x[i] = xv[k] + a * pv[k];
      r[i] = rv[k] + (-a) * qv[k];
    }
  }
}

// Second half fused with the direction update: z_i = (G t)_i is formed in
// registers and p = z + (rho1/rho) p (p = z on the first step), as k_cg_p.
template <typename T, typename V, int LAY = kRowBundled>
__global__ void __launch_bounds__(kRowBlock) k_factor_p(int n, RBMat<T, V> G, const T* __restrict__ t,
                                                         T* __restrict__ p, const CGState* st, int first) {
  if (st->done) return;
  const T beta = first ? T(0) : (T)(st->rho1 / st->rho);
  if constexpr (LAY == kRowUniform) {
    // bundles in pairs (uni_sum2)
    const int nb = (n + kWave - 1) / kWave;
    const BundleWalk bw = bundle_walk_units((nb + 1) / 2);
    for (int u = bw.first; u < bw.end; u += bw.stride) {
      const int b0 = 2 * u, b1 = min(2 * u + 1, nb - 1);
      const int row0 = b0 * kBundle + (threadIdx.x & (kWave - 1)), row1 = row0 + kBundle;
      const T p0 = p[min(row0, n - 1)], p1 = p[min(row1, n - 1)];
      T z0, z1;
      uni_sum2<T>(G, b0, b1, [&](int c) { return t[c]; }, z0, z1);
      if (row0 < n) p[row0] = first ? z0 : z0 + beta * p0;
      if (row1 < n) p[row1] = first ? z1 : z1 + beta * p1;
    }
  } else {
    const BundleWalk bw = bundle_walk(n);
    for (int b = bw.first; b < bw.end; b += bw.stride) {
      const int row = b * kBundle + (threadIdx.x & (kWave - 1));
      const T pold = p[min(row, n - 1)];
      const T z = rb_factor_sum<T, LAY>(G, b, [&](int c) { return t[c]; });
      if (row < n) p[row] = first ? z : z + beta * pold;
    }
  }
}

// alpha = rho1 / pq ; x += alpha p ; r += (-alpha) q ; rr = r.r ; stop test.
// DIAG: also z = invD r and rho1' = r.z for the next iteration.
template <typename T, bool DIAG>
__global__ void __launch_bounds__(kBlock) k_cg_xr(int n, const T* __restrict__ p, const T* __restrict__ q,
                                                   T* __restrict__ x, T* __restrict__ r, const T* __restrict__ invd,
                                                   T* __restrict__ z, CGState* st, RedWork w) {
  if (st->done) return;
  __shared__ double lds4[4];
  const double alpha = st->rho1 / st->pq;
  const T a = (T)alpha;
  T pv[kEPT], qv[kEPT], xv[kEPT], rv[kEPT], dv[kEPT];
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    const bool ok = i < n;
    pv[k] = ok ? p[i] : T(0);
    qv[k] = ok ? q[i] : T(0);
    xv[k] = ok ? x[i] : T(0);
    rv[k] = ok ? r[i] : T(0);
    if (DIAG) dv[k] = ok ? invd[i] : T(0);
  }
  double srr = 0.0, srz = 0.0;
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i < n) {
      x[i] = xv[k] + a * pv[k];
      const T rn = rv[k] + (-a) * qv[k];
      r[i] = rn;
      srr += (double)rn * (double)rn;
      if (DIAG) {
        const T zn = dv[k] * rn;
        z[i] = zn;
        srz += (double)rn * (double)zn;
      }
    }
  }
  double v[2];
  v[0] = block_sum(srr, lds4);
  __syncthreads();
  v[1] = DIAG ? block_sum(srz, lds4) : 0.0;
  store_partials<2>(v, w);
}

// ---- mixed precision helpers ------------------------------------------------------------
// x64 += (double) x32 ; x32 = 0  (fold the FP32 increment into the FP64 solution)
__global__ void __launch_bounds__(kBlock) k_fold_increment(int n, float* __restrict__ x32, double* __restrict__ x) {
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i < n) {
      x[i] = x[i] + (double)x32[i];
      x32[i] = 0.0f;
    }
  }
}
// narrow copy (FP64 -> FP32, round to nearest)
__global__ void __launch_bounds__(kBlock) k_demote(int n, const double* __restrict__ a, float* __restrict__ b) {
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i < n) b[i] = (float)a[i];
  }
}

// ---- launchers -------------------------------------------------------------------------------
static inline dim3 egrid(long n) { return dim3((unsigned)ept_blocks(n)); }

template <typename T>
void launch_residual(hipStream_t s, RedWork w, int n, RBMat<T> A, const T* x, const T* b, T* r, double* out) {
  const dim3 g((unsigned)rb_blocks(n, A));
  by_layout(A, [&](auto L) {
    hipLaunchKernelGGL((k_residual<T, decltype(L)::value>), g, dim3(kRowBlock), 0, s, n, A, x, b, r, w);
  });
  launch_finalize(s, w, rb_blocks(n, A), 1, kFinStore, out, nullptr);
}

template <typename T>
void launch_cheb_step(hipStream_t s, RedWork w, int n, RBMat<T> A, const T* xin, const T* yin, T* yout, T* p,
                      const T* invd, double a, double b, bool first, bool last, CGState* st,
                      const int* done_flag) {
  dim3 g((unsigned)rb_blocks(n, A)), blk(kRowBlock);
  const int* done = st ? &st->done : done_flag;
  const int rho = st ? 1 : 0;
  auto by_ell = [&](auto first_c, auto last_c) {
    by_layout(A, [&](auto lay_c) {
      hipLaunchKernelGGL((k_cheb_step<T, decltype(first_c)::value, decltype(last_c)::value, decltype(lay_c)::value>),
                         g, blk, 0, s, n, A, xin, yin, yout, p, invd, (T)a, (T)b, done, rho, w);
    });
  };
  if (first && last) by_ell(std::true_type{}, std::true_type{});
  else if (first) by_ell(std::true_type{}, std::false_type{});
  else if (last) by_ell(std::false_type{}, std::true_type{});
  else by_ell(std::false_type{}, std::false_type{});
  if (last && st) launch_finalize(s, w, rb_blocks(n, A), 1, kFinRho1, nullptr, st);
}

// rho1 = r . z into the CG state (a preconditioner applied by separate launches)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_rz(int n, const T* __restrict__ r, const T* __restrict__ z,
                                                CGState* st, RedWork w) {
  if (st->done) return;
  __shared__ double lds4[4];
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i < n) s += (double)r[i] * (double)z[i];
  }
  double v[1] = {block_sum(s, lds4)};
  store_partials<1>(v, w);
}
template <typename T>
void launch_rz(hipStream_t s, RedWork w, int n, const T* r, const T* z, CGState* st) {
  hipLaunchKernelGGL(k_rz<T>, egrid(n), dim3(kBlock), 0, s, n, r, z, st, w);
  launch_finalize(s, w, ept_blocks(n), 1, kFinRho1, nullptr, st);
}

template <typename T>
long launch_neumann_apply(hipStream_t s, RedWork w, int n, RBMat<T> A, int order, double f, const T* x, T* tmp,
                          T* z, const int* done, const T* rdot) {
  const dim3 g((unsigned)rb_blocks(n, A)), blk(kRowBlock);
  const T* yin = x;
  for (int i = 0; i < order; ++i) {
    T* out = ((order - 1 - i) % 2 == 0) ? z : tmp;
    const bool last = i == order - 1;
    const T* rd = last ? rdot : (const T*)nullptr;
    by_layout(A, [&](auto L) {
      if (last)
        hipLaunchKernelGGL((k_neumann_step<T, true, decltype(L)::value>), g, blk, 0, s, n, A, x, yin, out, (T)f, done,
                           rd, w);
      else
        hipLaunchKernelGGL((k_neumann_step<T, false, decltype(L)::value>), g, blk, 0, s, n, A, x, yin, out, (T)f, done,
                           rd, w);
    });
    yin = out;
  }
  return (rdot && order > 0) ? rb_blocks(n, A) : 0;
}

template <typename T>
void launch_diag_apply(hipStream_t s, RedWork w, int n, const T* r, const T* invd, T* z, CGState* st) {
  hipLaunchKernelGGL(k_diag_apply<T>, egrid(n), dim3(kBlock), 0, s, n, r, invd, z, st, w);
  launch_finalize(s, w, ept_blocks(n), 1, kFinRho1, nullptr, st);
}

template <typename T>
void launch_cg_p(hipStream_t s, int n, const T* z, T* p, CGState* st, int first) {
  hipLaunchKernelGGL(k_cg_p<T>, egrid(n), dim3(kBlock), 0, s, n, z, p, st, first);
}

template <typename T>
void launch_cg_q(hipStream_t s, RedWork w, int n, RBMat<T> A, const T* p, T* q, CGState* st, hipEvent_t t0,
                 hipEvent_t t1) {
  const dim3 g((unsigned)rb_blocks(n, A));
  by_layout(A, [&](auto L) {
    if (t0 || t1)
      hipExtLaunchKernelGGL((k_cg_q<T, decltype(L)::value>), g, dim3(kRowBlock), 0, s, t0, t1, 0, n, A, p, q, st, w);
    else
      hipLaunchKernelGGL((k_cg_q<T, decltype(L)::value>), g, dim3(kRowBlock), 0, s, n, A, p, q, st, w);
  });
  launch_finalize(s, w, rb_blocks(n, A), 1, kFinPQ, nullptr, st);
}

// the q = A p, p.q kernel of the fused iteration on its own: y = A x, *out = x . y (idle: a CG
// state that is never done)
template <typename T>
void launch_spmv_dot_rb(hipStream_t s, RedWork w, int n, RBMat<T> A, const T* x, T* y, CGState* idle, double* out) {
  if (n <= 0) {
    launch_finalize(s, w, 0, 1, kFinStore, out, nullptr);
    return;
  }
  const dim3 g((unsigned)rb_blocks(n, A));
  by_layout(A, [&](auto L) {
    hipLaunchKernelGGL((k_cg_q<T, decltype(L)::value>), g, dim3(kRowBlock), 0, s, n, A, x, y, idle, w);
  });
  launch_finalize(s, w, rb_blocks(n, A), 1, kFinStore, out, nullptr);
}

template <typename T>
void launch_cg_xr(hipStream_t s, RedWork w, int n, const T* p, const T* q, T* x, T* r, const T* invd_or_null,
                  T* z_or_null, CGState* st) {
  if (invd_or_null)
    hipLaunchKernelGGL((k_cg_xr<T, true>), egrid(n), dim3(kBlock), 0, s, n, p, q, x, r, invd_or_null, z_or_null,
                       st, w);
  else
    hipLaunchKernelGGL((k_cg_xr<T, false>), egrid(n), dim3(kBlock), 0, s, n, p, q, x, r, invd_or_null, z_or_null,
                       st, w);
  launch_finalize(s, w, ept_blocks(n), 2, invd_or_null ? kFinXRDiag : kFinXR, nullptr, st);
}

template <typename T>
void launch_factor_apply(hipStream_t s, RedWork w, int n, RBMat<T> G, RBMat<T> GT, const T* r, T* tmp, T* z,
                         CGState* st) {
  by_layout(GT, [&](auto L) {
    hipLaunchKernelGGL((k_rb_spmv_guarded<T, decltype(L)::value>), dim3((unsigned)rb_blocks(n, GT)), dim3(kRowBlock), 0, s, n, GT, r, tmp,
                       (const CGState*)st);
  });
  by_layout(G, [&](auto L) {
    hipLaunchKernelGGL((k_rb_spmv_dot<T, decltype(L)::value>), dim3((unsigned)rb_blocks(n, G)), dim3(kRowBlock), 0, s, n, G, tmp, r, z, st,
                       w);
  });
  launch_finalize(s, w, rb_blocks(n, G), 1, kFinRho1, nullptr, st);
}

template <typename T, typename V>
void launch_factor_p(hipStream_t s, RedWork w, int n, RBMat<T, V> G, RBMat<T, V> GT, const T* r, T* t, T* p,
                     CGState* st, int first, int rho1_only) {
  by_layout(GT, [&](auto L) {
    hipLaunchKernelGGL((k_factor_t<T, V, decltype(L)::value>), dim3((unsigned)rb_blocks(n, GT)), dim3(kRowBlock), 0, s, n, GT, r, t, st, w);
  });
  // first step (or a resumed one): rho1 only; later: rho = rho1, ++iter and the stop test of the
  // previous update, rho1 = t.t
  launch_finalize(s, w, rb_blocks(n, GT), 2, (first || rho1_only) ? kFinRho1B : kFinXRDiag, nullptr, st);
  by_layout(G, [&](auto L) {
    hipLaunchKernelGGL((k_factor_p<T, V, decltype(L)::value>), dim3((unsigned)rb_blocks(n, G)), dim3(kRowBlock), 0, s, n, G, t, p,
                       (const CGState*)st, first);
  });
}
template <typename T>
void launch_cg_update(hipStream_t s, int n, const T* p, const T* q, T* x, T* r, CGState* st) {
  hipLaunchKernelGGL(k_cg_update<T>, egrid(n), dim3(kBlock), 0, s, n, p, q, x, r, (const CGState*)st);
}

void launch_fold_increment(hipStream_t s, int n, float* x32, double* x) {
  hipLaunchKernelGGL(k_fold_increment, egrid(n), dim3(kBlock), 0, s, n, x32, x);
}
void launch_demote(hipStream_t s, int n, const double* a, float* b) {
  hipLaunchKernelGGL(k_demote, egrid(n), dim3(kBlock), 0, s, n, a, b);
}

#define ALIEN_HIP_INST_KRYLOV(T)                                                                           \
  template void launch_residual<T>(hipStream_t, RedWork, int, RBMat<T>, const T*, const T*, T*, double*);  \
  template void launch_cheb_step<T>(hipStream_t, RedWork, int, RBMat<T>, const T*, const T*, T*, T*,      \
                                    const T*, double, double, bool, bool, CGState*, const int*);        \
  template void launch_diag_apply<T>(hipStream_t, RedWork, int, const T*, const T*, T*, CGState*);      \
  template void launch_rz<T>(hipStream_t, RedWork, int, const T*, const T*, CGState*);                  \
  template long launch_neumann_apply<T>(hipStream_t, RedWork, int, RBMat<T>, int, double, const T*, T*, T*,  \
                                        const int*, const T*);                                            \
  template void launch_cg_p<T>(hipStream_t, int, const T*, T*, CGState*, int);                       \
  template void launch_factor_apply<T>(hipStream_t, RedWork, int, RBMat<T>, RBMat<T>, const T*, T*, T*,  \
                                       CGState*);                                                        \
  template void launch_factor_p<T, T>(hipStream_t, RedWork, int, RBMat<T>, RBMat<T>, const T*, T*, T*,  \
                                      CGState*, int, int);                                               \
  template void launch_cg_q<T>(hipStream_t, RedWork, int, RBMat<T>, const T*, T*, CGState*, hipEvent_t,      \
                               hipEvent_t);                                                             \
  template void launch_spmv_dot_rb<T>(hipStream_t, RedWork, int, RBMat<T>, const T*, T*, CGState*, double*); \
  template void launch_cg_update<T>(hipStream_t, int, const T*, const T*, T*, T*, CGState*);            \
  template void launch_cg_xr<T>(hipStream_t, RedWork, int, const T*, const T*, T*, T*, const T*, T*,     \
                                CGState*);
ALIEN_HIP_INST_KRYLOV(double)
ALIEN_HIP_INST_KRYLOV(float)
// FP32 arithmetic on bfloat16 factor values (learned M in the mixed solver)
template void launch_factor_p<float, bf16_bits>(hipStream_t, RedWork, int, RBMat<float, bf16_bits>,
                                                RBMat<float, bf16_bits>, const float*, float*, float*, CGState*,
                                                int, int);

} // namespace alien_hip
