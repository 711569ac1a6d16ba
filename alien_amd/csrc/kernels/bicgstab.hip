// Fused BiCGStab (BiCGStab.h:151-300, right preconditioning). One step of the
// reference's while loop is
//
//   bicg_p   p = (p + (-omega) v) beta + r   (axpy, scal, axpy; p = r in the
//            first block), beta = (rho1/rho)(alpha/omega) from the device state;
//            Jacobi: phat = invD p
//   bicg_v   v = A phat, v.r0           -> finalize: alpha = rho1 / (v.r0)
//   bicg_s   s = r + (-alpha) v, s.s    -> finalize: Iteration::stop(s);
//            Jacobi: shat = invD s
//   bicg_t   t = A shat, t.s, t.t       -> finalize: omega = (t.s)/(t.t)
//   bicg_xr  x = (x + omega shat) + alpha phat, r = s + (-omega) t, r.r, r.r0
//            -> finalize: rho = rho1, ++iter, Iteration::stop(r), next rho1
//
// with every element-wise step rounded as the reference's separate BLAS-1
// calls (same operands and order, no FMA). Other preconditioners run between
// the launches (learned M: launch_factor_apply_flag). Every kernel returns at
// entry once the state is done, so a hipGraph of several steps can be
// replayed; a stop on s leaves `half` set and launch_bicg_half applies the
// reference's x += alpha phat of that exit.
#include "../alien_hip_internal.h"
#include "device_common.h"
#include "krylov_state.h"
#include "launch.h"
#include "rb_rows.h"

namespace alien_hip {

#define EPT_INDEX(k) (blockIdx.x * (long)(kBlock * kEPT) + (k) * kBlock + threadIdx.x)

template <typename T, bool DIAG>
__global__ void __launch_bounds__(kBlock) k_bicg_p(int n, const T* __restrict__ r, const T* __restrict__ v,
                                                    T* __restrict__ p, const T* __restrict__ invd,
                                                    T* __restrict__ phat, const BiCGState* st) {
  if (st->done) return;
  const bool first = st->first != 0;
  const T beta = first ? T(0) : (T)((st->rho1 / st->rho) * (st->alpha / st->omega));
  const T mom = first ? T(0) : (T)(-st->omega);
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i >= n) continue;
    T pn;
    if (first) {
      pn = r[i];
    } else {
      T t = p[i] + mom * v[i];
      t = beta * t;
      pn = t + T(1) * r[i];
    }
    p[i] = pn;
    if (DIAG) phat[i] = invd[i] * pn;
  }
}

// v = A phat ; v.r0
template <typename T, int LAY = kRowBundled>
__global__ void __launch_bounds__(kRowBlock) k_bicg_v(int n, RBMat<T> A, const T* __restrict__ phat,
                                                       const T* __restrict__ r0, T* __restrict__ v,
                                                       const BiCGState* st, RedWork w) {
  if (st->done) return;
  __shared__ double lds4[kRowBlock / kWave];
  double s = 0.0;
  const BundleWalk bw = bundle_walk(n);
  for (int b = bw.first; b < bw.end; b += bw.stride) {
    const int row = b * kBundle + (threadIdx.x & (kWave - 1));
    const T r0v = r0[min(row, n - 1)];
    const T t = rb_sum<T, LAY>(A, b, [&](int c) { return phat[c]; });
    if (row < n) {
      v[row] = t;
      s += (double)t * (double)r0v;
    }
  }
  double vv[1] = {block_sum<kRowBlock>(s, lds4)};
  store_partials<1>(vv, w);
}

// s = r + (-alpha) v ; s.s ; Jacobi: shat = invD s
template <typename T, bool DIAG>
__global__ void __launch_bounds__(kBlock) k_bicg_s(int n, const T* __restrict__ r, const T* __restrict__ v,
                                                    T* __restrict__ s, const T* __restrict__ invd,
                                                    T* __restrict__ shat, const BiCGState* st, RedWork w) {
  if (st->done) return;
  __shared__ double lds4[4];
  const T ma = (T)(-st->alpha);
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i >= n) continue;
    const T sv = r[i] + ma * v[i];
    s[i] = sv;
    acc += (double)sv * (double)sv;
    if (DIAG) shat[i] = invd[i] * sv;
  }
  double vv[1] = {block_sum(acc, lds4)};
  store_partials<1>(vv, w);
}

// t = A shat ; t.s, t.t
template <typename T, int LAY = kRowBundled>
__global__ void __launch_bounds__(kRowBlock) k_bicg_t(int n, RBMat<T> A, const T* __restrict__ shat,
                                                       const T* __restrict__ sv, T* __restrict__ t,
                                                       const BiCGState* st, RedWork w) {
  if (st->done) return;
  __shared__ double lds4[kRowBlock / kWave];
  double ts = 0.0, tt = 0.0;
  const BundleWalk bw = bundle_walk(n);
  for (int b = bw.first; b < bw.end; b += bw.stride) {
    const int row = b * kBundle + (threadIdx.x & (kWave - 1));
    const T srow = sv[min(row, n - 1)];
    const T tv = rb_sum<T, LAY>(A, b, [&](int c) { return shat[c]; });
    if (row < n) {
      t[row] = tv;
      ts += (double)tv * (double)srow;
      tt += (double)tv * (double)tv;
    }
  }
  double vv[2];
  vv[0] = block_sum<kRowBlock>(ts, lds4);
  __syncthreads();
  vv[1] = block_sum<kRowBlock>(tt, lds4);
  store_partials<2>(vv, w);
}

// x = (x + omega shat) + alpha phat ; r = s + (-omega) t ; r.r, r.r0
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bicg_xr(int n, const T* __restrict__ shat, const T* __restrict__ phat,
                                                     const T* __restrict__ s, const T* __restrict__ t,
                                                     const T* __restrict__ r0, T* __restrict__ x,
                                                     T* __restrict__ r, const BiCGState* st, RedWork w) {
  if (st->done) return;
  __shared__ double lds4[4];
  const T om = (T)st->omega, al = (T)st->alpha, mom = (T)(-st->omega);
  double rr = 0.0, rr0 = 0.0;
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i >= n) continue;
    T xv = x[i] + om * shat[i];
    xv = xv + al * phat[i];
    x[i] = xv;
    const T rv = s[i] + mom * t[i];
    r[i] = rv;
    rr += (double)rv * (double)rv;
    rr0 += (double)rv * (double)r0[i];
  }
  double vv[2];
  vv[0] = block_sum(rr, lds4);
  __syncthreads();
  vv[1] = block_sum(rr0, lds4);
  store_partials<2>(vv, w);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_bicg_half(int n, const T* __restrict__ phat, T* __restrict__ x,
                                                       const BiCGState* st) {
  if (!st->half) return;
  const T al = (T)st->alpha;
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i < n) x[i] = x[i] + al * phat[i];
  }
}

// y = A x on the row-bundled layout, skipped once *done is set
template <typename T, int LAY = kRowBundled>
__global__ void __launch_bounds__(kRowBlock) k_rb_spmv_flag(int n, RBMat<T> A, const T* __restrict__ x,
                                                             T* __restrict__ y, const int* done) {
  if (*done) return;
  const BundleWalk bw = bundle_walk(n);
  for (int b = bw.first; b < bw.end; b += bw.stride) {
    const T t = rb_factor_sum<T, LAY>(A, b, [&](int c) { return x[c]; });
    const int row = b * kBundle + (threadIdx.x & (kWave - 1));
    if (row < n) y[row] = t;
  }
}

static inline dim3 egrid(long n) { return dim3((unsigned)ept_blocks(n)); }

template <typename T>
void launch_bicg_p(hipStream_t s, int n, const T* r, const T* v, T* p, const T* invd, T* phat, BiCGState* st) {
  if (invd)
    hipLaunchKernelGGL((k_bicg_p<T, true>), egrid(n), dim3(kBlock), 0, s, n, r, v, p, invd, phat,
                       (const BiCGState*)st);
  else
    hipLaunchKernelGGL((k_bicg_p<T, false>), egrid(n), dim3(kBlock), 0, s, n, r, v, p, invd, phat,
                       (const BiCGState*)st);
}
template <typename T>
void launch_bicg_v(hipStream_t s, RedWork w, int n, RBMat<T> A, const T* phat, const T* r0, T* v, BiCGState* st) {
  const dim3 g((unsigned)rb_blocks(n, A));
  by_layout(A, [&](auto L) {
    hipLaunchKernelGGL((k_bicg_v<T, decltype(L)::value>), g, dim3(kRowBlock), 0, s, n, A, phat, r0, v,
                       (const BiCGState*)st, w);
  });
  launch_finalize_bicg(s, w, rb_blocks(n, A), 1, kBFinAlpha, st);
}
template <typename T>
void launch_bicg_s(hipStream_t s, RedWork w, int n, const T* r, const T* v, T* sv, const T* invd, T* shat,
                   BiCGState* st) {
  if (invd)
    hipLaunchKernelGGL((k_bicg_s<T, true>), egrid(n), dim3(kBlock), 0, s, n, r, v, sv, invd, shat,
                       (const BiCGState*)st, w);
  else
    hipLaunchKernelGGL((k_bicg_s<T, false>), egrid(n), dim3(kBlock), 0, s, n, r, v, sv, invd, shat,
                       (const BiCGState*)st, w);
  launch_finalize_bicg(s, w, ept_blocks(n), 1, kBFinS, st);
}
template <typename T>
void launch_bicg_t(hipStream_t s, RedWork w, int n, RBMat<T> A, const T* shat, const T* sv, T* t, BiCGState* st) {
  const dim3 g((unsigned)rb_blocks(n, A));
  by_layout(A, [&](auto L) {
    hipLaunchKernelGGL((k_bicg_t<T, decltype(L)::value>), g, dim3(kRowBlock), 0, s, n, A, shat, sv, t,
                       (const BiCGState*)st, w);
  });
  launch_finalize_bicg(s, w, rb_blocks(n, A), 2, kBFinOmega, st);
}
template <typename T>
void launch_bicg_xr(hipStream_t s, RedWork w, int n, const T* shat, const T* phat, const T* sv, const T* t,
                    const T* r0, T* x, T* r, BiCGState* st) {
  hipLaunchKernelGGL(k_bicg_xr<T>, egrid(n), dim3(kBlock), 0, s, n, shat, phat, sv, t, r0, x, r,
                     (const BiCGState*)st, w);
  launch_finalize_bicg(s, w, ept_blocks(n), 2, kBFinXR, st);
}
template <typename T>
void launch_bicg_half(hipStream_t s, int n, const T* phat, T* x, const BiCGState* st) {
  hipLaunchKernelGGL(k_bicg_half<T>, egrid(n), dim3(kBlock), 0, s, n, phat, x, st);
}
template <typename T>
void launch_factor_apply_flag(hipStream_t s, int n, RBMat<T> G, RBMat<T> GT, const T* x, T* tmp, T* y,
                              const int* done) {
  by_layout(GT, [&](auto L) {
    hipLaunchKernelGGL((k_rb_spmv_flag<T, decltype(L)::value>), dim3((unsigned)rb_blocks(n, GT)), dim3(kRowBlock), 0,
                       s, n, GT, x, tmp, done);
  });
  by_layout(G, [&](auto L) {
    hipLaunchKernelGGL((k_rb_spmv_flag<T, decltype(L)::value>), dim3((unsigned)rb_blocks(n, G)), dim3(kRowBlock), 0,
                       s, n, G, (const T*)tmp, y, done);
  });
}

#define ALIEN_HIP_INST_BICG(T)                                                                              \
  template void launch_bicg_p<T>(hipStream_t, int, const T*, const T*, T*, const T*, T*, BiCGState*);      \
  template void launch_bicg_v<T>(hipStream_t, RedWork, int, RBMat<T>, const T*, const T*, T*, BiCGState*); \
  template void launch_bicg_s<T>(hipStream_t, RedWork, int, const T*, const T*, T*, const T*, T*,         \
                                 BiCGState*);                                                             \
  template void launch_bicg_t<T>(hipStream_t, RedWork, int, RBMat<T>, const T*, const T*, T*, BiCGState*); \
  template void launch_bicg_xr<T>(hipStream_t, RedWork, int, const T*, const T*, const T*, const T*,      \
                                  const T*, T*, T*, BiCGState*);                                          \
  template void launch_bicg_half<T>(hipStream_t, int, const T*, T*, const BiCGState*);                    \
  template void launch_factor_apply_flag<T>(hipStream_t, int, RBMat<T>, RBMat<T>, const T*, T*, T*,       \
                                            const int*);
ALIEN_HIP_INST_BICG(double)
ALIEN_HIP_INST_BICG(float)

} // namespace alien_hip
