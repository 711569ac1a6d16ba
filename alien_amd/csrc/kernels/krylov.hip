// Fused Krylov kernels: one Chebyshev-preconditioned CG iteration in six
// launches, with every step coefficient kept on the device.
//
// The reference (CG.h:111-145 with ChebyshevPreconditioner::_algo2,
// ChebyshevPreconditioner.h:392-434) issues, per iteration, 4 SpMVs, 3 dot
// products read back by the host and 19 BLAS-1 launches. Here:
//
//   cheb_step  x3  row kernel: t = (A y)_i, r = x_i - t, z = invD_i r,
//                  p_i = b p_i + z, y'_i = y_i + a p_i  (+ rho1 = r.z on the last)
//   cg_p           p = z + (rho1/rho) p        (p = z on the first iteration)
//   cg_q           q = A p, pq = p.q
//   cg_xr          x += alpha p, r -= alpha q, rr = r.r, stop test, ++iter
//
// Each element-wise step rounds exactly like the reference's separate BLAS-1
// calls (same operands, same order, no FMA), so apart from the summation
// order of the three dot products the iteration is the reference's.
//
// The stop test of Iteration::stop (Iteration.h:83-90) runs in cg_xr's last
// workgroup, which also counts the iteration; once `done` is set every later
// kernel returns at entry, so a hipGraph holding several iterations can be
// replayed without overshooting and the iteration count stays exact.
//
// Rows are processed in RB-CSR bundles: one wavefront per 64-row bundle,
// grid-stride over bundles with a grid that depends on n only (deterministic
// reductions).
#include "../alien_hip_internal.h"
#include "device_common.h"
#include "krylov_state.h"
#include "launch.h"

namespace alien_hip {

template <typename T>
__device__ __forceinline__ T rb_sum(long base, int w, const int* __restrict__ bc, const T* __restrict__ bv,
                                    const T* __restrict__ x) {
  T acc = 0;
  int j = 0;
  constexpr int U = 7;
  for (; j + U <= w; j += U) {
    int c[U];
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      c[u] = __builtin_nontemporal_load(bc + base + (long)(j + u) * kBundle);
      v[u] = __builtin_nontemporal_load(bv + base + (long)(j + u) * kBundle);
    }
    T xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = (c[u] >= 0) ? x[c[u]] : T(0);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (c[u] >= 0) acc = acc + v[u] * xv[u];
  }
  for (; j < w; ++j) {
    const int c = __builtin_nontemporal_load(bc + base + (long)j * kBundle);
    const T v = __builtin_nontemporal_load(bv + base + (long)j * kBundle);
    if (c >= 0) acc = acc + v * x[c];
  }
  return acc;
}

#define BUNDLE_LOOP(n)                                                                        \
  const int lane = threadIdx.x & (kWave - 1);                                                 \
  const int nb = ((n) + kBundle - 1) / kBundle;                                               \
  const int wave0 = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * kBlock + threadIdx.x) / kWave)); \
  const int nwaves = gridDim.x * (kBlock / kWave);                                            \
  for (int b = wave0; b < nb; b += nwaves)

// r = b - A x  (copy(b,r); mult(A,x,p); axpy(-1,p,r)); out = r.r
template <typename T>
__global__ void __launch_bounds__(kBlock) k_residual(int n, RBMat<T> A, const T* __restrict__ x,
                                                      const T* __restrict__ rhs, T* __restrict__ r,
                                                      RedWork w, double* __restrict__ out) {
  __shared__ double lds4[4];
  double s = 0.0;
  BUNDLE_LOOP(n) {
    const int s0 = A.bo[b];
    const int wdt = A.bo[b + 1] - s0;
    const T t = rb_sum<T>((long)s0 * kBundle + lane, wdt, A.bc, A.bv, x);
    const int row = b * kBundle + lane;
    if (row < n) {
      const T v = rhs[row] + T(-1) * t;
      r[row] = v;
      s += (double)v * (double)v;
    }
  }
  double v[1] = {block_sum(s, lds4)};
  grid_reduce_finish<1>(v, w, [&](int, double tot) { *out = tot; });
}

// One Chebyshev step (ChebyshevPreconditioner::_algo2). FIRST: y_in = x,
// p = z, y' = x + a z. Otherwise p = b p + z, y' = y + a p. LAST also
// accumulates rho1 = x . y' into the CG state.
template <typename T, bool FIRST, bool LAST>
__global__ void __launch_bounds__(kBlock) k_cheb_step(int n, RBMat<T> A, const T* __restrict__ xin,
                                                       const T* __restrict__ yin, T* __restrict__ yout,
                                                       T* __restrict__ p, const T* __restrict__ invd, T a, T bcoef,
                                                       CGState* st, RedWork w) {
  if (st && st->done) return;
  __shared__ double lds4[4];
  double s = 0.0;
  const T* src = FIRST ? xin : yin;
  BUNDLE_LOOP(n) {
    const int s0 = A.bo[b];
    const int wdt = A.bo[b + 1] - s0;
    const T t = rb_sum<T>((long)s0 * kBundle + lane, wdt, A.bc, A.bv, src);
    const int row = b * kBundle + lane;
    if (row < n) {
      const T xi = xin[row];
      const T rr = xi + T(-1) * t;
      const T z = invd[row] * rr;
      T pn;
      T yo;
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
      if (LAST) s += (double)xi * (double)yo;
    }
  }
  if (LAST) {
    double v[1] = {block_sum(s, lds4)};
    grid_reduce_finish<1>(v, w, [&](int, double tot) { st->rho1 = tot; });
  }
}

// Jacobi apply for the Diag preconditioner: z = invD .* r, rho1 = r . z
template <typename T>
__global__ void __launch_bounds__(kBlock) k_diag_apply(int n, const T* __restrict__ r, const T* __restrict__ invd,
                                                        T* __restrict__ z, CGState* st, RedWork w) {
  if (st->done) return;
  __shared__ double lds4[4];
  double s = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const T v = invd[i] * r[i];
    z[i] = v;
    s += (double)r[i] * (double)v;
  }
  double v[1] = {block_sum(s, lds4)};
  grid_reduce_finish<1>(v, w, [&](int, double tot) { st->rho1 = tot; });
}

// p = z + (rho1/rho) p  (axpy(alpha,p,z); copy(z,p)), or p = z on the first step
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_p(int n, const T* __restrict__ z, T* __restrict__ p, CGState* st,
                                                  int first) {
  if (st->done) return;
  if (first) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) p[i] = z[i];
    return;
  }
  const T beta = (T)(st->rho1 / st->rho);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) p[i] = z[i] + beta * p[i];
}

// q = A p ; pq = p . q ; a zero pq raises the breakdown flag
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_q(int n, RBMat<T> A, const T* __restrict__ p, T* __restrict__ q,
                                                  CGState* st, RedWork w) {
  if (st->done) return;
  __shared__ double lds4[4];
  double s = 0.0;
  BUNDLE_LOOP(n) {
    const int s0 = A.bo[b];
    const int wdt = A.bo[b + 1] - s0;
    const T t = rb_sum<T>((long)s0 * kBundle + lane, wdt, A.bc, A.bv, p);
    const int row = b * kBundle + lane;
    if (row < n) {
      q[row] = t;
      s += (double)p[row] * (double)t;
    }
  }
  double v[1] = {block_sum(s, lds4)};
  grid_reduce_finish<1>(v, w, [&](int, double tot) {
    st->pq = tot;
    if (tot == 0) {
      st->breakdown = 1;
      st->done = 1;
    }
  });
}

// alpha = rho1 / pq ; x += alpha p ; r += (-alpha) q ; rr = r.r ; stop test
template <typename T, typename TX>
__global__ void __launch_bounds__(kBlock) k_cg_xr(int n, const T* __restrict__ p, const T* __restrict__ q,
                                                   TX* __restrict__ x, T* __restrict__ r, CGState* st, RedWork w) {
  if (st->done) return;
  __shared__ double lds4[4];
  const double alpha = st->rho1 / st->pq;
  const T a = (T)alpha;
  double s = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] = x[i] + (TX)a * (TX)p[i];
    const T v = r[i] + (-a) * q[i];
    r[i] = v;
    s += (double)v * (double)v;
  }
  double v[1] = {block_sum(s, lds4)};
  grid_reduce_finish<1>(v, w, [&](int, double tot) {
    st->rho = st->rho1;
    st->it += 1;
    if (st->it >= st->max_it) {
      st->done = 1; // Iteration::stop returns before the dot; status unchanged
    } else {
      st->rr = tot;
      st->status = (tot < st->crit) ? 1 : 0;
      st->done = st->status;
    }
  });
}

// ---- mixed precision helpers ------------------------------------------------------------
// r32 = (float)(r64 * inv_scale) with inv_scale = 1/sqrt(rr) read from the device
__global__ void __launch_bounds__(kBlock) k_demote_scaled(int n, const double* __restrict__ r,
                                                           const double* __restrict__ rr, float* __restrict__ r32) {
  const double inv = 1.0 / sqrt(*rr);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) r32[i] = (float)(r[i] * inv);
}
// x64 += scale * (double)d32 with scale = sqrt(rr)
__global__ void __launch_bounds__(kBlock) k_promote_add(int n, const float* __restrict__ d,
                                                         const double* __restrict__ rr, double* __restrict__ x) {
  const double sc = sqrt(*rr);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) x[i] = x[i] + sc * (double)d[i];
}
// narrow copy (FP64 -> FP32)
__global__ void __launch_bounds__(kBlock) k_demote(int n, const double* __restrict__ a, float* __restrict__ b) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) b[i] = (float)a[i];
}

// ---- launchers -------------------------------------------------------------------------------
template <typename T>
void launch_residual(hipStream_t s, RedWork w, int n, RBMat<T> A, const T* x, const T* b, T* r, double* out) {
  hipLaunchKernelGGL(k_residual<T>, dim3(bundle_grid(n)), dim3(kBlock), 0, s, n, A, x, b, r, w, out);
}

template <typename T>
void launch_cheb_step(hipStream_t s, RedWork w, int n, RBMat<T> A, const T* xin, const T* yin, T* yout, T* p,
                      const T* invd, double a, double b, bool first, bool last, CGState* st) {
  dim3 g(bundle_grid(n)), blk(kBlock);
  if (first && last)
    hipLaunchKernelGGL((k_cheb_step<T, true, true>), g, blk, 0, s, n, A, xin, yin, yout, p, invd, (T)a, (T)b, st, w);
  else if (first)
    hipLaunchKernelGGL((k_cheb_step<T, true, false>), g, blk, 0, s, n, A, xin, yin, yout, p, invd, (T)a, (T)b, st, w);
  else if (last)
    hipLaunchKernelGGL((k_cheb_step<T, false, true>), g, blk, 0, s, n, A, xin, yin, yout, p, invd, (T)a, (T)b, st, w);
  else
    hipLaunchKernelGGL((k_cheb_step<T, false, false>), g, blk, 0, s, n, A, xin, yin, yout, p, invd, (T)a, (T)b, st, w);
}

template <typename T>
void launch_diag_apply(hipStream_t s, RedWork w, int n, const T* r, const T* invd, T* z, CGState* st) {
  hipLaunchKernelGGL(k_diag_apply<T>, dim3(red_grid(n)), dim3(kBlock), 0, s, n, r, invd, z, st, w);
}

template <typename T>
void launch_cg_p(hipStream_t s, int n, const T* z, T* p, CGState* st, int first) {
  hipLaunchKernelGGL(k_cg_p<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, z, p, st, first);
}

template <typename T>
void launch_cg_q(hipStream_t s, RedWork w, int n, RBMat<T> A, const T* p, T* q, CGState* st) {
  hipLaunchKernelGGL(k_cg_q<T>, dim3(bundle_grid(n)), dim3(kBlock), 0, s, n, A, p, q, st, w);
}

template <typename T, typename TX>
void launch_cg_xr(hipStream_t s, RedWork w, int n, const T* p, const T* q, TX* x, T* r, CGState* st) {
  hipLaunchKernelGGL((k_cg_xr<T, TX>), dim3(red_grid(n)), dim3(kBlock), 0, s, n, p, q, x, r, st, w);
}

void launch_demote_scaled(hipStream_t s, int n, const double* r, const double* rr, float* r32) {
  hipLaunchKernelGGL(k_demote_scaled, dim3(grid_for(n)), dim3(kBlock), 0, s, n, r, rr, r32);
}
void launch_promote_add(hipStream_t s, int n, const float* d, const double* rr, double* x) {
  hipLaunchKernelGGL(k_promote_add, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d, rr, x);
}
void launch_demote(hipStream_t s, int n, const double* a, float* b) {
  hipLaunchKernelGGL(k_demote, dim3(grid_for(n)), dim3(kBlock), 0, s, n, a, b);
}

#define ALIEN_HIP_INST_KRYLOV(T)                                                                       \
  template void launch_residual<T>(hipStream_t, RedWork, int, RBMat<T>, const T*, const T*, T*, double*); \
  template void launch_cheb_step<T>(hipStream_t, RedWork, int, RBMat<T>, const T*, const T*, T*, T*,     \
                                    const T*, double, double, bool, bool, CGState*);                   \
  template void launch_diag_apply<T>(hipStream_t, RedWork, int, const T*, const T*, T*, CGState*);     \
  template void launch_cg_p<T>(hipStream_t, int, const T*, T*, CGState*, int);                         \
  template void launch_cg_q<T>(hipStream_t, RedWork, int, RBMat<T>, const T*, T*, CGState*);
ALIEN_HIP_INST_KRYLOV(double)
ALIEN_HIP_INST_KRYLOV(float)
template void launch_cg_xr<double, double>(hipStream_t, RedWork, int, const double*, const double*, double*,
                                           double*, CGState*);
template void launch_cg_xr<float, float>(hipStream_t, RedWork, int, const float*, const float*, float*, float*,
                                         CGState*);

} // namespace alien_hip
