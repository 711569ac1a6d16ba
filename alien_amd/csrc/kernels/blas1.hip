// BLAS-1 kernels of the HIP backend: the vector operations that
// CBLASMPIKernel (simple_csr/algebra/CBLASMPIKernel.h:37-125) and the SYCL
// KernelInternal (sycl/algebra/SYCLKernelInternal.h:110-246) provide.
//
// All element-wise operations round exactly like the sequential reference
// loops (built with -ffp-contract=off: no fused multiply-add). Each lane
// handles kEPT elements spaced one workgroup apart (coalesced), loading all of
// them before using any. Dot products accumulate in FP64 in one launch and
// leave the scalar in device memory (partials + launch_finalize).
#include "../alien_hip_internal.h"
#include "device_common.h"
#include "launch.h"

namespace alien_hip {

#define EPT_INDEX(k) (blockIdx.x * (long)(kBlock * kEPT) + (k) * kBlock + threadIdx.x)

template <typename T>
__global__ void __launch_bounds__(kBlock) k_copy(int n, const T* __restrict__ x, T* __restrict__ y) {
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i < n) y[i] = x[i];
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_axpy(int n, T a, const T* __restrict__ x, T* __restrict__ y) {
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i < n) y[i] = y[i] + a * x[i];
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_scal(int n, T a, T* __restrict__ x) {
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i < n) x[i] = a * x[i];
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_assign(int n, T a, T* __restrict__ x) {
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i < n) x[i] = a;
  }
}

// z may alias x or y (pointwiseMult(invD, y, y) in the Chebyshev power method):
// each element is read and written by the same lane, so aliasing is harmless.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_pmult(int n, const T* x, const T* y, T* z) {
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    if (i < n) z[i] = x[i] * y[i];
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_dot(int n, const T* __restrict__ x, const T* __restrict__ y,
                                                 RedWork w) {
  __shared__ double lds4[4];
  T xv[kEPT], yv[kEPT];
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    xv[k] = (i < n) ? x[i] : T(0);
    yv[k] = (i < n) ? y[i] : T(0);
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kEPT; ++k) s += (double)xv[k] * (double)yv[k];
  double v[1] = {block_sum(s, lds4)};
  store_partials<1>(v, w);
}

// y += a x ; out = y . y
template <typename T>
__global__ void __launch_bounds__(kBlock) k_axpy_dot(int n, T a, const T* __restrict__ x, T* __restrict__ y,
                                                      RedWork w) {
  __shared__ double lds4[4];
  T xv[kEPT], yv[kEPT];
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    xv[k] = (i < n) ? x[i] : T(0);
    yv[k] = (i < n) ? y[i] : T(0);
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kEPT; ++k) {
    const long i = EPT_INDEX(k);
    const T v = yv[k] + a * xv[k];
    if (i < n) {
      y[i] = v;
      s += (double)v * (double)v;
    }
  }
  double v[1] = {block_sum(s, lds4)};
  store_partials<1>(v, w);
}

// ---- launchers --------------------------------------------------------------
static inline dim3 ept_grid(long n) { return dim3((unsigned)ept_blocks(n)); }

template <typename T>
void launch_copy(hipStream_t st, int n, const T* x, T* y) {
  if (n > 0) hipLaunchKernelGGL(k_copy<T>, ept_grid(n), dim3(kBlock), 0, st, n, x, y);
}
template <typename T>
void launch_axpy(hipStream_t st, int n, T a, const T* x, T* y) {
  if (n > 0) hipLaunchKernelGGL(k_axpy<T>, ept_grid(n), dim3(kBlock), 0, st, n, a, x, y);
}
template <typename T>
void launch_scal(hipStream_t st, int n, T a, T* x) {
  if (n > 0) hipLaunchKernelGGL(k_scal<T>, ept_grid(n), dim3(kBlock), 0, st, n, a, x);
}
template <typename T>
void launch_assign(hipStream_t st, int n, T a, T* x) {
  if (n > 0) hipLaunchKernelGGL(k_assign<T>, ept_grid(n), dim3(kBlock), 0, st, n, a, x);
}
template <typename T>
void launch_pmult(hipStream_t st, int n, const T* x, const T* y, T* z) {
  if (n > 0) hipLaunchKernelGGL(k_pmult<T>, ept_grid(n), dim3(kBlock), 0, st, n, x, y, z);
}
template <typename T>
void launch_dot(hipStream_t st, RedWork w, int n, const T* x, const T* y, double* out) {
  hipLaunchKernelGGL(k_dot<T>, ept_grid(n), dim3(kBlock), 0, st, n, x, y, w);
  launch_finalize(st, w, ept_blocks(n), 1, kFinStore, out, nullptr);
}
template <typename T>
void launch_axpy_dot(hipStream_t st, RedWork w, int n, T a, const T* x, T* y, double* out) {
  hipLaunchKernelGGL(k_axpy_dot<T>, ept_grid(n), dim3(kBlock), 0, st, n, a, x, y, w);
  launch_finalize(st, w, ept_blocks(n), 1, kFinStore, out, nullptr);
}

#define ALIEN_HIP_INST_BLAS1(T)                                                         \
  template void launch_copy<T>(hipStream_t, int, const T*, T*);                         \
  template void launch_axpy<T>(hipStream_t, int, T, const T*, T*);                      \
  template void launch_scal<T>(hipStream_t, int, T, T*);                                \
  template void launch_assign<T>(hipStream_t, int, T, T*);                              \
  template void launch_pmult<T>(hipStream_t, int, const T*, const T*, T*);              \
  template void launch_dot<T>(hipStream_t, RedWork, int, const T*, const T*, double*);  \
  template void launch_axpy_dot<T>(hipStream_t, RedWork, int, T, const T*, T*, double*);
ALIEN_HIP_INST_BLAS1(double)
ALIEN_HIP_INST_BLAS1(float)

// y[k] = x[idx[k]]: the values of a permuted copy of a matrix
__global__ void __launch_bounds__(kBlock) k_gather_f64(long n, const int* __restrict__ idx,
                                                        const double* __restrict__ x, double* __restrict__ y) {
  const long i = blockIdx.x * (long)kBlock + threadIdx.x;
  if (i < n) y[i] = x[idx[i]];
}
void launch_gather_f64(hipStream_t st, long n, const int* idx, const double* x, double* y) {
  if (n > 0)
    hipLaunchKernelGGL(k_gather_f64, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, n, idx, x, y);
}

} // namespace alien_hip
