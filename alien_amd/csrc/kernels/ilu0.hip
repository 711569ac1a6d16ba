// ILU(0) on level-set schedules.
//
// LUFactorisationAlgo (src/core/alien/expression/krylov/ILU0Preconditioner.h:71-231)
// factorises and solves row after row on the host. Row i only depends on the
// rows named by its strictly-lower entries (factorisation and L sweep) or its
// strictly-upper entries (U sweep), so rows of one level of that DAG run in
// parallel, one lane per row, and levels run in order (one launch each, all
// on one stream, capturable in a hipGraph). Inside a row the operations are
// the reference's, in the reference's order, with separate multiply and add,
// so the factors and the sweeps are bit-exact with the sequential host code
// for any schedule that respects the DAG (level sets of the natural order,
// or the colour classes of a multicolour permutation).
#include "../alien_hip_internal.h"
#include "../alien/hip/HIPContext.h"
#include "device_common.h"
#include "launch.h"

namespace alien_hip {

// IKJ update of row i. For each strictly-lower entry a_ik (ascending k):
// a_ik /= a_kk, then a_ij -= a_ik * a_kj for every later entry j of row i
// whose column also appears in row k. The reference finds matches with a
// scatter map of row k's columns; both rows are sorted, so a merge finds the
// same matches in the same order.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_ilu0_factor_level(int count, const int* __restrict__ rows,
                                                               const int* __restrict__ ro,
                                                               const int* __restrict__ cols,
                                                               const int* __restrict__ dcol,
                                                               T* __restrict__ lu) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  const int i = rows[t];
  const int rb = ro[i], re = ro[i + 1], di = dcol[i];
  for (int k = rb; k < di; ++k) {
    const int kr = cols[k];
    const T aik = lu[k] / lu[dcol[kr]];
    lu[k] = aik;
    int p = ro[kr];
    const int pe = ro[kr + 1];
    for (int j = k + 1; j < re; ++j) {
      const int cj = cols[j];
      while (p < pe && cols[p] < cj) ++p;
      if (p == pe) break;
      if (cols[p] == cj) lu[j] = lu[j] - aik * lu[p];
    }
  }
}

// One fixed-point sweep of FILU0's factorisation (FLUFactorisationAlgo::
// factorizeMultiIter / _factorizeRow, FILU0Preconditioner.h:146-170, 328-357):
// row i restarts from A's values and runs the IKJ update of ILU(0) with the
// other rows' entries (pivot a_kk and a_kj) read from the previous sweep, so
// all rows are independent: one lane per row. Row 0 keeps A's values.
__global__ void __launch_bounds__(kBlock) k_filu0_sweep(int n, const int* __restrict__ ro,
                                                         const int* __restrict__ cols,
                                                         const int* __restrict__ dcol,
                                                         const double* __restrict__ a,
                                                         const double* __restrict__ prev, double* __restrict__ lu) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int rb = ro[i], re = ro[i + 1], di = dcol[i];
  for (int k = rb; k < re; ++k) lu[k] = a[k];
  if (i == 0) return;
  for (int k = rb; k < di; ++k) {
    const int kr = cols[k];
    const double aik = lu[k] / prev[dcol[kr]];
    lu[k] = aik;
    int p = ro[kr];
    const int pe = ro[kr + 1];
    for (int j = k + 1; j < re; ++j) {
      const int cj = cols[j];
      while (p < pe && cols[p] < cj) ++p;
      if (p == pe) break;
      if (cols[p] == cj) lu[j] = lu[j] - aik * prev[p];
    }
  }
}

void launch_filu0_sweep(hipStream_t st, int n, const int* ro, const int* cols, const int* dcol, const double* a,
                        const double* prev, double* lu) {
  if (n > 0)
    hipLaunchKernelGGL(k_filu0_sweep, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, n, ro, cols,
                       dcol, a, prev, lu);
}

// L sweep of one level: t_i = y_i - sum_{k<dcol} l_ik t_k (solveL, :181-198)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_ilu0_lsolve_level(int count, const int* __restrict__ rows,
                                                               const int* __restrict__ ro,
                                                               const int* __restrict__ cols,
                                                               const int* __restrict__ dcol,
                                                               const T* __restrict__ lu,
                                                               const T* __restrict__ y, T* __restrict__ t,
                                                               const int* __restrict__ done) {
  if (done && *done) return;
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= count) return;
  const int i = rows[q];
  T v = y[i];
  const int e = dcol[i];
  for (int k = ro[i]; k < e; ++k) v = v - lu[k] * t[cols[k]];
  t[i] = v;
}

// U sweep of one level: x_i = (t_i - sum_{k>dcol} u_ik x_k) / u_ii (solveU, :200-231)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_ilu0_usolve_level(int count, const int* __restrict__ rows,
                                                               const int* __restrict__ ro,
                                                               const int* __restrict__ cols,
                                                               const int* __restrict__ dcol,
                                                               const T* __restrict__ lu,
                                                               const T* __restrict__ t, T* __restrict__ x,
                                                               const int* __restrict__ done) {
  if (done && *done) return;
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= count) return;
  const int i = rows[q];
  const int dk = dcol[i];
  T v = t[i];
  const int e = ro[i + 1];
  for (int k = dk + 1; k < e; ++k) v = v - lu[k] * x[cols[k]];
  x[i] = v / lu[dk];
}

// Consecutive small levels [l0, l1) in one workgroup: the levels in order,
// their rows strided over the workgroup's lanes, a barrier between levels (the
// rows of level l read rows of earlier levels only). The per-row operations are
// the per-level kernels'.
constexpr int kMergeRows = kBlock; // levels of at most one row per lane are merged
template <typename T, bool UPPER>
__global__ void __launch_bounds__(kBlock) k_ilu0_sweep_levels(int l0, int l1, const int* __restrict__ ptr,
                                                               const int* __restrict__ rows,
                                                               const int* __restrict__ ro,
                                                               const int* __restrict__ cols,
                                                               const int* __restrict__ dcol,
                                                               const T* __restrict__ lu, const T* in, T* out,
                                                               const int* __restrict__ done) {
  if (done && *done) return;
  for (int l = l0; l < l1; ++l) {
    for (int q = ptr[l] + (int)threadIdx.x; q < ptr[l + 1]; q += kBlock) {
      const int i = rows[q];
      if (UPPER) {
        const int dk = dcol[i];
        T v = in[i];
        const int e = ro[i + 1];
        for (int k = dk + 1; k < e; ++k) v = v - lu[k] * out[cols[k]];
        out[i] = v / lu[dk];
      } else {
        T v = in[i];
        const int e = dcol[i];
        for (int k = ro[i]; k < e; ++k) v = v - lu[k] * out[cols[k]];
        out[i] = v;
      }
    }
    __syncthreads();
  }
}

// ---- level-ordered sweeps (natural order) ------------------------------------
// The same per-row operations as k_ilu0_lsolve_level / k_ilu0_usolve_level, on the
// level-ordered copy of the factor (LevelSweep): a row's entries and values are read
// at [s * n + q], coalesced across the lanes of a level, and only the solution
// vector is gathered. Entries run in ascending column order and a missing entry
// (column -1) is skipped, so each row performs the reference's operations in its
// order.
// Levels of at most this many rows run in the one-workgroup kernel. Measured on the 128^3
// convection-diffusion system (382 levels per sweep, BiCGStab, 8 iterations): 0 (one launch per
// level) 0.051 s, 512 0.047 s, 2048 0.054 s, 4096 0.082 s; a launch costs about 3.9 us per level,
// a level in the workgroup a little less while it fits one chunk of rows.
long g_ilu_wg_rows = 512;

int level_sweep_slots(int n, const int* ro, const int* dcol, bool upper) {
  int k = 0;
  for (int i = 0; i < n; ++i) k = std::max(k, upper ? ro[i + 1] - dcol[i] - 1 : dcol[i] - ro[i]);
  // 1..8 as they are; 9..16 rounded up to 12 or 16 (the extra slots hold column -1), so that
  // the 27-point stencil's 13 takes the level-ordered copy with two more instantiations, not eight
  if (k >= 1 && k <= 8) return k;
  if (k > 8 && k <= 12) return 12;
  if (k > 12 && k <= kLevelMaxSlots) return 16;
  return 0;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_level_sweep_fill(int n, int k, int upper, const int* __restrict__ row,
                                                              const int* __restrict__ ro,
                                                              const int* __restrict__ cols,
                                                              const int* __restrict__ dcol,
                                                              const double* __restrict__ lu, int* __restrict__ col,
                                                              T* __restrict__ val, T* __restrict__ diag) {
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= n) return;
  const int i = row[q];
  const int b = upper ? dcol[i] + 1 : ro[i];
  const int e = upper ? ro[i + 1] : dcol[i];
  for (int s = 0; s < k; ++s) {
    const int kk = b + s;
    const bool on = kk < e;
    if (col) col[(long)s * n + q] = on ? cols[kk] : -1;
    val[(long)s * n + q] = on ? (T)lu[kk] : T(0);
  }
  if (upper) diag[q] = (T)lu[dcol[i]];
}

template <typename T>
void launch_level_sweep_fill(hipStream_t st, int n, int k, bool upper, const int* row, const int* ro, const int* cols,
                             const int* dcol, const double* lu, int* col, T* val, T* diag) {
  if (n > 0)
    hipLaunchKernelGGL(k_level_sweep_fill<T>, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, n, k,
                       upper ? 1 : 0, row, ro, cols, dcol, lu, col, val, diag);
}

// One row's level-ordered data and its new value (lower: t_i; upper: x_i) from `in`
// and the already final entries of `out`.
template <typename T, bool UPPER, int K>
struct LevelRow {
  int i;
  int c[K];
  T a[K];
  T d;
  __device__ __forceinline__ void fetch(const LevelSweep<T>& S, int q) {
    i = S.row[q];
#pragma unroll
    for (int s = 0; s < K; ++s) {
      c[s] = S.col[(long)s * S.n + q];
      a[s] = S.val[(long)s * S.n + q];
    }
    if (UPPER) d = S.diag[q];
  }
  __device__ __forceinline__ T value(const T* __restrict__ in, const T* out) const {
    T xv[K];
    T v = in[i];
#pragma unroll
    for (int s = 0; s < K; ++s) xv[s] = out[c[s] >= 0 ? c[s] : i]; // a missing entry loads its own row, unused
#pragma unroll
    for (int s = 0; s < K; ++s)
      if (c[s] >= 0) v = v - a[s] * xv[s];
    if (UPPER) v = v / d;
    return v;
  }
};

// one level, one lane per row (the stop flag is read beside the row's loads and only
// withholds the store)
template <typename T, bool UPPER, int K>
__global__ void __launch_bounds__(kBlock) k_level_sweep_one(int q0, int q1, LevelSweep<T> S, const T* __restrict__ in,
                                                             T* out, const int* __restrict__ done) {
  const int q = q0 + (int)(blockIdx.x * kBlock + threadIdx.x);
  if (q >= q1) return;
  const int stop = done ? *done : 0;
  LevelRow<T, UPPER, K> R;
  R.fetch(S, q);
  const T v = R.value(in, out);
  if (!stop) out[R.i] = v;
}

// Consecutive levels [l0, l1) in one workgroup of kLevelWgBlock lanes, a barrier
// between levels (a level's rows read earlier levels only, so a row may be stored
// as soon as it is computed). The first kLevelBatch x kLevelWgBlock rows of the next
// level are fetched before the barrier, overlapping the stores' completion.
constexpr int kLevelWgBlock = 1024;
constexpr int kLevelBatch = 2;
template <typename T, bool UPPER, int K>
__global__ void __launch_bounds__(kLevelWgBlock) k_level_sweep_run(int l0, int l1, const int* __restrict__ ptr,
                                                                    LevelSweep<T> S, const T* __restrict__ in, T* out,
                                                                    const int* __restrict__ done) {
  if (done && *done) return;
  // rows fetched ahead per lane: kLevelBatch, one for the wide sweeps (the registers of 1024 lanes)
  constexpr int kB = K > 8 ? 1 : kLevelBatch;
  constexpr int kChunk = kLevelWgBlock * kB;
  const int tid = (int)threadIdx.x;
  LevelRow<T, UPPER, K> R[kB];
  int q0 = ptr[l0], q1 = ptr[l0 + 1];
#pragma unroll
  for (int j = 0; j < kB; ++j)
    if (q0 + j * kLevelWgBlock + tid < q1) R[j].fetch(S, q0 + j * kLevelWgBlock + tid);
  for (int l = l0; l < l1; ++l) {
#pragma unroll
    for (int j = 0; j < kB; ++j)
      if (q0 + j * kLevelWgBlock + tid < q1) out[R[j].i] = R[j].value(in, out);
    for (int q = q0 + kChunk + tid; q < q1; q += kLevelWgBlock) { // rows beyond the first chunk
      LevelRow<T, UPPER, K> X;
      X.fetch(S, q);
      out[X.i] = X.value(in, out);
    }
    if (l + 1 < l1) {
      q0 = q1;
      q1 = ptr[l + 2];
#pragma unroll
      for (int j = 0; j < kB; ++j)
        if (q0 + j * kLevelWgBlock + tid < q1) R[j].fetch(S, q0 + j * kLevelWgBlock + tid);
    }
    __syncthreads();
  }
}

template <typename T, bool UPPER, int K>
void level_sweep_k(hipStream_t st, int nlev, const int* ptr, const int* ptr_dev, const LevelSweep<T>& S, const T* in,
                   T* out, const int* done) {
  int l = 0;
  while (l < nlev) {
    int m = l;
    while (m < nlev && ptr[m + 1] - ptr[m] <= g_ilu_wg_rows) ++m;
    if (m - l >= 2) {
      hipLaunchKernelGGL((k_level_sweep_run<T, UPPER, K>), dim3(1), dim3(kLevelWgBlock), 0, st, l, m, ptr_dev, S, in,
                         out, done);
      l = m;
      continue;
    }
    const int c = ptr[l + 1] - ptr[l];
    if (c > 0)
      hipLaunchKernelGGL((k_level_sweep_one<T, UPPER, K>), dim3((unsigned)((c + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                         st, ptr[l], ptr[l + 1], S, in, out, done);
    ++l;
  }
}

template <typename T, bool UPPER>
void level_sweep(hipStream_t st, int nlev, const int* ptr, const int* ptr_dev, const LevelSweep<T>& S, const T* in,
                 T* out, const int* done) {
  switch (S.k) {
  case 1: level_sweep_k<T, UPPER, 1>(st, nlev, ptr, ptr_dev, S, in, out, done); break;
  case 2: level_sweep_k<T, UPPER, 2>(st, nlev, ptr, ptr_dev, S, in, out, done); break;
  case 3: level_sweep_k<T, UPPER, 3>(st, nlev, ptr, ptr_dev, S, in, out, done); break;
  case 4: level_sweep_k<T, UPPER, 4>(st, nlev, ptr, ptr_dev, S, in, out, done); break;
  case 5: level_sweep_k<T, UPPER, 5>(st, nlev, ptr, ptr_dev, S, in, out, done); break;
  case 6: level_sweep_k<T, UPPER, 6>(st, nlev, ptr, ptr_dev, S, in, out, done); break;
  case 7: level_sweep_k<T, UPPER, 7>(st, nlev, ptr, ptr_dev, S, in, out, done); break;
  case 8: level_sweep_k<T, UPPER, 8>(st, nlev, ptr, ptr_dev, S, in, out, done); break;
  case 12: level_sweep_k<T, UPPER, 12>(st, nlev, ptr, ptr_dev, S, in, out, done); break;
  case 16: level_sweep_k<T, UPPER, 16>(st, nlev, ptr, ptr_dev, S, in, out, done); break;
  default: throw Error(ALIEN_HIP_EINVAL, "level-ordered ILU0 sweep: unsupported slot count");
  }
}

// ---- sync-free sweeps ------------------------------------------------------------
// One launch per sweep instead of one per level: persistent waves take chunks of at most 64 x rpl
// positions of one level from an atomic ticket, in level order, each lane working on rpl rows of
// its chunk in turn, and each row waits for the entries it reads. Those are written and polled
// with agent-scope relaxed atomics (coherent across the XCDs, whose L2s are not), over a
// signalling-NaN sentinel that the fill kernel writes first: no flag is needed, a value is ready
// when it is not the sentinel.
// Progress: a row only waits on rows of earlier levels, i.e. of chunks taken earlier by waves that
// are running, and no chunk spans two levels (the rows of a chunk never wait on each other).
// Every wait is bounded (g_ilu_sf_spins polls); one that runs out raises the error flag, the
// remaining waits give up within 256 polls, and the redo kernel queued after the sweep computes
// it again level by level.
// Each row performs the level kernels' operations in their order, so the result is theirs.
long g_ilu_sf_avg_rows = 32768, g_ilu_sf_blocks = 0, g_ilu_sf_rpl = 0, g_ilu_sf_spins = 1L << 16;
int g_ilu_sync_free = 1;

void level_chunks(int nlev, const int* ptr, int rows, std::vector<int>& cq, std::vector<int>& cn) {
  cq.clear();
  cn.clear();
  for (int l = 0; l < nlev; ++l)
    for (int q = ptr[l]; q < ptr[l + 1]; q += rows) {
      cq.push_back(q);
      cn.push_back(std::min(rows, ptr[l + 1] - q));
    }
}

void level_sf_shape(long n, long nlev, int& rpl, int& blocks) {
  // Measured on 7-point systems (BiCGStab, one launch per sweep): 128^3 (5,490 rows per level)
  // is fastest with one row per lane on 64 workgroups (2 rows: 0.039 s against 0.029 s), 256^3
  // (21,900 rows per level) with two rows per lane on 128-192 workgroups (0.145 s; one row per
  // lane 0.177-0.192 s, four 0.182-0.211 s). Both keep about three levels' rows in flight.
  const long avg = nlev > 0 ? n / nlev : n;
  rpl = g_ilu_sf_rpl == 1 || g_ilu_sf_rpl == 2 || g_ilu_sf_rpl == 4 ? (int)g_ilu_sf_rpl : (avg > 8192 ? 2 : 1);
  if (g_ilu_sf_blocks > 0) {
    blocks = (int)g_ilu_sf_blocks;
  } else {
    const long per_block = (long)kBlock * rpl; // rows one workgroup holds at once
    const long b = (3 * avg + per_block - 1) / per_block;
    blocks = (int)std::min(256L, std::max(64L, (b + kXcds - 1) / kXcds * kXcds));
  }
}

template <typename T> struct SfBits;
template <> struct SfBits<double> {
  typedef unsigned long long U;
  static constexpr U kSentinel = 0x7ff4deadbeef0001ull; // a signalling NaN: never an arithmetic result
  static __device__ __forceinline__ U bits(double v) { return (U)__double_as_longlong(v); }
  static __device__ __forceinline__ double value(U b) { return __longlong_as_double((long long)b); }
};
template <> struct SfBits<float> {
  typedef unsigned int U;
  static constexpr U kSentinel = 0x7fa5deadu;
  static __device__ __forceinline__ U bits(float v) { return __float_as_uint(v); }
  static __device__ __forceinline__ float value(U b) { return __uint_as_float(b); }
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_level_sf_fill(int n, T* out, int* work, const int* __restrict__ done) {
  if (done && *done) return;
  typedef typename SfBits<T>::U U;
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock)
    reinterpret_cast<U*>(out)[i] = SfBits<T>::kSentinel;
  if (blockIdx.x == 0 && threadIdx.x == 0) work[0] = 0; // ticket
}

template <typename T>
__device__ __forceinline__ T sf_wait(const T* p, int* err, long limit) {
  typedef typename SfBits<T>::U U;
  const U* q = reinterpret_cast<const U*>(p);
  U b = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  long spins = 0;
  while (b == SfBits<T>::kSentinel) {
    if (++spins > limit) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    if ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    __builtin_amdgcn_s_sleep(2);
    b = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return SfBits<T>::value(b);
}

template <typename T, bool UPPER, int K, int RPL>
__global__ void __launch_bounds__(kBlock) k_level_sf(LevelSweep<T> S, const T* __restrict__ in, T* out, int* work,
                                                      long limit, const int* __restrict__ done) {
  if (done && *done) return;
  typedef typename SfBits<T>::U U;
  const int lane = threadIdx.x & (kWave - 1);
  int* ticket = work;
  int* err = work + 1;
  for (;;) {
    int c = 0;
    if (lane == 0) c = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // every lane is active here, so the first is lane 0: the chunk index becomes wave-uniform
    // (a scalar register) and its bounds come through the scalar cache
    c = __builtin_amdgcn_readfirstlane(c);
    if (c >= S.nchunks) break;
    const int q0 = S.cq[c], m = S.cn[c];
    LevelRow<T, UPPER, K> R[RPL];
#pragma unroll
    for (int r = 0; r < RPL; ++r)
      if (lane + r * kWave < m) R[r].fetch(S, q0 + lane + r * kWave);
#pragma unroll
    for (int r = 0; r < RPL; ++r)
      if (lane + r * kWave < m) {
        T v = in[R[r].i];
        if constexpr (K > 1) {
          // Wait first for the nearest entry (the highest column of the lower part, the lowest of
          // the upper part: the latest level the row reads), after which the others are almost
          // always there, then read those together; the sum keeps the slot order.
          int sn = -1;
#pragma unroll
          for (int s = 0; s < K; ++s)
            if (R[r].c[s] >= 0 && (!UPPER || sn < 0)) sn = s;
          U b[K];
#pragma unroll
          for (int s = 0; s < K; ++s)
            if (s == sn) b[s] = SfBits<T>::bits(sf_wait<T>(out + R[r].c[s], err, limit));
#pragma unroll
          for (int s = 0; s < K; ++s)
            if (s != sn && R[r].c[s] >= 0)
              b[s] = __hip_atomic_load(reinterpret_cast<const U*>(out) + R[r].c[s], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
          for (int s = 0; s < K; ++s)
            if (R[r].c[s] >= 0) {
              if (b[s] == SfBits<T>::kSentinel) b[s] = SfBits<T>::bits(sf_wait<T>(out + R[r].c[s], err, limit));
              v = v - R[r].a[s] * SfBits<T>::value(b[s]);
            }
        } else {
#pragma unroll
          for (int s = 0; s < K; ++s)
            if (R[r].c[s] >= 0) v = v - R[r].a[s] * sf_wait<T>(out + R[r].c[s], err, limit);
        }
        if (UPPER) v = v / R[r].d;
        __hip_atomic_store(reinterpret_cast<U*>(out) + R[r].i, SfBits<T>::bits(v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
  }
}

// The redo of a sync-free sweep whose error flag is set (a wait ran out of polls), as a kernel so
// that no host round trip follows each sweep: one workgroup takes the levels in order, a barrier
// between them, with the level kernels' per-row operations; it then clears the flag. Without the
// flag it returns at once.
template <typename T, bool UPPER, int K>
__global__ void __launch_bounds__(kLevelWgBlock) k_level_sf_redo(int nlev, const int* __restrict__ ptr,
                                                                  LevelSweep<T> S, const T* __restrict__ in, T* out,
                                                                  int* work, const int* __restrict__ done) {
  if (done && *done) return;
  if (__hip_atomic_load(work + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
  for (int l = 0; l < nlev; ++l) {
    for (int q = ptr[l] + (int)threadIdx.x; q < ptr[l + 1]; q += kLevelWgBlock) {
      LevelRow<T, UPPER, K> X;
      X.fetch(S, q);
      out[X.i] = X.value(in, out);
    }
    __syncthreads();
  }
  __syncthreads(); // every lane has read the flag
  if (threadIdx.x == 0) __hip_atomic_store(work + 1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T, bool UPPER, int K>
void level_sf_k(hipStream_t st, int nlev, const int* ptr_dev, const LevelSweep<T>& S, const T* in, T* out, int* work,
                const int* done) {
  const dim3 g((unsigned)std::max(1, S.blocks)), b(kBlock);
  const long lim = g_ilu_sf_spins;
  switch (S.rpl) {
  case 1: hipLaunchKernelGGL((k_level_sf<T, UPPER, K, 1>), g, b, 0, st, S, in, out, work, lim, done); break;
  case 2: hipLaunchKernelGGL((k_level_sf<T, UPPER, K, 2>), g, b, 0, st, S, in, out, work, lim, done); break;
  case 4: hipLaunchKernelGGL((k_level_sf<T, UPPER, K, 4>), g, b, 0, st, S, in, out, work, lim, done); break;
  default: throw Error(ALIEN_HIP_EINVAL, "sync-free ILU0 sweep: rows per lane must be 1, 2 or 4");
  }
  hipLaunchKernelGGL((k_level_sf_redo<T, UPPER, K>), dim3(1), dim3(kLevelWgBlock), 0, st, nlev, ptr_dev, S, in, out,
                     work, done);
}

template <typename T, bool UPPER>
void level_sweep_sf(hipStream_t st, int nlev, const int* ptr_dev, const LevelSweep<T>& S, const T* in, T* out,
                    int* work, const int* done) {
  hipLaunchKernelGGL(k_level_sf_fill<T>, dim3((unsigned)grid_for(S.n, 4)), dim3(kBlock), 0, st, S.n, out, work, done);
  switch (S.k) {
  case 1: level_sf_k<T, UPPER, 1>(st, nlev, ptr_dev, S, in, out, work, done); break;
  case 2: level_sf_k<T, UPPER, 2>(st, nlev, ptr_dev, S, in, out, work, done); break;
  case 3: level_sf_k<T, UPPER, 3>(st, nlev, ptr_dev, S, in, out, work, done); break;
  case 4: level_sf_k<T, UPPER, 4>(st, nlev, ptr_dev, S, in, out, work, done); break;
  case 5: level_sf_k<T, UPPER, 5>(st, nlev, ptr_dev, S, in, out, work, done); break;
  case 6: level_sf_k<T, UPPER, 6>(st, nlev, ptr_dev, S, in, out, work, done); break;
  case 7: level_sf_k<T, UPPER, 7>(st, nlev, ptr_dev, S, in, out, work, done); break;
  case 8: level_sf_k<T, UPPER, 8>(st, nlev, ptr_dev, S, in, out, work, done); break;
  case 12: level_sf_k<T, UPPER, 12>(st, nlev, ptr_dev, S, in, out, work, done); break;
  case 16: level_sf_k<T, UPPER, 16>(st, nlev, ptr_dev, S, in, out, work, done); break;
  default: throw Error(ALIEN_HIP_EINVAL, "sync-free ILU0 sweep: unsupported slot count");
  }
}

template <typename T>
void launch_ilu0_solve_lev(hipStream_t st, int nl, const int* lptr_host, const int* lptr_dev, LevelSweep<T> L, int nu,
                           const int* uptr_host, const int* uptr_dev, LevelSweep<T> U, const T* y, T* t, T* x,
                           const int* done, int* sf_work) {
  if (sf_work && g_ilu_sync_free && L.nchunks > 0 && U.nchunks > 0) {
    // per sweep: sentinel fill, the persistent sweep, the redo kernel (a no-op unless a wait ran
    // out of polls); no host round trip, so a solve can capture it in a graph
    level_sweep_sf<T, false>(st, nl, lptr_dev, L, y, t, sf_work, done);
    level_sweep_sf<T, true>(st, nu, uptr_dev, U, t, x, sf_work, done);
    return;
  }
  level_sweep<T, false>(st, nl, lptr_host, lptr_dev, L, y, t, done);
  level_sweep<T, true>(st, nu, uptr_host, uptr_dev, U, t, x, done);
}

// ---- multicolour sweeps over the split factor --------------------------------
// The factor LU of the colour-permuted matrix is stored as two row-bundled
// parts (64 rows per bundle, slot (bo[b] + j) * 64 + lane, column -1 = padding)
// holding each row's strictly-lower and strictly-upper entries in ascending
// column order, and the diagonal u_ii. The rows of one colour class are
// independent and every row they read belongs to a colour swept before (the
// lower sweep takes the colours in order, the upper sweep in reverse), so one
// launch per colour and sweep computes solveL / solveU
// (ILU0Preconditioner.h:181-231) with the reference's operations in its order:
// t_i = y_i - l_ic t_c ... (ascending c), x_i = (t_i - u_ic x_c ...) / u_ii.
// The same layout of FILU0's factor (natural order) serves its Jacobi sweeps:
// addLMult / addUMult with alpha = -1 (FILU0Preconditioner.h:226-230, 297-303)
// add (-a) x = -(a x) term by term, the subtraction above, reading the
// previous sweep's vector instead of the one being written.

// One thread per row: its entries [ro, dcol) into the lower part, (dcol, ro+1)
// into the upper part, padding up to the bundle width; u_ii into diag. cols
// are written when lbc is not null (the structure is shared by FP64 and FP32).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_ilu0_split_fill(int n, const int* __restrict__ ro,
                                                             const int* __restrict__ cols,
                                                             const int* __restrict__ dcol,
                                                             const double* __restrict__ lu,
                                                             const int* __restrict__ lbo, const int* __restrict__ ubo,
                                                             int* __restrict__ lbc, int* __restrict__ ubc,
                                                             T* __restrict__ lbv, T* __restrict__ ubv,
                                                             T* __restrict__ diag) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int b = i / kWave, lane = i % kWave;
  const int k0 = ro[i], kd = dcol[i], k1 = ro[i + 1];
  diag[i] = (T)lu[kd];
  const int lw = lbo[b + 1] - lbo[b], uw = ubo[b + 1] - ubo[b];
  for (int j = 0; j < lw; ++j) {
    const long sl = (long)(lbo[b] + j) * kWave + lane;
    const bool on = k0 + j < kd;
    if (lbc) lbc[sl] = on ? cols[k0 + j] : -1;
    lbv[sl] = on ? (T)lu[k0 + j] : T(0);
  }
  for (int j = 0; j < uw; ++j) {
    const long sl = (long)(ubo[b] + j) * kWave + lane;
    const bool on = kd + 1 + j < k1;
    if (ubc) ubc[sl] = on ? cols[kd + 1 + j] : -1;
    ubv[sl] = on ? (T)lu[kd + 1 + j] : T(0);
  }
}

// Rows [r0, r1), one wavefront per bundle from bundle b0: out_i = in_i -
// sum_c a_ic src_c over the part's entries, then (KIND) as is (lower), divided
// by diag_i (ILU(0) upper, and the lower sweep of the last colour, which has no
// upper entries) or multiplied by diag_i holding the inverse diagonal (FILU0
// upper, pointwiseMult after addUMult). Columns below split are gathered from
// ysrc instead of src (the rows of colour 0, whose lower sweep is the identity).
// The slots are read in groups of kSweepGroup (columns, then values and the
// gathered values) before the group's subtractions, which keep their order.
// Multicolour: src = the vector being swept (rows of earlier colours); FILU0:
// src = the previous sweep. With rdot (the sweeps that produce z in CG) each
// workgroup also stores its partial of rdot . out at partial[poff +
// blockIdx.x], so that the launches together leave the partials of r . z for
// one finalize. Returns at entry once *done.
constexpr int kSweepGroup = 7;
enum SweepKind { kSweepLower = 0, kSweepUpperDiv = 1, kSweepUpperInvMul = 2, kSweepLowerDiv = 3 };
template <typename T, int KIND>
__global__ void __launch_bounds__(kRowBlock) k_ilu0_split_sweep(int b0, int r0, int r1, const int* __restrict__ bo,
                                                                 const int* __restrict__ bc,
                                                                 const T* __restrict__ bv,
                                                                 const T* __restrict__ diag,
                                                                 const T* __restrict__ in, const T* src, T* out,
                                                                 const T* __restrict__ ysrc, int split,
                                                                 const int* __restrict__ done,
                                                                 const T* __restrict__ rdot, RedWork red, int poff) {
  if (done && *done) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int b = b0 + (int)((blockIdx.x * (long)kRowBlock + threadIdx.x) / kWave);
  const bool wave_on = (long)b * kWave < r1;
  const int row = b * kWave + lane;
  const bool act = wave_on && row >= r0 && row < r1;
  T res = T(0);
  if (wave_on) {
    const int s0 = bo[b], w = bo[b + 1] - s0;
    T acc = act ? in[row] : T(0);
    for (int j0 = 0; j0 < w; j0 += kSweepGroup) {
      int c[kSweepGroup];
      T v[kSweepGroup], xv[kSweepGroup];
#pragma unroll
      for (int u = 0; u < kSweepGroup; ++u)
        c[u] = (act && j0 + u < w) ? bc[(long)(s0 + j0 + u) * kWave + lane] : -1;
#pragma unroll
      for (int u = 0; u < kSweepGroup; ++u) {
        v[u] = c[u] >= 0 ? bv[(long)(s0 + j0 + u) * kWave + lane] : T(0);
        xv[u] = c[u] < 0 ? T(0) : c[u] < split ? ysrc[c[u]] : src[c[u]];
      }
#pragma unroll
      for (int u = 0; u < kSweepGroup; ++u)
        if (c[u] >= 0) acc = acc - v[u] * xv[u];
    }
    if (act) {
      res = (KIND == kSweepUpperDiv || KIND == kSweepLowerDiv) ? acc / diag[row]
            : KIND == kSweepUpperInvMul                         ? diag[row] * acc
                                                                : acc;
      out[row] = res;
    }
  }
  if (KIND != kSweepLower && rdot) {
    __shared__ double lds4[kRowWaves];
    const double sv = act ? (double)rdot[row] * (double)res : 0.0;
    const double tot = block_sum<kRowBlock>(sv, lds4);
    if (threadIdx.x == 0) red.partial[poff + blockIdx.x] = tot;
  }
}

template <typename T>
void launch_ilu0_split_fill(hipStream_t st, int n, const int* ro, const int* cols, const int* dcol, const double* lu,
                            const int* lbo, const int* ubo, int* lbc, int* ubc, T* lbv, T* ubv, T* diag) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_ilu0_split_fill<T>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, n, ro, cols, dcol,
                     lu, lbo, ubo, lbc, ubc, lbv, ubv, diag);
}

int ilu0_split_partials(int ncolours, const int* cptr_host) {
  int tot = 0;
  for (int c = 0; c < ncolours; ++c) {
    const int r0 = cptr_host[c], r1 = cptr_host[c + 1];
    if (r1 > r0) tot += ((r1 - 1) / kWave - r0 / kWave + 1 + kRowWaves - 1) / kRowWaves;
  }
  return tot;
}

// Colour 0 comes first in the permuted order, so its rows have no strictly-
// lower entries (t = y there: no launch, later gathers of those rows read y),
// and the last colour has no strictly-upper entries (x = t / u_ii: its lower
// sweep divides and writes x directly). Launches: lower sweeps of colours 1 ..
// nc-1, then upper sweeps of colours nc-2 .. 0 (colour 0 reads t = y); one
// colour: x = y / u_ii.
template <typename T>
int launch_ilu0_split_apply(hipStream_t st, int ncolours, const int* cptr_host, ILUSplit<T> F, const T* y, T* t,
                            T* x, const int* done, const T* rdot, RedWork red) {
  if (ncolours < 1) return 0;
  int poff = 0;
  const int split = cptr_host[1]; // rows of colour 0
  auto geom = [&](int c, int& b0, int& nblk) {
    const int r0 = cptr_host[c], r1 = cptr_host[c + 1];
    if (r1 <= r0) return false;
    b0 = r0 / kWave;
    nblk = ((r1 - 1) / kWave - b0 + 1 + kRowWaves - 1) / kRowWaves;
    return true;
  };
  int b0, nblk;
  const int last = ncolours - 1;
  for (int c = 1; c < ncolours; ++c) {
    if (!geom(c, b0, nblk)) continue;
    if (c == last) {
      hipLaunchKernelGGL((k_ilu0_split_sweep<T, kSweepLowerDiv>), dim3(nblk), dim3(kRowBlock), 0, st, b0,
                         cptr_host[c], cptr_host[c + 1], F.lbo, F.lbc, F.lbv, F.diag, y, (const T*)t, x, y, split,
                         done, rdot, red, poff);
      if (rdot) poff += nblk;
    } else {
      hipLaunchKernelGGL((k_ilu0_split_sweep<T, kSweepLower>), dim3(nblk), dim3(kRowBlock), 0, st, b0,
                         cptr_host[c], cptr_host[c + 1], F.lbo, F.lbc, F.lbv, F.diag, y, (const T*)t, t, y, split,
                         done, (const T*)nullptr, red, 0);
    }
  }
  for (int c = ncolours == 1 ? 0 : last - 1; c >= 0; --c) {
    if (!geom(c, b0, nblk)) continue;
    hipLaunchKernelGGL((k_ilu0_split_sweep<T, kSweepUpperDiv>), dim3(nblk), dim3(kRowBlock), 0, st, b0,
                       cptr_host[c], cptr_host[c + 1], F.ubo, F.ubc, F.ubv, F.diag, c == 0 ? y : (const T*)t,
                       (const T*)x, x, y, 0, done, rdot, red, poff);
    if (rdot) poff += nblk;
  }
  return poff;
}

// FILU0 solve (FILU0Preconditioner.h:236-256, device branch): niter >= 1 Jacobi
// sweeps with the lower part from mx = y, then niter with the upper part from
// z = mx, each upper sweep scaled by the inverse diagonal (F.diag). t1 ends as
// mx; t2 and z alternate so that the last upper sweep writes z.
template <typename T>
int launch_filu0_split_apply(hipStream_t st, int n, int niter, ILUSplit<T> F, const T* y, T* t1, T* t2, T* z,
                             const int* done, const T* rdot, RedWork red) {
  const int nb = (n + kWave - 1) / kWave;
  const dim3 g((unsigned)((nb + kRowWaves - 1) / kRowWaves));
  // lower: src_0 = y, outputs alternate t2 / t1 ending in t1
  const T* src = y;
  for (int it = 0; it < niter; ++it) {
    T* out = ((niter - 1 - it) % 2 == 0) ? t1 : t2;
    hipLaunchKernelGGL((k_ilu0_split_sweep<T, kSweepLower>), g, dim3(kRowBlock), 0, st, 0, 0, n, F.lbo, F.lbc, F.lbv,
                       F.diag, y, src, out, (const T*)nullptr, 0, done, (const T*)nullptr, red, 0);
    src = out;
  }
  // upper: src_0 = mx = t1, outputs alternate t2 / z ending in z
  src = t1;
  for (int it = 0; it < niter; ++it) {
    T* out = ((niter - 1 - it) % 2 == 0) ? z : t2;
    const bool last = it == niter - 1;
    hipLaunchKernelGGL((k_ilu0_split_sweep<T, kSweepUpperInvMul>), g, dim3(kRowBlock), 0, st, 0, 0, n, F.ubo, F.ubc,
                       F.ubv, F.diag, (const T*)t1, src, out, (const T*)nullptr, 0, done,
                       last ? rdot : (const T*)nullptr, red, 0);
    src = out;
  }
  return (rdot && niter > 0) ? (int)g.x : 0;
}

template <typename T>
void launch_ilu0_factor(hipStream_t st, int nlevels, const int* ptr_host, const int* rows, const int* ro,
                        const int* cols, const int* dcol, T* lu) {
  for (int l = 0; l < nlevels; ++l) {
    const int c = ptr_host[l + 1] - ptr_host[l];
    if (c > 0)
      hipLaunchKernelGGL(k_ilu0_factor_level<T>, dim3((c + kBlock - 1) / kBlock), dim3(kBlock), 0, st, c,
                         rows + ptr_host[l], ro, cols, dcol, lu);
  }
}

template <typename T>
void launch_ilu0_solve(hipStream_t st, int nl, const int* lptr_host, const int* lrows, int nu,
                       const int* uptr_host, const int* urows, const int* ro, const int* cols, const int* dcol,
                       const T* lu, const T* y, T* t, T* x, const int* done, const int* lptr_dev,
                       const int* uptr_dev) {
  // per level one launch, or (with the level pointers on the device) one
  // single-workgroup launch per run of at least two levels of <= kMergeRows rows
  auto sweep = [&](bool upper, int nlev, const int* ptr, const int* ptr_dev, const int* rows) {
    int l = 0;
    while (l < nlev) {
      int m = l;
      if (ptr_dev)
        while (m < nlev && ptr[m + 1] - ptr[m] <= kMergeRows) ++m;
      if (m - l >= 2) {
        if (upper)
          hipLaunchKernelGGL((k_ilu0_sweep_levels<T, true>), dim3(1), dim3(kBlock), 0, st, l, m, ptr_dev, rows, ro,
                             cols, dcol, lu, (const T*)t, x, done);
        else
          hipLaunchKernelGGL((k_ilu0_sweep_levels<T, false>), dim3(1), dim3(kBlock), 0, st, l, m, ptr_dev, rows, ro,
                             cols, dcol, lu, y, t, done);
        l = m;
        continue;
      }
      const int c = ptr[l + 1] - ptr[l];
      if (c > 0) {
        if (upper)
          hipLaunchKernelGGL(k_ilu0_usolve_level<T>, dim3((c + kBlock - 1) / kBlock), dim3(kBlock), 0, st, c,
                             rows + ptr[l], ro, cols, dcol, lu, t, x, done);
        else
          hipLaunchKernelGGL(k_ilu0_lsolve_level<T>, dim3((c + kBlock - 1) / kBlock), dim3(kBlock), 0, st, c,
                             rows + ptr[l], ro, cols, dcol, lu, y, t, done);
      }
      ++l;
    }
  };
  sweep(false, nl, lptr_host, lptr_dev, lrows);
  sweep(true, nu, uptr_host, uptr_dev, urows);
}

#define ALIEN_HIP_INST_ILU0(T)                                                                            \
  template void launch_level_sweep_fill<T>(hipStream_t, int, int, bool, const int*, const int*, const int*,  \
                                           const int*, const double*, int*, T*, T*);                       \
  template void launch_ilu0_solve_lev<T>(hipStream_t, int, const int*, const int*, LevelSweep<T>, int,      \
                                         const int*, const int*, LevelSweep<T>, const T*, T*, T*,           \
                                         const int*, int*);                                                 \
  template void launch_ilu0_factor<T>(hipStream_t, int, const int*, const int*, const int*, const int*,   \
                                      const int*, T*);                                                    \
  template void launch_ilu0_solve<T>(hipStream_t, int, const int*, const int*, int, const int*,           \
                                     const int*, const int*, const int*, const int*, const T*, const T*,  \
                                     T*, T*, const int*, const int*, const int*);                         \
  template void launch_ilu0_split_fill<T>(hipStream_t, int, const int*, const int*, const int*,           \
                                          const double*, const int*, const int*, int*, int*, T*, T*, T*);  \
  template int launch_ilu0_split_apply<T>(hipStream_t, int, const int*, ILUSplit<T>, const T*, T*, T*,    \
                                          const int*, const T*, RedWork);                                  \
  template int launch_filu0_split_apply<T>(hipStream_t, int, int, ILUSplit<T>, const T*, T*, T*, T*,       \
                                           const int*, const T*, RedWork);
ALIEN_HIP_INST_ILU0(double)
ALIEN_HIP_INST_ILU0(float)

} // namespace alien_hip
