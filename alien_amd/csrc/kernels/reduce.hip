// Second stage of the deterministic grid reductions: one workgroup of 1024
// lanes sums the per-workgroup partials left by a reducing kernel (lane t
// takes partials t, t+1024, ... in order, then a fixed tree), and applies the
// totals: store a scalar, or advance the device-resident CG state.
#include "../alien_hip_internal.h"
#include "device_common.h"
#include "krylov_state.h"
#include "launch.h"

namespace alien_hip {

long g_row_blocks_f64 = kUncapped;
long g_row_blocks_f32 = 4096;

constexpr int kFin = 1024;

__global__ void __launch_bounds__(kFin) k_finalize(RedWork w, int nblk, int nslot, int op, double* out,
                                                    CGState* st) {
  if (op != kFinStore && st->done) return; // the producing kernel returned early too
  __shared__ double lds[kMaxSlots][kFin / kWave];
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), wid = tid / kWave;
  // lane t sums partials t, t + kFin, t + 2 kFin, ... (fixed order); 8 loads
  // are issued before they are added
  constexpr int U = 8;
  for (int s = 0; s < nslot; ++s) {
    const double* p = w.partial + (long)s * w.cap;
    double acc = 0.0;
    int b = tid;
    for (; b + (U - 1) * kFin < nblk; b += U * kFin) {
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = p[b + u * kFin];
#pragma unroll
      for (int u = 0; u < U; ++u) acc += v[u];
    }
    for (; b < nblk; b += kFin) acc += p[b];
    acc = wave_sum(acc);
    if (lane == 0) lds[s][wid] = acc;
  }
  __syncthreads();
  if (tid != 0) return;
  double tot[kMaxSlots];
  for (int s = 0; s < nslot; ++s) {
    double t = 0.0;
    for (int q = 0; q < kFin / kWave; ++q) t += lds[s][q];
    tot[s] = t;
  }
  switch (op) {
  case kFinStore:
    for (int s = 0; s < nslot; ++s) out[s] = tot[s];
    break;
  case kFinRho1:
    st->rho1 = tot[0];
    break;
  case kFinPQ:
    st->pq = tot[0];
    if (tot[0] == 0) { // CG breakdown (p.q == 0), handled by the host
      st->breakdown = 1;
      st->done = 1;
    }
    break;
  case kFinXR:
  case kFinXRDiag:
    // end of a CG update (CG.h:140-144) and Iteration::stop (Iteration.h:83-90)
    st->rho = st->rho1;
    st->it += 1;
    if (st->it >= st->max_it) {
      st->done = 1; // stop() returns before computing the dot; status unchanged
    } else {
      st->rr = tot[0];
      st->status = (tot[0] < st->crit) ? 1 : 0;
      st->done = st->status;
    }
    if (op == kFinXRDiag) st->rho1 = tot[1]; // r.z of the Jacobi apply fused into the update
    break;
  default:
    break;
  }
}

void launch_finalize(hipStream_t s, RedWork w, long nblk, int nslot, int op, double* out, CGState* st) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFin), 0, s, w, (int)nblk, nslot, op, out, st);
}

} // namespace alien_hip
