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
long g_row_blocks_uni = 16384;
long g_row_blocks_uni_wide = 32768;

constexpr int kFin = 1024;

// Sums of nslot partial arrays (nblk entries each) by one workgroup of kFin
// lanes: lane t takes partials t, t + kFin, ... in order (8 loads issued before
// they are added), then a fixed tree. The totals are valid on lane 0 only,
// which is the lane that returns true.
__device__ bool finalize_totals(RedWork w, int nblk, int nslot, double (&tot)[kMaxSlots]) {
  __shared__ double lds[kMaxSlots][kFin / kWave];
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), wid = tid / kWave;
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
  if (tid != 0) return false;
  for (int s = 0; s < nslot; ++s) {
    double t = 0.0;
    for (int q = 0; q < kFin / kWave; ++q) t += lds[s][q];
    tot[s] = t;
  }
  return true;
}

__global__ void __launch_bounds__(kFin) k_finalize(RedWork w, int nblk, int nslot, int op, double* out,
                                                    CGState* st) {
  if (op != kFinStore && st->done) return; // the producing kernel returned early too
  double tot[kMaxSlots];
  if (!finalize_totals(w, nblk, nslot, tot)) return;
  switch (op) {
  case kFinStore:
    for (int s = 0; s < nslot; ++s) out[s] = tot[s];
    break;
  case kFinRho1:
    st->rho1 = tot[0];
    break;
  case kFinRho1B: // rho1 from the second slot (the first holds r.r, unused here)
    st->rho1 = tot[1];
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

// BiCGStab steps (BiCGStab.h:151-300): the totals of each reduction and the
// decisions the reference takes on the host after it.
__global__ void __launch_bounds__(kFin) k_finalize_bicg(RedWork w, int nblk, int nslot, int op, BiCGState* st) {
  if (st->done) return;
  double tot[kMaxSlots];
  if (!finalize_totals(w, nblk, nslot, tot)) return;
  switch (op) {
  case kBFinAlpha: // alpha = rho1 / (v.r0); NullValueException("alpha") on 0 (:207, :250)
    if (tot[0] == 0) {
      st->breakdown = 1;
      st->done = 1;
    } else {
      st->alpha = st->rho1 / tot[0];
    }
    break;
  case kBFinS: // Iteration::stop(s): half-step exit, counted only in the first block (:194-199, :259-263)
    if (st->it >= st->max_it) {
      st->done = 1;
      st->half = 1;
    } else {
      st->rr = tot[0];
      st->rrmax = fmax(st->rrmax, tot[0]);
      if (tot[0] < st->crit) {
        st->status = 1;
        st->done = 1;
        st->half = 1;
      }
    }
    if (st->half && st->first) st->it += 1;
    break;
  case kBFinOmega: // omega = (t.s) / (t.t); t.t == 0: stop(r) in the first block, stop(s) in the loop (:210-218)
    if (tot[1] == 0) {
      const bool stop = st->first && (st->it >= st->max_it || st->rr0 < st->crit);
      if (stop) {
        if (st->it < st->max_it) st->status = 1;
        st->it += 1;
        st->half = 1;
      } else {
        st->breakdown = 2;
      }
      st->done = 1;
    } else {
      st->omega = tot[0] / tot[1];
    }
    break;
  case kBFinXR: // rho = rho1; ++iter; while (!stop(r)); next rho1 = r.r0 (:224-230, :232-234)
    st->rho = st->rho1;
    st->rho1 = tot[1];
    st->it += 1;
    st->first = 0;
    if (st->it >= st->max_it) {
      st->done = 1;
    } else {
      st->rr = tot[0];
      st->rrmax = fmax(st->rrmax, tot[0]);
      st->status = (tot[0] < st->crit) ? 1 : 0;
      st->done = st->status;
      if (!st->done && st->rho == 0) { // NullValueException("rho") at the next step (:240)
        st->breakdown = 3;
        st->done = 1;
      }
    }
    break;
  default:
    break;
  }
}

void launch_finalize_bicg(hipStream_t s, RedWork w, long nblk, int nslot, int op, BiCGState* st) {
  hipLaunchKernelGGL(k_finalize_bicg, dim3(1), dim3(kFin), 0, s, w, (int)nblk, nslot, op, st);
}

void launch_finalize(hipStream_t s, RedWork w, long nblk, int nslot, int op, double* out, CGState* st) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFin), 0, s, w, (int)nblk, nslot, op, out, st);
}

} // namespace alien_hip
