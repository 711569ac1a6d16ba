// Whole-solve driver of the hip backend: test_krylov's run(alg)
// (examples/test_krylov.cpp:500-665) for --kernel hip, plus the fused,
// graph-captured and mixed-precision variants of the CG path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <type_traits>
#include <vector>

#include "alien/hip/HIPBackEnd.h"
#include "alien/hip/krylov/HIPFILU0Preconditioner.h"
#include "alien/hip/krylov/HIPILU0Preconditioner.h"
#include "alien/hip/krylov/Krylov.h"
#include "alien_hip_internal.h"
#include "kernels/device_common.h"
#include "kernels/gnn.h"
#include "kernels/krylov_state.h"
#include "kernels/launch.h"

namespace alien_hip {

namespace {

// Identity preconditioner (test_krylov has none; used for unpreconditioned runs).
template <typename AlgebraT>
struct IdentityPreconditioner {
  void init() {}
  template <typename Alg>
  void solve(Alg& alg, const typename AlgebraT::Vector& x, typename AlgebraT::Vector& y) const {
    alg.copy(x, y);
  }
};

struct EventTimer {
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t s;
  explicit EventTimer(hipStream_t st) : s(st) {
    ALIEN_HIP_CHECK(hipEventCreate(&a));
    ALIEN_HIP_CHECK(hipEventCreate(&b));
  }
  ~EventTimer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  void start() { ALIEN_HIP_CHECK(hipEventRecord(a, s)); }
  double stop() {
    ALIEN_HIP_CHECK(hipEventRecord(b, s));
    ALIEN_HIP_CHECK(hipEventSynchronize(b));
    float ms = 0;
    ALIEN_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    return ms * 1e-3;
  }
};

// Host-side validation of an input CSR before anything is uploaded: offsets
// monotone, columns in range and strictly ascending inside each row, and a
// stored diagonal in every row (dcol must point at a_ii for the inverse
// diagonal and the ILU0 pivots).
void validate_csr(int n, const int* ro, const int* cols, bool need_diag) {
  if (n < 0 || !ro || (n > 0 && !cols)) throw Error(ALIEN_HIP_EINVAL, "CSR: NULL arrays");
  if (ro[0] != 0) throw Error(ALIEN_HIP_EINVAL, "CSR: row_offset[0] != 0");
  for (int i = 0; i < n; ++i) {
    if (ro[i + 1] < ro[i]) throw Error(ALIEN_HIP_EINVAL, "CSR: row_offset not monotone");
    bool diag = false;
    for (int k = ro[i]; k < ro[i + 1]; ++k) {
      if (cols[k] < 0 || cols[k] >= n) throw Error(ALIEN_HIP_EINVAL, "CSR: column out of range");
      if (k > ro[i] && cols[k] <= cols[k - 1]) throw Error(ALIEN_HIP_EINVAL, "CSR: columns not strictly ascending");
      diag |= (cols[k] == i);
    }
    if (need_diag && !diag) throw Error(ALIEN_HIP_EINVAL, "CSR: missing diagonal entry");
  }
}

// Symmetric permutation P A P^T with rows (and columns) renumbered so that
// new row r is old row perm[r]; columns re-sorted ascending inside each row.
void permute_csr(int n, const int* ro, const int* cols, const double* vals, const std::vector<int>& perm,
                 std::vector<int>& pro, HostBuf<int>& pcols, HostBuf<double>& pvals) {
  HostBuf<int> inv;
  inv.alloc(n);
  parallel_ranges(n, [&](long a, long b) {
    for (long r = a; r < b; ++r) inv[perm[r]] = (int)r;
  });
  pro.assign(n + 1, 0);
  for (int r = 0; r < n; ++r) pro[r + 1] = pro[r] + (ro[perm[r] + 1] - ro[perm[r]]);
  pcols.alloc(ro[n]);
  pvals.alloc(ro[n]);
  // rows are independent: host threads over row ranges, insertion sort for short rows
  parallel_ranges(n, [&](long a, long b) {
    std::vector<std::pair<int, double>> row;
    for (long r = a; r < b; ++r) {
      const int o = perm[r];
      row.clear();
      for (int k = ro[o]; k < ro[o + 1]; ++k) row.emplace_back(inv[cols[k]], vals[k]);
      if (row.size() <= 32) {
        for (size_t q = 1; q < row.size(); ++q)
          for (size_t t = q; t > 0 && row[t].first < row[t - 1].first; --t) std::swap(row[t], row[t - 1]);
      } else {
        std::sort(row.begin(), row.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
      }
      for (size_t q = 0; q < row.size(); ++q) {
        pcols[pro[r] + q] = row[q].first;
        pvals[pro[r] + q] = row[q].second;
      }
    }
  });
}

// rows sorted by colour, natural order inside a colour; cptr[c] = first permuted
// row of colour c (nc + 1 entries)
std::vector<int> colour_permutation(int n, const int* ro, const int* cols, std::vector<int>& cptr) {
  std::vector<int> color(n);
  int nc = host_color(n, ro, cols, color.data());
  std::vector<int> perm(n);
  std::vector<int> cnt(nc + 1, 0);
  for (int i = 0; i < n; ++i) ++cnt[color[i] + 1];
  for (int c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
  cptr = cnt;
  for (int i = 0; i < n; ++i) perm[cnt[color[i]]++] = i;
  return perm;
}

using Alien::HIPInternalLinearAlgebra;
using Alien::HIPMatrix;
using Alien::HIPVector;
typedef HIPInternalLinearAlgebra Alg;

// ---- reference mode: the Krylov templates over the hip algebra ----------------------------
template <typename Precond>
void run_reference(Alg& alg, const alien_hip_solve_params* p, Precond& P, const HIPMatrix& A, const HIPVector& b,
                   HIPVector& x, alien_hip_solve_result* res, EventTimer& tm) {
  Alien::Iteration<Alg> it(alg, b, p->tol, p->max_iter);
  tm.start();
  try {
    const bool future = p->mode == ALIEN_HIP_MODE_REFERENCE_FUTURE;
    if (p->solver == ALIEN_HIP_CG) {
      Alien::CG<Alg> s(alg);
      if (future) s.solve2(P, it, A, b, x);
      else s.solve(P, it, A, b, x);
    } else {
      Alien::BiCGStab<Alg> s(alg);
      if (future) s.solve2(P, it, A, b, x);
      else s.solve(P, it, A, b, x);
    }
  } catch (const Alg::NullValueException&) {
    res->error = 1;
  }
  res->solve_s = tm.stop();
  res->iterations = it();
  res->status = it.getStatus() ? 1 : 0;
  res->value = it.getValue();
}

// ---- fused CG ------------------------------------------------------------------------------
enum PrecKind {
  kPrecDiag = 0,
  kPrecCheb = 1,
  kPrecFactor = 2,
  kPrecILU0 = 3,
  kPrecFILU0 = 4,
  kPrecNeumann = 5,
  kPrecILU0Level = 6
};
// Natural-order ILU(0) on its level-set schedules (the reference's ILU0 bit for bit):
// one launch per level and sweep over the CSR of A and the factor in CSR order.
struct LevelILU {
  int nl = 0, nu = 0;
  const int* lptr = nullptr; // host
  const int* uptr = nullptr; // host
  const int* lrows = nullptr;
  const int* urows = nullptr;
  const int* lptr_dev = nullptr;
  const int* uptr_dev = nullptr;
  const int* ro = nullptr;
  const int* cols = nullptr;
  const int* dcol = nullptr;
  const double* lu64 = nullptr;
  const float* lu32 = nullptr;
  // the factor in level order (both sweeps), when the preconditioner built it
  bool ordered = false;
  LevelSweep<double> l64{}, u64{};
  LevelSweep<float> l32{}, u32{};
  int* sf_work = nullptr; // the sync-free sweeps' ticket and error flag, when built
};
// The sweeps run sync-free (a few launches per apply, no host call): a solve may capture them.
inline bool level_sync_free(const LevelILU& L) {
  return L.ordered && L.sf_work && g_ilu_sync_free && L.l64.nchunks > 0 && L.u64.nchunks > 0;
}
struct PrecondSpec {
  PrecKind kind = kPrecDiag;
  int order = 0;
  std::vector<double> a, b; // Chebyshev step coefficients
  // factored approximate inverse M = G G^T: G and G^T row-bundled, in the
  // precision of the solver using them (one pair set)
  RBMat<double> g64{}, gt64{};
  RBMat<float> g32{}, gt32{};
  // FP32 solver with bfloat16 factor values (used instead of g32 / gt32 when `half`)
  RBMat<float, bf16_bits> g16{}, gt16{};
  bool half = false;
  // multicolour ILU(0): the colour classes of the permuted rows and the split
  // factor in the solver's precision; FILU0: the split factor (diag = its
  // inverse diagonal) and the Jacobi sweeps per part
  std::vector<int> cptr;
  LevelILU lev;
  int filu_iters = 0;
  double neumann_factor = 0; // Neumann: f of z = f sum_{k<=order} (I - f A)^k r
  ILUSplit<double> lu64{};
  ILUSplit<float> lu32{};
  template <typename T> RBMat<T> g() const;
  template <typename T> RBMat<T> gt() const;
  template <typename T> ILUSplit<T> lu() const;
};

// Storage of the split ILU(0) factor: bundle widths from the host CSR (longest
// strictly-lower / strictly-upper part of the bundle's rows), columns filled
// with the first values.
struct ILUSplitStore {
  DeviceArray<int> lbo, ubo, lbc, ubc;
  DeviceArray<double> lbv64, ubv64, d64;
  DeviceArray<float> lbv32, ubv32, d32;
  size_t lslots = 0, uslots = 0;
  void build(hipStream_t st, int n, const int* ro, const int* dcol) {
    const int nb = (n + kBundle - 1) / kBundle;
    std::vector<int> hl(nb + 1), hu(nb + 1);
    long al = 0, au = 0;
    for (int b = 0; b < nb; ++b) {
      hl[b] = (int)al;
      hu[b] = (int)au;
      int wl = 0, wu = 0;
      for (int i = b * kBundle; i < std::min(n, (b + 1) * kBundle); ++i) {
        wl = std::max(wl, dcol[i] - ro[i]);
        wu = std::max(wu, ro[i + 1] - dcol[i] - 1);
      }
      al += wl;
      au += wu;
      if (std::max(al, au) * kBundle > 0x7fffffffL)
        throw Error(ALIEN_HIP_EINVAL, "ILU0 split factor exceeds int32 indexing");
    }
    hl[nb] = (int)al;
    hu[nb] = (int)au;
    lslots = (size_t)al * kBundle;
    uslots = (size_t)au * kBundle;
    lbo.upload(hl.data(), hl.size(), st);
    ubo.upload(hu.data(), hu.size(), st);
    lbc.resize(lslots);
    ubc.resize(uslots);
  }
  template <typename T>
  ILUSplit<T> fill(hipStream_t st, const HIPMatrix& A, const double* lu, DeviceArray<T>& lv, DeviceArray<T>& uv,
                   DeviceArray<T>& d) {
    const bool cols = lv.empty() && d32.empty() && d64.empty(); // first fill writes the columns
    lv.resize(lslots);
    uv.resize(uslots);
    d.resize(A.getNRows());
    launch_ilu0_split_fill<T>(st, A.getNRows(), A.d_ro.data(), A.d_cols.data(), A.d_dcol.data(), lu, lbo.data(),
                              ubo.data(), cols ? lbc.data() : nullptr, cols ? ubc.data() : nullptr, lv.data(),
                              uv.data(), d.data());
    check_launch("ILU0 split fill");
    return ILUSplit<T>{lbo.data(), lbc.data(), lv.data(), ubo.data(), ubc.data(), uv.data(), d.data()};
  }
};

// t = G^T r, p = G t + beta p in the factor's storage precision
template <typename T>
void factor_p(hipStream_t s, RedWork w, int n, const PrecondSpec& P, const T* r, T* t, T* p, CGState* st,
              int first, int rho1_only = 0) {
  if constexpr (std::is_same<T, float>::value) {
    if (P.half) {
      launch_factor_p<float, bf16_bits>(s, w, n, P.g16, P.gt16, r, t, p, st, first, rho1_only);
      return;
    }
  }
  launch_factor_p<T, T>(s, w, n, P.g<T>(), P.gt<T>(), r, t, p, st, first, rho1_only);
}
template <> RBMat<double> PrecondSpec::g<double>() const { return g64; }
template <> RBMat<double> PrecondSpec::gt<double>() const { return gt64; }
template <> RBMat<float> PrecondSpec::g<float>() const { return g32; }
template <> RBMat<float> PrecondSpec::gt<float>() const { return gt32; }
template <typename T>
void level_ilu_apply(hipStream_t s, const LevelILU& L, const T* y, T* t, T* x, const int* done) {
  if (L.ordered) {
    if constexpr (std::is_same<T, float>::value)
      launch_ilu0_solve_lev<float>(s, L.nl, L.lptr, L.lptr_dev, L.l32, L.nu, L.uptr, L.uptr_dev, L.u32, y, t, x, done,
                                   L.sf_work);
    else
      launch_ilu0_solve_lev<double>(s, L.nl, L.lptr, L.lptr_dev, L.l64, L.nu, L.uptr, L.uptr_dev, L.u64, y, t, x,
                                    done, L.sf_work);
    return;
  }
  const T* lu;
  if constexpr (std::is_same<T, float>::value) lu = L.lu32;
  else lu = L.lu64;
  launch_ilu0_solve<T>(s, L.nl, L.lptr, L.lrows, L.nu, L.uptr, L.urows, L.ro, L.cols, L.dcol, lu, y, t, x, done,
                       L.lptr_dev, L.uptr_dev);
}
template <> ILUSplit<double> PrecondSpec::lu<double>() const { return lu64; }
template <> ILUSplit<float> PrecondSpec::lu<float>() const { return lu32; }

// Progress watch of a solve with the learned preconditioner: stalled once the residual is not
// finite, or its norm has not halved (rr below a quarter) since the last point of progress within
// `window` iterations. CG with an SPD M keeps reducing its error even where its residual norm
// wanders, so only an M that does not help leaves the residual above its earlier level for a whole
// window; a stricter rule (a tenfold drop per 400 iterations) handed healthy solves of strongly
// anisotropic 256^3 systems to the fallback.
struct ProgressWatch {
  int window = 0;
  double best = INFINITY;
  int it_best = 0;
  bool started = false;
  // feed one residual (rr = r.r) at iteration it; true when stalled
  bool stalled(double rr, int it) {
    if (!std::isfinite(rr)) return true;
    if (!started || rr < 0.25 * best) {
      started = true;
      best = rr;
      it_best = it;
      return false;
    }
    return it - it_best >= window;
  }
  // a true residual measured after a stall signal: progress resets the watch
  bool confirms_progress(double rr, int it) {
    if (!(rr < 0.25 * best)) return false;
    best = rr;
    it_best = it;
    return true;
  }
};

template <typename T>
struct FusedCG {
  alien_hip_ctx* ctx;
  int n;
  RBMat<T> A;
  const T* invd;
  PrecondSpec P;
  // r, z: residual and preconditioned residual; p, q: direction and A p;
  // cp: the Chebyshev sweep's own direction vector
  DeviceArray<T> r, z, p, q, y0, y1, cp, tmp, tmp2;
  DeviceArray<CGState> st;
  CGState* h_st = nullptr; // pinned

  FusedCG(alien_hip_ctx* c, int n_, RBMat<T> A_, const T* invd_, const PrecondSpec& P_)
  : ctx(c), n(n_), A(A_), invd(invd_), P(P_) {
    ctx->reserve_for(n);
    r.resize(n); p.resize(n); q.resize(n);
    if (P.kind != kPrecFactor) z.resize(n);
    if (P.kind == kPrecCheb) cp.resize(n);
    if (P.kind == kPrecCheb && P.order > 1) { y0.resize(n); y1.resize(n); }
    if (P.kind == kPrecFactor || P.kind == kPrecILU0 || P.kind == kPrecFILU0 || P.kind == kPrecNeumann ||
        P.kind == kPrecILU0Level)
      tmp.resize(n);
    if (P.kind == kPrecFILU0) tmp2.resize(n);
    if (P.kind == kPrecILU0) ctx->reserve_red(ilu0_split_partials((int)P.cptr.size() - 1, P.cptr.data()));
    st.resize(1);
    static_assert(sizeof(CGState) <= alien_hip_ctx::kPinnedBytes, "state mirror exceeds a pinned block");
    h_st = static_cast<CGState*>(ctx->take_pinned());
    if (!h_st) ALIEN_HIP_CHECK(hipHostMalloc((void**)&h_st, sizeof(CGState), hipHostMallocDefault));
  }
  ~FusedCG() {
    release_graph();
    for (hipEvent_t e : tev) (void)hipEventDestroy(e);
    if (h_st && !ctx->give_pinned(h_st)) (void)hipHostFree(h_st);
  }

  // Step-kernel timing (ctx->step_timing): iteration slot i of an advance launches its q = A p
  // kernel with start and end events tev[2 i], tev[2 i + 1] (hipExtLaunchKernel: the events time
  // the kernel itself); the durations of the slots that ran an iteration are added to the
  // context's totals after the advance's state read.
  std::vector<hipEvent_t> tev;
  void ensure_timing_events(int k) {
    while ((int)tev.size() < 2 * k) {
      hipEvent_t e;
      ALIEN_HIP_CHECK(hipEventCreate(&e));
      tev.push_back(e);
    }
  }
  void timed_cg_q(hipStream_t s, int tslot) {
    if (tslot >= 0)
      launch_cg_q<T>(s, ctx->red(), n, A, p.data(), q.data(), st.data(), tev[2 * tslot], tev[2 * tslot + 1]);
    else
      launch_cg_q<T>(s, ctx->red(), n, A, p.data(), q.data(), st.data());
  }
  void collect_timing(int ran) {
    for (int i = 0; i < ran; ++i) {
      float ms = 0.f;
      ALIEN_HIP_CHECK(hipEventElapsedTime(&ms, tev[2 * i], tev[2 * i + 1]));
      ctx->step_us += 1e3 * (double)ms;
      ++ctx->step_launches;
    }
  }

  // z = M r and rho1 = r.z into the state
  void precond(const T* rin, T* zout) {
    hipStream_t s = ctx->stream;
    if (P.kind == kPrecDiag) {
      launch_diag_apply<T>(s, ctx->red(), n, rin, invd, zout, st.data());
      return;
    }
    if (P.kind == kPrecFactor) {
      launch_factor_apply<T>(s, ctx->red(), n, P.g<T>(), P.gt<T>(), rin, tmp.data(), zout, st.data());
      return;
    }
    if (P.kind == kPrecILU0) {
      // rho1 = r . z from the upper sweeps' partials
      const int np = launch_ilu0_split_apply<T>(s, (int)P.cptr.size() - 1, P.cptr.data(), P.lu<T>(), rin,
                                                tmp.data(), zout, &st.data()->done, rin, ctx->red());
      launch_finalize(s, ctx->red(), np, 1, kFinRho1, nullptr, st.data());
      return;
    }
    if (P.kind == kPrecFILU0) {
      const int np = launch_filu0_split_apply<T>(s, n, P.filu_iters, P.lu<T>(), rin, tmp.data(), tmp2.data(), zout,
                                                 &st.data()->done, rin, ctx->red());
      launch_finalize(s, ctx->red(), np, 1, kFinRho1, nullptr, st.data());
      return;
    }
    if (P.kind == kPrecNeumann) {
      const long np = launch_neumann_apply<T>(s, ctx->red(), n, A, P.order, P.neumann_factor, rin, tmp.data(), zout,
                                              &st.data()->done, rin);
      launch_finalize(s, ctx->red(), np, 1, kFinRho1, nullptr, st.data());
      return;
    }
    if (P.kind == kPrecILU0Level) {
      level_ilu_apply<T>(s, P.lev, rin, tmp.data(), zout, &st.data()->done);
      launch_rz<T>(s, ctx->red(), n, rin, zout, st.data());
      return;
    }
    const int m = P.order;
    T* bufs[2] = {y0.data(), y1.data()};
    const T* yin = rin;
    for (int i = 0; i < m; ++i) {
      T* yout = (i == m - 1) ? zout : bufs[i % 2];
      launch_cheb_step<T>(s, ctx->red(), n, A, rin, yin, yout, cp.data(), invd, P.a[i], P.b[i], i == 0,
                          i == m - 1, st.data());
      yin = yout;
    }
  }
  // One CG update: z = M r, p = z (+ beta p), q = A p, x, r, stop test. With
  // the Diag preconditioner cg_xr already produced z and rho1 for this
  // iteration, except on the first one.
  // (tslot >= 0: timed as step-timing slot tslot)
  void iteration(T* x, bool first, int tslot = -1) {
    hipStream_t s = ctx->stream;
    const bool diag = P.kind == kPrecDiag;
    if (P.kind == kPrecFactor) {
      // z = G (G^T r) never stored: rho1 = |G^T r|^2 and p = G t + beta p; the stop test of this
      // update runs in the next iteration's first finalize, so the update has no reduction
      factor_p<T>(s, ctx->red(), n, P, r.data(), tmp.data(), p.data(), st.data(), first ? 1 : 0);
      timed_cg_q(s, tslot);
      launch_cg_update<T>(s, n, p.data(), q.data(), x, r.data(), st.data());
      return;
    } else {
      if (!diag || first) precond(r.data(), z.data());
      launch_cg_p<T>(s, n, z.data(), p.data(), st.data(), first ? 1 : 0);
    }
    timed_cg_q(s, tslot);
    launch_cg_xr<T>(s, ctx->red(), n, p.data(), q.data(), x, r.data(), diag ? invd : nullptr,
                    diag ? z.data() : nullptr, st.data());
  }
  // After r has been replaced from outside (reliable update) at the end of an
  // iteration: refresh z and rho1 (Jacobi; the learned M recomputes them in its
  // next step).
  void residual_replaced() {
    if (P.kind == kPrecDiag) precond(r.data(), z.data());
  }
  // Learned M only: the solve stopped inside an iteration (the stop test runs in
  // the next step's first kernel, after rho and the count advanced) and r was
  // replaced; finish that iteration with rho1 recomputed from the new r.
  void resume_inside_step(T* x) {
    hipStream_t s = ctx->stream;
    factor_p<T>(s, ctx->red(), n, P, r.data(), tmp.data(), p.data(), st.data(), 0, 1);
    launch_cg_q<T>(s, ctx->red(), n, A, p.data(), q.data(), st.data());
    launch_cg_update<T>(s, n, p.data(), q.data(), x, r.data(), st.data());
    check_launch("learned CG resumed step");
  }
  CGState read_state() {
    ALIEN_HIP_CHECK(hipMemcpyAsync(h_st, st.data(), sizeof(CGState), hipMemcpyDeviceToHost, ctx->stream));
    ALIEN_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return *h_st;
  }

  void write_state(const CGState& h) {
    *h_st = h;
    ALIEN_HIP_CHECK(hipMemcpyAsync(st.data(), h_st, sizeof(CGState), hipMemcpyHostToDevice, ctx->stream));
    ALIEN_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }

  // k loop iterations captured once in a hipGraph (re-captured if x changes)
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  T* g_x = nullptr;
  int g_k = 0;
  void release_graph() {
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    ge = nullptr;
    g = nullptr;
  }
  void launch_graph(T* x, int k) {
    if (!ge || g_x != x || g_k != k) {
      release_graph();
      ALIEN_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      for (int i = 0; i < k; ++i) iteration(x, false);
      ALIEN_HIP_CHECK(hipStreamEndCapture(ctx->stream, &g));
      ALIEN_HIP_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
      g_x = x;
      g_k = k;
    }
    ALIEN_HIP_CHECK(hipGraphLaunch(ge, ctx->stream));
  }
  // k more iterations (fewer if the device stop flag is raised), then the state. With step
  // timing on, the iterations are launched directly (events recorded by a graph's nodes carry
  // no timestamps here): the same kernels in the same order on the same stream.
  CGState advance(T* x, bool graph, int k) {
    const bool timed = ctx->step_timing;
    const int it0 = h_st->it; // the state last read or written
    if (timed) ensure_timing_events(k);
    if (graph && !timed) {
      launch_graph(x, k);
    } else {
      for (int i = 0; i < k; ++i) iteration(x, false, timed ? i : -1);
      check_launch("fused CG iteration");
    }
    const CGState cur = read_state();
    // slots after the stop only returned at entry: count those that ran an iteration
    if (timed) collect_timing(std::max(0, std::min(k, cur.it - it0)));
    return cur;
  }

  // Solve A x = b from the given x (x_is_zero: skip the initial residual SpMV).
  // Returns the final state; bb = (b,b). With stall_window > 0 the solve also
  // stops (*stalled = true) when a ProgressWatch of that window says so.
  CGState solve(const T* b, T* x, bool x_is_zero, double tol, int max_it, bool graph, int poll_every,
                double* bb_out, int stall_window = 0, bool* stalled = nullptr) {
    hipStream_t s = ctx->stream;
    const int slot = ctx->take_slot();
    // Iteration ctor: (b,b) and the criteria
    launch_dot<T>(s, ctx->red(), n, b, b, ctx->dslot(slot));
    const double bb = ctx->read_slot(slot);
    if (bb_out) *bb_out = bb;
    CGState h{};
    h.crit = tol * tol * bb;
    h.max_it = max_it;
    h.rr = h.crit + 1;
    h.status = (bb == 0) ? 1 : 0;
    if (bb == 0) return h; // nullRhs
    write_state(h);
    if (x_is_zero) launch_copy<T>(s, n, b, r.data());
    else launch_residual<T>(s, ctx->red(), n, A, x, b, r.data(), ctx->dslot(slot));
    iteration(x, true);
    check_launch("fused CG first iteration");
    CGState cur = read_state();
    // without a graph, poll after every iteration (the reference's host-side stop test)
    const int k = graph ? std::max(1, poll_every) : 1;
    ProgressWatch watch{stall_window};
    while (!cur.done) {
      cur = advance(x, graph, k);
      if (stall_window <= 0) continue;
      if (!std::isfinite(cur.rr) || (watch.stalled(cur.rr, cur.it) && !cur.done)) {
        *stalled = true;
        break;
      }
    }
    return cur;
  }
};

PrecondSpec cheb_spec(const Alien::ChebyshevPreconditioner<Alg>& C) {
  PrecondSpec ps;
  ps.kind = kPrecCheb;
  ps.order = C.order();
  ps.a.resize(ps.order);
  ps.b.resize(ps.order);
  C.coefficients(ps.a.data(), ps.b.data());
  return ps;
}

// CG breakdown (p.q == 0): Iteration::stop(r) decides between a clean exit
// with one more iteration counted and NullValueException (CG.h:96-104).
template <typename T>
void finish_breakdown(FusedCG<T>& F, CGState& s, int* error) {
  if (!s.breakdown) return;
  if (s.it >= s.max_it) {
    s.it += 1;
    return;
  }
  const int slot = F.ctx->take_slot();
  launch_dot<T>(F.ctx->stream, F.ctx->red(), F.n, F.r.data(), F.r.data(), F.ctx->dslot(slot));
  const double rr = F.ctx->read_slot(slot);
  s.rr = rr;
  s.status = rr < s.crit;
  if (s.status) s.it += 1;
  else *error = 1;
}

// Mixed-precision CG with reliable updates (Sleijpen & van der Vorst, "Reliable
// updated residuals in hybrid Bi-CG methods", Computing 56 (1996); the
// mixed-precision form of Clark et al., "Solving lattice QCD systems of
// equations using mixed precision solvers on GPUs", CPC 181 (2010)).
// The CG recurrence runs in FP32 (matrix, preconditioner, vectors; dots and the
// step coefficients in FP64) on the increment x32 of the FP64 solution. Each
// time the recursive residual has dropped by delta since the last update (or
// the FP32 stop test fires), the increment is folded into x64, the true
// residual b - A x64 is recomputed in FP64 and replaces the FP32 residual, and
// CG continues with its search direction: no restart, so the Krylov space
// built so far is kept. Convergence is judged on the FP64 true residual only.
// With stall_window > 0 the solve gives up (stalled = true, x = the last
// reliable update's solution) once the FP32 residual is not finite, or a
// ProgressWatch of that window signals and the true residual then shows no
// progress either.
struct MixedResult {
  int it = 0, updates = 0, status = 0, error = 0;
  bool stalled = false;      // learned M: no progress within the stall window (ILU0 takes over)
  bool fp32_stalled = false; // the FP32 inner iteration stopped making progress (FP64 takes over)
  double rel = 0;
};

// Mixed BiCGStab's FP32 stagnation rule: the FP64 true residual, checked at
// each reliable update, must fall to a quarter of its best value within
// max(kMinGroup, kGroupFactor x the longest such interval so far) iterations.
// FP32 storage of A carries a relative error of ~6e-8 per entry, and when that
// times the condition number nears 1 the inner iteration stops reducing the
// true residual: FP64 then finishes the solve.
constexpr int kMinGroup = 250;
constexpr int kGroupFactor = 3;
// squared-norm thresholds of mixed BiCGStab's restart rule
constexpr double kRestartGap = 2.25;
constexpr double kRestartPeak = 100.0;
// A true residual above kRiseFactor x the lowest so far (squared norms: 100x in
// norm) at a reliable update means the FP32 recurrence diverged (its rounding
// at a residual peak broke the Bi-CG relations): the iteration resumes from the
// best iterate, at most kMaxRises times before FP64 takes over from it (with the
// learned M, which may not suit a nonsymmetric operator, ILU0 takes over at once).
constexpr double kRiseFactor = 1e4;
constexpr int kMaxRises = 1;

MixedResult mixed_reliable_cg(alien_hip_ctx* ctx, int n, RBMat<double> A64, RBMat<float> A32, const float* invd32,
                              const PrecondSpec& ps, const double* b, double* x, double tol, int max_it,
                              double delta, int poll_every, bool graph, int stall_window = 0) {
  MixedResult out;
  hipStream_t s = ctx->stream;
  const int slot = ctx->take_slot();
  launch_dot<double>(s, ctx->red(), n, b, b, ctx->dslot(slot));
  const double bb = ctx->read_slot(slot);
  if (bb == 0) {
    out.status = 1;
    return out;
  }
  const double crit = tol * tol * bb;
  DeviceArray<double> r64(n);
  DeviceArray<float> x32(n);
  ALIEN_HIP_CHECK(hipMemsetAsync(x32.data(), 0, sizeof(float) * n, s));
  FusedCG<float> F(ctx, n, A32, invd32, ps);
  const int rslot = ctx->take_slot();
  launch_residual<double>(s, ctx->red(), n, A64, x, b, r64.data(), ctx->dslot(rslot));
  double rr64 = ctx->read_slot(rslot);
  out.rel = std::sqrt(rr64 / bb);
  if (rr64 < crit) {
    out.status = 1;
    return out;
  }
  launch_demote(s, n, r64.data(), F.r.data());
  CGState h{};
  h.crit = crit;
  h.max_it = max_it;
  h.rr = rr64;
  F.write_state(h);
  F.iteration(x32.data(), true);
  check_launch("mixed CG first iteration");
  CGState cur = F.read_state();
  double rr_last = rr64;
  const double d2 = delta * delta;
  const int k = graph ? std::max(1, poll_every) : 1;
  ProgressWatch watch{stall_window};
  watch.stalled(rr64, 0);
  for (;;) {
    if (!std::isfinite(cur.rr)) {
      // the increment is lost; x holds the last reliable update
      if (stall_window > 0) out.stalled = true;
      else out.fp32_stalled = true;
      break;
    }
    // a stall signal forces a reliable update: the true residual decides
    const bool stall = stall_window > 0 && watch.stalled(cur.rr, cur.it);
    const bool update = cur.done || cur.rr < d2 * rr_last || stall;
    if (!update) {
      cur = F.advance(x32.data(), graph, k);
      continue;
    }
    // reliable update: x64 += x32, x32 = 0, r64 = b - A x64 in FP64
    launch_fold_increment(s, n, x32.data(), x);
    launch_residual<double>(s, ctx->red(), n, A64, x, b, r64.data(), ctx->dslot(rslot));
    rr64 = ctx->read_slot(rslot);
    ++out.updates;
    out.rel = std::sqrt(rr64 / bb);
    if (rr64 < crit) {
      out.status = 1;
      break;
    }
    if (cur.breakdown) {
      out.error = 1;
      break;
    }
    if (cur.it >= max_it || !(rr64 == rr64)) break;
    if (stall && !watch.confirms_progress(rr64, cur.it)) {
      out.stalled = true;
      break;
    }
    launch_demote(s, n, r64.data(), F.r.data());
    rr_last = rr64;
    const bool inside = cur.done && ps.kind == kPrecFactor; // FP32 stop test fired at a step's start
    cur.done = 0;
    cur.status = 0;
    cur.rr = rr64;
    F.write_state(cur);
    if (inside) F.resume_inside_step(x32.data());
    else F.residual_replaced();
    cur = F.advance(x32.data(), graph, k);
  }
  out.it = cur.it;
  return out;
}

// Fused BiCGStab (kernels/bicgstab.hip) with the Diag or the learned
// preconditioner; the reference's call sequence, with the step scalars and
// the stop/breakdown decisions on the device.
template <typename T>
struct FusedBiCG {
  alien_hip_ctx* ctx;
  int n;
  RBMat<T> A;
  const T* invd;
  PrecondSpec P;
  DeviceArray<T> r, r0, p, phat, v, sv, shat, t, tmp, tmp2, cp;
  DeviceArray<BiCGState> st;
  BiCGState* h_st = nullptr; // pinned

  FusedBiCG(alien_hip_ctx* c, int n_, RBMat<T> A_, const T* invd_, const PrecondSpec& P_)
  : ctx(c), n(n_), A(A_), invd(invd_), P(P_) {
    ctx->reserve_for(n);
    for (auto* a : {&r, &r0, &p, &phat, &v, &sv, &shat, &t}) a->resize(n);
    if (P.kind == kPrecFactor || P.kind == kPrecILU0 || P.kind == kPrecFILU0 || P.kind == kPrecNeumann ||
        P.kind == kPrecCheb || P.kind == kPrecILU0Level)
      tmp.resize(n);
    if (P.kind == kPrecFILU0 || P.kind == kPrecCheb) tmp2.resize(n);
    if (P.kind == kPrecCheb) cp.resize(n);
    st.resize(1);
    static_assert(sizeof(BiCGState) <= alien_hip_ctx::kPinnedBytes, "state mirror exceeds a pinned block");
    h_st = static_cast<BiCGState*>(ctx->take_pinned());
    if (!h_st) ALIEN_HIP_CHECK(hipHostMalloc((void**)&h_st, sizeof(BiCGState), hipHostMallocDefault));
  }
  ~FusedBiCG() {
    release_graph();
    if (h_st && !ctx->give_pinned(h_st)) (void)hipHostFree(h_st);
  }
  const int* done_flag() { return &st.data()->done; }
  void precond(const T* in, T* out) {
    if (P.kind == kPrecILU0)
      launch_ilu0_split_apply<T>(ctx->stream, (int)P.cptr.size() - 1, P.cptr.data(), P.lu<T>(), in, tmp.data(),
                                 out, done_flag(), nullptr, ctx->red());
    else if (P.kind == kPrecFILU0)
      launch_filu0_split_apply<T>(ctx->stream, n, P.filu_iters, P.lu<T>(), in, tmp.data(), tmp2.data(), out,
                                  done_flag(), nullptr, ctx->red());
    else if (P.kind == kPrecNeumann)
      launch_neumann_apply<T>(ctx->stream, ctx->red(), n, A, P.order, P.neumann_factor, in, tmp.data(), out,
                              done_flag(), nullptr);
    else if (P.kind == kPrecILU0Level)
      level_ilu_apply<T>(ctx->stream, P.lev, in, tmp.data(), out, done_flag());
    else if (P.kind == kPrecCheb) {
      // the steps of FusedCG::precond, stop flag from this state, no rho1
      T* bufs[2] = {tmp.data(), tmp2.data()};
      const T* yin = in;
      for (int i = 0; i < P.order; ++i) {
        T* yout = (i == P.order - 1) ? out : bufs[i % 2];
        launch_cheb_step<T>(ctx->stream, ctx->red(), n, A, in, yin, yout, cp.data(), invd, P.a[i], P.b[i], i == 0,
                            i == P.order - 1, nullptr, done_flag());
        yin = yout;
      }
    } else
      launch_factor_apply_flag<T>(ctx->stream, n, P.g<T>(), P.gt<T>(), in, tmp.data(), out, done_flag());
  }
  // one step of the while loop (or the first block, per the state's `first`)
  void step(T* x) {
    hipStream_t s = ctx->stream;
    const bool diag = P.kind == kPrecDiag;
    const T* dinv = diag ? invd : nullptr;
    launch_bicg_p<T>(s, n, r.data(), v.data(), p.data(), dinv, phat.data(), st.data());
    if (!diag) precond(p.data(), phat.data());
    launch_bicg_v<T>(s, ctx->red(), n, A, phat.data(), r0.data(), v.data(), st.data());
    launch_bicg_s<T>(s, ctx->red(), n, r.data(), v.data(), sv.data(), dinv, shat.data(), st.data());
    // level-set ILU0 with one launch per level (never captured in a graph): a stop on s skips the
    // second apply's hundreds of per-level launches, which would otherwise each return at once
    if (P.kind == kPrecILU0Level && !level_sync_free(P.lev) && read_state().done) return;
    if (!diag) precond(sv.data(), shat.data());
    launch_bicg_t<T>(s, ctx->red(), n, A, shat.data(), sv.data(), t.data(), st.data());
    launch_bicg_xr<T>(s, ctx->red(), n, shat.data(), phat.data(), sv.data(), t.data(), r0.data(), x, r.data(),
                      st.data());
  }
  BiCGState read_state() {
    ALIEN_HIP_CHECK(hipMemcpyAsync(h_st, st.data(), sizeof(BiCGState), hipMemcpyDeviceToHost, ctx->stream));
    ALIEN_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return *h_st;
  }
  void write_state(const BiCGState& h) {
    *h_st = h;
    ALIEN_HIP_CHECK(hipMemcpyAsync(st.data(), h_st, sizeof(BiCGState), hipMemcpyHostToDevice, ctx->stream));
    ALIEN_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  T* g_x = nullptr;
  int g_k = 0;
  void release_graph() {
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    ge = nullptr;
    g = nullptr;
  }
  BiCGState advance(T* x, bool graph, int k) {
    if (graph) {
      if (!ge || g_x != x || g_k != k) {
        release_graph();
        ALIEN_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < k; ++i) step(x);
        ALIEN_HIP_CHECK(hipStreamEndCapture(ctx->stream, &g));
        ALIEN_HIP_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        g_x = x;
        g_k = k;
      }
      ALIEN_HIP_CHECK(hipGraphLaunch(ge, ctx->stream));
    } else {
      for (int i = 0; i < k; ++i) step(x);
      check_launch("fused BiCGStab step");
    }
    return read_state();
  }
  // (re)start from r (already in place): r0 = r, rho1 = r.r0, first block next
  void start(double rr, double crit, int it, int max_it) {
    launch_copy<T>(ctx->stream, n, r.data(), r0.data());
    BiCGState h{};
    h.rho1 = rr;
    h.rr = rr;
    h.rr0 = rr;
    h.rrmax = rr;
    h.crit = crit;
    h.it = it;
    h.max_it = max_it;
    h.first = 1;
    write_state(h);
  }
  // Solve from the given x; stall_window > 0 stops (*stalled) on a non-finite
  // residual or no tenfold drop within the window, as FusedCG::solve.
  BiCGState solve(const T* b, T* x, double tol, int max_it, bool graph, int poll_every, double* bb_out,
                  int stall_window = 0, bool* stalled = nullptr) {
    hipStream_t s = ctx->stream;
    const int slot = ctx->take_slot();
    launch_dot<T>(s, ctx->red(), n, b, b, ctx->dslot(slot));
    const double bb = ctx->read_slot(slot);
    if (bb_out) *bb_out = bb;
    BiCGState h{};
    if (bb == 0) { // Iteration::nullRhs
      h.status = 1;
      return h;
    }
    launch_residual<T>(s, ctx->red(), n, A, x, b, r.data(), ctx->dslot(slot));
    const double rr = ctx->read_slot(slot);
    start(rr, tol * tol * bb, 0, max_it);
    const int k = graph ? std::max(1, poll_every) : 1;
    BiCGState cur = advance(x, graph, 1);
    ProgressWatch watch{stall_window};
    watch.stalled(rr, 0);
    while (!cur.done) {
      cur = advance(x, graph, k);
      if (stall_window <= 0) continue;
      if (!std::isfinite(cur.rr) || (watch.stalled(cur.rr, cur.it) && !cur.done)) {
        *stalled = true;
        break;
      }
    }
    launch_bicg_half<T>(s, n, phat.data(), x, st.data());
    return cur;
  }
};

// Mixed-precision BiCGStab with reliable updates (Sleijpen & van der Vorst 1996,
// the reference for the scheme): the FP32 recurrence updates an increment
// x32; when the recursive residual has dropped by delta below the largest
// residual since the last update (their M(r) rule: a peak of the erratic
// BiCGStab residual leaves an FP32 rounding error of ~6e-8 x the peak in the
// recursive residual, so it must be replaced soon after the peak has passed),
// or the FP32 solve stops, x32 is folded into x64 and r = b - A x64 (FP64)
// replaces the FP32 residual. The shadow residual r0 and the search direction
// are kept (rho1 = r.r0 recomputed) unless the replaced residual differs from
// the recursive one by more than kRestartGap, or the group's peak exceeded
// kRestartPeak x the last true residual (squared norms): the Bi-CG relations
// are then too far off and the recurrence restarts from the true residual
// (r0 = r, first block), as it does after a stop on s, a breakdown or an FP32
// stop that the true residual does not confirm. Convergence is judged on the
// FP64 true residual only.
MixedResult mixed_reliable_bicg(alien_hip_ctx* ctx, int n, RBMat<double> A64, RBMat<float> A32,
                                const float* invd32, const PrecondSpec& ps, const double* b, double* x,
                                double tol, int max_it, double delta, int poll_every, bool graph,
                                int stall_window = 0) {
  MixedResult out;
  hipStream_t s = ctx->stream;
  const int slot = ctx->take_slot();
  launch_dot<double>(s, ctx->red(), n, b, b, ctx->dslot(slot));
  const double bb = ctx->read_slot(slot);
  if (bb == 0) {
    out.status = 1;
    return out;
  }
  const double crit = tol * tol * bb;
  DeviceArray<double> r64(n);
  DeviceArray<float> x32(n);
  ALIEN_HIP_CHECK(hipMemsetAsync(x32.data(), 0, sizeof(float) * n, s));
  FusedBiCG<float> F(ctx, n, A32, invd32, ps);
  const int rslot = ctx->take_slot();
  auto true_res = [&]() {
    launch_residual<double>(s, ctx->red(), n, A64, x, b, r64.data(), ctx->dslot(rslot));
    return ctx->read_slot(rslot);
  };
  double rr64 = true_res();
  out.rel = std::sqrt(rr64 / bb);
  if (rr64 < crit) {
    out.status = 1;
    return out;
  }
  auto restart = [&](int it) {
    launch_demote(s, n, r64.data(), F.r.data());
    const int sl = ctx->take_slot();
    launch_dot<float>(s, ctx->red(), n, F.r.data(), F.r.data(), ctx->dslot(sl));
    F.start(ctx->read_slot(sl), crit, it, max_it);
  };
  restart(0);
  BiCGState cur = F.advance(x32.data(), graph, 1);
  double rr_last = rr64;
  int restarts = 0, it_best = 0, gap_max = 0, rises = 0;
  double rr_best = rr64;
  // the iterate of the lowest true residual so far: hand-overs and divergence restarts resume from it
  double rr_min = rr64;
  DeviceArray<double> x_min(n);
  launch_copy<double>(s, n, x, x_min.data());
  auto resume_min = [&]() {
    launch_copy<double>(s, n, x_min.data(), x);
    rr64 = true_res();
    out.rel = std::sqrt(rr64 / bb);
  };
  const double d2 = delta * delta;
  const int k = graph ? std::max(1, poll_every) : 1;
  ProgressWatch watch{stall_window};
  watch.stalled(rr64, 0);
  for (;;) {
    if (!std::isfinite(cur.rr) || !std::isfinite(cur.alpha) || !std::isfinite(cur.omega)) {
      // the FP32 recursion overflowed: its increment is lost
      resume_min();
      if (stall_window > 0) out.stalled = true;
      else out.fp32_stalled = true;
      break;
    }
    if (stall_window == 0 && !cur.done && cur.it - it_best >= std::max(kMinGroup, kGroupFactor * gap_max) &&
        cur.it < max_it) {
      // stagnated: keep the increment only if it lowers the true residual below the best
      launch_fold_increment(s, n, x32.data(), x);
      if (true_res() < rr_min) ++out.updates;
      else resume_min();
      out.fp32_stalled = true;
      break;
    }
    const bool stall = stall_window > 0 && watch.stalled(cur.rr, cur.it);
    const bool update = cur.done || cur.rr < d2 * std::max(rr_last, cur.rrmax) || stall;
    if (!update) {
      cur = F.advance(x32.data(), graph, k);
      continue;
    }
    launch_bicg_half<float>(s, n, F.phat.data(), x32.data(), F.st.data());
    launch_fold_increment(s, n, x32.data(), x);
    rr64 = true_res();
    ++out.updates;
    out.rel = std::sqrt(rr64 / bb);
    if (rr64 < crit) {
      out.status = 1;
      break;
    }
    if (cur.it >= max_it || !(rr64 == rr64)) break;
    if (stall && !watch.confirms_progress(rr64, cur.it)) {
      if (rr64 > rr_min) resume_min();
      out.stalled = true;
      break;
    }
    if (rr64 < rr_min) {
      rr_min = rr64;
      launch_copy<double>(s, n, x, x_min.data());
    } else if (rr64 > kRiseFactor * rr_min) {
      // the FP32 recurrence diverged from its best iterate: resume from that iterate, or hand over
      // from it (to FP64 after kMaxRises such restarts; the learned M's to ILU0 at once)
      if (ctx->trace)
        std::fprintf(stderr, "[alien_hip] mixed BiCGStab it %d: true %.3e, %.3e at the best iterate: %s\n", cur.it,
                     std::sqrt(rr64 / bb), std::sqrt(rr_min / bb),
                     stall_window > 0 ? "ILU0" : rises < kMaxRises ? "resume" : "FP64");
      resume_min();
      if (stall_window > 0) {
        out.stalled = true;
        break;
      }
      if (++rises > kMaxRises) {
        out.fp32_stalled = true;
        break;
      }
      rr_last = rr64;
      restart(cur.it);
      cur = F.advance(x32.data(), graph, k);
      continue;
    }
    const bool far = rr64 > kRestartGap * cur.rr || cur.rrmax > kRestartPeak * rr_last;
    if (ctx->trace)
      std::fprintf(stderr, "[alien_hip] mixed BiCGStab it %d: true %.3e recursive %.3e peak %.3e%s\n", cur.it,
                   std::sqrt(rr64 / bb), std::sqrt(cur.rr / bb), std::sqrt(cur.rrmax / bb),
                   cur.done || far ? " restart" : "");
    rr_last = rr64;
    if (rr64 < 0.25 * rr_best) {
      gap_max = std::max(gap_max, cur.it - it_best);
      rr_best = rr64;
      it_best = cur.it;
    }
    if (cur.done || far) { // stop on s, FP32 stop not confirmed, breakdown, or Bi-CG relations lost
      if (cur.breakdown && ++restarts > 10) {
        out.error = 1;
        break;
      }
      restart(cur.it);
      cur = F.advance(x32.data(), graph, k);
      continue;
    }
    // reliable update: r <- b - A x64, rho1 <- r.r0
    launch_demote(s, n, r64.data(), F.r.data());
    const int sl = ctx->take_slot();
    launch_dot<float>(s, ctx->red(), n, F.r.data(), F.r0.data(), ctx->dslot(sl));
    cur.rho1 = ctx->read_slot(sl);
    cur.rr = rr64;
    cur.rrmax = rr64;
    F.write_state(cur);
    cur = F.advance(x32.data(), graph, k);
  }
  out.it = cur.it;
  return out;
}

double true_residual(Alg& alg, const HIPMatrix& A, const HIPVector& b, const HIPVector& x) {
  // CSR SpMV (bit-exact with _seqMult), FP64
  HIPVector r;
  r.init(A.resource(), true);
  alien_hip::launch_spmv_csr<double>(alg.stream(), A.getNRows(), A.d_ro.data(), A.d_cols.data(), A.d_vals.data(),
                                     x.data(), r.data());
  alg.scal(-1., r);
  alg.axpy(1., b, r);
  const double rr = alg.dot(r, r), bb = alg.dot(b, b);
  return bb > 0 ? std::sqrt(rr / bb) : 0.0;
}

} // namespace

} // namespace alien_hip

namespace alien_hip {
namespace {

// Parameter checks shared by the solve entry points.
void validate_params(const alien_hip_solve_params* p) {
  if (p->mode < ALIEN_HIP_MODE_REFERENCE || p->mode > ALIEN_HIP_MODE_REFERENCE_FUTURE)
    throw Error(ALIEN_HIP_EINVAL, "solve: unknown mode");
  if (p->solver != ALIEN_HIP_CG && p->solver != ALIEN_HIP_BICGS) throw Error(ALIEN_HIP_EINVAL, "solve: unknown solver");
  const bool fused = p->mode == ALIEN_HIP_MODE_FUSED || p->mode == ALIEN_HIP_MODE_GRAPH;
  const bool mixed = p->mode == ALIEN_HIP_MODE_MIXED;
  const bool learned = p->precond == ALIEN_HIP_PREC_LEARNED;
  const bool bicg = p->solver == ALIEN_HIP_BICGS;
  const bool ilu = p->precond == ALIEN_HIP_PREC_ILU0;
  const bool filu = p->precond == ALIEN_HIP_PREC_FILU0;
  const bool neumann = p->precond == ALIEN_HIP_PREC_NEUMANN;
  if ((fused || mixed) && !bicg && p->precond != ALIEN_HIP_PREC_CHEB && p->precond != ALIEN_HIP_PREC_DIAG &&
      !ilu && !filu && !neumann && !learned)
    throw Error(ALIEN_HIP_EINVAL, "fused/graph/mixed modes implement CG with the Diag, Chebyshev, ILU0, FILU0, "
                                  "Neumann or learned preconditioner");
  if ((fused || mixed) && bicg && p->precond != ALIEN_HIP_PREC_DIAG && p->precond != ALIEN_HIP_PREC_CHEB && !ilu &&
      !filu && !neumann && !learned)
    throw Error(ALIEN_HIP_EINVAL, "fused/graph/mixed modes implement BiCGStab with the Diag, Chebyshev, ILU0, "
                                  "FILU0, Neumann or learned preconditioner");
  if ((fused || mixed) && neumann && p->poly_order < 1)
    throw Error(ALIEN_HIP_EINVAL, "fused/graph/mixed Neumann needs poly_order >= 1");
  if ((fused || mixed) && filu && p->filu_solver_niter < 1)
    throw Error(ALIEN_HIP_EINVAL, "fused/graph/mixed FILU0 needs filu_solver_niter >= 1");
  if (ilu && p->ilu_schedule != 0 && p->ilu_schedule != 1)
    throw Error(ALIEN_HIP_EINVAL, "ILU0: ilu_schedule must be 0 (level sets) or 1 (multicolour)");
  if (learned && !(fused || mixed))
    throw Error(ALIEN_HIP_EINVAL, "the learned preconditioner runs in the fused, graph or mixed modes");
  if (learned && !p->gnn_theta) throw Error(ALIEN_HIP_EINVAL, "learned preconditioner: gnn_theta is NULL");
  if (p->stall_window < 0) throw Error(ALIEN_HIP_EINVAL, "solve_csr: stall_window < 0");
  if (p->precond == ALIEN_HIP_PREC_FILU0 && (p->filu_factor_niter < 0 || p->filu_solver_niter < 0))
    throw Error(ALIEN_HIP_EINVAL, "FILU0: negative iteration counts");
  if (p->factor_storage != 0 && p->factor_storage != 1)
    throw Error(ALIEN_HIP_EINVAL, "solve_csr: factor_storage must be 0 (FP32) or 1 (bfloat16)");
  if (p->precond == ALIEN_HIP_PREC_CHEB && p->poly_order < 1)
    throw Error(ALIEN_HIP_EINVAL, "Chebyshev order must be >= 1");
  if (p->max_iter < 0 || !(p->tol >= 0)) throw Error(ALIEN_HIP_EINVAL, "solve: bad tol / max_iter");
}

// ALIEN_HIP_TRACE: wall time of each set-up phase on stderr (the stream drained at each mark)
struct SetupTrace {
  alien_hip_ctx* ctx;
  std::chrono::steady_clock::time_point last = std::chrono::steady_clock::now();
  explicit SetupTrace(alien_hip_ctx* c) : ctx(c) {}
  void operator()(const char* what) {
    if (!ctx->trace) return;
    ALIEN_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "alien_hip set-up: %-28s %8.2f ms\n", what,
                 std::chrono::duration<double, std::milli>(now - last).count());
    last = now;
  }
};

void solve_on(alien_hip_ctx* ctx, const alien_hip_solve_params* p, HIPMatrix& A, const double* b_in, double* x_io,
              alien_hip_solve_result* res, EventTimer& tm, SetupTrace& phase, const std::vector<int>* colours,
              std::unique_ptr<gnn::Engine>* kept_engine);

// The learned M's fallback: multicolour ILU0 (the fused ILU0 of the colour-permuted matrix P A P^T,
// two launches per sweep on a 7-point operator) from the iterate x, for at most max_iter
// iterations, in FP64 (FUSED / GRAPH) or mixed. Colouring and permutation on the host from A's
// host CSR; x (device) receives the result. The reference's natural-order ILU0 runs one launch per
// level (766 per sweep at 256^3, ~7 ms per CG iteration) and made a fallback slower than the solve
// it rescued (a 192^3 two-unknown BiCGStab: 56 s).
alien_hip_solve_result ilu0_multicolour_from(alien_hip_ctx* ctx, const alien_hip_solve_params* p, HIPMatrix& A,
                                             const double* b_dev, double* x_dev, int max_iter, bool mixed) {
  const int n = A.getNRows();
  std::vector<int> cptr, pro;
  HostBuf<int> pcols;
  HostBuf<double> pvals;
  const std::vector<int> perm = colour_permutation(n, A.kcol(), A.cols(), cptr);
  permute_csr(n, A.kcol(), A.cols(), A.hostValues(), perm, pro, pcols, pvals);
  std::vector<double> bh(n), xh(n), pb(n), px(n);
  ALIEN_HIP_CHECK(hipMemcpyAsync(bh.data(), b_dev, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  ALIEN_HIP_CHECK(hipMemcpyAsync(xh.data(), x_dev, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  ALIEN_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  for (int r = 0; r < n; ++r) {
    pb[r] = bh[perm[r]];
    px[r] = xh[perm[r]];
  }
  alien_hip_solve_result r{};
  {
    Alien::CSRHostView view{n, pro.data(), pcols.data(), pvals.data()};
    HIPMatrix PA(ctx, view, ALIEN_HIP_LAYOUT_CSR, /*borrow_host=*/true);
    PA.ensureFused(p->layout);
    alien_hip_solve_params q = *p;
    q.precond = ALIEN_HIP_PREC_ILU0;
    q.ilu_schedule = 1;
    q.max_iter = max_iter;
    q.gnn_theta = nullptr;
    q.stall_window = 0;
    q.mode = mixed ? ALIEN_HIP_MODE_MIXED
                   : (p->mode == ALIEN_HIP_MODE_MIXED ? ALIEN_HIP_MODE_GRAPH : p->mode);
    EventTimer tm2(ctx->stream);
    tm2.start();
    SetupTrace phase2(ctx);
    solve_on(ctx, &q, PA, pb.data(), px.data(), &r, tm2, phase2, &cptr, nullptr);
  }
  for (int k = 0; k < n; ++k) xh[perm[k]] = px[k];
  ALIEN_HIP_CHECK(hipMemcpyAsync(x_dev, xh.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  ALIEN_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return r;
}

// The solve proper on a converted matrix: preconditioner set-up, Krylov loop,
// true residual, solution copied to x_io. tm was started at the beginning of
// the set-up; phase marks the set-up trace. kept_engine (a kept matrix's): the
// learned preconditioner's network engine, reused by later solves on the same
// structure (its factor layouts are built once) and created there when empty.
void solve_on(alien_hip_ctx* ctx, const alien_hip_solve_params* p, HIPMatrix& A, const double* b_in, double* x_io,
              alien_hip_solve_result* res, EventTimer& tm, SetupTrace& phase,
              const std::vector<int>* colours, std::unique_ptr<gnn::Engine>* kept_engine = nullptr) {
  const int n = A.getNRows();
  const bool fused = p->mode == ALIEN_HIP_MODE_FUSED || p->mode == ALIEN_HIP_MODE_GRAPH;
  const bool mixed = p->mode == ALIEN_HIP_MODE_MIXED;
  const bool learned = p->precond == ALIEN_HIP_PREC_LEARNED;
  const bool bicg = p->solver == ALIEN_HIP_BICGS;
  Alg alg(ctx);
  HIPVector b, x;
  alg.allocate(A.resource(), b, x);
  b.setValues(b_in);
  x.setValues(x_io);

  if (!fused && !mixed) {
    switch (p->precond) {
    case ALIEN_HIP_PREC_DIAG: {
      Alien::DiagPreconditioner<Alg> P(alg, A);
      P.init();
      res->setup_s = tm.stop();
      run_reference(alg, p, P, A, b, x, res, tm);
      break;
    }
    case ALIEN_HIP_PREC_CHEB: {
      Alien::ChebyshevPreconditioner<Alg> P(alg, A, p->poly_factor, p->poly_order, p->poly_max_iter);
      P.init();
      res->cheb_alpha = P.alpha();
      res->cheb_beta = P.beta();
      res->setup_s = tm.stop();
      run_reference(alg, p, P, A, b, x, res, tm);
      break;
    }
    case ALIEN_HIP_PREC_ILU0: {
      Alien::ILU0Preconditioner<Alg> P(alg, A);
      if (colours && !colours->empty()) P.setColourSchedule(*colours);
      P.init();
      res->setup_s = tm.stop();
      run_reference(alg, p, P, A, b, x, res, tm);
      break;
    }
    case ALIEN_HIP_PREC_NEUMANN: {
      Alien::NeumannPolyPreconditioner<Alg> P(alg, A, p->poly_factor, p->poly_order, p->poly_max_iter);
      P.init();
      res->setup_s = tm.stop();
      run_reference(alg, p, P, A, b, x, res, tm);
      break;
    }
    case ALIEN_HIP_PREC_FILU0: {
      Alien::FILU0Preconditioner<Alg> P(alg, A);
      P.setParameter("nb-factor-iter", p->filu_factor_niter);
      P.setParameter("nb-solver-iter", p->filu_solver_niter);
      P.init();
      res->setup_s = tm.stop();
      run_reference(alg, p, P, A, b, x, res, tm);
      break;
    }
    default: {
      IdentityPreconditioner<Alg> P;
      res->setup_s = tm.stop();
      run_reference(alg, p, P, A, b, x, res, tm);
    }
    }
  } else {
    // preconditioner set-up shared by the fused paths (FP64, as the reference)
    std::unique_ptr<Alien::ChebyshevPreconditioner<Alg>> cheb;
    std::unique_ptr<Alien::DiagPreconditioner<Alg>> diag;
    PrecondSpec ps;
    const double* invd64;
    if (p->precond == ALIEN_HIP_PREC_CHEB) {
      cheb.reset(new Alien::ChebyshevPreconditioner<Alg>(alg, A, p->poly_factor, p->poly_order, p->poly_max_iter));
      cheb->init();
      res->cheb_alpha = cheb->alpha();
      res->cheb_beta = cheb->beta();
      ps = cheb_spec(*cheb);
      invd64 = cheb->invDiag().data();
    } else {
      // Diag, and the learned preconditioner's fallback
      diag.reset(new Alien::DiagPreconditioner<Alg>(alg, A));
      diag->init();
      invd64 = diag->invDiag().data();
    }
    // learned preconditioner: network inference on A (read in place), then G
    // and G^T row-bundled in the solver's precision
    std::unique_ptr<gnn::Engine> net_own; // the engine of a one-shot solve
    DeviceArray<float> theta, gv32, gtv32;
    DeviceArray<double> gv64, gtv64;
    DeviceArray<bf16_bits> gv16, gtv16;
    // multicolour ILU(0): factorised in FP64 on the level-set schedule of the
    // colour-permuted matrix (the reference's factor), then split into its
    // strictly-lower and strictly-upper row-bundled parts
    std::unique_ptr<Alien::ILU0Preconditioner<Alg>> ilu;
    ILUSplitStore lus;
    auto ilu_fp64 = [&](PrecondSpec& q) {
      q.lu64 = lus.fill<double>(ctx->stream, A, ilu->luValues(), lus.lbv64, lus.ubv64, lus.d64);
    };
    DeviceArray<float> lev_lu32, lev_l32, lev_u32, lev_d32;
    if (p->precond == ALIEN_HIP_PREC_ILU0 && !(colours && !colours->empty())) {
      // natural order: the level-set factor and sweeps of the REFERENCE mode
      ilu.reset(new Alien::ILU0Preconditioner<Alg>(alg, A));
      ilu->init();
      phase("ILU0 level-set factorisation");
      ps.kind = kPrecILU0Level;
      LevelILU& L = ps.lev;
      L.nl = ilu->lowerLevels();
      L.nu = ilu->upperLevels();
      L.lptr = ilu->lowerPtr();
      L.uptr = ilu->upperPtr();
      L.lrows = ilu->lowerRows();
      L.urows = ilu->upperRows();
      L.lptr_dev = ilu->lowerPtrDevice();
      L.uptr_dev = ilu->upperPtrDevice();
      L.ro = A.d_ro.data();
      L.cols = A.d_cols.data();
      L.dcol = A.d_dcol.data();
      L.lu64 = ilu->luValues();
      L.ordered = ilu->levelOrdered();
      if (L.ordered) {
        L.l64 = ilu->lowerSweep();
        L.u64 = ilu->upperSweep();
        L.sf_work = ilu->sfWork();
      }
      if (mixed && L.ordered) {
        ilu->fillSweepsF32(ctx->stream, lev_l32, lev_u32, lev_d32, L.l32, L.u32);
      } else if (mixed) {
        lev_lu32.resize(A.getNnz());
        launch_demote(ctx->stream, A.getNnz(), ilu->luValues(), lev_lu32.data());
        L.lu32 = lev_lu32.data();
      }
    } else if (p->precond == ALIEN_HIP_PREC_ILU0) {
      ilu.reset(new Alien::ILU0Preconditioner<Alg>(alg, A));
      ilu->setColourSchedule(*colours);
      ilu->init();
      phase("ILU0 factorisation");
      ps.kind = kPrecILU0;
      ps.cptr = *colours;
      lus.build(ctx->stream, n, A.kcol(), A.dcol());
      if (mixed) ps.lu32 = lus.fill<float>(ctx->stream, A, ilu->luValues(), lus.lbv32, lus.ubv32, lus.d32);
      else ilu_fp64(ps);
      phase("ILU0 split factor");
    }
    // FILU0: the REFERENCE-mode factor (exact or fixed-point sweeps) split the
    // same way, its inverse diagonal in place of the diagonal
    std::unique_ptr<Alien::FILU0Preconditioner<Alg>> filu;
    DeviceArray<float> finvd32;
    auto filu_fp64 = [&](PrecondSpec& q) {
      q.lu64 = lus.fill<double>(ctx->stream, A, filu->luValues(), lus.lbv64, lus.ubv64, lus.d64);
      q.lu64.diag = filu->invDiag().data();
    };
    if (p->precond == ALIEN_HIP_PREC_FILU0) {
      filu.reset(new Alien::FILU0Preconditioner<Alg>(alg, A));
      filu->setParameter("nb-factor-iter", p->filu_factor_niter);
      filu->setParameter("nb-solver-iter", p->filu_solver_niter);
      filu->init();
      phase("FILU0 factorisation");
      ps.kind = kPrecFILU0;
      ps.filu_iters = p->filu_solver_niter;
      lus.build(ctx->stream, n, A.kcol(), A.dcol());
      if (mixed) {
        ps.lu32 = lus.fill<float>(ctx->stream, A, filu->luValues(), lus.lbv32, lus.ubv32, lus.d32);
        finvd32.resize(n);
        launch_demote(ctx->stream, n, filu->invDiag().data(), finvd32.data());
        ps.lu32.diag = finvd32.data();
      } else {
        filu_fp64(ps);
      }
      phase("FILU0 split factor");
    }
    if (p->precond == ALIEN_HIP_PREC_NEUMANN) {
      Alien::NeumannPolyPreconditioner<Alg> np(alg, A, p->poly_factor, p->poly_order, p->poly_max_iter);
      np.init(); // the factor (estimated when poly_factor = 0)
      ps.kind = kPrecNeumann;
      ps.order = p->poly_order;
      ps.neumann_factor = np.factor();
      phase("Neumann factor");
    }
    // Level-set ILU0 with one launch per level and sweep: a graph of k iterations
    // would hold k x 2 x levels nodes and cost more to instantiate than it saves,
    // so those solves launch directly; the sync-free sweeps are captured.
    const bool capture = ps.kind != kPrecILU0Level || level_sync_free(ps.lev);
    int stall_window = 0;
    if (learned) {
      gnn::SystemView v;
      v.n = n;
      v.ro = A.kcol();
      v.cols = A.cols();
      v.dcol = A.dcol();
      v.d_ro = A.d_ro.data();
      v.d_cols = A.d_cols.data();
      v.d_dcol = A.d_dcol.data();
      v.d_vals = A.d_vals.data();
      if (A.isRBDIA() && !A.isPairDIA()) {
        v.rb_bo = A.bundleOffsets().data();
        v.d_rb_doff = A.d_doff.data();
        v.d_rb_dmask = A.d_dmask.data();
      }
      phase("Jacobi set-up");
      gnn::Engine* eng = nullptr;
      if (kept_engine && *kept_engine && (*kept_engine)->context() == ctx && (*kept_engine)->n() == n) {
        eng = kept_engine->get();
        eng->reset_system(v);
      } else {
        net_own.reset(new gnn::Engine(ctx, v));
        eng = net_own.get();
        if (kept_engine) *kept_engine = std::move(net_own); // owned by the kept matrix from here
      }
      theta.upload(p->gnn_theta, gnn::kNParam, ctx->stream);
      phase("network engine");
      gnn::Engine* net = eng;
      net->infer(theta.data());
      phase("network inference");
      net->build_factor_layout();
      phase("factor layout");
      ps.kind = kPrecFactor;
      const size_t ng = net->g_slots(), ngt = net->gt_slots();
      if (mixed && !bicg && p->factor_storage == 1) {
        gv16.resize(ng);
        gtv16.resize(ngt);
        net->fill_factor<bf16_bits>(gv16.data(), gtv16.data());
        ps.g16 = net->g_mat<float, bf16_bits>(gv16.data());
        ps.gt16 = net->gt_mat<float, bf16_bits>(gtv16.data());
        ps.half = true;
      } else if (mixed) {
        gv32.resize(ng);
        gtv32.resize(ngt);
        net->fill_factor<float>(gv32.data(), gtv32.data());
        ps.g32 = net->g_mat<float>(gv32.data());
        ps.gt32 = net->gt_mat<float>(gtv32.data());
      } else {
        gv64.resize(ng);
        gtv64.resize(ngt);
        net->fill_factor<double>(gv64.data(), gtv64.data());
        ps.g64 = net->g_mat<double>(gv64.data());
        ps.gt64 = net->gt_mat<double>(gtv64.data());
      }
      net->release_network_buffers();
      phase("factor fill + release");
      // default window: 400 iterations, growing with the grid edge N^(1/3) (CG's iteration
      // counts on these operators do): 2048 at 256^3
      stall_window = p->stall_window > 0 ? p->stall_window
                                         : std::max(400, (int)(8.0 * std::cbrt((double)n)));
    }
    if (!mixed) {
      RBMat<double> rb = A.rb64();
      const bool graph = p->mode == ALIEN_HIP_MODE_GRAPH && capture;
      DeviceArray<double> x0;
      if (learned) {
        x0.resize(n);
        launch_copy<double>(ctx->stream, n, x.data(), x0.data());
      }
      res->setup_s = tm.stop();
      tm.start();
      bool stalled = false;
      if (bicg) {
        BiCGState s;
        {
          FusedBiCG<double> F(ctx, n, rb, invd64, ps);
          s = F.solve(b.data(), x.data(), p->tol, p->max_iter, graph, p->poll_every, nullptr, stall_window,
                      &stalled);
        }
        if (stalled) {
          // multicolour ILU0 from the last finite iterate, for the remaining iterations
          if (!std::isfinite(s.rr)) launch_copy<double>(ctx->stream, n, x0.data(), x.data());
          const alien_hip_solve_result r2 =
              ilu0_multicolour_from(ctx, p, A, b.data(), x.data(), std::max(0, p->max_iter - s.it), false);
          res->solve_s = tm.stop();
          res->fallback = 1 | r2.fallback;
          res->iterations = s.it + r2.iterations;
          res->status = r2.status;
          res->error = r2.error;
          res->value = r2.value;
        } else {
          res->solve_s = tm.stop();
          res->iterations = s.it;
          res->status = s.status;
          res->error = s.breakdown ? 1 : 0; // NullValueException
          double bb = alg.dot(b, b);
          res->value = bb > 0 ? std::sqrt(s.rr) / std::sqrt(bb) : 0.0;
        }
        res->true_rel_res = true_residual(alg, A, b, x);
        x.copyValuesTo(x_io);
        ALIEN_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return;
      }
      CGState s;
      {
        FusedCG<double> F(ctx, n, rb, invd64, ps);
        s = F.solve(b.data(), x.data(), false, p->tol, p->max_iter, graph, p->poll_every, nullptr, stall_window,
                    &stalled);
        if (!stalled) finish_breakdown(F, s, &res->error);
      }
      if (stalled) {
        // multicolour ILU0 from the last finite iterate, for the remaining iterations
        if (!std::isfinite(s.rr)) launch_copy<double>(ctx->stream, n, x0.data(), x.data());
        const alien_hip_solve_result r2 =
            ilu0_multicolour_from(ctx, p, A, b.data(), x.data(), std::max(0, p->max_iter - s.it), false);
        res->solve_s = tm.stop();
        res->fallback = 1 | r2.fallback;
        res->iterations = s.it + r2.iterations;
        res->status = r2.status;
        res->error = r2.error;
        res->value = r2.value;
      } else {
        res->solve_s = tm.stop();
        res->iterations = s.it;
        res->status = s.status;
        double bb = alg.dot(b, b);
        res->value = bb > 0 ? std::sqrt(s.rr) / std::sqrt(bb) : 0.0;
      }
    } else {
      // FP32 matrix, preconditioner and Krylov vectors, FP64 dots; FP64 defect correction
      A.ensureRB32();
      phase("FP32 matrix");
      DeviceArray<float> invd32(n);
      launch_demote(ctx->stream, n, invd64, invd32.data());
      RBMat<float> rb32 = A.rb32();
      RBMat<double> rb64 = A.rb64();
      res->setup_s = tm.stop();
      tm.start();
      const double delta = p->inner_tol > 0 ? p->inner_tol : 0.1;
      auto run = bicg ? mixed_reliable_bicg : mixed_reliable_cg;
      MixedResult m = run(ctx, n, rb64, rb32, invd32.data(), ps, b.data(), x.data(), p->tol, p->max_iter, delta,
                          p->poll_every, capture, stall_window);
      if (m.stalled) {
        // mixed multicolour ILU0 from the last reliable update, for the remaining iterations
        const alien_hip_solve_result r2 =
            ilu0_multicolour_from(ctx, p, A, b.data(), x.data(), std::max(0, p->max_iter - m.it), true);
        res->fallback = 1 | r2.fallback;
        m.it += r2.iterations;
        m.updates += r2.outer;
        m.status = r2.status;
        m.error = r2.error;
        m.rel = r2.value;
        m.stalled = false;
      }
      if (m.fp32_stalled) {
        // FP64 from the last reliable update, for the remaining iterations, with the FP64 form of
        // the inner preconditioner (ILU0 for the learned M, its stall fallback)
        res->fallback |= 2;
        const int used = m.it;
        const int left = std::max(0, p->max_iter - used);
        PrecondSpec p64; // Jacobi, or (F)ILU0 with the FP64 factor
        if (ps.kind == kPrecILU0) {
          p64.kind = kPrecILU0;
          p64.cptr = ps.cptr;
          ilu_fp64(p64);
        } else if (ps.kind == kPrecFILU0) {
          p64.kind = kPrecFILU0;
          p64.filu_iters = ps.filu_iters;
          filu_fp64(p64);
        } else if (ps.kind == kPrecNeumann || ps.kind == kPrecCheb || ps.kind == kPrecILU0Level) {
          p64 = ps; // the FP64 solver applies them to A, invd and the FP64 factor
        }
        if (ps.kind == kPrecFactor) {
          // the learned M's fallback in FP64: multicolour ILU0
          const alien_hip_solve_result r2 = ilu0_multicolour_from(ctx, p, A, b.data(), x.data(), left, false);
          res->fallback |= 1 | r2.fallback;
          m.it = used + r2.iterations;
          m.status = r2.status;
          m.error = r2.error;
          m.rel = r2.value;
        } else if (bicg) {
          FusedBiCG<double> F(ctx, n, rb64, invd64, p64);
          const BiCGState s = F.solve(b.data(), x.data(), p->tol, left, capture,
                                      p->poll_every, nullptr);
          m.it = used + s.it;
          m.status = s.status;
          m.error = s.breakdown ? 1 : 0;
          m.rel = std::sqrt(s.rr / alg.dot(b, b));
        } else {
          FusedCG<double> F(ctx, n, rb64, invd64, p64);
          CGState s = F.solve(b.data(), x.data(), false, p->tol, left, capture,
                              p->poll_every, nullptr);
          finish_breakdown(F, s, &m.error);
          m.it = used + s.it;
          m.status = s.status;
          m.rel = std::sqrt(s.rr / alg.dot(b, b));
        }
      }
      res->solve_s = tm.stop();
      res->iterations = m.it;
      res->outer = m.updates;
      res->status = m.status;
      res->error = m.error;
      res->value = m.rel;
    }
  }
  res->true_rel_res = true_residual(alg, A, b, x);
  x.copyValuesTo(x_io);
}

} // namespace
} // namespace alien_hip

using namespace alien_hip;

extern "C" int alien_hip_solve_csr(alien_hip_ctx* ctx, const alien_hip_solve_params* p, int n, const int* ro,
                                   const int* cols, const double* vals, const double* bh, double* xh,
                                   alien_hip_solve_result* res) {
  if (!ctx || !p || !res || !bh || !xh || !vals) return fail(ALIEN_HIP_EINVAL, "solve_csr: NULL argument");
  std::memset(res, 0, sizeof(*res));
  try {
    const bool need_diag = p->precond != ALIEN_HIP_PREC_NONE;
    validate_csr(n, ro, cols, need_diag);
    validate_params(p);
    const bool fused = p->mode == ALIEN_HIP_MODE_FUSED || p->mode == ALIEN_HIP_MODE_GRAPH;
    const bool mixed = p->mode == ALIEN_HIP_MODE_MIXED;
    EventTimer tm(ctx->stream);
    tm.start(); // set-up time includes the colouring and the permutation
    SetupTrace phase(ctx);

    // multicolour ILU0: solve the permuted system P A P^T (P x) = P b
    std::vector<int> perm;
    std::vector<int> pro;
    HostBuf<int> pcols;
    HostBuf<double> pvals;
    std::vector<double> pb(n), px(n);
    const int* Aro = ro;
    const int* Acols = cols;
    const double* Avals = vals;
    const double* b_in = bh;
    double* x_io = xh;
    std::vector<int> cptr;
    if (p->precond == ALIEN_HIP_PREC_ILU0 && p->ilu_schedule == 1) {
      perm = colour_permutation(n, ro, cols, cptr);
      permute_csr(n, ro, cols, vals, perm, pro, pcols, pvals);
      for (int r = 0; r < n; ++r) { pb[r] = bh[perm[r]]; px[r] = xh[perm[r]]; }
      Aro = pro.data(); Acols = pcols.data(); Avals = pvals.data();
      b_in = pb.data(); x_io = px.data();
    }

    if (!perm.empty()) phase("colouring + permutation");
    Alien::CSRHostView view{n, Aro, Acols, Avals};
    HIPMatrix A(ctx, view, (fused || mixed) ? ALIEN_HIP_LAYOUT_CSR : p->layout, /*borrow_host=*/true);
    phase("matrix upload");
    if (fused || mixed) A.ensureFused(p->layout);
    phase("row-bundled layout");
    solve_on(ctx, p, A, b_in, x_io, res, tm, phase, &cptr);
    if (!perm.empty())
      for (int r = 0; r < n; ++r) xh[perm[r]] = px[r];
    ALIEN_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  } catch (const Error& e) {
    return fail(e.code, e.what());
  } catch (const std::exception& e) {
    return fail(ALIEN_HIP_EINVAL, e.what());
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Converted matrix kept between solves (Alien's MultiMatrixImpl keeps the hip
// matrix and re-converts on a new timestamp: MultiMatrixImpl.cc:226-248). The
// structure (device CSR, row-bundled layouts) is built once; new values are
// uploaded and refilled into the kept layouts.
// ---------------------------------------------------------------------------
struct alien_hip_matrix {
  alien_hip_ctx* ctx = nullptr;
  alien_hip::HIPMatrix A;
  bool diag_ok = false; // every row stores its diagonal entry
  // the learned preconditioner's network engine (factor layouts of A's structure), kept between solves
  std::unique_ptr<alien_hip::gnn::Engine> engine;
  // multicolour ILU0: the colour-permuted matrix P A P^T, built by the first
  // such solve and kept with A (its values gathered from A's on each update)
  std::unique_ptr<alien_hip::HIPMatrix> PA;
  std::vector<int> perm, cptr;
  alien_hip::DeviceArray<int> d_map; // PA's entry k = A's entry d_map[k]

  void build_coloured() {
    const int n = A.getNRows();
    const int* ro = A.kcol();
    const int* cols = A.cols();
    perm = alien_hip::colour_permutation(n, ro, cols, cptr);
    // permute the entry indices, then read the values through them
    const long nnz = A.getNnz();
    alien_hip::HostBuf<double> idx, pidx, pvals;
    alien_hip::HostBuf<int> pcols, map;
    idx.alloc(nnz);
    alien_hip::parallel_ranges(nnz, [&](long a, long b) {
      for (long k = a; k < b; ++k) idx[k] = (double)k;
    });
    std::vector<int> pro;
    alien_hip::permute_csr(n, ro, cols, idx.data(), perm, pro, pcols, pidx);
    map.alloc(nnz);
    pvals.alloc(nnz);
    const double* v = A.hostValues();
    alien_hip::parallel_ranges(nnz, [&](long a, long b) {
      for (long k = a; k < b; ++k) {
        map[k] = (int)pidx[k];
        pvals[k] = v[map[k]];
      }
    });
    std::unique_ptr<alien_hip::HIPMatrix> P(new alien_hip::HIPMatrix);
    Alien::CSRHostView view{n, pro.data(), pcols.data(), pvals.data()};
    P->build(ctx, view, ALIEN_HIP_LAYOUT_CSR, /*borrow_host=*/false);
    d_map.upload(map.data(), (size_t)nnz, ctx->stream);
    ALIEN_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    PA = std::move(P);
  }
};

extern "C" int alien_hip_matrix_create(alien_hip_ctx* ctx, int n, const int* ro, const int* cols,
                                       const double* vals, alien_hip_matrix** out) {
  if (!ctx || !out || !vals) return fail(ALIEN_HIP_EINVAL, "matrix_create: NULL argument");
  *out = nullptr;
  try {
    validate_csr(n, ro, cols, false);
    std::unique_ptr<alien_hip_matrix> m(new alien_hip_matrix);
    m->ctx = ctx;
    Alien::CSRHostView view{n, ro, cols, vals};
    m->A.build(ctx, view, ALIEN_HIP_LAYOUT_CSR, /*borrow_host=*/false);
    const int* d = m->A.dcol();
    m->diag_ok = true;
    for (int i = 0; i < n && m->diag_ok; ++i) m->diag_ok = d[i] < ro[i + 1] && cols[d[i]] == i;
    *out = m.release();
  } catch (const Error& e) {
    return fail(e.code, e.what());
  } catch (const std::exception& e) {
    return fail(ALIEN_HIP_EINVAL, e.what());
  }
  return 0;
}

extern "C" int alien_hip_matrix_update_values(alien_hip_matrix* m, const double* vals) {
  if (!m || (!vals && m->A.getNnz() > 0)) return fail(ALIEN_HIP_EINVAL, "matrix_update_values: NULL argument");
  try {
    m->A.updateValues(vals);
    if (m->PA) m->PA->updateValuesGather(m->d_map.data(), m->A.d_vals.data());
  } catch (const Error& e) {
    return fail(e.code, e.what());
  } catch (const std::exception& e) {
    return fail(ALIEN_HIP_EINVAL, e.what());
  }
  return 0;
}

extern "C" int alien_hip_matrix_destroy(alien_hip_matrix* m) {
  delete m;
  return 0;
}

extern "C" int alien_hip_solve(alien_hip_ctx* ctx, const alien_hip_solve_params* p, alien_hip_matrix* m,
                               const double* bh, double* xh, alien_hip_solve_result* res) {
  if (!ctx || !p || !m || !res || !bh || !xh) return fail(ALIEN_HIP_EINVAL, "solve: NULL argument");
  std::memset(res, 0, sizeof(*res));
  try {
    validate_params(p);
    if (p->precond != ALIEN_HIP_PREC_NONE && !m->diag_ok)
      throw Error(ALIEN_HIP_EINVAL, "CSR: missing diagonal entry");
    const bool fused = p->mode == ALIEN_HIP_MODE_FUSED || p->mode == ALIEN_HIP_MODE_GRAPH;
    const bool mixed = p->mode == ALIEN_HIP_MODE_MIXED;
    const bool coloured = p->precond == ALIEN_HIP_PREC_ILU0 && p->ilu_schedule == 1;
    const int n = m->A.getNRows();
    ctx->reserve_for(n);
    EventTimer tm(ctx->stream);
    tm.start();
    SetupTrace phase(ctx);
    // multicolour ILU0 solves P A P^T (P x) = P b on the kept permuted matrix
    if (coloured && !m->PA) {
      m->build_coloured();
      phase("colour permutation");
    }
    HIPMatrix& A = coloured ? *m->PA : m->A;
    std::vector<double> pb, px;
    const double* b_in = bh;
    double* x_io = xh;
    if (coloured) {
      pb.resize(n);
      px.resize(n);
      for (int r = 0; r < n; ++r) {
        pb[r] = bh[m->perm[r]];
        px[r] = xh[m->perm[r]];
      }
      b_in = pb.data();
      x_io = px.data();
    }
    if (fused || mixed) A.ensureFused(p->layout);
    else A.setSpmvLayout(p->layout);
    phase("row-bundled layout");
    solve_on(ctx, p, A, b_in, x_io, res, tm, phase, coloured ? &m->cptr : nullptr, coloured ? nullptr : &m->engine);
    if (coloured)
      for (int r = 0; r < n; ++r) xh[m->perm[r]] = px[r];
    ALIEN_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  } catch (const Error& e) {
    return fail(e.code, e.what());
  } catch (const std::exception& e) {
    return fail(ALIEN_HIP_EINVAL, e.what());
  }
  return 0;
}
