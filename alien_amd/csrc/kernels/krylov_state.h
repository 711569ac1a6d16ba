// Device-resident state of the fused Krylov iterations and the launchers of
// krylov.hip.
#pragma once

#include <hip/hip_runtime.h>

namespace alien_hip {

struct RedWork;

// bfloat16 bits (the upper half of an FP32 value): value storage of the learned
// factor in the FP32 solver
typedef unsigned short bf16_bits;

// Row-bundled operand (rb_rows.h): RB-CSR when doff is NULL, RB-DIA otherwise.
// T is the arithmetic type, V the stored value type (T, or bf16_bits for T = float).
template <typename T, typename V = T>
struct RBMat {
  const int* bo; // bundle offsets in 64-entry slots, nbundles + 1
  const int* bc; // RB-CSR: columns, -1 = padding
  const V* bv;   // values
  const int* doff = nullptr;                  // RB-DIA: column offset of each bundle slot
  const unsigned long long* dmask = nullptr;  // RB-DIA: rows of the bundle holding that slot
  // RB-DIA of a matrix whose rows hold column pairs (c, c + 1), c even (two unknowns
  // per cell, interleaved): a slot is one pair at c = (row & ~1) + doff, its two
  // values in planes 2 s and 2 s + 1
  int pair = 0;
  // Alien's BEllPack (sliced ELL, SYCLBEllPackMatrix.cc:69-170) when > 0: the number of
  // 64-row bundles per slice (slice height / 64); bo then holds the slice offsets (bro),
  // bc / bv the BEllPack columns and values (padding: column 0, value 0)
  int ell = 0;
  // Uniform RB-DIA when > 0: every bundle holds the same `uni` slots, at the column offsets
  // uoff[0..uni) (ascending), bundle b's at slots b uni .. b uni + uni - 1, absent entries
  // stored as zeros. The row kernels then take the offsets from here (kernel arguments)
  // instead of loading bo / doff / dmask per bundle; bo, doff and dmask still describe the
  // same matrix as an ordinary RB-DIA. nrows: the matrix's rows (gathers are clamped to it).
  int uni = 0;
  int nrows = 0;
  int uoff[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};
constexpr int kUniMaxSlots = 8;

// Row-kernel instance per layout class (a template argument: each kernel's code holds only its
// layout's row sum)
enum RowLayout { kRowBundled = 0, kRowEll = 1, kRowUniform = 2 };

// Scalars of one CG solve; lives in device memory, read by the host only to
// poll `done`. Mirrors the locals of CG::solve (rho, rho1, alpha) and the
// fields of Iteration (iteration count, criteria value, status).
struct CGState {
  double rho;   // previous r.z
  double rho1;  // current r.z
  double pq;    // p.q
  double rr;    // last computed r.r (Iteration's m_value)
  double crit;  // tol^2 (b,b)
  int it;       // iterations done
  int max_it;
  int done;     // set when the stop test passes, max_it is reached or on breakdown
  int status;   // Iteration::getStatus()
  int breakdown;// p.q == 0
  int pad[3];
};

// Scalars of one BiCGStab solve (BiCGStab.h:151-300), device-resident like CGState.
struct BiCGState {
  double rho;    // rho of the previous step
  double rho1;   // r.r0 of the current step
  double alpha;  // rho1 / (v.r0)
  double omega;  // (t.s) / (t.t)
  double rr;     // last (r,r) or (s,s) of a stop test (Iteration's m_value)
  double crit;   // tol^2 (b,b)
  double rr0;    // (r,r) of the initial residual (first-block breakdown rule)
  double rrmax;  // largest (r,r) or (s,s) of a stop test since the last (re)start
  int it, max_it;
  int done;      // stop test passed, max_it reached or breakdown
  int status;    // Iteration::getStatus()
  int breakdown; // 0, 1 = v.r0 == 0 (alpha), 2 = t.t == 0 (beta), 3 = rho == 0
  int half;      // stopped on s: x += alpha phat is still to be applied
  int first;     // the first block (before the while loop), counted differently (BiCGStab.h:194-199)
  int pad;
};

// BiCGStab step kernels (bicgstab.hip). DIAG: Jacobi fused (phat = invD p, shat = invD s).
template <typename T>
void launch_bicg_p(hipStream_t, int n, const T* r, const T* v, T* p, const T* invd_or_null, T* phat,
                   BiCGState* st);
template <typename T>
void launch_bicg_v(hipStream_t, RedWork w, int n, RBMat<T> A, const T* phat, const T* r0, T* v, BiCGState* st);
template <typename T>
void launch_bicg_s(hipStream_t, RedWork w, int n, const T* r, const T* v, T* s, const T* invd_or_null, T* shat,
                   BiCGState* st);
template <typename T>
void launch_bicg_t(hipStream_t, RedWork w, int n, RBMat<T> A, const T* shat, const T* s, T* t, BiCGState* st);
template <typename T>
void launch_bicg_xr(hipStream_t, RedWork w, int n, const T* shat, const T* phat, const T* s, const T* t,
                    const T* r0, T* x, T* r, BiCGState* st);
// x += alpha phat when the solve stopped on s (half step)
template <typename T>
void launch_bicg_half(hipStream_t, int n, const T* phat, T* x, const BiCGState* st);
// y = M x with M = G G^T (tmp = G^T x), skipped once done is set
template <typename T>
void launch_factor_apply_flag(hipStream_t, int n, RBMat<T> G, RBMat<T> GT, const T* x, T* tmp, T* y,
                              const int* done);

template <typename T>
void launch_residual(hipStream_t, RedWork w, int n, RBMat<T> A, const T* x, const T* b, T* r, double* out);
// one Chebyshev step; with st (CG) the stop flag is st->done and the last step
// leaves rho1 = x . y in the state, without (BiCGStab) the flag is done_flag
template <typename T>
void launch_cheb_step(hipStream_t, RedWork w, int n, RBMat<T> A, const T* xin, const T* yin, T* yout, T* p,
                      const T* invd, double a, double b, bool first, bool last, CGState* st,
                      const int* done_flag = nullptr);
template <typename T>
void launch_diag_apply(hipStream_t, RedWork w, int n, const T* r, const T* invd, T* z, CGState* st);
// rho1 = r . z into the CG state (after a preconditioner applied by separate launches)
template <typename T>
void launch_rz(hipStream_t, RedWork w, int n, const T* r, const T* z, CGState* st);
// Neumann polynomial apply z = f sum_{k<=order} (I - f A)^k x in order >= 1 row
// kernels (tmp: workspace); with rdot the last leaves the partials of rdot . z,
// whose count is returned (0 without rdot). Returns early once *done.
template <typename T>
long launch_neumann_apply(hipStream_t, RedWork w, int n, RBMat<T> A, int order, double f, const T* x, T* tmp, T* z,
                          const int* done, const T* rdot);
template <typename T>
void launch_cg_p(hipStream_t, int n, const T* z, T* p, CGState* st, int first);
// (t0, t1 non-null: the launch records the kernel's start and end in them, hipExtLaunchKernel)
template <typename T>
void launch_cg_q(hipStream_t, RedWork w, int n, RBMat<T> A, const T* p, T* q, CGState* st,
                 hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
// y = A x and *out = x . y by the kernel of launch_cg_q (idle: a state that is never done)
template <typename T>
void launch_spmv_dot_rb(hipStream_t, RedWork w, int n, RBMat<T> A, const T* x, T* y, CGState* idle, double* out);
// invd_or_null != NULL fuses the Diag preconditioner (z = invD r, rho1 = r.z)
template <typename T>
void launch_cg_xr(hipStream_t, RedWork w, int n, const T* p, const T* q, T* x, T* r, const T* invd_or_null,
                  T* z_or_null, CGState* st);
// z = G (G^T r) and rho1 = r.z (factored approximate inverse M = G G^T)
// x += alpha p, r -= alpha q with no reduction (learned-M iteration)
template <typename T>
void launch_cg_update(hipStream_t, int n, const T* p, const T* q, T* x, T* r, CGState* st);
// M = G G^T fused with the direction update: t = G^T r, rho1 = t.t, p = G t + (rho1/rho) p;
// after the first step its finalize also completes the previous update (++iter, stop test on r.r)
template <typename T, typename V>
void launch_factor_p(hipStream_t, RedWork w, int n, RBMat<T, V> G, RBMat<T, V> GT, const T* r, T* t, T* p,
                     CGState* st, int first, int rho1_only = 0);
template <typename T>
void launch_factor_apply(hipStream_t, RedWork w, int n, RBMat<T> G, RBMat<T> GT, const T* r, T* tmp, T* z,
                         CGState* st);
void launch_fold_increment(hipStream_t, int n, float* x32, double* x);
void launch_demote(hipStream_t, int n, const double* a, float* b);

} // namespace alien_hip
