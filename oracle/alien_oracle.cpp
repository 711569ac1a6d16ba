/*
 * alien_oracle.cpp -- TEST INFRASTRUCTURE ONLY (see alien_oracle.h).
 *
 * A plain restatement of Alien's SimpleCSR Krylov path in FP64 sequential
 * loops. Built with g++ -O2 -ffp-contract=off so every multiply and add is
 * rounded separately, in the order the reference's loops imply.
 *
 * The solvers follow the reference's call sequence step for step (the
 * sequence of BLAS-1 / SpMV calls decides the rounding, so it must stay the
 * same); the code is a restatement, not a copy.
 */
#include "alien_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::vector<double> Vec;

struct NullValue : std::runtime_error {
  explicit NullValue(const char* w) : std::runtime_error(w) {}
};

/* Row-block ranges: rank r owns [first(r), first(r+1)) as in the contiguous
 * row partition of examples/test_krylov.cpp:175-184. */
static inline int block_first(int n, int nt, int r) {
  long q = n / nt, rem = n % nt;
  return (int)(r * q + std::min<long>(r, rem));
}

/* CSR matrix view + FP64 algebra with the SimpleCSRInternalLinearAlgebra
 * method set (SimpleCSRInternalLinearAlgebra.cc:74-257, CBLASMPIKernel.h:37-125). */
struct Algebra {
  int n;
  const int* ro;
  const int* cols;
  const double* vals;
  std::vector<int> dcol;
  int nt;

  Algebra(int n_, const int* ro_, const int* cols_, const double* vals_, int nt_)
  : n(n_), ro(ro_), cols(cols_), vals(vals_), dcol(n_), nt(std::max(1, nt_)) {
    orc_dcol(n, ro, cols, dcol.data());
  }

  void allocate(Vec& v) const { v.assign(n, 0.0); } /* SimpleCSRVector::init zero-fills */

  void copy(const Vec& x, Vec& y) const {
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
    for (int i = 0; i < n; ++i) y[i] = x[i];
  }
  void axpy(double a, const Vec& x, Vec& y) const { /* y <- y + a x */
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
    for (int i = 0; i < n; ++i) y[i] = y[i] + a * x[i];
  }
  void scal(double a, Vec& x) const {
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
    for (int i = 0; i < n; ++i) x[i] = a * x[i];
  }
  void pointwiseMult(const Vec& x, const Vec& y, Vec& z) const {
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
    for (int i = 0; i < n; ++i) z[i] = x[i] * y[i];
  }
  void assign(Vec& x, double a) const { std::fill(x.begin(), x.end(), a); }
  double dot(const Vec& x, const Vec& y) const {
    if (nt == 1) return orc_dot(n, x.data(), y.data());
    /* one partial per row block, summed in block order (the mpAllReduce of
     * CBLASMPIKernel::dot, CBLASMPIKernel.h:88-99, with ranks in order) */
    std::vector<double> part(nt, 0.0);
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
      int r = omp_get_thread_num();
#else
      int r = 0;
#endif
      int b = block_first(n, nt, r), e = block_first(n, nt, r + 1);
      double s = 0.0;
      for (int i = b; i < e; ++i) s = s + x[i] * y[i];
      part[r] = s;
    }
    double s = 0.0;
    for (int r = 0; r < nt; ++r) s = s + part[r];
    return s;
  }
  double norm2(const Vec& x) const { return std::sqrt(dot(x, x)); } /* CBLASMPIKernel.h:114-125 */
  void mult(const Vec& x, Vec& y) const {
    if (nt == 1) {
      orc_spmv_csr(n, ro, cols, vals, x.data(), y.data());
      return;
    }
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int i = 0; i < n; ++i) {
      double t = 0.0;
      for (int k = ro[i]; k < ro[i + 1]; ++k) t = t + vals[k] * x[cols[k]];
      y[i] = t;
    }
  }
  void computeInvDiag(Vec& d) const { orc_invdiag_csr(n, dcol.data(), vals, d.data()); }
};

/* Iteration (src/core/alien/expression/krylov/Iteration.h:44-117) */
struct Iteration {
  const Algebra& alg;
  int max_it;
  int it = 0;
  double nrm2_b, crit, sqrt_nrm2_b, value;
  bool status;
  Iteration(const Algebra& a, const Vec& b, double tol, int maxit) : alg(a), max_it(maxit) {
    nrm2_b = alg.dot(b, b);
    crit = tol * tol * nrm2_b;
    sqrt_nrm2_b = std::sqrt(nrm2_b);
    value = crit + 1;
    status = (nrm2_b == 0);
  }
  bool nullRhs() const { return nrm2_b == 0.; }
  bool stop(const Vec& r) {
    if (it >= max_it) return true;
    value = alg.dot(r, r);
    status = value < crit;
    return status;
  }
  void inc() { ++it; }
  double getValue() const { return sqrt_nrm2_b == 0 ? 0. : std::sqrt(value) / sqrt_nrm2_b; }
};

struct Precond {
  virtual ~Precond() {}
  virtual void solve(const Algebra& alg, const Vec& x, Vec& y) = 0;
};

/* DiagPreconditioner (DiagPreconditioner.h:49-67) */
struct DiagPrec : Precond {
  Vec inv_diag;
  explicit DiagPrec(const Algebra& alg) {
    alg.allocate(inv_diag);
    alg.assign(inv_diag, 1.);
    alg.computeInvDiag(inv_diag);
  }
  void solve(const Algebra& alg, const Vec& x, Vec& y) override { alg.pointwiseMult(inv_diag, x, y); }
};

struct NonePrec : Precond {
  void solve(const Algebra& alg, const Vec& x, Vec& y) override { alg.copy(x, y); }
};

/* ChebyshevPreconditioner, non-Saad branch (ChebyshevPreconditioner.h:73-202, 392-444) */
struct ChebPrec : Precond {
  double factor, alpha = 0, beta = 0;
  int order, max_it;
  Vec inv_diag, my, mr, mp, mz;

  ChebPrec(const Algebra& alg, double factor_, int order_, int max_it_)
  : factor(factor_), order(order_), max_it(max_it_) {
    alg.allocate(my); alg.allocate(mr);
    alg.allocate(mp); alg.allocate(mz);
    alg.allocate(inv_diag);
    alg.assign(inv_diag, 1.);
    alg.computeInvDiag(inv_diag);
    computeBeta(alg);
    computeAlpha(alg);
    if (factor != 0.) alpha = beta / factor; /* :99-100 overwrites computeAlpha's result */
  }

  /* start vector: default-seeded mt19937 through uniform_real_distribution(-1,1),
   * drawn in index order (:122-129) */
  static void random_start(const Algebra& alg, Vec& x) {
    std::mt19937 gen;
    std::uniform_real_distribution<> dis(-1., 1.);
    for (int i = 0; i < alg.n; ++i) x[i] = dis(gen);
  }

  void computeBeta(const Algebra& alg) { /* :118-156 */
    Vec x; alg.allocate(x);
    random_start(alg, x);
    double nrm = alg.norm2(x);
    alg.scal(1 / nrm, x);
    for (int i = 0; i < max_it; ++i) {
      alg.mult(x, my);
      alg.pointwiseMult(inv_diag, my, my);
      double xAx = alg.dot(my, x);
      double xdx = alg.dot(x, x);
      double nrme = alg.norm2(my);
      alg.scal(1 / nrme, my);
      alg.copy(my, x);
      beta = xAx / xdx;
    }
  }
  void computeAlpha(const Algebra& alg) { /* :158-202 */
    Vec x; alg.allocate(x);
    random_start(alg, x);
    double nrm = alg.norm2(x);
    alg.scal(1 / nrm, x);
    for (int i = 0; i < max_it; ++i) {
      alg.mult(x, my);
      alg.pointwiseMult(inv_diag, my, my);
      alg.axpy(-beta, x, my);
      double xAx = alg.dot(my, x);
      double xdx = alg.dot(x, x);
      alpha = xAx / xdx;
      double nrme = alg.norm2(my);
      alg.scal(1 / nrme, my);
      alg.copy(my, x);
    }
    alpha += beta;
  }
  void solve(const Algebra& alg, const Vec& x, Vec& y) override { /* _algo2 :392-434 */
    const double d = (beta + alpha) / 2;
    const double c = (beta - alpha) / 2;
    alg.copy(x, y);
    alg.mult(y, my);
    alg.copy(x, mr);
    alg.axpy(-1., my, mr);
    alg.pointwiseMult(inv_diag, mr, mz);
    alg.copy(mz, mp);
    double a = 2 / d;
    alg.axpy(a, mp, y);
    for (int i = 1; i < order; ++i) {
      alg.mult(y, my);
      alg.copy(x, mr);
      alg.axpy(-1., my, mr);
      alg.pointwiseMult(inv_diag, mr, mz);
      double b = a * (c / 2.) * (c / 2.);
      a = 1. / (d - b);
      alg.scal(b, mp);
      alg.axpy(1., mz, mp);
      alg.axpy(a, mp, y);
    }
  }
};

/* ILU0Preconditioner / LUFactorisationAlgo, sequential branch (ILU0Preconditioner.h:50-248).
 * With nthreads > 1 the factor is block-Jacobi over the row blocks (bjacobi=true, :95-119). */
struct ILU0Prec : Precond {
  int n;
  const int* ro;
  const int* cols;
  std::vector<int> dcol;
  std::vector<int> nloc_end; /* per row: end of the entries inside the row's block */
  Vec lu, tmp;
  ILU0Prec(const Algebra& alg) : n(alg.n), ro(alg.ro), cols(alg.cols), dcol(alg.dcol) {
    lu.assign(alg.vals, alg.vals + ro[n]);
    tmp.assign(n, 0.);
    if (alg.nt == 1) {
      orc_ilu0_factor(n, ro, cols, dcol.data(), lu.data());
    } else {
      factor_bjacobi(alg.nt);
    }
  }
  /* block-Jacobi: only entries whose column lies in the row's block take part */
  void factor_bjacobi(int nt) {
    std::vector<int> owner_first(nt + 1);
    for (int r = 0; r <= nt; ++r) owner_first[r] = block_first(n, nt, r);
    std::vector<int> blk(n);
    for (int r = 0; r < nt; ++r)
      for (int i = owner_first[r]; i < owner_first[r + 1]; ++i) blk[i] = r;
    /* drop couplings between blocks: zero them in the factor, they are never used */
    for (int i = 0; i < n; ++i)
      for (int k = ro[i]; k < ro[i + 1]; ++k)
        if (blk[cols[k]] != blk[i]) lu[k] = 0.0;
    orc_ilu0_factor(n, ro, cols, dcol.data(), lu.data());
  }
  void solve(const Algebra&, const Vec& y, Vec& x) override {
    orc_ilu0_solve(n, ro, cols, dcol.data(), lu.data(), y.data(), x.data());
  }
};

/* CG::solve (CG.h:58-150) */
int cg_solve(const Algebra& alg, Precond& P, Iteration& iter, const Vec& b, Vec& x) {
  if (iter.nullRhs()) return 0;
  double rho = 0, rho1 = 0, alpha = 0;
  Vec p, z, q, r;
  alg.allocate(p); alg.allocate(z); alg.allocate(q); alg.allocate(r);
  alg.copy(b, r);
  alg.mult(x, p);
  alg.axpy(-1., p, r);

  P.solve(alg, r, z);
  rho1 = alg.dot(r, z);
  alg.copy(z, p);
  alg.mult(p, q);
  alpha = alg.dot(p, q);
  if (alpha == 0) {
    if (iter.stop(r)) { iter.inc(); return 0; }
    throw NullValue("alpha");
  }
  alpha = rho1 / alpha;
  alg.axpy(alpha, p, x);
  alg.axpy(-alpha, q, r);
  rho = rho1;
  iter.inc();

  while (!iter.stop(r)) {
    P.solve(alg, r, z);
    rho1 = alg.dot(r, z);
    alpha = rho1 / rho;
    alg.axpy(alpha, p, z);
    alg.copy(z, p);
    alg.mult(p, q);
    alpha = alg.dot(q, p);
    if (alpha == 0) {
      if (iter.stop(r)) { iter.inc(); return 0; }
      throw NullValue("alpha");
    }
    alpha = rho1 / alpha;
    alg.axpy(alpha, p, x);
    alg.axpy(-alpha, q, r);
    rho = rho1;
    iter.inc();
  }
  return 0;
}

/* BiCGStab::solve (BiCGStab.h:151-300) */
int bicgs_solve(const Algebra& alg, Precond& P, Iteration& iter, const Vec& b, Vec& x) {
  if (iter.nullRhs()) return 0;
  double rho = 0, rho1 = 0, alpha = 0, beta = 0, omega = 0;
  Vec p, phat, s, shat, t, v, r, r0;
  for (Vec* w : {&p, &phat, &s, &shat, &t, &v, &r, &r0}) alg.allocate(*w);
  alg.copy(b, r);
  alg.mult(x, r0);
  alg.axpy(-1., r0, r);
  alg.copy(r, r0);
  alg.copy(r, p);
  rho1 = alg.dot(r, r0);

  P.solve(alg, p, phat);
  alg.mult(phat, v);
  alpha = alg.dot(v, r0);
  if (alpha == 0) throw NullValue("alpha");
  alpha = rho1 / alpha;
  alg.copy(r, s);
  alg.axpy(-alpha, v, s);
  if (iter.stop(s)) {
    iter.inc();
    alg.axpy(alpha, phat, x);
    return 0;
  }
  P.solve(alg, s, shat);
  alg.mult(shat, t);
  omega = alg.dot(t, s);
  beta = alg.dot(t, t);
  if (beta == 0) {
    if (iter.stop(r)) {
      iter.inc();
      alg.axpy(alpha, phat, x);
      return 0;
    }
    throw NullValue("beta");
  }
  omega = omega / beta;
  alg.axpy(omega, shat, x);
  alg.axpy(alpha, phat, x);
  alg.copy(s, r);
  alg.axpy(-omega, t, r);
  rho = rho1;
  iter.inc();

  while (!iter.stop(r)) {
    rho1 = alg.dot(r, r0);
    beta = (rho1 / rho) * (alpha / omega);
    alg.axpy(-omega, v, p);
    alg.scal(beta, p);
    alg.axpy(1., r, p);
    if (rho == 0) throw NullValue("rho");
    P.solve(alg, p, phat);
    alg.mult(phat, v);
    alpha = alg.dot(v, r0);
    if (alpha == 0) throw NullValue("alpha");
    alpha = rho1 / alpha;
    alg.copy(r, s);
    alg.axpy(-alpha, v, s);
    if (iter.stop(s)) { /* no ++iter on this exit (:259-263) */
      alg.axpy(alpha, phat, x);
      return 0;
    }
    P.solve(alg, s, shat);
    alg.mult(shat, t);
    omega = alg.dot(t, s);
    beta = alg.dot(t, t);
    if (beta == 0) {
      if (iter.stop(s)) {
        alg.axpy(alpha, phat, x);
        return 0;
      }
      throw NullValue("beta");
    }
    omega = omega / beta;
    alg.axpy(omega, shat, x);
    alg.axpy(alpha, phat, x);
    alg.copy(s, r);
    alg.axpy(-omega, t, r);
    rho = rho1;
    iter.inc();
  }
  return 0;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

} // namespace

extern "C" {

void orc_dcol(int nrows, const int* ro, const int* cols, int* dcol) {
  for (int i = 0; i < nrows; ++i) {
    int idx = ro[i];
    for (int k = ro[i]; k < ro[i + 1]; ++k) {
      if (cols[k] < i) ++idx;
      else break;
    }
    dcol[i] = idx;
  }
}

void orc_spmv_csr(int nrows, const int* ro, const int* cols, const double* vals,
                  const double* x, double* y) {
  for (int i = 0; i < nrows; ++i) {
    double t = 0.;
    for (int k = ro[i]; k < ro[i + 1]; ++k) t += vals[k] * x[cols[k]];
    y[i] = t;
  }
}

void orc_spmv_bsr(int nbrows, int bs, const int* ro, const int* cols, const double* vals,
                  const double* x, double* y) {
  for (int i = 0; i < nbrows; ++i) {
    double* yi = y + (long)i * bs;
    for (int e = 0; e < bs; ++e) yi[e] = 0.;
    for (int k = ro[i]; k < ro[i + 1]; ++k) {
      const double* m = vals + (long)k * bs * bs;
      const double* xj = x + (long)cols[k] * bs;
      for (int e = 0; e < bs; ++e)
        for (int u = 0; u < bs; ++u) yi[e] += m[e + bs * u] * xj[u];
    }
  }
}

void orc_spmv_vblock(int nbrows, const int* ro, const int* cols, const int* block_cols, const int* block_size,
                     const int* block_offset, const double* vals, const double* x, double* y) {
  std::vector<double> tmpy;
  for (int irow = 0; irow < nbrows; ++irow) {
    const int bsr = block_size[irow];
    tmpy.assign(bsr, 0.);
    for (int j = ro[irow]; j < ro[irow + 1]; ++j) {
      const int col = cols[j];
      const int bsc = block_size[col];
      const double* xb = x + block_offset[col];
      const double* m = vals + block_cols[j];
      for (int kr = 0; kr < bsr; ++kr)
        for (int kc = 0; kc < bsc; ++kc) tmpy[kr] += m[kr * bsc + kc] * xb[kc];
    }
    for (int kr = 0; kr < bsr; ++kr) y[block_offset[irow] + kr] = tmpy[kr];
  }
}

void orc_addlmult(int nrows, const int* ro, const int* cols, const int* dcol,
                  const double* vals, double alpha, const double* x, double* y) {
  for (int i = 0; i < nrows; ++i) {
    double t = y[i];
    for (int k = ro[i]; k < dcol[i]; ++k) t += alpha * vals[k] * x[cols[k]];
    y[i] = t;
  }
}

void orc_addumult(int nrows, const int* ro, const int* cols, const int* dcol,
                  const double* vals, double alpha, const double* x, double* y) {
  for (int i = 0; i < nrows; ++i) {
    double t = y[i];
    for (int k = dcol[i] + 1; k < ro[i + 1]; ++k) t += alpha * vals[k] * x[cols[k]];
    y[i] = t;
  }
}

void orc_invdiag_csr(int nrows, const int* dcol, const double* vals, double* y) {
  for (int i = 0; i < nrows; ++i) y[i] = 1. / vals[dcol[i]];
}

int orc_bellpack_nslices(int nrows, int slice) { return (nrows + slice - 1) / slice; }

void orc_bellpack_bro(int nrows, int slice, const int* ro, int* bro) {
  int ns = orc_bellpack_nslices(nrows, slice);
  int off = 0;
  for (int s = 0; s < ns; ++s) {
    bro[s] = off;
    int mx = 0;
    for (int i = s * slice; i < std::min(nrows, (s + 1) * slice); ++i) mx = std::max(mx, ro[i + 1] - ro[i]);
    off += mx;
  }
  bro[ns] = off;
}

void orc_bellpack_fill(int nrows, int slice, const int* ro, const int* cols, const double* vals,
                       const int* bro, int* bcols, double* bvals) {
  int ns = orc_bellpack_nslices(nrows, slice);
  /* rows past nrows in the last slice are never written by the reference kernels;
   * its buffers are value-initialised, so they read 0 */
  std::memset(bcols, 0, sizeof(int) * (size_t)bro[ns] * slice);
  if (bvals) std::memset(bvals, 0, sizeof(double) * (size_t)bro[ns] * slice);
  for (int i = 0; i < nrows; ++i) {
    int s = i / slice, l = i % slice;
    long base = (long)bro[s] * slice;
    int w = bro[s + 1] - bro[s];
    int len = ro[i + 1] - ro[i];
    for (int j = 0; j < w; ++j) {
      long e = base + (long)j * slice + l;
      if (j < len) {
        bcols[e] = cols[ro[i] + j];
        if (bvals) bvals[e] = vals[ro[i] + j];
      } else {
        bcols[e] = 0;
        if (bvals) bvals[e] = 0.;
      }
    }
  }
}

void orc_bellpack_masks(int nrows, int slice, const int* ro, const int* dcol, const int* bro,
                        uint8_t* lmask, uint8_t* umask) {
  int ns = orc_bellpack_nslices(nrows, slice);
  std::memset(lmask, 0, (size_t)bro[ns] * slice);
  std::memset(umask, 0, (size_t)bro[ns] * slice);
  for (int i = 0; i < nrows; ++i) {
    int s = i / slice, l = i % slice;
    long base = (long)bro[s] * slice;
    int len = ro[i + 1] - ro[i];
    int dpos = dcol[i] - ro[i];
    for (int j = 0; j < len; ++j) {
      long e = base + (long)j * slice + l;
      lmask[e] = (j < dpos) ? 1 : 0;
      umask[e] = (j > dpos) ? 1 : 0;
    }
  }
}

void orc_bellpack_spmv(int nrows, int slice, const int* bro, const int* bcols, const double* bvals,
                       const double* x, double* y) {
  for (int i = 0; i < nrows; ++i) {
    int s = i / slice, l = i % slice;
    long base = (long)bro[s] * slice;
    int w = bro[s + 1] - bro[s];
    double t = 0.;
    for (int j = 0; j < w; ++j) {
      long e = base + (long)j * slice + l;
      t += bvals[e] * x[bcols[e]];
    }
    y[i] = t;
  }
}

void orc_bellpack_invdiag(int nrows, int slice, const int* bro, const int* bcols,
                          const double* bvals, double* y) {
  for (int i = 0; i < nrows; ++i) {
    int s = i / slice, l = i % slice;
    long base = (long)bro[s] * slice;
    int w = bro[s + 1] - bro[s];
    for (int j = 0; j < w; ++j) {
      long e = base + (long)j * slice + l;
      if (bcols[e] == i && bvals[e] != 0) y[i] = 1. / bvals[e];
    }
  }
}

double orc_dot(int n, const double* x, const double* y) {
  double s = 0.;
  for (int i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

void orc_ilu0_factor(int n, const int* ro, const int* cols, const int* dcol, double* v) {
  std::vector<int> work(n, -1);
  for (int i = 1; i < n; ++i) {
    for (int k = ro[i]; k < dcol[i]; ++k) {
      int kr = cols[k];
      double aik = v[k] / v[dcol[kr]];
      v[k] = aik;
      for (int l = ro[kr]; l < ro[kr + 1]; ++l) work[cols[l]] = l;
      for (int j = k + 1; j < ro[i + 1]; ++j) {
        int kj = work[cols[j]];
        if (kj != -1) v[j] -= aik * v[kj];
      }
      for (int l = ro[kr]; l < ro[kr + 1]; ++l) work[cols[l]] = -1;
    }
  }
}

void orc_filu0_factor(int n, const int* ro, const int* cols, const int* dcol, const double* a, int niter,
                      double* lu) {
  std::vector<double> prev(ro[n]);
  std::vector<int> work(n, -1);
  std::copy(a, a + ro[n], lu);
  for (int it = 0; it < niter; ++it) {
    std::copy(lu, lu + ro[n], prev.begin()); /* values0 = current factor */
    std::copy(a, a + ro[n], lu);             /* values = A */
    for (int i = 1; i < n; ++i) {            /* _factorizeRow (:328-357) */
      for (int k = ro[i]; k < dcol[i]; ++k) {
        const int kr = cols[k];
        const double aik = lu[k] / prev[dcol[kr]];
        lu[k] = aik;
        for (int l = ro[kr]; l < ro[kr + 1]; ++l) work[cols[l]] = l;
        for (int j = k + 1; j < ro[i + 1]; ++j) {
          const int kj = work[cols[j]];
          if (kj != -1) lu[j] -= aik * prev[kj];
        }
        for (int l = ro[kr]; l < ro[kr + 1]; ++l) work[cols[l]] = -1;
      }
    }
  }
}

void orc_ilu0_solve(int n, const int* ro, const int* cols, const int* dcol, const double* lu,
                    const double* y, double* x) {
  std::vector<double> t(n);
  for (int i = 0; i < n; ++i) { /* solveL :181-198 */
    double v = y[i];
    for (int k = ro[i]; k < dcol[i]; ++k) v -= lu[k] * t[cols[k]];
    t[i] = v;
  }
  for (int i = n - 1; i >= 0; --i) { /* solveU :200-231 */
    int dk = dcol[i];
    double v = t[i];
    for (int k = dk + 1; k < ro[i + 1]; ++k) v -= lu[k] * x[cols[k]];
    x[i] = v / lu[dk];
  }
}

int orc_levels_lower(int n, const int* ro, const int* cols, const int* dcol, int* level) {
  int nl = 0;
  for (int i = 0; i < n; ++i) {
    int lv = 0;
    for (int k = ro[i]; k < dcol[i]; ++k) lv = std::max(lv, level[cols[k]] + 1);
    level[i] = lv;
    nl = std::max(nl, lv + 1);
  }
  return n > 0 ? nl : 0;
}

int orc_levels_upper(int n, const int* ro, const int* cols, const int* dcol, int* level) {
  int nl = 0;
  for (int i = n - 1; i >= 0; --i) {
    int lv = 0;
    for (int k = dcol[i] + 1; k < ro[i + 1]; ++k) lv = std::max(lv, level[cols[k]] + 1);
    level[i] = lv;
    nl = std::max(nl, lv + 1);
  }
  return n > 0 ? nl : 0;
}

int orc_color_greedy(int n, const int* ro, const int* cols, int* color) {
  /* symmetrised adjacency: i ~ j if a_ij or a_ji is stored */
  std::vector<std::vector<int>> adj(n);
  for (int i = 0; i < n; ++i)
    for (int k = ro[i]; k < ro[i + 1]; ++k) {
      int j = cols[k];
      if (j != i) { adj[i].push_back(j); adj[j].push_back(i); }
    }
  int nc = 0;
  std::vector<int> used;
  for (int i = 0; i < n; ++i) color[i] = -1;
  for (int i = 0; i < n; ++i) {
    used.assign(nc + 1, 0);
    for (int j : adj[i])
      if (color[j] >= 0) used[color[j]] = 1;
    int c = 0;
    while (used[c]) ++c;
    color[i] = c;
    nc = std::max(nc, c + 1);
  }
  return nc;
}

/* FILU0Preconditioner / FLUFactorisationAlgo (FILU0Preconditioner.h:35-370), the device branch of its
 * triangular solves (on_host_only = false, :226-230 and :297-303): Jacobi sweeps through the masked
 * products addLMult / addUMult and the inverse diagonal of the factor. */
struct FILU0Prec : Precond {
  int n, nsolve;
  const int* ro;
  const int* cols;
  std::vector<int> dcol;
  Vec lu, inv_diag, mx, xk;
  FILU0Prec(const Algebra& alg, int nfactor, int nsolve_) : n(alg.n), nsolve(nsolve_), ro(alg.ro), cols(alg.cols),
                                                            dcol(alg.dcol) {
    lu.assign(ro[n], 0.0);
    if (nfactor == 0) { /* BaseType::factorize (:88-90) */
      std::copy(alg.vals, alg.vals + ro[n], lu.begin());
      orc_ilu0_factor(n, ro, cols, dcol.data(), lu.data());
    } else {
      orc_filu0_factor(n, ro, cols, dcol.data(), alg.vals, nfactor, lu.data());
    }
    inv_diag.assign(n, 1.0); /* assign(m_inv_diag, 1); computeInvDiag (:94-96) */
    orc_invdiag_csr(n, dcol.data(), lu.data(), inv_diag.data());
    mx.assign(n, 0.0);
    xk.assign(n, 0.0);
  }
  /* solve (:236-256): L sweeps then U sweeps */
  void solve(const Algebra& alg, const Vec& x, Vec& y) override {
    alg.copy(x, mx);
    for (int it = 0; it < nsolve; ++it) { /* solveL, device branch (:226-230) */
      alg.copy(mx, xk);
      alg.copy(x, mx);
      orc_addlmult(n, ro, cols, dcol.data(), lu.data(), -1., xk.data(), mx.data());
    }
    alg.copy(mx, y);
    for (int it = 0; it < nsolve; ++it) { /* solveU, device branch (:297-303) */
      alg.copy(y, xk);
      alg.copy(mx, y);
      orc_addumult(n, ro, cols, dcol.data(), lu.data(), -1., xk.data(), y.data());
      alg.pointwiseMult(inv_diag, y, y);
    }
  }
};

/* NeumannPolyPreconditioner (NeumannPolyPreconditioner.h:85-175), factor != 0 (the
 * factor == 0 power method starts from std::random_device: not restated) */
struct NeumannPrec : Precond {
  double factor;
  int order;
  Vec my, myy;
  NeumannPrec(const Algebra& alg, double f, int o) : factor(f), order(o) {
    my.assign(alg.n, 0.0);
    myy.assign(alg.n, 0.0);
  }
  void solve(const Algebra& alg, const Vec& x, Vec& y) override {
    alg.copy(x, y);
    for (int i = 0; i < order; ++i) {
      alg.mult(y, myy); /* evalG */
      alg.copy(y, my);
      alg.axpy(-factor, myy, my);
      alg.axpy(1., x, my);
      alg.copy(my, y);
    }
    alg.scal(factor, y);
  }
};

/* Factored approximate inverse M = G G^T applied as the learned preconditioner is:
 * t = G^T x, y = G t, each a sequential CSR row sum (_seqMult order). */
struct FactorPrec : Precond {
  int n;
  const int *gro, *gcols, *tro, *tcols;
  const double *gvals, *tvals;
  Vec t;
  FactorPrec(int n_, const orc_solve_params* p)
  : n(n_), gro(p->g_ro), gcols(p->g_cols), tro(p->gt_ro), tcols(p->gt_cols), gvals(p->g_vals), tvals(p->gt_vals),
    t(n_, 0.0) {}
  void solve(const Algebra&, const Vec& x, Vec& y) override {
    orc_spmv_csr(n, tro, tcols, tvals, x.data(), t.data());
    orc_spmv_csr(n, gro, gcols, gvals, t.data(), y.data());
  }
};

int orc_solve(const orc_solve_params* p, int n, const int* ro, const int* cols, const double* vals,
              const double* bptr, double* xptr, orc_solve_result* res) {
  std::memset(res, 0, sizeof(*res));
  Algebra alg(n, ro, cols, vals, p->nthreads);
  Vec b(bptr, bptr + n), x(xptr, xptr + n);
  double t0 = now_s();
  Precond* P = nullptr;
  ChebPrec* cheb = nullptr;
  switch (p->precond) {
  case ORC_PREC_DIAG: P = new DiagPrec(alg); break;
  case ORC_PREC_CHEB:
    cheb = new ChebPrec(alg, p->poly_factor, p->poly_order, p->poly_max_iter);
    P = cheb;
    res->cheb_alpha = cheb->alpha;
    res->cheb_beta = cheb->beta;
    break;
  case ORC_PREC_ILU0: P = new ILU0Prec(alg); break;
  case ORC_PREC_FILU0: P = new FILU0Prec(alg, p->filu_factor_niter, p->filu_solver_niter); break;
  case ORC_PREC_NEUMANN: P = new NeumannPrec(alg, p->poly_factor, p->poly_order); break;
  case ORC_PREC_FACTOR:
    if (!p->g_ro || !p->gt_ro) return -1;
    P = new FactorPrec(n, p);
    break;
  default: P = new NonePrec(); break;
  }
  double t1 = now_s();
  Iteration iter(alg, b, p->tol, p->max_iter);
  try {
    if (p->solver == ORC_CG) cg_solve(alg, *P, iter, b, x);
    else bicgs_solve(alg, *P, iter, b, x);
  } catch (const NullValue&) {
    res->error = 1;
  }
  double t2 = now_s();
  delete P;
  res->iterations = iter.it;
  res->status = iter.status ? 1 : 0;
  res->value = iter.getValue();
  res->setup_s = t1 - t0;
  res->solve_s = t2 - t1;
  Vec r(n);
  alg.mult(x, r);
  double rr = 0., bb = 0.;
  for (int i = 0; i < n; ++i) {
    double d = bptr[i] - r[i];
    rr += d * d;
    bb += bptr[i] * bptr[i];
  }
  res->true_rel_res = bb > 0 ? std::sqrt(rr / bb) : 0.;
  std::memcpy(xptr, x.data(), sizeof(double) * n);
  return res->error;
}

void orc_cheb_bounds(int n, const int* ro, const int* cols, const double* vals, double factor, int max_it,
                     double* alpha, double* beta) {
  Algebra alg(n, ro, cols, vals, 1);
  ChebPrec P(alg, factor, 1, max_it);
  *alpha = P.alpha;
  *beta = P.beta;
}

void orc_cheb_apply(int n, const int* ro, const int* cols, const double* vals, double alpha, double beta,
                    int order, const double* xp, double* yp) {
  Algebra alg(n, ro, cols, vals, 1);
  ChebPrec P(alg, 0., order, 0); /* no power iteration: the bounds are set below */
  P.alpha = alpha;
  P.beta = beta;
  Vec x(xp, xp + n), y(n, 0.);
  P.solve(alg, x, y);
  std::memcpy(yp, y.data(), sizeof(double) * n);
}

double orc_time_cg_iters(int n, const int* ro, const int* cols, const double* vals, const double* bptr,
                         int precond, int poly_order, int n_iter, int nthreads) {
  Algebra alg(n, ro, cols, vals, nthreads);
  /* preconditioner set-up is not timed; for Chebyshev a short power method stands in */
  Precond* P = nullptr;
  if (precond == ORC_PREC_CHEB) P = new ChebPrec(alg, 0.5, poly_order, 2);
  else if (precond == ORC_PREC_DIAG) P = new DiagPrec(alg);
  else P = new NonePrec();
  Vec b(bptr, bptr + n), x(n, 0.), p(n), z(n), q(n), r(n);
  alg.copy(b, r);
  double rho = 1.0;
  double t0 = now_s();
  for (int it = 0; it < n_iter; ++it) { /* the loop body of CG.h:111-145 */
    P->solve(alg, r, z);
    double rho1 = alg.dot(r, z);
    double a = rho1 / rho;
    alg.axpy(a, p, z);
    alg.copy(z, p);
    alg.mult(p, q);
    double pq = alg.dot(q, p);
    a = rho1 / pq;
    alg.axpy(a, p, x);
    alg.axpy(-a, q, r);
    rho = rho1;
    volatile double rr = alg.dot(r, r); /* the stop test's dot */
    (void)rr;
  }
  double t = (now_s() - t0) / std::max(1, n_iter);
  delete P;
  return t;
}

double orc_time_spmv(int n, const int* ro, const int* cols, const double* vals, const double* xp,
                     double* yp, int reps, int nthreads) {
  Algebra alg(n, ro, cols, vals, nthreads);
  Vec x(xp, xp + n), y(n);
  alg.mult(x, y);
  double t0 = now_s();
  for (int r = 0; r < reps; ++r) alg.mult(x, y);
  double t = (now_s() - t0) / std::max(1, reps);
  std::memcpy(yp, y.data(), sizeof(double) * n);
  return t;
}

} // extern "C"
