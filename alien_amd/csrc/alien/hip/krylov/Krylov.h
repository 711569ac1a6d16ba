// Krylov solvers and preconditioners, generic on an algebra with Alien's duck
// type (copy/axpy/scal/dot/mult/pointwiseMult/computeInvDiag/assign/exec,
// allocate/free, FutureType, NullValueException).
//
// The numerical recipe of each class is the reference's, step for step,
// because the order of the algebra calls fixes the rounding and the
// iteration counts that the HIP backend must reproduce:
//   Iteration              src/core/alien/expression/krylov/Iteration.h:44-117
//   CG::solve              CG.h:58-150
//   BiCGStab::solve        BiCGStab.h:151-300
//   DiagPreconditioner     DiagPreconditioner.h:49-67
//   ChebyshevPreconditioner ChebyshevPreconditioner.h:73-202 (init, power method), 392-444 (_algo2)
// The classes carry the reference's names and constructor arguments so that
// test_krylov-style drivers compile unchanged against this header.
#pragma once

#include <cmath>
#include <random>
#include <string>

namespace Alien {

class ITraceMng;

template <typename AlgebraT>
class Iteration {
 public:
  typedef AlgebraT AlgebraType;
  typedef typename AlgebraType::Vector VectorType;
  typedef typename AlgebraType::FutureType FutureType;
  typedef double ValueType;

  Iteration(AlgebraType& alg, const VectorType& b, ValueType tol, int max_iter, ITraceMng* = nullptr)
  : m_alg(alg), m_max_it(max_iter), m_tol(tol), m_future(m_value) {
    m_alg.dot(b, b, m_future);
    m_nrm2_b = m_future.get();
    m_criteria = m_tol * m_tol * m_nrm2_b;
    m_sqrt_nrm2_b = std::sqrt(m_nrm2_b);
    m_value = m_criteria + 1;
    m_status = (m_nrm2_b == 0);
  }
  bool nullRhs() const { return m_nrm2_b == 0.; }
  bool first() const { return m_iter == 0; }
  // iter >= max: stop without a dot; else status = (r,r) < tol^2 (b,b)
  bool stop(const VectorType& r) {
    if (m_iter >= m_max_it) return true;
    m_alg.dot(r, r, m_future);
    m_status = m_future.get() < m_criteria;
    return m_status;
  }
  void operator++() { ++m_iter; }
  ValueType getValue() const { return m_sqrt_nrm2_b == 0 ? 0. : std::sqrt(m_value) / m_sqrt_nrm2_b; }
  int operator()() const { return m_iter; }
  bool getStatus() const { return m_status; }
  // used by the fused solvers, which evaluate the stop test on the device
  void setState(int it, double value, bool status) {
    m_iter = it;
    m_value = value;
    m_status = status;
  }
  ValueType criteria() const { return m_criteria; }
  int maxIter() const { return m_max_it; }
  ValueType normB2() const { return m_nrm2_b; }

 private:
  AlgebraType& m_alg;
  int m_max_it = 0;
  ValueType m_tol = 0;
  int m_iter = 0;
  ValueType m_value = 0;
  FutureType m_future;
  ValueType m_criteria = 0, m_nrm2_b = 0, m_sqrt_nrm2_b = 0;
  bool m_status = false;
};

template <typename AlgebraT>
class CG {
 public:
  typedef AlgebraT AlgebraType;
  typedef typename AlgebraType::Matrix MatrixType;
  typedef typename AlgebraType::Vector VectorType;
  typedef double ValueType;

  explicit CG(AlgebraType& alg, ITraceMng* = nullptr) : m_alg(alg) {}
  void setOutputLevel(int) {}

  // preconditioned CG (Saad); the first update is peeled out of the loop
  template <typename PrecondT, typename IterT>
  int solve(PrecondT& M, IterT& it, const MatrixType& A, const VectorType& b, VectorType& x) {
    if (it.nullRhs()) return 0;
    ValueType rho = 0, rho1 = 0, alpha = 0;
    VectorType p, z, q, r;
    m_alg.allocate(AlgebraType::resource(A), p, z, q, r);
    // r = b - A x
    m_alg.copy(b, r);
    m_alg.mult(A, x, p);
    m_alg.axpy(-1., p, r);
    // first step: p = M^-1 r
    m_alg.exec(M, r, z);
    rho1 = m_alg.dot(r, z);
    m_alg.copy(z, p);
    m_alg.mult(A, p, q);
    alpha = m_alg.dot(p, q);
    if (alpha == 0) return breakdown(it, r, p, z, q);
    alpha = rho1 / alpha;
    m_alg.axpy(alpha, p, x);
    m_alg.axpy(-alpha, q, r);
    rho = rho1;
    ++it;
    while (!it.stop(r)) {
      m_alg.exec(M, r, z);
      rho1 = m_alg.dot(r, z);
      alpha = rho1 / rho;
      m_alg.axpy(alpha, p, z); // p = z + (rho1/rho) p, formed in z then copied
      m_alg.copy(z, p);
      m_alg.mult(A, p, q);
      alpha = m_alg.dot(q, p);
      if (alpha == 0) return breakdown(it, r, p, z, q);
      alpha = rho1 / alpha;
      m_alg.axpy(alpha, p, x);
      m_alg.axpy(-alpha, q, r);
      rho = rho1;
      ++it;
    }
    m_alg.free(p, z, q, r);
    return 0;
  }

  // The same iteration with the dot products taken as FutureType: both dots of
  // the first step are launched before the first read-back (CG::solve2,
  // CG.h:152-244). Same arithmetic as solve, same iterates.
  template <typename PrecondT, typename IterT>
  int solve2(PrecondT& M, IterT& it, const MatrixType& A, const VectorType& b, VectorType& x) {
    typedef typename AlgebraType::FutureType FutureType;
    if (it.nullRhs()) return 0;
    ValueType rho = 0, rho1 = 0, alpha = 0;
    FutureType frho1(rho1), falpha(alpha);
    VectorType p, z, q, r;
    m_alg.allocate(AlgebraType::resource(A), p, z, q, r);
    m_alg.copy(b, r);
    m_alg.mult(A, x, p);
    m_alg.axpy(-1., p, r);
    m_alg.exec(M, r, z);
    m_alg.copy(z, p);
    m_alg.mult(A, p, q);
    m_alg.dot(r, p, frho1);
    m_alg.dot(p, q, falpha);
    if (falpha.get() == 0) return breakdown(it, r, p, z, q);
    alpha = frho1.get() / alpha;
    m_alg.axpy(alpha, p, x);
    m_alg.axpy(-alpha, q, r);
    rho = rho1;
    ++it;
    while (!it.stop(r)) {
      m_alg.exec(M, r, z);
      m_alg.dot(r, z, frho1);
      alpha = frho1.get() / rho;
      m_alg.axpy(alpha, p, z);
      m_alg.copy(z, p);
      m_alg.mult(A, p, q);
      m_alg.dot(p, q, falpha);
      if (falpha.get() == 0) return breakdown(it, r, p, z, q);
      alpha = rho1 / alpha;
      m_alg.axpy(alpha, p, x);
      m_alg.axpy(-alpha, q, r);
      rho = rho1;
      ++it;
    }
    m_alg.free(p, z, q, r);
    return 0;
  }

 private:
  template <typename IterT>
  int breakdown(IterT& it, VectorType& r, VectorType& p, VectorType& z, VectorType& q) {
    if (it.stop(r)) {
      ++it;
      m_alg.free(p, z, q, r);
      return 0;
    }
    m_alg.free(p, z, q, r);
    throw typename AlgebraType::NullValueException("alpha");
  }
  AlgebraType& m_alg;
};

template <typename AlgebraT>
class BiCGStab {
 public:
  typedef AlgebraT AlgebraType;
  typedef typename AlgebraType::Matrix MatrixType;
  typedef typename AlgebraType::Vector VectorType;
  typedef double ValueType;

  explicit BiCGStab(AlgebraType& alg, ITraceMng* = nullptr) : m_alg(alg) {}
  void setOutputLevel(int) {}

  // right-preconditioned BiCGStab; r0 is the initial residual
  template <typename PrecondT, typename IterT>
  int solve(PrecondT& M, IterT& it, const MatrixType& A, const VectorType& b, VectorType& x) {
    if (it.nullRhs()) return 0;
    typedef typename AlgebraType::NullValueException NullValue;
    ValueType rho = 0, rho1 = 0, alpha = 0, beta = 0, omega = 0;
    VectorType p, phat, s, shat, t, v, r, r0;
    m_alg.allocate(AlgebraType::resource(A), p, phat, s, shat, t, v, r, r0);
    struct Release {
      AlgebraType& a;
      VectorType &p, &phat, &s, &shat, &t, &v, &r, &r0;
      ~Release() { a.free(p, phat, s, shat, t, v, r, r0); }
    } release{m_alg, p, phat, s, shat, t, v, r, r0};

    m_alg.copy(b, r);
    m_alg.mult(A, x, r0);
    m_alg.axpy(-1., r0, r);
    m_alg.copy(r, r0);
    m_alg.copy(r, p);
    rho1 = m_alg.dot(r, r0);

    m_alg.exec(M, p, phat);
    m_alg.mult(A, phat, v);
    alpha = m_alg.dot(v, r0);
    if (alpha == 0) throw NullValue("alpha");
    alpha = rho1 / alpha;
    m_alg.copy(r, s);
    m_alg.axpy(-alpha, v, s);
    if (it.stop(s)) { // half-step exit of the first sweep counts one iteration
      ++it;
      m_alg.axpy(alpha, phat, x);
      return 0;
    }
    m_alg.exec(M, s, shat);
    m_alg.mult(A, shat, t);
    omega = m_alg.dot(t, s);
    beta = m_alg.dot(t, t);
    if (beta == 0) {
      if (it.stop(r)) {
        ++it;
        m_alg.axpy(alpha, phat, x);
        return 0;
      }
      throw NullValue("beta");
    }
    omega = omega / beta;
    m_alg.axpy(omega, shat, x);
    m_alg.axpy(alpha, phat, x);
    m_alg.copy(s, r);
    m_alg.axpy(-omega, t, r);
    rho = rho1;
    ++it;

    while (!it.stop(r)) {
      rho1 = m_alg.dot(r, r0);
      beta = (rho1 / rho) * (alpha / omega);
      m_alg.axpy(-omega, v, p); // p = r + beta (p - omega v)
      m_alg.scal(beta, p);
      m_alg.axpy(1., r, p);
      if (rho == 0) throw NullValue("rho");
      m_alg.exec(M, p, phat);
      m_alg.mult(A, phat, v);
      alpha = m_alg.dot(v, r0);
      if (alpha == 0) throw NullValue("alpha");
      alpha = rho1 / alpha;
      m_alg.copy(r, s);
      m_alg.axpy(-alpha, v, s);
      if (it.stop(s)) { // half-step exit inside the loop does not count an iteration
        m_alg.axpy(alpha, phat, x);
        return 0;
      }
      m_alg.exec(M, s, shat);
      m_alg.mult(A, shat, t);
      omega = m_alg.dot(t, s);
      beta = m_alg.dot(t, t);
      if (beta == 0) {
        if (it.stop(s)) {
          m_alg.axpy(alpha, phat, x);
          return 0;
        }
        throw NullValue("beta");
      }
      omega = omega / beta;
      m_alg.axpy(omega, shat, x);
      m_alg.axpy(alpha, phat, x);
      m_alg.copy(s, r);
      m_alg.axpy(-omega, t, r);
      rho = rho1;
      ++it;
    }
    return 0;
  }

  // The same iteration with FutureType dots (BiCGStab::solve2, BiCGStab.h:302-457):
  // t.s and t.t are launched together, one read-back each when needed. Same
  // arithmetic and exits as solve.
  template <typename PrecondT, typename IterT>
  int solve2(PrecondT& M, IterT& it, const MatrixType& A, const VectorType& b, VectorType& x) {
    typedef typename AlgebraType::FutureType FutureType;
    typedef typename AlgebraType::NullValueException NullValue;
    if (it.nullRhs()) return 0;
    ValueType rho = 0, rho1 = 0, alpha = 0, beta = 0, omega = 0;
    FutureType frho1(rho1), falpha(alpha), fbeta(beta), fomega(omega);
    VectorType p, phat, s, shat, t, v, r, r0;
    m_alg.allocate(AlgebraType::resource(A), p, phat, s, shat, t, v, r, r0);
    struct Release {
      AlgebraType& a;
      VectorType &p, &phat, &s, &shat, &t, &v, &r, &r0;
      ~Release() { a.free(p, phat, s, shat, t, v, r, r0); }
    } release{m_alg, p, phat, s, shat, t, v, r, r0};

    m_alg.copy(b, r);
    m_alg.mult(A, x, r0);
    m_alg.axpy(-1., r0, r);
    m_alg.copy(r, r0);
    m_alg.copy(r, p);
    m_alg.dot(r, r0, frho1);
    m_alg.exec(M, p, phat);
    m_alg.mult(A, phat, v);
    m_alg.dot(v, r0, falpha);
    if (falpha.get() == 0) throw NullValue("alpha");
    alpha = frho1.get() / alpha;
    m_alg.copy(r, s);
    m_alg.axpy(-alpha, v, s);
    if (it.stop(s)) {
      ++it;
      m_alg.axpy(alpha, phat, x);
      return 0;
    }
    m_alg.exec(M, s, shat);
    m_alg.mult(A, shat, t);
    m_alg.dot(t, s, fomega);
    m_alg.dot(t, t, fbeta);
    if (fbeta.get() == 0) {
      if (it.stop(r)) {
        ++it;
        m_alg.axpy(alpha, phat, x);
        return 0;
      }
      throw NullValue("beta");
    }
    omega = fomega.get() / beta;
    m_alg.axpy(omega, shat, x);
    m_alg.axpy(alpha, phat, x);
    m_alg.copy(s, r);
    m_alg.axpy(-omega, t, r);
    rho = rho1;
    ++it;

    while (!it.stop(r)) {
      m_alg.dot(r, r0, frho1);
      beta = (frho1.get() / rho) * (alpha / omega);
      m_alg.axpy(-omega, v, p);
      m_alg.scal(beta, p);
      m_alg.axpy(1., r, p);
      if (rho == 0) throw NullValue("rho");
      m_alg.exec(M, p, phat);
      m_alg.mult(A, phat, v);
      m_alg.dot(v, r0, falpha);
      if (falpha.get() == 0) throw NullValue("alpha");
      alpha = rho1 / alpha;
      m_alg.copy(r, s);
      m_alg.axpy(-alpha, v, s);
      if (it.stop(s)) {
        m_alg.axpy(alpha, phat, x);
        return 0;
      }
      m_alg.exec(M, s, shat);
      m_alg.mult(A, shat, t);
      m_alg.dot(t, s, fomega);
      m_alg.dot(t, t, fbeta);
      if (fbeta.get() == 0) {
        if (it.stop(s)) {
          m_alg.axpy(alpha, phat, x);
          return 0;
        }
        throw NullValue("beta");
      }
      omega = fomega.get() / beta;
      m_alg.axpy(omega, shat, x);
      m_alg.axpy(alpha, phat, x);
      m_alg.copy(s, r);
      m_alg.axpy(-omega, t, r);
      rho = rho1;
      ++it;
    }
    return 0;
  }

 private:
  AlgebraType& m_alg;
};

template <typename AlgebraT>
class DiagPreconditioner {
 public:
  typedef AlgebraT AlgebraType;
  typedef typename AlgebraType::Matrix MatrixType;
  typedef typename AlgebraType::Vector VectorType;

  DiagPreconditioner(AlgebraType& alg, const MatrixType& A) : m_alg(alg), m_A(A) {}
  void init() {
    m_alg.allocate(AlgebraType::resource(m_A), m_inv_diag);
    m_alg.assign(m_inv_diag, 1.);
    m_alg.computeInvDiag(m_A, m_inv_diag);
  }
  void update() {}
  template <typename Alg>
  void solve(Alg& alg, const VectorType& x, VectorType& y) const {
    alg.pointwiseMult(m_inv_diag, x, y);
  }
  const VectorType& invDiag() const { return m_inv_diag; }

 private:
  AlgebraType& m_alg;
  const MatrixType& m_A;
  VectorType m_inv_diag;
};

// Neumann polynomial preconditioner (NeumannPolyPreconditioner.h:85-175):
// y = factor * sum_{k<=order} (I - factor A)^k x, evaluated as the reference
// does (y = x; order times: m_y = y; m_y += -factor (A y); m_y += x; y = m_y;
// then y *= factor). With factor == 0 the factor is 1 / |A|_2 from a power
// method whose start vector comes from std::random_device, as in the
// reference (non-deterministic; test_krylov passes a non-zero factor).
template <typename AlgebraT>
class NeumannPolyPreconditioner {
 public:
  typedef AlgebraT AlgebraType;
  typedef typename AlgebraType::Matrix MatrixType;
  typedef typename AlgebraType::Vector VectorType;

  NeumannPolyPreconditioner(AlgebraType& alg, const MatrixType& A, double factor, int order, int factor_max_iter)
  : m_alg(alg), m_A(A), m_factor(factor), m_order(order), m_factor_max_iter(factor_max_iter) {}
  void init() {
    m_alg.allocate(AlgebraType::resource(m_A), m_y, m_yy);
    m_alg.assign(m_y, 0.);
    m_alg.assign(m_yy, 0.);
    if (m_factor == 0) computeFactor();
  }
  void update() {}
  template <typename Alg>
  void solve(Alg& alg, const VectorType& x, VectorType& y) const {
    VectorType& my = const_cast<VectorType&>(m_y);
    VectorType& myy = const_cast<VectorType&>(m_yy);
    alg.copy(x, y);
    for (int i = 0; i < m_order; ++i) {
      alg.mult(m_A, y, myy); // evalG: yy = A y ; m_y = y - factor yy
      alg.copy(y, my);
      alg.axpy(-m_factor, myy, my);
      alg.axpy(1., x, my);
      alg.copy(my, y);
    }
    alg.scal(m_factor, y);
  }
  double factor() const { return m_factor; }

 private:
  void computeFactor() {
    VectorType x;
    m_alg.allocate(AlgebraType::resource(m_A), x);
    std::random_device rd;
    std::mt19937 gen(rd());
    std::uniform_real_distribution<> dis(-1., 1.);
    m_alg.assign(x, [&](int) { return dis(gen); });
    m_alg.mult(m_A, x, m_y);
    double nk = m_alg.norm2(m_y);
    for (int i = 0; i < m_factor_max_iter; ++i) {
      m_alg.copy(m_y, x);
      m_alg.mult(m_A, x, m_y);
      const double nn = m_alg.norm2(m_y);
      m_factor = nn / nk;
      nk = nn;
    }
    m_factor = 1. / m_factor;
    m_alg.free(x);
  }
  AlgebraType& m_alg;
  const MatrixType& m_A;
  double m_factor;
  int m_order, m_factor_max_iter;
  VectorType m_y, m_yy;
};

// Chebyshev polynomial preconditioner, non-Saad variant. Quirks kept on
// purpose (they decide the iteration counts): the interval is [alpha, beta]
// with alpha overwritten by beta/factor when factor != 0; the apply starts
// from y = x, not from 0; the power-method start vector is a default-seeded
// mt19937 through uniform_real_distribution(-1, 1), drawn on the host in
// index order.
template <typename AlgebraT>
class ChebyshevPreconditioner {
 public:
  typedef AlgebraT AlgebraType;
  typedef typename AlgebraType::Matrix MatrixType;
  typedef typename AlgebraType::Vector VectorType;
  typedef double ValueType;

  ChebyshevPreconditioner(AlgebraType& alg, const MatrixType& A, ValueType factor, int order,
                          int factor_max_iter, ITraceMng* = nullptr)
  : m_alg(alg), m_A(A), m_factor(factor), m_order(order), m_max_it(factor_max_iter) {}
  ~ChebyshevPreconditioner() { m_alg.free(m_inv_diag, m_y, m_r, m_p, m_z); }
  void setOutputLevel(int) {}

  void init() {
    m_alg.allocate(AlgebraType::resource(m_A), m_y, m_r, m_p, m_z);
    m_alg.allocate(AlgebraType::resource(m_A), m_inv_diag);
    m_alg.assign(m_inv_diag, 1.);
    m_alg.computeInvDiag(m_A, m_inv_diag);
    if (m_beta == 0) computeBeta();
    if (m_alpha == 0) computeAlpha();
    if (m_factor != 0.) m_alpha = m_beta / m_factor;
    m_delta = (m_beta - m_alpha) / 2;
    m_theta = (m_beta + m_alpha) / 2;
  }

  // largest eigenvalue of D^-1 A by the power method
  void computeBeta() {
    VectorType x;
    m_alg.allocate(AlgebraType::resource(m_A), x);
    randomStart(x);
    for (int i = 0; i < m_max_it; ++i) {
      m_alg.mult(m_A, x, m_y);
      m_alg.pointwiseMult(m_inv_diag, m_y, m_y);
      ValueType xAx = m_alg.dot(m_y, x);
      ValueType xdx = m_alg.dot(x, x);
      ValueType nrm = m_alg.norm2(m_y);
      m_alg.scal(1 / nrm, m_y);
      m_alg.copy(m_y, x);
      m_beta = xAx / xdx;
    }
    m_alg.free(x);
  }
  // smallest eigenvalue, by the power method on D^-1 A - beta I
  void computeAlpha() {
    VectorType x;
    m_alg.allocate(AlgebraType::resource(m_A), x);
    randomStart(x);
    for (int i = 0; i < m_max_it; ++i) {
      m_alg.mult(m_A, x, m_y);
      m_alg.pointwiseMult(m_inv_diag, m_y, m_y);
      m_alg.axpy(-m_beta, x, m_y);
      ValueType xAx = m_alg.dot(m_y, x);
      ValueType xdx = m_alg.dot(x, x);
      m_alpha = xAx / xdx;
      ValueType nrm = m_alg.norm2(m_y);
      m_alg.scal(1 / nrm, m_y);
      m_alg.copy(m_y, x);
    }
    m_alpha += m_beta;
    m_alg.free(x);
  }

  // degree-m Chebyshev sweep on A y = x from y = x
  template <typename Alg>
  void solve(Alg& alg, const VectorType& x, VectorType& y) const {
    (void)alg;
    const ValueType d = (m_beta + m_alpha) / 2;
    const ValueType c = (m_beta - m_alpha) / 2;
    m_alg.copy(x, y);
    m_alg.mult(m_A, y, m_y);
    m_alg.copy(x, m_r);
    m_alg.axpy(-1., m_y, m_r);
    m_alg.pointwiseMult(m_inv_diag, m_r, m_z);
    m_alg.copy(m_z, m_p);
    ValueType a = 2 / d;
    m_alg.axpy(a, m_p, y);
    for (int i = 1; i < m_order; ++i) {
      m_alg.mult(m_A, y, m_y);
      m_alg.copy(x, m_r);
      m_alg.axpy(-1., m_y, m_r);
      m_alg.pointwiseMult(m_inv_diag, m_r, m_z);
      ValueType b = a * (c / 2.) * (c / 2.);
      a = 1. / (d - b);
      m_alg.scal(b, m_p);
      m_alg.axpy(1., m_z, m_p);
      m_alg.axpy(a, m_p, y);
    }
  }

  ValueType alpha() const { return m_alpha; }
  ValueType beta() const { return m_beta; }
  int order() const { return m_order; }
  const VectorType& invDiag() const { return m_inv_diag; }

  // step coefficients of the sweep: a_0 = 2/d, then (b_i, a_i) for i >= 1
  void coefficients(ValueType* a_out, ValueType* b_out) const {
    const ValueType d = (m_beta + m_alpha) / 2;
    const ValueType c = (m_beta - m_alpha) / 2;
    ValueType a = 2 / d;
    a_out[0] = a;
    b_out[0] = 0;
    for (int i = 1; i < m_order; ++i) {
      ValueType b = a * (c / 2.) * (c / 2.);
      a = 1. / (d - b);
      a_out[i] = a;
      b_out[i] = b;
    }
  }

 private:
  void randomStart(VectorType& x) {
    std::mt19937 gen;
    std::uniform_real_distribution<> dis(-1., 1.);
    m_alg.assign(x, [&](std::size_t) { return dis(gen); });
    ValueType nrm = m_alg.norm2(x);
    m_alg.scal(1 / nrm, x);
  }

  AlgebraType& m_alg;
  const MatrixType& m_A;
  ValueType m_factor = 0, m_alpha = 0, m_beta = 0, m_theta = 0, m_delta = 0;
  int m_order, m_max_it;
  VectorType m_inv_diag;
  mutable VectorType m_y, m_r, m_p, m_z;
};

} // namespace Alien
