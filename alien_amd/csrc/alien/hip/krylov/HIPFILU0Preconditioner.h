// FILU0 preconditioner of the hip backend: the "iterative" ILU(0) of
// FILU0Preconditioner / FLUFactorisationAlgo
// (src/core/alien/expression/krylov/FILU0Preconditioner.h:35-437).
//
// Factorisation: nb-factor-iter = 0 is the exact ILU(0) factor (here the
// level-set factorisation of ILU0Preconditioner<HIPInternalLinearAlgebra>,
// bit-exact with the sequential one); otherwise that many fixed-point sweeps,
// each restarting from A's values and reading the previous sweep for the other
// rows (one lane per row). Solves: the reference's device branch (:226-230,
// :297-303), nb-solver-iter Jacobi sweeps per triangular factor through the
// masked products addLMult / addUMult and the factor's inverse diagonal,
// called in the reference's order on the algebra.
#pragma once

#include <string>
#include <vector>

#include "../HIPBackEnd.h"
#include "HIPILU0Preconditioner.h"
#include "Krylov.h"

namespace Alien {

template <typename AlgebraT>
class FILU0Preconditioner;

template <>
class FILU0Preconditioner<HIPInternalLinearAlgebra> {
 public:
  typedef HIPInternalLinearAlgebra AlgebraType;
  typedef HIPMatrix MatrixType;
  typedef HIPVector VectorType;

  FILU0Preconditioner(AlgebraType& alg, const MatrixType& A, ITraceMng* = nullptr) : m_alg(alg), m_A(A) {}

  void setParameter(std::string param, int value) {
    if (param.compare("nb-factor-iter") == 0) m_nb_factor_iter = value;
    else if (param.compare("nb-solver-iter") == 0) m_nb_solver_iter = value;
  }
  void setParameter(std::string param, double value) {
    if (param.compare("tol") == 0) m_tol = value;
  }

  void init() {
    const int n = m_A.getNRows();
    const long nnz = m_A.getNnz();
    hipStream_t st = m_alg.stream();
    d_lu.resize(nnz);
    if (m_nb_factor_iter == 0) {
      ILU0Preconditioner<HIPInternalLinearAlgebra> exact(m_alg, m_A);
      exact.init();
      ALIEN_HIP_CHECK(hipMemcpyAsync(d_lu.data(), exact.luValues(), sizeof(double) * nnz, hipMemcpyDeviceToDevice,
                                     st));
      ALIEN_HIP_CHECK(hipStreamSynchronize(st));
    } else {
      alien_hip::DeviceArray<double> prev(nnz);
      ALIEN_HIP_CHECK(hipMemcpyAsync(d_lu.data(), m_A.d_vals.data(), sizeof(double) * nnz, hipMemcpyDeviceToDevice,
                                     st));
      for (int it = 0; it < m_nb_factor_iter; ++it) {
        ALIEN_HIP_CHECK(hipMemcpyAsync(prev.data(), d_lu.data(), sizeof(double) * nnz, hipMemcpyDeviceToDevice, st));
        alien_hip::launch_filu0_sweep(st, n, m_A.d_ro.data(), m_A.d_cols.data(), m_A.d_dcol.data(),
                                      m_A.d_vals.data(), prev.data(), d_lu.data());
        alien_hip::check_launch("filu0_sweep");
      }
      ALIEN_HIP_CHECK(hipStreamSynchronize(st));
    }
    m_alg.allocate(AlgebraType::resource(m_A), m_inv_diag, m_x, m_xk);
    m_alg.assign(m_inv_diag, 1.);
    alien_hip::launch_invdiag_csr(st, n, m_A.d_dcol.data(), d_lu.data(), m_inv_diag.data());
    alien_hip::check_launch("filu0 invdiag");
    if (m_A.spmvLayout() == kLayoutBEllPack) {
      // the factor as a BEllPack matrix of A's profile (the reference clones A in its own
      // layout, FILU0Preconditioner.h:106-110): LU scattered to A's slots, A's L/U masks
      MatrixType& A = const_cast<MatrixType&>(m_A);
      A.ensureBEllPack(true);
      d_lu_bell.resize((size_t)A.bellpackRowOffsets().back() * HIPMatrix::kBEllPackSlice);
      alien_hip::launch_bellpack_fill<double>(st, n, HIPMatrix::kBEllPackSlice, A.d_ro.data(), A.d_cols.data(),
                                              d_lu.data(), nullptr, A.d_bro.data(), nullptr, d_lu_bell.data(),
                                              nullptr, nullptr);
      alien_hip::check_launch("filu0 bellpack factor");
      ALIEN_HIP_CHECK(hipStreamSynchronize(st));
    }
  }

  // solve (:236-256): x -> m_x by L sweeps, m_x -> y by U sweeps
  void solve(AlgebraType& alg, const VectorType& x, VectorType& y) const {
    VectorType& mx = const_cast<VectorType&>(m_x);
    VectorType& xk = const_cast<VectorType&>(m_xk);
    alg.copy(x, mx);
    for (int it = 0; it < m_nb_solver_iter; ++it) { // solveL, device branch
      alg.copy(mx, xk);
      alg.copy(x, mx);
      addMult(alg, 0, xk, mx);
    }
    alg.copy(mx, y);
    for (int it = 0; it < m_nb_solver_iter; ++it) { // solveU, device branch
      alg.copy(y, xk);
      alg.copy(mx, y);
      addMult(alg, 1, xk, y);
      alg.pointwiseMult(m_inv_diag, y, y);
    }
  }
  void solve(const VectorType& x, VectorType& y) const { solve(m_alg, x, y); }

  const double* luValues() const { return d_lu.data(); }
  const VectorType& invDiag() const { return m_inv_diag; }

 private:
  // algebra.addLMult / addUMult(-1, LU, xk, y): A's structure with the factor's values; on a
  // BEllPack matrix through the masks (MatrixInternal::addLMult / addUMult,
  // SYCLBEllPackMatrix.cc:769-911), else the CSR loops (SimpleCSRMatrixMultT.h:193-233)
  void addMult(AlgebraType& alg, int upper, const VectorType& xk, VectorType& y) const {
    if (!d_lu_bell.empty())
      alien_hip::launch_addmult_bellpack(alg.stream(), m_A.getNRows(), HIPMatrix::kBEllPackSlice, -1.,
                                         m_A.d_bro.data(), m_A.d_bellc.data(), d_lu_bell.data(),
                                         upper ? m_A.d_umask.data() : m_A.d_lmask.data(), xk.data(), y.data());
    else
      alien_hip::launch_addmult_csr(alg.stream(), m_A.getNRows(), upper, -1., m_A.d_ro.data(), m_A.d_cols.data(),
                                    m_A.d_dcol.data(), d_lu.data(), xk.data(), y.data());
    alien_hip::check_launch(upper ? "filu0 addUMult" : "filu0 addLMult");
  }

  AlgebraType& m_alg;
  const MatrixType& m_A;
  int m_nb_factor_iter = 0;
  int m_nb_solver_iter = 0;
  double m_tol = 0;
  alien_hip::DeviceArray<double> d_lu;
  alien_hip::DeviceArray<double> d_lu_bell; // the factor in A's BEllPack slots (BEllPack layout only)
  VectorType m_inv_diag, m_x, m_xk;
};

} // namespace Alien
