// ILU(0) preconditioner of the hip backend.
//
// Same factor and same sweeps as LUFactorisationAlgo / ILU0Preconditioner
// (src/core/alien/expression/krylov/ILU0Preconditioner.h:50-323), which work
// on host pointers (y.data()) and so cannot run on device vectors. Here the
// factorisation and both triangular sweeps run on the GPU over level-set
// schedules built once on the host; the results are bit-exact with the
// sequential reference because each row performs the same operations in the
// same order, after all the rows it reads are final.
#pragma once

#include <vector>

#include "../HIPBackEnd.h"
#include "Krylov.h"

namespace Alien {

template <typename AlgebraT>
class ILU0Preconditioner;

template <>
class ILU0Preconditioner<HIPInternalLinearAlgebra> {
 public:
  typedef HIPInternalLinearAlgebra AlgebraType;
  typedef HIPMatrix MatrixType;
  typedef HIPVector VectorType;

  ILU0Preconditioner(AlgebraType& alg, const MatrixType& A, ITraceMng* = nullptr) : m_alg(alg), m_A(A) {}
  // Colour classes of a colour-permuted matrix (colour c = rows [cptr[c],
  // cptr[c+1]), no two coupled rows in one class): used by init() as the
  // schedules, the lower DAG's in increasing colour order, the upper DAG's in
  // decreasing order, in place of the level sets.
  void setColourSchedule(const std::vector<int>& cptr) { m_colours = cptr; }
  void init() {
    const int n = m_A.getNRows();
    hipStream_t st = m_alg.stream();
    if (!m_colours.empty()) {
      const int nc = (int)m_colours.size() - 1;
      m_nl = m_nu = nc;
      m_lptr = m_colours;
      std::vector<int> rows(n);
      for (int i = 0; i < n; ++i) rows[i] = i;
      d_lrows.upload(rows.data(), rows.size(), st);
      m_uptr.assign(nc + 1, 0);
      int pos = 0;
      for (int l = 0; l < nc; ++l) {
        const int c = nc - 1 - l;
        for (int r = m_colours[c]; r < m_colours[c + 1]; ++r) rows[pos++] = r;
        m_uptr[l + 1] = pos;
      }
      d_urows.upload(rows.data(), rows.size(), st);
      factorize(st);
      buildLevelSweeps(st, false); // the colour classes as the levels
      return;
    }
    std::vector<int> level(n);
    m_nl = alien_hip::host_levels(n, m_A.kcol(), m_A.cols(), m_A.dcol(), 0, level.data());
    m_lptr.resize(m_nl + 1);
    std::vector<int> rows(n);
    alien_hip::group_by_level(n, level.data(), m_nl, m_lptr.data(), rows.data());
    d_lrows.upload(rows.data(), rows.size(), st);
    m_nu = alien_hip::host_levels(n, m_A.kcol(), m_A.cols(), m_A.dcol(), 1, level.data());
    m_uptr.resize(m_nu + 1);
    alien_hip::group_by_level(n, level.data(), m_nu, m_uptr.data(), rows.data());
    d_urows.upload(rows.data(), rows.size(), st);
    factorize(st);
    buildLevelSweeps(st, true);
  }
  void update() {}
  void solve(AlgebraType& alg, const VectorType& y, VectorType& x) const {
    if (levelOrdered())
      alien_hip::launch_ilu0_solve_lev<double>(alg.stream(), m_nl, m_lptr.data(), d_lptr.data(), m_lsw, m_nu,
                                               m_uptr.data(), d_uptr.data(), m_usw, y.data(),
                                               const_cast<double*>(m_t.data()), x.data(), nullptr, sfWork());
    else
      alien_hip::launch_ilu0_solve<double>(alg.stream(), m_nl, m_lptr.data(), d_lrows.data(), m_nu, m_uptr.data(),
                                           d_urows.data(), m_A.d_ro.data(), m_A.d_cols.data(), m_A.d_dcol.data(),
                                           d_lu.data(), y.data(), const_cast<double*>(m_t.data()), x.data(),
                                           nullptr, d_lptr.data(), d_uptr.data());
    alien_hip::check_launch("ilu0_solve");
  }
  // The factor copied in schedule order (LevelSweep: level sets, or colour classes) for both
  // sweeps, built when every strictly-lower and strictly-upper row length is at most
  // kLevelMaxSlots.
  bool levelOrdered() const { return m_lsw.k > 0 && m_usw.k > 0; }
  const alien_hip::LevelSweep<double>& lowerSweep() const { return m_lsw; }
  const alien_hip::LevelSweep<double>& upperSweep() const { return m_usw; }
  // the sync-free sweeps' ticket and error flag (null when the work lists are not built); one
  // solve at a time per preconditioner
  int* sfWork() const { return m_lsw.nchunks > 0 ? const_cast<int*>(d_sfwork.data()) : nullptr; }
  // The same sweeps with FP32 values (columns shared), filled from the factor into val / diag
  void fillSweepsF32(hipStream_t st, alien_hip::DeviceArray<float>& lval, alien_hip::DeviceArray<float>& uval,
                     alien_hip::DeviceArray<float>& udiag, alien_hip::LevelSweep<float>& L,
                     alien_hip::LevelSweep<float>& U) const {
    const int n = m_A.getNRows();
    lval.resize((size_t)m_lsw.k * n);
    uval.resize((size_t)m_usw.k * n);
    udiag.resize(n);
    alien_hip::launch_level_sweep_fill<float>(st, n, m_lsw.k, false, d_lrows.data(), m_A.d_ro.data(),
                                              m_A.d_cols.data(), m_A.d_dcol.data(), d_lu.data(), nullptr,
                                              lval.data(), nullptr);
    alien_hip::launch_level_sweep_fill<float>(st, n, m_usw.k, true, d_urows.data(), m_A.d_ro.data(),
                                              m_A.d_cols.data(), m_A.d_dcol.data(), d_lu.data(), nullptr,
                                              uval.data(), udiag.data());
    alien_hip::check_launch("ilu0 level sweeps (FP32)");
    L = {n, m_lsw.k, m_lsw.row, m_lsw.col, lval.data(), nullptr, m_lsw.nchunks, m_lsw.cq, m_lsw.cn, m_lsw.rpl,
         m_lsw.blocks};
    U = {n, m_usw.k, m_usw.row, m_usw.col, uval.data(), udiag.data(), m_usw.nchunks, m_usw.cq, m_usw.cn, m_usw.rpl,
         m_usw.blocks};
  }
  void solve(const VectorType& y, VectorType& x) const { solve(m_alg, y, x); }

  int lowerLevels() const { return m_nl; }
  int upperLevels() const { return m_nu; }
  // the schedules: host level pointers, device row lists
  const int* lowerPtr() const { return m_lptr.data(); }
  const int* upperPtr() const { return m_uptr.data(); }
  const int* lowerRows() const { return d_lrows.data(); }
  const int* upperRows() const { return d_urows.data(); }
  // the level pointers on the device (the sweeps merge runs of small levels)
  const int* lowerPtrDevice() const { return d_lptr.data(); }
  const int* upperPtrDevice() const { return d_uptr.data(); }
  const double* luValues() const { return d_lu.data(); }

 private:
  void factorize(hipStream_t st) {
    d_lptr.upload(m_lptr.data(), m_lptr.size(), st);
    d_uptr.upload(m_uptr.data(), m_uptr.size(), st);
    d_lu.resize(m_A.getNnz());
    ALIEN_HIP_CHECK(hipMemcpyAsync(d_lu.data(), m_A.d_vals.data(), sizeof(double) * m_A.getNnz(),
                                   hipMemcpyDeviceToDevice, st));
    alien_hip::launch_ilu0_factor<double>(st, m_nl, m_lptr.data(), d_lrows.data(), m_A.d_ro.data(),
                                          m_A.d_cols.data(), m_A.d_dcol.data(), d_lu.data());
    alien_hip::check_launch("ilu0_factor");
    m_alg.allocate(AlgebraType::resource(m_A), m_t);
    ALIEN_HIP_CHECK(hipStreamSynchronize(st));
  }
  void buildLevelSweeps(hipStream_t st, bool levels) {
    const int n = m_A.getNRows();
    const int kl = alien_hip::level_sweep_slots(n, m_A.kcol(), m_A.dcol(), false);
    const int ku = alien_hip::level_sweep_slots(n, m_A.kcol(), m_A.dcol(), true);
    if (kl == 0 || ku == 0) return; // the CSR-order kernels
    d_lcol.resize((size_t)kl * n);
    d_lval.resize((size_t)kl * n);
    d_ucol.resize((size_t)ku * n);
    d_uval.resize((size_t)ku * n);
    d_udiag.resize(n);
    alien_hip::launch_level_sweep_fill<double>(st, n, kl, false, d_lrows.data(), m_A.d_ro.data(), m_A.d_cols.data(),
                                               m_A.d_dcol.data(), d_lu.data(), d_lcol.data(), d_lval.data(), nullptr);
    alien_hip::launch_level_sweep_fill<double>(st, n, ku, true, d_urows.data(), m_A.d_ro.data(), m_A.d_cols.data(),
                                               m_A.d_dcol.data(), d_lu.data(), d_ucol.data(), d_uval.data(),
                                               d_udiag.data());
    alien_hip::check_launch("ilu0 level sweeps");
    ALIEN_HIP_CHECK(hipStreamSynchronize(st));
    m_lsw = {n, kl, d_lrows.data(), d_lcol.data(), d_lval.data(), nullptr};
    m_usw = {n, ku, d_urows.data(), d_ucol.data(), d_uval.data(), d_udiag.data()};
    // the sync-free work lists: level sets only (colour classes are few and large), and only
    // where the levels are small on average
    const long nlev = std::max(m_nl, m_nu);
    if (levels && alien_hip::g_ilu_sync_free && nlev > 0 && n <= alien_hip::g_ilu_sf_avg_rows * nlev) {
      int rpl = 1, blocks = 64;
      alien_hip::level_sf_shape(n, nlev, rpl, blocks);
      std::vector<int> q, c;
      alien_hip::level_chunks(m_nl, m_lptr.data(), alien_hip::kWave * rpl, q, c);
      d_lcq.upload(q.data(), q.size(), st);
      d_lcn.upload(c.data(), c.size(), st);
      m_lsw.nchunks = (int)q.size();
      m_lsw.rpl = m_usw.rpl = rpl;
      m_lsw.blocks = m_usw.blocks = blocks;
      alien_hip::level_chunks(m_nu, m_uptr.data(), alien_hip::kWave * rpl, q, c);
      d_ucq.upload(q.data(), q.size(), st);
      d_ucn.upload(c.data(), c.size(), st);
      m_usw.nchunks = (int)q.size();
      const int zero[2] = {0, 0};
      d_sfwork.upload(zero, 2, st);
      ALIEN_HIP_CHECK(hipStreamSynchronize(st));
      m_lsw.cq = d_lcq.data();
      m_lsw.cn = d_lcn.data();
      m_usw.cq = d_ucq.data();
      m_usw.cn = d_ucn.data();
    }
  }
  AlgebraType& m_alg;
  const MatrixType& m_A;
  alien_hip::LevelSweep<double> m_lsw, m_usw;
  alien_hip::DeviceArray<int> d_lcol, d_ucol;
  alien_hip::DeviceArray<int> d_lcq, d_lcn, d_ucq, d_ucn, d_sfwork;
  alien_hip::DeviceArray<double> d_lval, d_uval, d_udiag;
  int m_nl = 0, m_nu = 0;
  std::vector<int> m_lptr, m_uptr, m_colours;
  alien_hip::DeviceArray<int> d_lrows, d_urows, d_lptr, d_uptr;
  alien_hip::DeviceArray<double> d_lu;
  VectorType m_t;
};

} // namespace Alien
