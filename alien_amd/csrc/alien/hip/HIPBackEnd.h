// The `hip` backend of Alien: backend tag, AlgebraTraits, vector, matrix and
// the internal linear algebra whose method set is the duck type that Alien's
// Krylov templates call (src/core/alien/expression/krylov/*.h).
//
// Mirrors:
//   tag + traits       SYCLBackEnd.h:66-97 / SimpleCSRBackEnd.h:62-94
//   vector             SYCLVector (sycl/data/SYCLVector.h:47-150), SimpleCSRVector.h:96-163
//   matrix             SYCLBEllPackMatrix (sycl/data/SYCLBEllPackMatrix.h:63-225)
//   algebra            SYCLInternalLinearAlgebra (sycl/algebra/SYCLInternalLinearAlgebra.h:46-141)
//
// Inside an Alien build (ALIEN_HIP_HAVE_ALGEBRA_TRAITS defined, Alien's include
// directories on the path) the vector and matrix derive from Alien's
// IVectorImpl / IMatrixImpl and are constructed by MultiVectorImpl /
// MultiMatrixImpl (`new matrix_type(this)`, MultiMatrixImpl.h:318), then filled
// by the SimpleCSR -> hip converters through build() / setValues(); the
// algebras derive from IInternalLinearAlgebra / IInternalLinearAlgebraExpr.
// Standalone (this repository's library and tests) the same classes are built
// directly from a host CSR view, and the few Alien names the traits use are
// declared below.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../alien_hip_internal.h"
#include "../../kernels/launch.h"
#include "HIPContext.h"

#ifdef ALIEN_HIP_HAVE_ALGEBRA_TRAITS
#include <alien/core/backend/BackEnd.h>
#include <alien/core/backend/IInternalLinearAlgebraExprT.h>
#include <alien/core/backend/IInternalLinearAlgebraT.h>
#include <alien/core/impl/IMatrixImpl.h>
#include <alien/core/impl/IVectorImpl.h>
#include <alien/core/impl/MultiMatrixImpl.h>
#include <alien/core/impl/MultiVectorImpl.h>
#else
namespace Arccore {
namespace MessagePassing {
class IMessagePassingMng;
} // namespace MessagePassing
} // namespace Arccore
#endif

namespace Alien {

#ifndef ALIEN_HIP_HAVE_ALGEBRA_TRAITS
// Declared by Alien in alien/core/backend/BackEnd.h (BackEndId is Arccore's String there,
// constructible from the same literal) and alien/utils/Precomp.h.
typedef std::string BackEndId;
template <typename Tag> struct AlgebraTraits;
using Arccore::MessagePassing::IMessagePassingMng;
class MultiMatrixImpl;
class MultiVectorImpl;
#endif

namespace BackEnd {
namespace tag {
struct hip {};
} // namespace tag
} // namespace BackEnd

template <typename T> class HIPVectorT;
class HIPMatrix;
class HIPInternalLinearAlgebra;
class HIPInternalLinearAlgebraExpr;

// Backend traits (SYCLBackEnd.h:76-96, SimpleCSRBackEnd.h:62-94): inside Alien the
// algebra types are its virtual interfaces, which the hip algebras implement.
template <>
struct AlgebraTraits<BackEnd::tag::hip> {
  typedef HIPMatrix matrix_type;
  typedef HIPVectorT<double> vector_type;
#ifdef ALIEN_HIP_HAVE_ALGEBRA_TRAITS
  typedef IInternalLinearAlgebra<HIPMatrix, HIPVectorT<double>> algebra_type;
  typedef IInternalLinearAlgebraExpr<HIPMatrix, HIPVectorT<double>> algebra_expr_type;
#else
  typedef HIPInternalLinearAlgebra algebra_type;
  typedef HIPInternalLinearAlgebraExpr algebra_expr_type;
#endif
  typedef double value_type;
  // a new algebra on the library's default context (one HIP stream per process); the
  // caller owns it, as LinearAlgebra<tag> owns the one algebra_factory returns
  static algebra_type* algebra_factory(IMessagePassingMng* p_mng = nullptr);
  static algebra_expr_type* algebra_expr_factory(IMessagePassingMng* p_mng = nullptr);
  static BackEndId name() { return "hip"; }
};

// "Resource" of a vector: its size and the context that owns its stream
// (Alien uses the row VectorDistribution, SimpleCSRInternalLinearAlgebra.cc:74-78).
struct HIPResource {
  int n = 0;
  alien_hip_ctx* ctx = nullptr;
};

template <typename T>
class HIPVectorT
#ifdef ALIEN_HIP_HAVE_ALGEBRA_TRAITS
: public IVectorImpl
#endif
{
 public:
  typedef T ValueType;
  static const bool on_host_only = false;

#ifdef ALIEN_HIP_HAVE_ALGEBRA_TRAITS
  HIPVectorT() : IVectorImpl(nullptr, AlgebraTraits<BackEnd::tag::hip>::name()) {}
  // created by MultiVectorImpl, which then calls init(distribution, true)
  explicit HIPVectorT(const MultiVectorImpl* multi_impl)
  : IVectorImpl(multi_impl, AlgebraTraits<BackEnd::tag::hip>::name()) {}
  HIPVectorT(const HIPResource& r) : HIPVectorT() { init(r, true); }
  // IVectorImpl::init: the local rows of the distribution, on the default context
  void init(const VectorDistribution& dist, bool need_alloc) override {
    init(HIPResource{dist.localSize(), alien_hip::default_ctx()}, need_alloc);
  }
  void clear() override { release(); }
#else
  HIPVectorT() = default;
  explicit HIPVectorT(const MultiVectorImpl*) {}
  HIPVectorT(const HIPResource& r) { init(r, true); }
  void clear() { release(); }
#endif

  // zero-filled like SimpleCSRVector::init (SimpleCSRVector.h:111-131)
  void init(const HIPResource& r, bool need_alloc) {
    m_ctx = r.ctx;
    m_n = r.n;
    if (need_alloc) {
      m_data.resize((size_t)m_n);
      if (m_n) ALIEN_HIP_CHECK(hipMemsetAsync(m_data.data(), 0, sizeof(T) * m_n, m_ctx->stream));
    }
  }
  void release() {
    m_data.reset();
    m_n = 0;
  }
  int scalarizedLocalSize() const { return m_n; }
  int getAllocSize() const { return m_n; }
  T* getDataPtr() { return m_data.data(); }
  const T* getDataPtr() const { return m_data.data(); }
  T* data() { return m_data.data(); }
  const T* data() const { return m_data.data(); }
  alien_hip_ctx* ctx() const { return m_ctx; }
  HIPResource resource() const { return HIPResource{m_n, m_ctx}; }

  void setValues(const T* host) {
    if (m_n) ALIEN_HIP_CHECK(hipMemcpyAsync(m_data.data(), host, sizeof(T) * m_n, hipMemcpyHostToDevice, m_ctx->stream));
    ALIEN_HIP_CHECK(hipStreamSynchronize(m_ctx->stream));
  }
  void copyValuesTo(T* host) const {
    if (m_n) ALIEN_HIP_CHECK(hipMemcpyAsync(host, m_data.data(), sizeof(T) * m_n, hipMemcpyDeviceToHost, m_ctx->stream));
    ALIEN_HIP_CHECK(hipStreamSynchronize(m_ctx->stream));
  }
  // y[i] = f(i) evaluated on the host, in index order, then uploaded: the
  // order of the host-side random stream is the reference's
  // (SYCLVector::apply, sycl/data/SYCLVector.h:145-150).
  template <typename Lambda>
  void apply(const Lambda& f) {
    std::vector<T> h(m_n);
    for (int i = 0; i < m_n; ++i) h[i] = f(i);
    setValues(h.data());
  }

 private:
  alien_hip::DeviceArray<T> m_data;
  int m_n = 0;
  alien_hip_ctx* m_ctx = nullptr;
};
typedef HIPVectorT<double> HIPVector;

// Host CSR input (Alien's SimpleCSR profile + values, ascending columns).
struct CSRHostView {
  int nrows = 0;
  const int* row_offset = nullptr;
  const int* cols = nullptr;
  const double* vals = nullptr;
};

enum HIPLayout { kLayoutCSR = ALIEN_HIP_LAYOUT_CSR, kLayoutBEllPack = ALIEN_HIP_LAYOUT_BELLPACK,
                 kLayoutRBCSR = ALIEN_HIP_LAYOUT_RBCSR, kLayoutRBDIA = ALIEN_HIP_LAYOUT_RBDIA };

class HIPMatrix
#ifdef ALIEN_HIP_HAVE_ALGEBRA_TRAITS
: public IMatrixImpl
#endif
{
 public:
  typedef double ValueType;
  typedef BackEnd::tag::hip TagType;
  static const bool on_host_only = false;
  static constexpr int kBEllPackSlice = 1024; // BEllPackStructInfo<1024,int>

#ifdef ALIEN_HIP_HAVE_ALGEBRA_TRAITS
  HIPMatrix() : IMatrixImpl(nullptr, AlgebraTraits<BackEnd::tag::hip>::name()) {}
  // MultiMatrixImpl::getImpl (MultiMatrixImpl.h:318) creates the empty matrix; the
  // SimpleCSR -> hip converter fills it with build()
  explicit HIPMatrix(const MultiMatrixImpl* multi_impl)
  : IMatrixImpl(multi_impl, AlgebraTraits<BackEnd::tag::hip>::name()) {}
  HIPMatrix(alien_hip_ctx* ctx, const CSRHostView& a, int spmv_layout = kLayoutRBCSR, bool borrow_host = false)
  : HIPMatrix() {
    build(ctx, a, spmv_layout, borrow_host);
  }
  void clear() override { release(); }
#else
  HIPMatrix() = default;
  explicit HIPMatrix(const MultiMatrixImpl*) {}
  HIPMatrix(alien_hip_ctx* ctx, const CSRHostView& a, int spmv_layout = kLayoutRBCSR, bool borrow_host = false) {
    build(ctx, a, spmv_layout, borrow_host);
  }
  void clear() { release(); }
#endif
  HIPMatrix(const HIPMatrix&) = delete;
  HIPMatrix& operator=(const HIPMatrix&) = delete;

  // A new matrix with the same profile and values (owning its host copy), in the same
  // SpMV layout: the cloneTo of Alien's matrix impls (SYCLBEllPackMatrix.h:123).
  HIPMatrix* cloneTo(const MultiMatrixImpl* multi) const {
    std::unique_ptr<HIPMatrix> c(new HIPMatrix(multi));
    if (m_ctx) {
      CSRHostView v{m_n, h_ro, h_cols, hostValues()};
      c->build(m_ctx, v, m_layout, /*borrow_host=*/false);
    }
    return c.release();
  }
  // Alien's notification that the values were changed through data() on the host
  // (SYCLBEllPackMatrix::notifyChanges): upload them and refill the layouts.
  void notifyChanges() {
    if (m_ctx && h_vals == m_vals.data()) updateValues(m_vals.data());
  }
  void endUpdate() { notifyChanges(); }
  int getAllocSize() const { return m_n; }
  // host values (the matrix's own copy): Alien's data() of the host-side profile
  double* data() {
    syncHostValues();
    return m_vals.empty() ? nullptr : m_vals.data();
  }
  // frees the device arrays and the host copy (IMatrixImpl::clear)
  void release() {
    d_ro.reset();
    d_cols.reset();
    d_dcol.reset();
    d_vals.reset();
    freeRB();
    d_bro.reset();
    d_bellc.reset();
    d_bellv.reset();
    d_bellv32.reset();
    d_lmask.reset();
    d_umask.reset();
    m_fused_ell = false;
    m_kcol.clear();
    m_cols.clear();
    m_dcol.clear();
    m_vals.clear();
    m_bo.clear();
    m_bro.clear();
    h_ro = h_cols = nullptr;
    h_vals = nullptr;
    m_n = m_nnz = 0;
    m_layout = kLayoutCSR;
  }

  // Upload a host CSR (ascending columns) and build the requested SpMV layout:
  // the body of the SimpleCSR -> hip matrix converter. The host profile is
  // copied, or read in place when borrow_host says the caller's arrays outlive
  // the matrix.
  void build(alien_hip_ctx* ctx, const CSRHostView& a, int spmv_layout = kLayoutRBCSR, bool borrow_host = false) {
    m_ctx = ctx;
    m_n = a.nrows;
    m_nnz = a.row_offset[m_n];
    m_ctx->reserve_for(m_n);
    if (borrow_host) {
      h_ro = a.row_offset;
      h_cols = a.cols;
      h_vals = a.vals;
    } else {
      m_kcol.assign(a.row_offset, a.row_offset + m_n + 1);
      m_cols.assign(a.cols, a.cols + m_nnz);
      m_vals.assign(a.vals, a.vals + m_nnz);
      h_ro = m_kcol.data();
      h_cols = m_cols.data();
      h_vals = m_vals.data();
    }
    hipStream_t st = m_ctx->stream;
    d_ro.upload(h_ro, m_n + 1, st);
    d_cols.upload(h_cols, m_nnz, st);
    // the diagonal positions computed on the device, the host profile's copy read back
    d_dcol.resize(m_n);
    alien_hip::launch_dcol(st, m_n, d_ro.data(), d_cols.data(), d_dcol.data());
    alien_hip::check_launch("dcol");
    m_dcol.resize(m_n);
    if (m_n)
      ALIEN_HIP_CHECK(hipMemcpyAsync(m_dcol.data(), d_dcol.data(), sizeof(int) * m_n, hipMemcpyDeviceToHost, st));
    d_vals.upload(h_vals, m_nnz, st);
    setSpmvLayout(spmv_layout);
    ALIEN_HIP_CHECK(hipStreamSynchronize(st));
  }

  int getNRows() const { return m_n; }
  int getNnz() const { return m_nnz; }
  bool isParallel() const { return false; }
  HIPResource resource() const { return HIPResource{m_n, m_ctx}; }
  alien_hip_ctx* ctx() const { return m_ctx; }
  int spmvLayout() const { return m_layout; }

  // host profile (CSRStructInfo: kcol/cols/dcol)
  const int* kcol() const { return h_ro; }
  const int* cols() const { return h_cols; }
  const int* dcol() const { return m_dcol.data(); }
  const double* hostValues() const {
    syncHostValues();
    return h_vals;
  }
  // The device CSR values were changed on `st` (by the expression algebra): the layouts
  // already built are refilled from them and the host copy is marked stale.
  void deviceValuesChanged(hipStream_t st) {
    ALIEN_HIP_CHECK(hipStreamSynchronize(st));
    valuesChanged();
  }
  // New values on the same structure (the matrix owns its host copy): the
  // device CSR values and every value layout already built are refilled.
  // The host copy is refreshed from the device only when a host-side builder
  // needs it (BEllPack).
  void updateValues(const double* vals) {
    if (h_vals != m_vals.data())
      throw alien_hip::Error(ALIEN_HIP_EINVAL, "updateValues: the matrix reads the caller's arrays");
    d_vals.upload(vals, m_nnz, m_ctx->stream);
    valuesChanged();
  }
  // New values gathered on the device, vals[k] = src[map[k]] (a permuted copy of
  // another matrix whose device values src holds)
  void updateValuesGather(const int* d_map, const double* d_src) {
    if (h_vals != m_vals.data())
      throw alien_hip::Error(ALIEN_HIP_EINVAL, "updateValues: the matrix reads the caller's arrays");
    alien_hip::launch_gather_f64(m_ctx->stream, m_nnz, d_map, d_src, d_vals.data());
    alien_hip::check_launch("gather values");
    valuesChanged();
  }

 private:
  void valuesChanged() {
    hipStream_t st = m_ctx->stream;
    m_host_vals_stale = true;
    if (!d_bv.empty()) {
      if (m_rbdia)
        fillDIA<double>(d_bv.data());
      else
        alien_hip::launch_rbcsr_fill<double>(st, m_n, d_ro.data(), d_cols.data(), d_vals.data(), d_bo.data(),
                                             d_bc.data(), d_bv.data());
    }
    if (!d_bv32.empty()) {
      if (m_rbdia)
        fillDIA<float>(d_bv32.data());
      else
        alien_hip::launch_rbcsr_fill<float>(st, m_n, d_ro.data(), d_cols.data(), d_vals.data(), d_bo.data(),
                                            d_bc.data(), d_bv32.data());
    }
    alien_hip::check_launch("updateValues refill");
    if (!d_bellv.empty())
      alien_hip::launch_bellpack_fill<double>(st, m_n, kBEllPackSlice, d_ro.data(), d_cols.data(), d_vals.data(),
                                              nullptr, d_bro.data(), nullptr, d_bellv.data(), nullptr, nullptr);
    if (!d_bellv32.empty())
      alien_hip::launch_bellpack_fill<float>(st, m_n, kBEllPackSlice, d_ro.data(), d_cols.data(), d_vals.data(),
                                             nullptr, d_bro.data(), nullptr, d_bellv32.data(), nullptr, nullptr);
    alien_hip::check_launch("updateValues BEllPack refill");
    ALIEN_HIP_CHECK(hipStreamSynchronize(st));
  }

 public:
  // host bundle offsets of the row-bundled layout (after ensureRB / ensureRBCSR)
  const std::vector<int>& bundleOffsets() const { return m_bo; }
  // host slice offsets of the BEllPack layout, nslices + 1 entries (after ensureBEllPack)
  const std::vector<int>& bellpackRowOffsets() const { return m_bro; }

  // Builds the layout on first use and selects it for mult().
  void setSpmvLayout(int layout) {
    if (layout == kLayoutRBCSR) ensureRBCSR();
    else if (layout == kLayoutRBDIA) ensureRB(true);
    else if (layout == kLayoutBEllPack) ensureBEllPack(false);
    else if (layout != kLayoutCSR) throw alien_hip::Error(ALIEN_HIP_EINVAL, "unknown SpMV layout");
    m_layout = layout;
  }

  // Row-bundled layout of the fused solvers: RB-DIA (no column stream, its
  // structure built on the device) when allowed and it stores at most
  // kRbdiaMaxGrowth times the RB-CSR slots, RB-CSR otherwise. FP32 values
  // follow with ensureRB32().
  void ensureRB(bool allow_dia) {
    if (!d_bo.empty() && (allow_dia || !m_rbdia)) return;
    if (!d_bo.empty()) freeRB();
    if (allow_dia) {
      const int nb = (m_n + alien_hip::kBundle - 1) / alien_hip::kBundle;
      std::vector<int> csr_bo(nb + 1);
      const bool csr_ok = alien_hip::rbcsr_bundle_offsets(m_n, h_ro, csr_bo.data()) == 0;
      hipStream_t st = m_ctx->stream;
      const long slots = alien_hip::rbdia_count_device(st, m_n, d_ro.data(), d_cols.data(),
                                                       alien_hip::kRbdiaMaxWidth, m_bo);
      if (slots >= 0 && (!csr_ok || slots <= alien_hip::kRbdiaMaxGrowth * csr_bo[nb])) {
        buildDIA(slots, false);
        return;
      }
      // rows of column pairs (two unknowns per cell, interleaved): one slot per pair
      const long pslots = alien_hip::rbdia_count_device(st, m_n, d_ro.data(), d_cols.data(),
                                                        alien_hip::kRbdiaMaxWidth, m_bo, true);
      if (pslots >= 0 && (!csr_ok || 2 * pslots <= alien_hip::kRbdiaMaxGrowth * csr_bo[nb])) {
        buildDIA(pslots, true);
        return;
      }
    }
    ensureRBCSR();
  }
  // RB-DIA from the bundle offsets in m_bo (slots per lane; pair mode: pairs per lane)
  void buildDIA(long slots, bool pair) {
    hipStream_t st = m_ctx->stream;
    d_bo.upload(m_bo.data(), m_bo.size(), st);
    d_doff.resize(slots);
    d_dmask.resize(slots);
    alien_hip::launch_rbdia_struct(st, m_n, d_ro.data(), d_cols.data(), d_bo.data(), d_doff.data(), d_dmask.data(),
                                   pair);
    alien_hip::check_launch("rbdia_struct");
    m_rbdia = true;
    m_pair = pair;
    m_uni = 0;
    if (!pair) slots = tryUniform(slots);
    d_bv.resize((size_t)slots * alien_hip::kBundle * (pair ? 2 : 1));
    fillDIA<double>(d_bv.data());
    settle();
  }
  // Uniform RB-DIA (RBMat::uni) when the matrix's distinct diagonal offsets number at most
  // kUniMaxSlots and padding every bundle to all of them adds at most 1/32 of the slots (a
  // stencil matrix: only its boundary bundles lack some offsets). Returns the slot count.
  long tryUniform(long slots) {
    hipStream_t st = m_ctx->stream;
    const int nb = (m_n + alien_hip::kBundle - 1) / alien_hip::kBundle;
    if (nb == 0 || slots <= 0) return slots;
    std::vector<int> doff(slots);
    ALIEN_HIP_CHECK(hipMemcpyAsync(doff.data(), d_doff.data(), slots * sizeof(int), hipMemcpyDeviceToHost, st));
    ALIEN_HIP_CHECK(hipStreamSynchronize(st));
    std::vector<int> u;
    for (int d : doff) {
      if (std::find(u.begin(), u.end(), d) != u.end()) continue;
      u.push_back(d);
      if ((int)u.size() > alien_hip::kUniMaxSlots) return slots;
    }
    std::sort(u.begin(), u.end());
    const long k = (long)u.size();
    if (k * nb > slots + slots / 32 || k * nb * alien_hip::kBundle > 0x7fffffffL) return slots;
    d_doff.resize(k * nb);
    d_dmask.resize(k * nb);
    alien_hip::launch_rbdia_uniform_struct(st, m_n, d_ro.data(), d_cols.data(), (int)k, u.data(), d_bo.data(),
                                           d_doff.data(), d_dmask.data());
    alien_hip::check_launch("rbdia_uniform_struct");
    for (int b = 0; b <= nb; ++b) m_bo[b] = (int)(b * k);
    m_uni = (int)k;
    for (int j = 0; j < (int)k; ++j) m_uoff[j] = u[j];
    return k * nb;
  }
  template <typename T>
  void setUniform(alien_hip::RBMat<T>& a) const {
    if (!m_uni) return;
    a.uni = m_uni;
    a.nrows = m_n;
    for (int j = 0; j < m_uni; ++j) a.uoff[j] = m_uoff[j];
  }
  template <typename T>
  void fillDIA(T* bv) {
    hipStream_t st = m_ctx->stream;
    if (m_pair)
      alien_hip::launch_rbdia2_fill<T>(st, m_n, d_ro.data(), d_cols.data(), d_vals.data(), d_bo.data(),
                                       d_doff.data(), bv);
    else
      alien_hip::launch_rbdia_fill<T>(st, m_n, d_ro.data(), d_cols.data(), d_vals.data(), d_bo.data(),
                                      d_doff.data(), bv);
    alien_hip::check_launch("rbdia_fill");
  }
  void ensureRB32() {
    if (m_fused_ell) {
      ensureBEllPack32();
      return;
    }
    if (d_bo.empty()) ensureRB(true);
    if (!d_bv32.empty()) return;
    d_bv32.resize(d_bv.size());
    if (m_rbdia)
      fillDIA<float>(d_bv32.data());
    else
      alien_hip::launch_rbcsr_fill<float>(m_ctx->stream, m_n, d_ro.data(), d_cols.data(), d_vals.data(),
                                          d_bo.data(), d_bc.data(), d_bv32.data());
    alien_hip::check_launch("row-bundled fill<float>");
    settle();
  }
  bool isRBDIA() const { return m_rbdia && !m_fused_ell; }
  bool isPairDIA() const { return m_rbdia && m_pair && !m_fused_ell; }
  // FP64 operand: the uniform structure read as an ordinary RB-DIA (its masks skip the padded
  // slots). Measured at 256^3: the FP64 row kernels on RB-DIA with one bundle per wave (q = A p
  // 197 us, a Chebyshev step 313 us) beat their uniform instances (216 us, 337 us), which pay off
  // for FP32 only (q = A p 100 us against 125 us).
  alien_hip::RBMat<double> rb64() const {
    if (m_fused_ell) return ellMat<double>(d_bellv.data());
    return alien_hip::RBMat<double>{d_bo.data(), m_rbdia ? nullptr : d_bc.data(), d_bv.data(),
                                    m_rbdia ? d_doff.data() : nullptr, m_rbdia ? d_dmask.data() : nullptr,
                                    m_pair ? 1 : 0};
  }
  alien_hip::RBMat<float> rb32() const {
    if (m_fused_ell) return ellMat<float>(d_bellv32.data());
    alien_hip::RBMat<float> a{d_bo.data(), m_rbdia ? nullptr : d_bc.data(), d_bv32.data(),
                              m_rbdia ? d_doff.data() : nullptr, m_rbdia ? d_dmask.data() : nullptr, m_pair ? 1 : 0};
    setUniform(a);
    return a;
  }
  int uniformSlots() const { return m_uni; }
  template <typename T>
  alien_hip::RBMat<T> ellMat(const T* bv) const {
    alien_hip::RBMat<T> a{d_bro.data(), d_bellc.data(), bv};
    a.ell = kBEllPackSlice / alien_hip::kBundle;
    return a;
  }
  void freeRB() {
    d_bo.reset();
    d_bc.reset();
    d_bv.reset();
    d_bv32.reset();
    d_doff.reset();
    d_dmask.reset();
    m_pair = false;
    m_rbdia = false;
    m_uni = 0;
  }

  void ensureRBCSR() {
    if (!d_bo.empty() && !m_rbdia) return;
    if (!d_bo.empty()) freeRB();
    const int nb = (m_n + alien_hip::kBundle - 1) / alien_hip::kBundle;
    m_bo.resize(nb + 1);
    if (alien_hip::rbcsr_bundle_offsets(m_n, h_ro, m_bo.data()) != 0)
      throw alien_hip::Error(ALIEN_HIP_EINVAL, "RB-CSR storage exceeds int32 indexing");
    hipStream_t st = m_ctx->stream;
    d_bo.upload(m_bo.data(), m_bo.size(), st);
    const size_t slots = (size_t)m_bo[nb] * alien_hip::kBundle;
    d_bc.resize(slots);
    d_bv.resize(slots);
    alien_hip::launch_rbcsr_fill<double>(st, m_n, d_ro.data(), d_cols.data(), d_vals.data(), d_bo.data(),
                                         d_bc.data(), d_bv.data());
    alien_hip::check_launch("rbcsr_fill");
    settle();
  }
  // Alien's BEllPack arrays (slice offsets from the host profile, columns, values and, on
  // request, the L/U masks scattered on the device from the device CSR)
  void ensureBEllPack(bool masks) {
    hipStream_t st = m_ctx->stream;
    const int ns = alien_hip::bellpack_nslices(m_n, kBEllPackSlice);
    if (d_bro.empty()) {
      m_bro.resize(ns + 1);
      alien_hip::bellpack_bro(m_n, kBEllPackSlice, h_ro, m_bro.data());
      if ((long)m_bro[ns] * kBEllPackSlice > 0x7fffffffL)
        throw alien_hip::Error(ALIEN_HIP_EINVAL, "BEllPack storage exceeds int32 indexing");
      const size_t total = (size_t)m_bro[ns] * kBEllPackSlice;
      d_bro.upload(m_bro.data(), m_bro.size(), st);
      d_bellc.resize(total);
      d_bellv.resize(total);
      alien_hip::launch_bellpack_fill<double>(st, m_n, kBEllPackSlice, d_ro.data(), d_cols.data(), d_vals.data(),
                                              nullptr, d_bro.data(), d_bellc.data(), d_bellv.data(), nullptr,
                                              nullptr);
      alien_hip::check_launch("bellpack_fill");
      settle();
    }
    if (masks && d_lmask.empty()) {
      const size_t total = (size_t)m_bro[ns] * kBEllPackSlice;
      d_lmask.resize(total);
      d_umask.resize(total);
      alien_hip::launch_bellpack_fill<double>(st, m_n, kBEllPackSlice, d_ro.data(), d_cols.data(), nullptr,
                                              d_dcol.data(), d_bro.data(), nullptr, nullptr, d_lmask.data(),
                                              d_umask.data());
      alien_hip::check_launch("bellpack_fill masks");
      settle();
    }
  }
  // FP32 copy of the BEllPack values (mixed-precision solvers on that layout)
  void ensureBEllPack32() {
    ensureBEllPack(false);
    if (!d_bellv32.empty() || d_bellv.empty()) return;
    d_bellv32.resize(d_bellv.size());
    alien_hip::launch_bellpack_fill<float>(m_ctx->stream, m_n, kBEllPackSlice, d_ro.data(), d_cols.data(),
                                           d_vals.data(), nullptr, d_bro.data(), nullptr, d_bellv32.data(), nullptr,
                                           nullptr);
    alien_hip::check_launch("bellpack_fill<float>");
    settle();
  }
  // Operand layout of the fused solvers: Alien's BEllPack read in place (the row kernels'
  // slice mode) when asked for, else the row-bundled layouts (ensureRB)
  void ensureFused(int layout) {
    m_fused_ell = layout == kLayoutBEllPack;
    if (m_fused_ell) ensureBEllPack(false);
    else ensureRB(layout != kLayoutRBCSR);
  }
  bool fusedBEllPack() const { return m_fused_ell; }

  // device arrays (public for the kernels' callers inside the backend)
  alien_hip::DeviceArray<int> d_ro, d_cols, d_dcol;
  alien_hip::DeviceArray<double> d_vals;
  alien_hip::DeviceArray<int> d_bo, d_bc;        // RB-CSR (d_bo, d_bv, d_bv32 shared with RB-DIA)
  alien_hip::DeviceArray<double> d_bv;
  alien_hip::DeviceArray<float> d_bv32;
  alien_hip::DeviceArray<int> d_doff;            // RB-DIA
  alien_hip::DeviceArray<unsigned long long> d_dmask;
  alien_hip::DeviceArray<int> d_bro, d_bellc;    // BEllPack
  alien_hip::DeviceArray<double> d_bellv;
  alien_hip::DeviceArray<float> d_bellv32;
  alien_hip::DeviceArray<uint8_t> d_lmask, d_umask;

 private:
  alien_hip_ctx* m_ctx = nullptr;
  bool m_rbdia = false;
  int m_uni = 0;     // uniform RB-DIA: slots per bundle (0: not uniform)
  int m_uoff[8] = {}; // and their column offsets
  bool m_pair = false; // RB-DIA in pair mode
  bool m_fused_ell = false; // the fused solvers read the BEllPack arrays
  int m_n = 0, m_nnz = 0;
  int m_layout = kLayoutCSR;
  std::vector<int> m_kcol, m_cols, m_dcol, m_bo, m_bro;
  std::vector<double> m_vals;
  // host values behind the device's after updateValues (read back on demand)
  void syncHostValues() const {
    if (!m_host_vals_stale) return;
    ALIEN_HIP_CHECK(hipMemcpyAsync(const_cast<double*>(m_vals.data()), d_vals.data(), sizeof(double) * m_nnz,
                                   hipMemcpyDeviceToHost, m_ctx->stream));
    ALIEN_HIP_CHECK(hipStreamSynchronize(m_ctx->stream));
    m_host_vals_stale = false;
  }
  mutable bool m_host_vals_stale = false;
  // The layouts are built on the stream of the matrix's own context, but a
  // kept matrix may be solved through any context on its device: each
  // one-time layout build drains its stream before the layout is used.
  void settle() const { ALIEN_HIP_CHECK(hipStreamSynchronize(m_ctx->stream)); }
  const int* h_ro = nullptr;   // host profile: the copies above or the caller's arrays
  const int* h_cols = nullptr;
  const double* h_vals = nullptr;
};

// Future of a device-resident scalar: dot(x, y, f) leaves the result on the
// device; get() copies it back and synchronises (SYCLInternal::Future,
// sycl/algebra/SYCLKernelInternal.h:32-84).
class HIPFuture {
 public:
  explicit HIPFuture(double& v) : m_value(v) {}
  double& operator()() { return m_value; }
  double operator()() const { return m_value; }
  void attach(alien_hip_ctx* c, int slot) {
    m_ctx = c;
    m_slot = slot;
    m_pending = true;
  }
  double get() {
    if (m_pending) {
      m_value = m_ctx->read_slot(m_slot);
      m_pending = false;
    }
    return m_value;
  }

 private:
  double& m_value;
  alien_hip_ctx* m_ctx = nullptr;
  int m_slot = 0;
  bool m_pending = false;
};

class HIPInternalLinearAlgebra
#ifdef ALIEN_HIP_HAVE_ALGEBRA_TRAITS
: public IInternalLinearAlgebra<HIPMatrix, HIPVectorT<double>>
#endif
{
 public:
  typedef BackEnd::tag::hip BackEndType;
  typedef HIPMatrix Matrix;
  typedef HIPVector Vector;
  typedef HIPResource ResourceType;
  typedef HIPFuture FutureType;

  class NullValueException : public std::runtime_error {
   public:
    explicit NullValueException(const std::string& what) : std::runtime_error(what) {}
  };
  class NotImplementedException : public std::runtime_error {
   public:
    explicit NotImplementedException(const std::string& what) : std::runtime_error(what) {}
  };

  explicit HIPInternalLinearAlgebra(alien_hip_ctx* ctx = nullptr) : m_ctx(ctx ? ctx : alien_hip::default_ctx()) {}

  alien_hip_ctx* ctx() const { return m_ctx; }
  hipStream_t stream() const { return m_ctx->stream; }

  static ResourceType resource(const Matrix& A) { return A.resource(); }
  void allocate(const ResourceType& r, Vector& v) {
    m_ctx->reserve_for(r.n);
    v.init(r, true);
  }
  template <typename T0, typename... T>
  void allocate(const ResourceType& r, T0& v0, T&... args) {
    allocate(r, v0);
    allocate(r, args...);
  }
  void free(Vector& v) { v.clear(); }
  template <typename T0, typename... T>
  void free(T0& v0, T&... args) {
    free(v0);
    free(args...);
  }

  // ---- IInternalLinearAlgebra (IInternalLinearAlgebraT.h:69-186) ----
  double norm0(const Vector&) const { throw NotImplementedException("HIPInternalLinearAlgebra::norm0"); }
  double norm1(const Vector&) const { throw NotImplementedException("HIPInternalLinearAlgebra::norm1"); }
  double norm2(const Vector& x) const { return std::sqrt(dot(x, x)); } // sqrt(dot), CBLASMPIKernel.h:114-125

  void mult(const Matrix& A, const Vector& x, Vector& y) const {
    switch (A.spmvLayout()) {
    case kLayoutRBCSR:
    case kLayoutRBDIA:
      alien_hip::launch_spmv_rb<double>(stream(), A.getNRows(), A.rb64(), x.data(), y.data());
      break;
    case kLayoutBEllPack:
      alien_hip::launch_spmv_bellpack<double>(stream(), A.getNRows(), HIPMatrix::kBEllPackSlice, A.d_bro.data(),
                                              A.d_bellc.data(), A.d_bellv.data(), x.data(), y.data());
      break;
    default:
      alien_hip::launch_spmv_csr<double>(stream(), A.getNRows(), A.d_ro.data(), A.d_cols.data(), A.d_vals.data(),
                                         x.data(), y.data());
    }
    alien_hip::check_launch("mult");
  }
  // y += alpha (strictly lower part of A) x: on a BEllPack matrix through its lower mask,
  // the reference's device path (MatrixInternal::addLMult, SYCLBEllPackMatrix.cc:769-840),
  // else the CSR loop (SimpleCSRMatrixMultT.h:193-212)
  void addLMult(double alpha, const Matrix& A, const Vector& x, Vector& y) const { addTriMult(0, alpha, A, x, y); }
  // y += alpha (strictly upper part of A) x (SYCLBEllPackMatrix.cc:842-911, SimpleCSRMatrixMultT.h:214-233)
  void addUMult(double alpha, const Matrix& A, const Vector& x, Vector& y) const { addTriMult(1, alpha, A, x, y); }
  void addTriMult(int upper, double alpha, const Matrix& A, const Vector& x, Vector& y) const {
    if (A.spmvLayout() == kLayoutBEllPack) {
      const_cast<Matrix&>(A).ensureBEllPack(true); // masks built once, on the first masked product
      alien_hip::launch_addmult_bellpack(stream(), A.getNRows(), HIPMatrix::kBEllPackSlice, alpha, A.d_bro.data(),
                                         A.d_bellc.data(), A.d_bellv.data(),
                                         upper ? A.d_umask.data() : A.d_lmask.data(), x.data(), y.data());
    } else {
      alien_hip::launch_addmult_csr(stream(), A.getNRows(), upper, alpha, A.d_ro.data(), A.d_cols.data(),
                                    A.d_dcol.data(), A.d_vals.data(), x.data(), y.data());
    }
    alien_hip::check_launch(upper ? "addUMult" : "addLMult");
  }
  // d_i = 1 / a[dcol_i] (SimpleCSRMatrixMultT::computeInvDiag, :494-509)
  void computeInvDiag(const Matrix& A, Vector& d) const {
    alien_hip::launch_invdiag_csr(stream(), A.getNRows(), A.d_dcol.data(), A.d_vals.data(), d.data());
    alien_hip::check_launch("computeInvDiag");
  }
  // y_i = y_i / a[dcol_i] (SimpleCSRMatrixMultT::multInvDiag, :478-492)
  void multInvDiag(const Matrix& A, Vector& y) const {
    Vector d;
    d.init(A.resource(), true);
    computeInvDiag(A, d);
    pointwiseMult(d, y, y);
  }

  void axpy(double a, const Vector& x, Vector& y) const {
    alien_hip::launch_axpy<double>(stream(), x.scalarizedLocalSize(), a, x.data(), y.data());
    alien_hip::check_launch("axpy");
  }
  void aypx(double, Vector&, const Vector&) const { throw NotImplementedException("HIPInternalLinearAlgebra::aypx"); }
  void copy(const Vector& x, Vector& y) const {
    alien_hip::launch_copy<double>(stream(), x.scalarizedLocalSize(), x.data(), y.data());
    alien_hip::check_launch("copy");
  }
  double dot(const Vector& x, const Vector& y) const {
    const int s = m_ctx->take_slot();
    alien_hip::launch_dot<double>(stream(), m_ctx->red(), x.scalarizedLocalSize(), x.data(), y.data(), m_ctx->dslot(s));
    alien_hip::check_launch("dot");
    return m_ctx->read_slot(s);
  }
  void dot(const Vector& x, const Vector& y, FutureType& f) const {
    const int s = m_ctx->take_slot();
    alien_hip::launch_dot<double>(stream(), m_ctx->red(), x.scalarizedLocalSize(), x.data(), y.data(), m_ctx->dslot(s));
    alien_hip::check_launch("dot(future)");
    f.attach(m_ctx, s);
  }
  void scal(double a, Vector& x) const {
    alien_hip::launch_scal<double>(stream(), x.scalarizedLocalSize(), a, x.data());
    alien_hip::check_launch("scal");
  }
  void diagonal(const Matrix&, Vector&) const { throw NotImplementedException("HIPInternalLinearAlgebra::diagonal"); }
  void reciprocal(Vector&) const { throw NotImplementedException("HIPInternalLinearAlgebra::reciprocal"); }
  void pointwiseMult(const Vector& x, const Vector& y, Vector& z) const {
    alien_hip::launch_pmult<double>(stream(), x.scalarizedLocalSize(), x.data(), y.data(), z.data());
    alien_hip::check_launch("pointwiseMult");
  }
  void assign(Vector& x, double a) const {
    alien_hip::launch_assign<double>(stream(), x.scalarizedLocalSize(), a, x.data());
    alien_hip::check_launch("assign");
  }
  template <typename LambdaT>
  void assign(Vector& x, const LambdaT& lambda) const {
    x.apply(lambda);
  }
  template <typename PrecondT>
  void exec(PrecondT& precond, const Vector& x, Vector& y) {
    precond.solve(*this, x, y);
  }

 private:
  alien_hip_ctx* m_ctx = nullptr;
};

// The expression algebra (IInternalLinearAlgebraExprT.h:54-279; SYCLInternalLinearAlgebraExpr,
// SYCLInternalLinearAlgebra.h:143-180): the vector operations of HIPInternalLinearAlgebra, the
// matrix operations on the device values of matrices of one profile (the layouts already built
// are refilled), and the same operations on host arrays.
class HIPInternalLinearAlgebraExpr
#ifdef ALIEN_HIP_HAVE_ALGEBRA_TRAITS
: public IInternalLinearAlgebraExpr<HIPMatrix, HIPVectorT<double>>
#endif
{
 public:
  typedef HIPMatrix Matrix;
  typedef HIPVector Vector;
#ifdef ALIEN_HIP_HAVE_ALGEBRA_TRAITS
  typedef UniqueArray<Real> HostArray;
  typedef Integer SizeType;
#else
  typedef std::vector<double> HostArray;
  typedef int SizeType;
#endif
  typedef HIPInternalLinearAlgebra::NotImplementedException NotImplementedException;

  explicit HIPInternalLinearAlgebraExpr(alien_hip_ctx* ctx = nullptr) : m_alg(ctx) {}

  double norm0(const Vector& x) const { return m_alg.norm0(x); }
  double norm1(const Vector& x) const { return m_alg.norm1(x); }
  double norm2(const Vector& x) const { return m_alg.norm2(x); }
  // Frobenius norm: sqrt of the dot product of the stored values
  double norm2(const Matrix& a) const {
    const int s = m_alg.ctx()->take_slot();
    m_alg.ctx()->reserve_for(a.getNnz());
    alien_hip::launch_dot<double>(m_alg.stream(), m_alg.ctx()->red(), a.getNnz(), a.d_vals.data(), a.d_vals.data(),
                                  m_alg.ctx()->dslot(s));
    alien_hip::check_launch("norm2(matrix)");
    return std::sqrt(m_alg.ctx()->read_slot(s));
  }
  void mult(const Matrix& a, const Vector& x, Vector& r) const { m_alg.mult(a, x, r); }
  void axpy(double alpha, const Vector& x, Vector& y) const { m_alg.axpy(alpha, x, y); }
  void aypx(double alpha, Vector& y, const Vector& x) const { m_alg.aypx(alpha, y, x); }
  void copy(const Vector& x, Vector& r) const { m_alg.copy(x, r); }
  // r's values = a's values (same profile)
  void copy(const Matrix& a, Matrix& r) const {
    sameProfile(a, r, "copy(matrix)");
    alien_hip::launch_copy<double>(m_alg.stream(), a.getNnz(), a.d_vals.data(), r.d_vals.data());
    alien_hip::check_launch("copy(matrix)");
    r.deviceValuesChanged(m_alg.stream());
  }
  // r's values += a's values (same profile)
  void add(const Matrix& a, Matrix& r) const {
    sameProfile(a, r, "add(matrix)");
    alien_hip::launch_axpy<double>(m_alg.stream(), a.getNnz(), 1.0, a.d_vals.data(), r.d_vals.data());
    alien_hip::check_launch("add(matrix)");
    r.deviceValuesChanged(m_alg.stream());
  }
  double dot(const Vector& x, const Vector& y) const { return m_alg.dot(x, y); }
  void scal(double alpha, Vector& x) const { m_alg.scal(alpha, x); }
  void scal(double alpha, Matrix& a) const {
    alien_hip::launch_scal<double>(m_alg.stream(), a.getNnz(), alpha, a.d_vals.data());
    alien_hip::check_launch("scal(matrix)");
    a.deviceValuesChanged(m_alg.stream());
  }
  void diagonal(const Matrix& a, Vector& x) const { m_alg.diagonal(a, x); }
  void reciprocal(Vector& x) const { m_alg.reciprocal(x); }
  void pointwiseMult(const Vector& x, const Vector& y, Vector& w) const { m_alg.pointwiseMult(x, y, w); }

  // host arrays: r = a x through the device (upload x, product, download r), the BLAS-1
  // operations as plain host loops
  void mult(const Matrix& a, const HostArray& x, HostArray& r) const {
    Vector dx(a.resource()), dr(a.resource());
    dx.setValues(x.data());
    m_alg.mult(a, dx, dr);
    dr.copyValuesTo(r.data());
  }
  void axpy(double alpha, const HostArray& x, HostArray& y) const {
    for (SizeType i = 0; i < (SizeType)x.size(); ++i) y[i] = y[i] + alpha * x[i];
  }
  void aypx(double, HostArray&, const HostArray&) const {
    throw NotImplementedException("HIPInternalLinearAlgebraExpr::aypx");
  }
  void copy(const HostArray& x, HostArray& r) const {
    for (SizeType i = 0; i < (SizeType)x.size(); ++i) r[i] = x[i];
  }
  double dot(SizeType local_size, const HostArray& x, const HostArray& y) const {
    double s = 0.0;
    for (SizeType i = 0; i < local_size; ++i) s = s + x[i] * y[i];
    return s;
  }
  void scal(double alpha, HostArray& x) const {
    for (SizeType i = 0; i < (SizeType)x.size(); ++i) x[i] = alpha * x[i];
  }

  HIPInternalLinearAlgebra& algebra() { return m_alg; }

 private:
  static void sameProfile(const Matrix& a, const Matrix& r, const char* what) {
    if (a.getNRows() != r.getNRows() || a.getNnz() != r.getNnz())
      throw alien_hip::Error(ALIEN_HIP_EINVAL, std::string(what) + ": matrices of different profiles");
  }
  HIPInternalLinearAlgebra m_alg;
};

inline AlgebraTraits<BackEnd::tag::hip>::algebra_type*
AlgebraTraits<BackEnd::tag::hip>::algebra_factory(IMessagePassingMng*) {
  return new HIPInternalLinearAlgebra();
}
inline AlgebraTraits<BackEnd::tag::hip>::algebra_expr_type*
AlgebraTraits<BackEnd::tag::hip>::algebra_expr_factory(IMessagePassingMng*) {
  return new HIPInternalLinearAlgebraExpr();
}

} // namespace Alien
