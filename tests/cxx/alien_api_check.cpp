// Drives the hip backend through its C++ interface the way Alien's own code
// does, and prints one JSON object per check on stdout for
// tests/test_gpu_alien_api.py to compare with the CPU oracle:
//  * AlgebraTraits<BackEnd::tag::hip>: name(), algebra_factory(),
//    algebra_expr_factory() (SYCLBackEnd.h:76-96);
//  * a matrix as MultiMatrixImpl::getImpl creates it (`new matrix_type(this)`,
//    MultiMatrixImpl.h:318) then filled by the SimpleCSR -> hip converter
//    (build), and cloneTo;
//  * examples/test_krylov.cpp's run(alg) (:500-665) with the factory's algebra:
//    CG + Chebyshev on the 7-point Laplacian, BiCGStab + FILU0 on the
//    convection-diffusion operator, on the CSR and the BEllPack layouts;
//  * the expression algebra: Frobenius norm, scal / copy / add of matrices and
//    the host-array product.
// Usage: alien_api_check <n>  (grid edge of the test systems)
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "alien/hip/HIPBackEnd.h"
#include "alien/hip/krylov/HIPFILU0Preconditioner.h"
#include "alien/hip/krylov/Krylov.h"

using namespace Alien;

namespace {

struct Csr {
  int n = 0;
  std::vector<int> ro, cols;
  std::vector<double> vals;
};

Csr generate(int kind, int n) {
  Csr a;
  const long nnz = alien_hip_gen_nnz(kind, n, n, n);
  a.n = n * n * n;
  a.ro.resize(a.n + 1);
  a.cols.resize(nnz);
  a.vals.resize(nnz);
  if (alien_hip_gen_csr(kind, n, n, n, nullptr, 0, a.ro.data(), a.cols.data(), a.vals.data()) != 0) {
    std::fprintf(stderr, "gen_csr: %s\n", alien_hip_last_error());
    std::exit(2);
  }
  return a;
}

void print_vec(const char* key, const std::vector<double>& v) {
  std::printf("\"%s\": [", key);
  for (size_t i = 0; i < v.size(); ++i) std::printf(i ? ", %.17g" : "%.17g", v[i]);
  std::printf("]");
}

// test_krylov's run(alg) for one solver / preconditioner; b = A x_true (x_true from
// the library's U[0,1) stream, seed 42), x0 = 0
template <typename Precond>
void run_case(const char* name, HIPInternalLinearAlgebra& alg, const HIPMatrix& A, Precond& P, bool bicgs,
              double tol, int max_iter) {
  const int n = A.getNRows();
  std::vector<double> xt(n), xh(n);
  alien_hip_fill_uniform01(42, n, xt.data());
  HIPVector xtrue(A.resource()), b(A.resource()), x(A.resource());
  xtrue.setValues(xt.data());
  alg.mult(A, xtrue, b);
  Iteration<HIPInternalLinearAlgebra> it(alg, b, tol, max_iter);
  int error = 0;
  try {
    if (bicgs) {
      BiCGStab<HIPInternalLinearAlgebra> s(alg);
      s.solve(P, it, A, b, x);
    } else {
      CG<HIPInternalLinearAlgebra> s(alg);
      s.solve(P, it, A, b, x);
    }
  } catch (const HIPInternalLinearAlgebra::NullValueException&) {
    error = 1;
  }
  x.copyValuesTo(xh.data());
  std::printf("{\"case\": \"%s\", \"iterations\": %d, \"status\": %d, \"error\": %d, \"value\": %.17g, ", name, it(),
              it.getStatus() ? 1 : 0, error, it.getValue());
  print_vec("x", xh);
  std::printf("}\n");
}

} // namespace

int main(int argc, char** argv) {
  const int edge = argc > 1 ? std::atoi(argv[1]) : 8;
  typedef AlgebraTraits<BackEnd::tag::hip> Traits;
  std::unique_ptr<Traits::algebra_type> alg(Traits::algebra_factory());
  std::unique_ptr<Traits::algebra_expr_type> expr(Traits::algebra_expr_factory());
  const std::string name = Traits::name();
  std::printf("{\"case\": \"traits\", \"name\": \"%s\"}\n", name.c_str());

  const Csr lap = generate(ALIEN_HIP_GEN_POISSON7, edge);
  const Csr cd = generate(ALIEN_HIP_GEN_CONVDIFF7, edge);
  for (int layout : {kLayoutCSR, kLayoutBEllPack}) {
    const char* lname = layout == kLayoutCSR ? "csr" : "bellpack";
    // MultiMatrixImpl creates the empty matrix, the converter fills it
    HIPMatrix A(static_cast<const MultiMatrixImpl*>(nullptr));
    A.build(alien_hip::default_ctx(), CSRHostView{lap.n, lap.ro.data(), lap.cols.data(), lap.vals.data()}, layout);
    ChebyshevPreconditioner<HIPInternalLinearAlgebra> cheb(*alg, A, 0.5, 3, 10);
    cheb.init();
    run_case((std::string("cg-chebyshev-") + lname).c_str(), *alg, A, cheb, false, 1e-8, 1000);

    HIPMatrix C(static_cast<const MultiMatrixImpl*>(nullptr));
    C.build(alien_hip::default_ctx(), CSRHostView{cd.n, cd.ro.data(), cd.cols.data(), cd.vals.data()}, layout);
    FILU0Preconditioner<HIPInternalLinearAlgebra> filu(*alg, C);
    filu.setParameter("nb-factor-iter", 2);
    filu.setParameter("nb-solver-iter", 3);
    filu.init();
    run_case((std::string("bicgstab-filu0-") + lname).c_str(), *alg, C, filu, true, 1e-8, 1000);

    // cloneTo: the same operator in the same layout
    std::unique_ptr<HIPMatrix> D(C.cloneTo(nullptr));
    DiagPreconditioner<HIPInternalLinearAlgebra> diag(*alg, *D);
    diag.init();
    run_case((std::string("bicgstab-diag-clone-") + lname).c_str(), *alg, *D, diag, true, 1e-8, 1000);
  }

  // expression algebra on the convection-diffusion matrix
  HIPMatrix A(static_cast<const MultiMatrixImpl*>(nullptr));
  A.build(alien_hip::default_ctx(), CSRHostView{cd.n, cd.ro.data(), cd.cols.data(), cd.vals.data()}, kLayoutRBCSR);
  std::unique_ptr<HIPMatrix> B(A.cloneTo(nullptr));
  const double f0 = expr->norm2(A);
  expr->scal(2.0, *B);                 // B = 2 A
  const double f1 = expr->norm2(*B);
  expr->add(A, *B);                    // B = 3 A
  std::vector<double> xs(cd.n), y3(cd.n);
  for (int i = 0; i < cd.n; ++i) xs[i] = 1.0 + (i % 7) * 0.25;
  std::vector<double> xv(xs.begin(), xs.end()), yv(cd.n);
  expr->mult(*B, xv, yv);              // host arrays, product on the device (RB-CSR)
  expr->copy(A, *B);                   // B = A
  const double f2 = expr->norm2(*B);
  std::printf("{\"case\": \"expr\", \"frob\": %.17g, \"frob_2a\": %.17g, \"frob_copy\": %.17g, ", f0, f1, f2);
  print_vec("y_3a", yv);
  std::printf("}\n");
  return 0;
}
