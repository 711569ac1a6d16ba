/*
 * alien_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of Alien's SimpleCSR Krylov path (FP64, sequential loops,
 * compiled with -ffp-contract=off so that "sequential order" is bit-exact).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker; the product path never does.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the arcaneframework/alien source tree).
 *
 * Parity pins (what checks each function against something other than its
 * own formula) are listed in tests/test_oracle.py; functions with none are
 * marked "parity unpinned" below.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- CSR profile ------------------------------------------------------- */
/* CSRStructInfo::computeUpperDiagOffset (src/core/alien/kernels/simple_csr/CSRStructInfo.h:291-309) */
void orc_dcol(int nrows, const int* row_offset, const int* cols, int* dcol);

/* ---- SpMV -------------------------------------------------------------- */
/* SimpleCSRMatrixMultT::_seqMult (simple_csr/algebra/SimpleCSRMatrixMultT.h:172-191) */
void orc_spmv_csr(int nrows, const int* row_offset, const int* cols, const double* vals,
                  const double* x, double* y);
/* SimpleCSRMatrixMultT::_seqMultBlock (SimpleCSRMatrixMultT.h:317-342), column-major blocks */
void orc_spmv_bsr(int nbrows, int bs, const int* row_offset, const int* cols, const double* vals,
                  const double* x, double* y);
/* SimpleCSRMatrixMultT::_seqMultVariableBlock (SimpleCSRMatrixMultT.h:439-476): block j of
 * block row i is size[i] x size[cols[j]], row-major from vals[block_cols[j]]; block i of x
 * and y starts at block_offset[i] */
void orc_spmv_vblock(int nbrows, const int* row_offset, const int* cols, const int* block_cols,
                     const int* block_size, const int* block_offset, const double* vals, const double* x,
                     double* y);
/* _seqAddLMult / _seqAddUMult (SimpleCSRMatrixMultT.h:193-233) */
void orc_addlmult(int nrows, const int* row_offset, const int* cols, const int* dcol,
                  const double* vals, double alpha, const double* x, double* y);
void orc_addumult(int nrows, const int* row_offset, const int* cols, const int* dcol,
                  const double* vals, double alpha, const double* x, double* y);
/* SimpleCSRMatrixMultT::computeInvDiag (SimpleCSRMatrixMultT.h:494-509) */
void orc_invdiag_csr(int nrows, const int* dcol, const double* vals, double* y);

/* ---- BEllPack (sliced ELL, slice height 1024) --------------------------- */
/* BEllPackStructInfo::computeBlockRowOffset (sycl/data/SYCLBEllPackMatrix.cc:69-88) */
int  orc_bellpack_nslices(int nrows, int slice);
void orc_bellpack_bro(int nrows, int slice, const int* row_offset, int* bro);
/* block_cols / values scatter (SYCLBEllPackMatrix.cc:127-170, 422-460); padding col 0, value 0 */
void orc_bellpack_fill(int nrows, int slice, const int* row_offset, const int* cols,
                       const double* vals, const int* bro, int* bcols, double* bvals);
/* lower/upper masks (SYCLBEllPackMatrix.cc:235-315) */
void orc_bellpack_masks(int nrows, int slice, const int* row_offset, const int* dcol,
                        const int* bro, uint8_t* lmask, uint8_t* umask);
/* MatrixInternal::mult (SYCLBEllPackMatrix.cc:623-688) */
void orc_bellpack_spmv(int nrows, int slice, const int* bro, const int* bcols, const double* bvals,
                       const double* x, double* y);
/* MatrixInternal::computeInvDiag (SYCLBEllPackMatrix.cc:930-986); y must be preset */
void orc_bellpack_invdiag(int nrows, int slice, const int* bro, const int* bcols,
                          const double* bvals, double* y);

/* ---- BLAS-1 (CBLASMPIKernel.h:37-125; sequential loops) ---------------- */
double orc_dot(int n, const double* x, const double* y); /* order: parity unpinned (CBLAS) */

/* ---- ILU0 (ILU0Preconditioner.h:71-231, sequential branch) -------------- */
void orc_ilu0_factor(int nrows, const int* row_offset, const int* cols, const int* dcol,
                     double* lu /* in: A values, out: LU */);
void orc_ilu0_solve(int nrows, const int* row_offset, const int* cols, const int* dcol,
                    const double* lu, const double* y, double* x);

/* ---- schedules (integer) ------------------------------------------------ */
/* level of row i in the lower DAG: 0 if no k<dcol entry, else 1+max(level[cols[k]]) */
int  orc_levels_lower(int nrows, const int* row_offset, const int* cols, const int* dcol, int* level);
/* level in the upper DAG, counted from the last row */
int  orc_levels_upper(int nrows, const int* row_offset, const int* cols, const int* dcol, int* level);
/* greedy first-fit colouring in natural row order over the symmetrised pattern */
int  orc_color_greedy(int nrows, const int* row_offset, const int* cols, int* color);

/* ---- solvers (Iteration.h, CG.h:58-150, BiCGStab.h:151-300,
 *               DiagPreconditioner.h, ChebyshevPreconditioner.h:73-202,392-444,
 *               ILU0Preconditioner.h) ------------------------------------- */
enum { ORC_CG = 0, ORC_BICGS = 1 };
enum { ORC_PREC_DIAG = 0, ORC_PREC_CHEB = 1, ORC_PREC_ILU0 = 2, ORC_PREC_NONE = 3, ORC_PREC_FILU0 = 4,
       ORC_PREC_NEUMANN = 5, ORC_PREC_FACTOR = 6 };

typedef struct {
  int    solver;        /* ORC_CG / ORC_BICGS */
  int    precond;       /* ORC_PREC_* */
  double tol;
  int    max_iter;
  double poly_factor;   /* Chebyshev: alpha = beta / factor when factor != 0 */
  int    poly_order;
  int    poly_max_iter; /* power-method iterations */
  int    nthreads;      /* >1: row-block partition, dots summed per block then in block order */
  int    filu_factor_niter; /* FILU0: fixed-point factorisation sweeps (0 = exact ILU0) */
  int    filu_solver_niter; /* FILU0: Jacobi sweeps per triangular solve */
  /* FACTOR: y = G (G^T x), G and G^T given as CSR (ascending columns), sequential row sums */
  const int* g_ro;
  const int* g_cols;
  const double* g_vals;
  const int* gt_ro;
  const int* gt_cols;
  const double* gt_vals;
} orc_solve_params;

typedef struct {
  int    iterations;    /* Iteration::operator()() */
  int    status;        /* Iteration::getStatus() */
  int    error;         /* 0 ok, 1 NullValueException */
  double value;         /* Iteration::getValue() (recursive residual ratio) */
  double true_rel_res;  /* ||b - A x|| / ||b|| recomputed after the solve */
  double cheb_alpha;    /* Chebyshev eigen bounds after init() */
  double cheb_beta;
  double setup_s;       /* steady_clock phase timers (StdTimer semantics) */
  double solve_s;
} orc_solve_result;

/* FLUFactorisationAlgo::factorizeMultiIter (FILU0Preconditioner.h:146-170): niter sweeps, each
 * restarting from A's values and reading the previous sweep's factor for the other rows */
void orc_filu0_factor(int n, const int* ro, const int* cols, const int* dcol, const double* a, int niter,
                      double* lu);

/* ChebyshevPreconditioner::init (ChebyshevPreconditioner.h:73-116): inverse diagonal, power
 * method for beta (computeBeta, :118-156) and for alpha on the shifted operator (computeAlpha,
 * :158-202), max_it iterations each, then alpha = beta / factor when factor != 0 (:99-100) */
void orc_cheb_bounds(int nrows, const int* row_offset, const int* cols, const double* vals, double factor,
                     int max_it, double* alpha, double* beta);
/* ChebyshevPreconditioner::_algo2 (:392-434) with the given bounds: y = M^-1 x, y starts at x */
void orc_cheb_apply(int nrows, const int* row_offset, const int* cols, const double* vals, double alpha,
                    double beta, int order, const double* x, double* y);

/* x is the initial guess on input and the solution on output */
int orc_solve(const orc_solve_params* p, int nrows, const int* row_offset, const int* cols,
              const double* vals, const double* b, double* x, orc_solve_result* res);

/* Time `n_iter` CG iterations (the loop body of CG.h:111-145, stop-test dot
 * included) with the given preconditioner and nthreads: bounded CPU-baseline
 * sample. Returns seconds per iteration. */
double orc_time_cg_iters(int nrows, const int* row_offset, const int* cols, const double* vals,
                         const double* b, int precond, int poly_order, int n_iter, int nthreads);
/* Time `reps` CSR SpMVs with nthreads; returns seconds per SpMV. */
double orc_time_spmv(int nrows, const int* row_offset, const int* cols, const double* vals,
                     const double* x, double* y, int reps, int nthreads);

#ifdef __cplusplus
}
#endif
