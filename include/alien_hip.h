/*
 * alien_hip.h -- C ABI of libalien_hip.so, the MI355X (gfx950) Krylov backend
 * that registers beside Alien's `simplecsr` and `sycl` kernels.
 *
 * Conventions
 *  - Every function returns int: 0 = success, otherwise a hipError_t value or
 *    ALIEN_HIP_EINVAL (-22). alien_hip_last_error() gives the message.
 *  - Device-side entry points take DEVICE pointers they do not own, sizes, and
 *    a context handle whose stream orders the work. Nothing here synchronises
 *    the host unless the comment says so.
 *  - Integers are int32 row/column indices (Alien's `Integer`,
 *    src/core/alien/kernels/simple_csr/CSRStructInfo.h:44); sizes are int.
 *  - Reductions write their result to device memory (device-resident scalar)
 *    and are deterministic: a fixed grid and a fixed combination order.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the arcaneframework/alien source tree).
 */
#ifndef ALIEN_HIP_H
#define ALIEN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALIEN_HIP_EINVAL (-22)

typedef struct alien_hip_ctx alien_hip_ctx;       /* stream + reduction workspace */
typedef struct alien_hip_matrix alien_hip_matrix; /* device matrix in all layouts */

/* ------------------------------------------------------------------------ */
/* Runtime / context                                                         */
/* ------------------------------------------------------------------------ */

/* Message of the last failing call on this thread (never NULL). */
const char* alien_hip_last_error(void);
/* Library version string. Does not touch the GPU. */
const char* alien_hip_version(void);

/* Context = one HIP stream (NULL = create a private one) plus the workspace
 * used by the reductions. Replaces SYCLInternal::KernelInternal, the queue
 * holder of SYCLInternalLinearAlgebra (sycl/algebra/SYCLInternalLinearAlgebra.h:137). */
int alien_hip_ctx_create(void* stream, alien_hip_ctx** out);
int alien_hip_ctx_destroy(alien_hip_ctx* ctx);
int alien_hip_ctx_synchronize(alien_hip_ctx* ctx);
void* alien_hip_ctx_stream(alien_hip_ctx* ctx);
/* Instrumentation for benchmarks; no counterpart in the reference. Writes the totals
 * so far (microseconds, launches) to the non-NULL outputs, then: on = 1 clears them and
 * makes the fused CG solvers on this context (FUSED, GRAPH and MIXED modes) launch their
 * q = A p kernel with start and end events (hipExtLaunchKernel, on the context's stream) and
 * add up its durations (the GRAPH and MIXED iterations are then launched directly, not
 * replayed from a captured graph); on = 0 stops. Time solves with it off. */
int alien_hip_ctx_step_timing(alien_hip_ctx* ctx, int on, double* total_us, long long* launches);

/* ------------------------------------------------------------------------ */
/* Host-side system generators and structure builders (no GPU involved)     */
/* ------------------------------------------------------------------------ */

enum {
  ALIEN_HIP_GEN_POISSON7 = 0,  /* 7-point Laplacian [6,-1x6], Dirichlet eliminated */
  ALIEN_HIP_GEN_CONVDIFF7 = 1, /* upwind convection-diffusion; params {D, Pe, vx, vy, vz} */
  ALIEN_HIP_GEN_LOGNORMAL7 = 2 /* log-normal diffusion; params {sigma, ax, ay, az} */
};
int64_t alien_hip_gen_nnz(int kind, int nx, int ny, int nz);
/* Fills row_offset[N+1], cols[nnz], vals[nnz] (ascending columns per row).
 * Stands in for the assembly done by examples/test_krylov.cpp:291-437. */
int alien_hip_gen_csr(int kind, int nx, int ny, int nz, const double* params, uint64_t seed,
                      int* row_offset, int* cols, double* vals);
/* 2 unknowns per cell, 2x2 column-major blocks (SimpleCSRMatrixMultT.h:317-342 layout);
 * params {sigma, coupling}; nnz blocks = alien_hip_gen_nnz(...); vals has 4*nnzb entries */
int alien_hip_gen_bsr2(int nx, int ny, int nz, const double* params, uint64_t seed,
                       int* row_offset, int* cols, double* vals);
/* splitmix64 stream -> U[0,1) (x_true of the workloads) */
int alien_hip_fill_uniform01(uint64_t seed, int64_t n, double* out);

/* CSRStructInfo::computeUpperDiagOffset (simple_csr/CSRStructInfo.h:291-309) */
int alien_hip_host_dcol(int nrows, const int* row_offset, const int* cols, int* dcol);
/* BEllPackStructInfo::computeBlockRowOffset (sycl/data/SYCLBEllPackMatrix.cc:69-88);
 * bro has nslices+1 entries, nslices = ceil(nrows/slice) */
int alien_hip_host_bellpack_bro(int nrows, int slice, const int* row_offset, int* bro);
/* block_cols / values scatter and L/U masks (SYCLBEllPackMatrix.cc:127-170, 235-315, 422-460);
 * arrays have bro[nslices]*slice entries; vals/bvals and masks may be NULL */
int alien_hip_host_bellpack_fill(int nrows, int slice, const int* row_offset, const int* cols,
                                 const double* vals, const int* bro, int* bcols, double* bvals,
                                 uint8_t* lmask, uint8_t* umask);
/* Level sets of the ILU0 triangular sweeps (replace the strictly sequential row
 * recurrences of LUFactorisationAlgo::solveL/solveU, ILU0Preconditioner.h:181-231).
 * upper = 0: lower DAG (row i after rows cols[k], k < dcol[i]); upper = 1: upper DAG.
 * Returns the number of levels in *nlevels. */
int alien_hip_host_levels(int nrows, const int* row_offset, const int* cols, const int* dcol,
                          int upper, int* level, int* nlevels);
/* Greedy first-fit colouring over the symmetrised pattern, natural row order.
 * Returns the number of colours in *ncolors. */
int alien_hip_host_color(int nrows, const int* row_offset, const int* cols, int* color, int* ncolors);

/* ------------------------------------------------------------------------ */
/* Device kernels: SpMV                                                       */
/* ------------------------------------------------------------------------ */

/* y = A x, CSR, one lane per row, sequential column order, separate multiply
 * and add: bit-exact with SimpleCSRMatrixMultT::_seqMult (SimpleCSRMatrixMultT.h:172-191). */
int alien_hip_spmv_csr_f64(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                           const double* vals, const double* x, double* y);
int alien_hip_spmv_csr_f32(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                           const float* vals, const float* x, float* y);
/* y = A x on Alien's sliced-ELL layout (slice height `slice`, padding col 0 / value 0):
 * MatrixInternal::mult (sycl/data/SYCLBEllPackMatrix.cc:623-688), its per-row sums in its
 * order (padding slots add 0 * x[0]). One wavefront per 64 rows of a slice when slice is a
 * multiple of 64, else one lane per row. */
int alien_hip_spmv_bellpack_f64(alien_hip_ctx* ctx, int nrows, int slice, const int* bro,
                                const int* bcols, const double* bvals, const double* x, double* y);
int alien_hip_spmv_bellpack_f32(alien_hip_ctx* ctx, int nrows, int slice, const int* bro,
                                const int* bcols, const float* bvals, const float* x, float* y);
/* The BEllPack arrays (bro[nslices] * slice entries; bro from alien_hip_host_bellpack_bro)
 * scattered on the device from a device CSR: the block_cols / values scatter and the L/U
 * masks of SYCLBEllPackMatrix.cc:127-170, 235-315, 422-460, equal to
 * alien_hip_host_bellpack_fill's. Any of bcols / bvals / lmask / umask may be NULL; dcol
 * (diagonal positions) is read for the masks only. */
int alien_hip_bellpack_fill_f64(alien_hip_ctx* ctx, int nrows, int slice, const int* row_offset,
                                const int* cols, const double* vals, const int* dcol, const int* bro,
                                int* bcols, double* bvals, uint8_t* lmask, uint8_t* umask);
int alien_hip_bellpack_fill_f32(alien_hip_ctx* ctx, int nrows, int slice, const int* row_offset,
                                const int* cols, const double* vals, const int* bro, int* bcols,
                                float* bvals);
/* y += alpha * (strict lower | strict upper) A x through the BEllPack masks:
 * MatrixInternal::addLMult / addUMult (SYCLBEllPackMatrix.cc:769-911) */
int alien_hip_addmult_bellpack_f64(alien_hip_ctx* ctx, int upper, double alpha, int nrows, int slice,
                                   const int* bro, const int* bcols, const double* bvals,
                                   const uint8_t* mask, const double* x, double* y);
/* CSR addLMult / addUMult (SimpleCSRMatrixMultT.h:193-233) */
int alien_hip_addmult_csr_f64(alien_hip_ctx* ctx, int upper, double alpha, int nrows,
                              const int* row_offset, const int* cols, const int* dcol,
                              const double* vals, const double* x, double* y);
/* y_blk(i) = sum_j B_ij x_blk(j), 2x2 column-major blocks:
 * SimpleCSRMatrixMultT::_seqMultBlock (SimpleCSRMatrixMultT.h:317-342) */
int alien_hip_spmv_bsr2_f64(alien_hip_ctx* ctx, int nbrows, const int* row_offset, const int* cols,
                            const double* vals, const double* x, double* y);
int alien_hip_spmv_bsr2_f32(alien_hip_ctx* ctx, int nbrows, const int* row_offset, const int* cols,
                            const float* vals, const float* x, float* y);
/* The same product over row-bundled 2x2 blocks ("RB-BSR2"): 64 block rows per
 * bundle, bundle offsets as alien_hip_rbcsr_bundle_offsets over the block row
 * offsets; slot s = bundle_off[b] + j holds block j of each row, column at
 * bcols[64 s + l] (-1 = padding) and value q (column-major, m[ieq + 2 iu]) at
 * bvals[64 (4 s + q) + l]. rbbsr2_fill converts a device 2x2 block CSR (FP64
 * values, 4 per block) into bcols / bvals (64 x bundle_off[nbundles] columns,
 * 4x that many values). Bit-exact with the spmv_bsr2 kernels. */
int alien_hip_rbbsr2_fill_f64(alien_hip_ctx* ctx, int nbrows, const int* row_offset, const int* cols,
                              const double* vals, const int* bundle_off, int* bcols, double* bvals);
int alien_hip_rbbsr2_fill_f32(alien_hip_ctx* ctx, int nbrows, const int* row_offset, const int* cols,
                              const double* vals, const int* bundle_off, int* bcols, float* bvals);
int alien_hip_spmv_rbbsr2_f64(alien_hip_ctx* ctx, int nbrows, const int* bundle_off, const int* bcols,
                              const double* bvals, const double* x, double* y);
int alien_hip_spmv_rbbsr2_f32(alien_hip_ctx* ctx, int nbrows, const int* bundle_off, const int* bcols,
                              const float* bvals, const float* x, float* y);
/* Variable-block CSR product: SimpleCSRMatrixMultT::_seqMultVariableBlock
 * (SimpleCSRMatrixMultT.h:439-476). Block row i holds block_size[i] scalar rows,
 * starting at block_offset[i] in x and y (block_offset = running sum of
 * block_size, VBlockImpl::offset); block j of a row (column cols[j]) is
 * block_size[i] x block_size[cols[j]], row-major, from vals[block_cols[j]]
 * (CSRStructInfo::getBlockCols). One lane per block row, the reference's order. */
int alien_hip_spmv_vblock_f64(alien_hip_ctx* ctx, int nbrows, const int* row_offset, const int* cols,
                              const int* block_cols, const int* block_size, const int* block_offset,
                              const double* vals, const double* x, double* y);
/* Host: the scalar CSR of a variable-block matrix. Row block_offset[i] + k lists,
 * block by block and column by column, the entries the product above sums for
 * it, in its order, so the CSR / RB-CSR / RB-DIA kernels compute the same sums
 * bit for bit. vblock_scalar_nnz gives the entry count; vblock_to_csr fills
 * srow_offset[n + 1], scols[nnz] and (when svals is not NULL) svals[nnz], with
 * n = block_offset[nbrows - 1] + block_size[nbrows - 1]; it fails when
 * block_offset is not the running sum of block_size or a block column is out of
 * range. */
int64_t alien_hip_vblock_scalar_nnz(int nbrows, const int* row_offset_host, const int* cols_host,
                                    const int* block_size_host);
int alien_hip_vblock_to_csr(int nbrows, const int* row_offset_host, const int* cols_host,
                            const int* block_cols_host, const int* block_size_host,
                            const int* block_offset_host, const double* vals_host, int* srow_offset,
                            int* scols, double* svals);
/* inv_diag_i = 1 / a[dcol_i]: SimpleCSRMatrixMultT::computeInvDiag (:494-509).
 * Every row must store its diagonal (dcol_i then points at a_ii). */
int alien_hip_invdiag_csr_f64(alien_hip_ctx* ctx, int nrows, const int* dcol, const double* vals,
                              double* y);
/* inv_diag_i = 1/a_ii only where a_ii is stored and nonzero (y keeps its value elsewhere):
 * MatrixInternal::computeInvDiag (SYCLBEllPackMatrix.cc:930-986) */
int alien_hip_invdiag_bellpack_f64(alien_hip_ctx* ctx, int nrows, int slice, const int* bro,
                                   const int* bcols, const double* bvals, double* y);

/* Row-bundled CSR ("RB-CSR"): rows in bundles of 64 (one wavefront), entries
 * interleaved by lane so each value/column load is one coalesced 512/256-byte
 * access; padding slots hold column -1 and are skipped, so the per-row sum is
 * the sequential CSR sum, bit-exact with _seqMult. Built on the device from CSR.
 * nbundles = ceil(nrows/64); bundle_off has nbundles+1 entries (in slots of 64). */
int alien_hip_rbcsr_bundle_offsets(int nrows, const int* row_offset_host, int* bundle_off_host);
int alien_hip_rbcsr_fill_f64(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                             const double* vals, const int* bundle_off, int* bcols, double* bvals);
int alien_hip_rbcsr_fill_f32(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                             const double* vals, const int* bundle_off, int* bcols, float* bvals);
int alien_hip_spmv_rbcsr_f64(alien_hip_ctx* ctx, int nrows, const int* bundle_off, const int* bcols,
                             const double* bvals, const double* x, double* y);
int alien_hip_spmv_rbcsr_f32(alien_hip_ctx* ctx, int nrows, const int* bundle_off, const int* bcols,
                             const float* bvals, const float* x, float* y);

/* Row-bundled diagonal offsets ("RB-DIA"): the bundles of RB-CSR, but slot j of
 * bundle b stands for the j-th smallest distinct column offset c - r among the
 * bundle's entries (doff[bundle_off[b] + j]); bit l of dmask[...] marks the
 * rows 64 b + l holding that entry (value 0 elsewhere). No column index is
 * streamed, only values; the per-row sum is still the sequential CSR sum,
 * bit-exact with _seqMult. Host structure in two calls: rbdia_count fills
 * bundle_off and returns the slot count (bundle_off[nbundles]), or -1 when a
 * bundle holds more than max_width distinct offsets or a row's columns are
 * not strictly ascending; rbdia_struct fills doff
 * and dmask (that many entries each). Values are filled on the device. */
int64_t alien_hip_rbdia_count(int nrows, const int* row_offset_host, const int* cols_host, int max_width,
                              int* bundle_off_host);
int alien_hip_rbdia_struct(int nrows, const int* row_offset_host, const int* cols_host, int max_width,
                           int* doff_host, uint64_t* dmask_host);
/* The same structure built on the device from a device CSR (one wavefront per
 * bundle): rbdia_count_device fills the HOST array bundle_off and returns the
 * slot count, or -1 as rbdia_count; rbdia_struct_device writes doff and dmask
 * (device arrays of that many entries) given the device copy of bundle_off.
 * HIPMatrix's RB-DIA conversion runs these. pair != 0: the pair mode for rows
 * made of column pairs (c, c + 1), c even (two unknowns per cell, interleaved;
 * -1 when a row is not): a slot is one pair at column (row & ~1) + doff, its two
 * values at bvals[64 (2 s) + l] and bvals[64 (2 s + 1) + l] (rbdia2_fill,
 * spmv_rbdia2); same sums as CSR bit for bit. */
int64_t alien_hip_rbdia_count_device(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                                     int max_width, int pair, int* bundle_off_host);
int alien_hip_rbdia_struct_device(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                                  int pair, const int* bundle_off, int* doff, uint64_t* dmask);
int alien_hip_rbdia2_fill_f64(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                              const double* vals, const int* bundle_off, const int* doff, double* bvals);
int alien_hip_rbdia2_fill_f32(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                              const double* vals, const int* bundle_off, const int* doff, float* bvals);
int alien_hip_spmv_rbdia2_f64(alien_hip_ctx* ctx, int nrows, const int* bundle_off, const int* doff,
                              const uint64_t* dmask, const double* bvals, const double* x, double* y);
int alien_hip_spmv_rbdia2_f32(alien_hip_ctx* ctx, int nrows, const int* bundle_off, const int* doff,
                              const uint64_t* dmask, const float* bvals, const float* x, float* y);
int alien_hip_rbdia_fill_f64(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                             const double* vals, const int* bundle_off, const int* doff, double* bvals);
int alien_hip_rbdia_fill_f32(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                             const double* vals, const int* bundle_off, const int* doff, float* bvals);
int alien_hip_spmv_rbdia_f64(alien_hip_ctx* ctx, int nrows, const int* bundle_off, const int* doff,
                             const uint64_t* dmask, const double* bvals, const double* x, double* y);
int alien_hip_spmv_rbdia_f32(alien_hip_ctx* ctx, int nrows, const int* bundle_off, const int* doff,
                             const uint64_t* dmask, const float* bvals, const float* x, float* y);
/* fused on RB-DIA: y = A x (the sequential CSR sum of _seqMult) and *result = x . y (FP64
 * accumulation), the q = A p, p.q step of CG::solve (CG.h:133-134) as the fused CG iteration
 * runs it (the same kernel). */
int alien_hip_spmv_dot_rbdia_f64(alien_hip_ctx* ctx, int nrows, const int* bundle_off, const int* doff,
                                 const uint64_t* dmask, const double* bvals, const double* x, double* y,
                                 double* result);
int alien_hip_spmv_dot_rbdia_f32(alien_hip_ctx* ctx, int nrows, const int* bundle_off, const int* doff,
                                 const uint64_t* dmask, const float* bvals, const float* x, float* y,
                                 double* result);

/* Uniform RB-DIA ("UDIA"), the layout the fused solvers use for stencil matrices: an RB-DIA whose
 * bundles all hold the same k <= 8 slots, at the k distinct diagonal offsets of the whole matrix
 * (ascending), bundle b's at slots b k .. b k + k - 1, absent entries stored as zeros. The kernels
 * take the k offsets as arguments and load no per-bundle descriptor: one value load and one gather
 * per slot, all independent (RB-DIA's bundle offsets and slot descriptors are two dependent loads
 * before them). Sums: the CSR sum's nonzero terms in its order, plus +0 for each absent entry
 * (equal values; only the sign of a zero sum can differ). Same SpMV as SimpleCSRMatrix's
 * (SimpleCSRMatrixMultT.h:42-76, _seqMult); no counterpart layout in the reference.
 * host_diagonals (no GPU): the distinct col - row of the CSR, ascending, into offsets (room for
 * max_k), *k_out = their number or -1 if more than max_k. rbdia_uniform_struct: bundle_off
 * (nbundles + 1), doff and dmask (nbundles k entries) of that layout, an ordinary RB-DIA of the
 * matrix (rbdia_fill_* then fills its values); offsets is a host array. */
int alien_hip_host_diagonals(int nrows, const int* row_offset, const int* cols, int max_k, int* offsets,
                             int* k_out);
int alien_hip_rbdia_uniform_struct(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols, int k,
                                   const int* offsets, int* bundle_off, int* doff, uint64_t* dmask);
int alien_hip_spmv_udia_f64(alien_hip_ctx* ctx, int nrows, int k, const int* offsets, const double* bvals,
                            const double* x, double* y);
int alien_hip_spmv_udia_f32(alien_hip_ctx* ctx, int nrows, int k, const int* offsets, const float* bvals,
                            const float* x, float* y);
/* the fused q = A p, p.q step (CG.h:133-134) on UDIA, as the fused CG iteration runs it */
int alien_hip_spmv_dot_udia_f64(alien_hip_ctx* ctx, int nrows, int k, const int* offsets, const double* bvals,
                                const double* x, double* y, double* result);
int alien_hip_spmv_dot_udia_f32(alien_hip_ctx* ctx, int nrows, int k, const int* offsets, const float* bvals,
                                const float* x, float* y, double* result);

/* ------------------------------------------------------------------------ */
/* Device kernels: BLAS-1 (CBLASMPIKernel.h:37-125 / SYCLKernelInternal.h:110-246) */
/* ------------------------------------------------------------------------ */
int alien_hip_copy_f64(alien_hip_ctx* ctx, int n, const double* x, double* y);          /* y = x */
int alien_hip_axpy_f64(alien_hip_ctx* ctx, int n, double a, const double* x, double* y); /* y += a x */
int alien_hip_scal_f64(alien_hip_ctx* ctx, int n, double a, double* x);                  /* x = a x */
int alien_hip_assign_f64(alien_hip_ctx* ctx, int n, double a, double* x);                /* x = a */
int alien_hip_pointwise_mult_f64(alien_hip_ctx* ctx, int n, const double* x, const double* y,
                                 double* z); /* z = x .* y */
/* *result (device) = x . y, FP64 accumulation, one pass. Replaces the six SYCL
 * dot variants (SYCLKernelInternal.h:261-1085) and CBLAS ddot. */
int alien_hip_dot_f64(alien_hip_ctx* ctx, int n, const double* x, const double* y, double* result);
/* same for FP32 vectors, accumulated in FP64 */
int alien_hip_dot_f32(alien_hip_ctx* ctx, int n, const float* x, const float* y, double* result);
/* fused: y += a x, then *result = y . y (axpy + stop-test dot in one pass) */
int alien_hip_axpy_dot_f64(alien_hip_ctx* ctx, int n, double a, const double* x, double* y,
                           double* result);
/* fused: y = A x (CSR, bit-exact with _seqMult), *result = y . x */
int alien_hip_spmv_dot_csr_f64(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                               const double* vals, const double* x, double* y, double* result);
/* Allocation-free variant of alien_hip_spmv_dot_csr_f64. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). It takes
 * a HIP stream (NULL = the null stream of the current device) instead of a
 * context and allocates nothing: the reduction's per-workgroup partials live
 * in a static 512 KiB device array that each GPU holds until the process ends.
 * nrows above ALIEN_HIP_NOALLOC_MAX_ROWS is rejected with ALIEN_HIP_EINVAL.
 * Every call on one GPU shares that array: calls that overlap on one GPU (on
 * different streams, or from different threads) race and must be ordered by
 * the caller. Results equal alien_hip_spmv_dot_csr_f64's bit for bit. */
#define ALIEN_HIP_NOALLOC_MAX_ROWS 16777216
int alien_hip_spmv_dot_csr_f64_noalloc(void* stream, int nrows, const int* row_offset, const int* cols,
                                       const double* vals, const double* x, double* y, double* result);

/* ------------------------------------------------------------------------ */
/* Device kernels: ILU0 on level-set schedules                               */
/* ------------------------------------------------------------------------ */
/* In-place ILU(0) factorisation, IKJ per row, rows of one level in parallel,
 * (every row must store its diagonal: the pivots are read at dcol),
 * levels in order: bit-exact with LUFactorisationAlgo::factorize, sequential
 * branch (ILU0Preconditioner.h:71-179, 156-178). level_rows lists the rows of
 * level l in level_ptr[l]..level_ptr[l+1] (built on the host from
 * alien_hip_host_levels with upper = 0). */
int alien_hip_ilu0_factor_f64(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                              const int* dcol, double* lu, int nlevels, const int* level_ptr_host,
                              const int* level_rows);
/* one fixed-point sweep of FILU0's factorisation (FLUFactorisationAlgo::factorizeMultiIter /
 * _factorizeRow, FILU0Preconditioner.h:146-170, 328-357): lu = A's values, then the IKJ update
 * of every row i >= 1 with the other rows read from prev (device pointers, CSR of A) */
int alien_hip_filu0_sweep_f64(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                              const int* dcol, const double* vals, const double* prev, double* lu);
/* L x1 = y then U x = x1 (solveL/solveU, ILU0Preconditioner.h:181-231), level by
 * level; t is an n-vector of workspace. lower/upper schedules as above. */
int alien_hip_ilu0_solve_f64(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                             const int* dcol, const double* lu, int nlev_l, const int* lptr_host,
                             const int* lrows, int nlev_u, const int* uptr_host, const int* urows,
                             const double* y, double* t, double* x);

/* ------------------------------------------------------------------------ */
/* Whole solves through the C++ backend (test_krylov's run(alg) with         */
/* --kernel hip): examples/test_krylov.cpp:500-665                           */
/* ------------------------------------------------------------------------ */
enum { ALIEN_HIP_CG = 0, ALIEN_HIP_BICGS = 1 };
/* Preconditioners per mode: REFERENCE and REFERENCE_FUTURE run all but
 * LEARNED; FUSED / GRAPH / MIXED run all of them with both solvers (ILU0 on
 * level sets, ilu_schedule = 0, launched directly in GRAPH mode; FILU0 with
 * filu_solver_niter >= 1; Neumann with poly_order >= 1). */
enum { ALIEN_HIP_PREC_DIAG = 0, ALIEN_HIP_PREC_CHEB = 1, ALIEN_HIP_PREC_ILU0 = 2,
       ALIEN_HIP_PREC_NONE = 3,
       ALIEN_HIP_PREC_LEARNED = 4, /* M = G G^T from the graph network (FUSED/GRAPH/MIXED modes) */
       ALIEN_HIP_PREC_FILU0 = 5, /* FILU0Preconditioner: fixed-point ILU(0) factor,
                                    Jacobi triangular sweeps */
       ALIEN_HIP_PREC_NEUMANN = 6 /* NeumannPolyPreconditioner: poly_factor,
                                     poly_order, poly_max_iter */ };
/* RB-CSR: row-bundled CSR; RB-DIA: row-bundled diagonal offsets (no column
 * stream; RB-CSR when it would store over 1.25x the slots). In the FUSED, GRAPH
 * and MIXED modes layout RBCSR forces RB-CSR, any other value selects RB-DIA. */
enum { ALIEN_HIP_LAYOUT_CSR = 0, ALIEN_HIP_LAYOUT_BELLPACK = 1, ALIEN_HIP_LAYOUT_RBCSR = 2,
       ALIEN_HIP_LAYOUT_RBDIA = 3 };
enum {
  ALIEN_HIP_MODE_REFERENCE = 0, /* the reference call sequence, one launch per algebra call */
  ALIEN_HIP_MODE_FUSED = 1,     /* fused kernels, device-resident scalars, host polls stop test */
  ALIEN_HIP_MODE_GRAPH = 2,     /* fused iteration captured in a hipGraph, batched stop polls */
  ALIEN_HIP_MODE_MIXED = 3,     /* FP32 inner solver inside FP64 iterative refinement */
  ALIEN_HIP_MODE_REFERENCE_FUTURE = 4 /* the templates' solve2 (CG.h:152-244, BiCGStab.h:302-457):
                                         the REFERENCE sequence with FutureType dots */
};

typedef struct {
  int    solver;        /* ALIEN_HIP_CG / ALIEN_HIP_BICGS */
  int    precond;       /* ALIEN_HIP_PREC_* */
  int    layout;        /* ALIEN_HIP_LAYOUT_* for the SpMV of the solve */
  int    mode;          /* ALIEN_HIP_MODE_* */
  double tol;           /* relative residual (Iteration: ||r||^2 < tol^2 ||b||^2) */
  int    max_iter;
  double poly_factor;   /* Chebyshev: test_krylov --poly-factor */
  int    poly_order;    /* --poly-order */
  int    poly_max_iter; /* --poly-factor-max-iter */
  int    ilu_schedule;  /* 0 = level sets (natural order, bit-exact), 1 = multicolour */
  double inner_tol;     /* MIXED: reliable-update threshold delta: the FP64 true residual
                           replaces the FP32 one once it has dropped by delta (CG: below the
                           last update's; BiCGStab: below the largest since then) (0 = 0.1) */
  int    inner_max_iter;/* unused (kept for layout) */
  int    max_outer;     /* unused (kept for layout) */
  int    poll_every;    /* GRAPH: iterations per graph launch / stop-flag poll */
  const float* gnn_theta; /* LEARNED: network parameters (alien_hip_gnn_nparams() floats) */
  int    stall_window;  /* LEARNED: iterations in which the residual norm does not halve
                           (from its level at the last such halving) before the ILU0 fallback
                           (multicolour ILU0 of the colour-permuted matrix, from the current
                           iterate, in the solve's mode); MIXED: the true residual of a forced
                           reliable update must confirm the stall (mixed BiCGStab also falls
                           back when the true residual rises 100x above its lowest, from the
                           iterate of the lowest one).
                           0 = max(400, 8 N^(1/3)). A non-finite residual falls back at once. */
  int    filu_factor_niter; /* FILU0: nb-factor-iter (0 = exact ILU(0) factor) */
  int    filu_solver_niter; /* FILU0: nb-solver-iter (test_krylov default 3) */
  int    factor_storage;    /* LEARNED, MIXED CG: G and G^T values stored as 0 = FP32, 1 = bfloat16
                               (FP32 arithmetic, half the factor bytes; 2% faster at 256^3) */
} alien_hip_solve_params;

typedef struct {
  int    iterations;    /* Iteration::operator()() (MIXED: total inner iterations) */
  int    status;        /* Iteration::getStatus() */
  int    error;         /* 0, or 1 = NullValueException */
  int    outer;         /* MIXED: refinement steps */
  double value;         /* Iteration::getValue() */
  double true_rel_res;  /* ||b - A x||_2 / ||b||_2, recomputed in FP64 on the GPU */
  double cheb_alpha;
  double cheb_beta;
  double setup_s;       /* conversion + preconditioner init, hipEvent-timed */
  double solve_s;       /* the Krylov solve, hipEvent-timed */
  int    fallback;      /* bit 0 (LEARNED): the learned M stalled and ILU0 took over;
                           bit 1 (MIXED): the FP32 inner iteration overflowed, or (BiCGStab)
                           diverged or stagnated, and the solve finished in FP64 from the last
                           reliable x (BiCGStab: the iterate of the lowest true residual) */
  int    pad_;
} alien_hip_solve_result;

/* Host CSR in, host x out (x holds the initial guess on entry). Uploads,
 * converts to the requested layout, runs, downloads x. Synchronous. */
/* A converted matrix kept between solves, as Alien's MultiMatrixImpl keeps the
 * hip matrix and converts again only on a new timestamp (MultiMatrixImpl.cc:226-248):
 * matrix_create uploads a host CSR (ascending columns; the arrays are copied and
 * may be freed on return); matrix_update_values replaces the values on the same
 * structure (host array of nnz values) and refills every layout already built;
 * alien_hip_solve runs alien_hip_solve_csr's solve on it (row-bundled layouts are
 * built at the first solve that needs them and kept). Multicolour ILU0
 * (ilu_schedule = 1) builds the colour-permuted matrix at its first solve and
 * keeps it; matrix_update_values then gathers its values on the device. A
 * matrix is used with contexts of the device it was created on, one solve at
 * a time. */
typedef struct alien_hip_matrix alien_hip_matrix;
int alien_hip_matrix_create(alien_hip_ctx* ctx, int nrows, const int* row_offset_host, const int* cols_host,
                            const double* vals_host, alien_hip_matrix** out);
int alien_hip_matrix_update_values(alien_hip_matrix* m, const double* vals_host);
int alien_hip_matrix_destroy(alien_hip_matrix* m);
int alien_hip_solve(alien_hip_ctx* ctx, const alien_hip_solve_params* p, alien_hip_matrix* m,
                    const double* b_host, double* x_host, alien_hip_solve_result* result);

int alien_hip_solve_csr(alien_hip_ctx* ctx, const alien_hip_solve_params* p, int nrows,
                        const int* row_offset, const int* cols, const double* vals,
                        const double* b, double* x, alien_hip_solve_result* res);

/* ------------------------------------------------------------------------ */
/* MatrixMarket system I/O (host). Reading follows MatrixMarketReader        */
/* (src/movesemantic/alien/move/impl/MatrixMarketReader.cc): "coordinate     */
/* real" matrices, non-"general" symmetry mirrored, repeated entries summed, */
/* CSR with ascending columns; "array real" one-column vectors. Writing      */
/* follows MatrixMarketSystemWriter::dump (src/refsemantic/alien/ref/        */
/* import_export/MatrixMarketSystemWriter.cc:34-140), values with 17 digits. */
/* ------------------------------------------------------------------------ */
/* Two calls: with row_offset (values) NULL only the sizes are returned. */
int alien_hip_mtx_read_csr(const char* path, int* nrows, int64_t* nnz, int* row_offset, int* cols, double* vals);
int alien_hip_mtx_read_vector(const char* path, int* n, double* values);
int alien_hip_mtx_write_csr(const char* path, int nrows, const int* row_offset, const int* cols, const double* vals,
                            const char* description);
int alien_hip_mtx_write_vector(const char* path, int n, const double* values, const char* description);

/* ------------------------------------------------------------------------ */
/* Learned preconditioner M = G G^T (no counterpart in the reference)        */
/* ------------------------------------------------------------------------ */
/* A graph network (3 message-passing layers, hidden width 32, bf16 weights on
 * the matrix cores, FP32 accumulation) maps an SPD matrix with a symmetric
 * pattern to a lower-triangular G on A's lower pattern. Parameters are one
 * FP32 vector of alien_hip_gnn_nparams() entries. All pointers are HOST
 * pointers; the calls upload, compute on the GPU and synchronise. */
int alien_hip_gnn_nparams(void);
/* deterministic initial parameters (host only): random layers, output heads at
 * the first-order inverse factor Ghat = I + E_lower */
int alien_hip_gnn_init(uint64_t seed, float* theta);
/* Ghat per CSR entry of A (strictly-upper entries 0), G = D^-1/2 Ghat */
int alien_hip_gnn_forward(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                          const double* vals, const float* theta, float* ghat);
/* loss = L1 + beta * probe term |V - Ghat Ghat^T Ahat V|^2 / (4 nrows) on 4 probe vectors V
 * (nrows x 4, row-major). L1 = the residual reduction after cg_steps preconditioned CG steps,
 * mean over the 4 columns of V as right-hand sides of log(|r_k|^2 / |r_0|^2) (cg_steps >= 1), or
 * Kaporin's condition number log(tr(Ghat^T Ahat Ghat)/N) - (2/N) sum log Ghat_ii (cg_steps = 0).
 * parts[0] = L1, parts[1] = probe term; grad has nparams entries */
int alien_hip_gnn_loss_grad(alien_hip_ctx* ctx, int nrows, const int* row_offset, const int* cols,
                            const double* vals, const float* theta, const float* probes, double beta,
                            int cg_steps, double* loss, double* parts, float* grad);
typedef struct {
  int      steps;            /* Adam steps */
  int      steps_per_system; /* steps on one generated system before drawing the next */
  int      n_min, n_max;     /* grid edge of the generated systems (log-normal diffusion, n^3) */
  double   sigma_min, sigma_max; /* log-normal sigma, uniform */
  double   aniso_max;        /* per-axis anisotropy, log-uniform in [1, aniso_max] */
  double   lr;               /* Adam learning rate */
  double   beta;             /* weight of the probe term */
  uint64_t seed;
  int      cg_steps;         /* L1 of alien_hip_gnn_loss_grad: CG steps, 0 = Kaporin */
  int      pad_;
} alien_hip_gnn_train_params;
/* trains theta in place on the GPU; loss_hist (may be NULL) gets one value per step */
int alien_hip_gnn_train(alien_hip_ctx* ctx, const alien_hip_gnn_train_params* p, float* theta,
                        double* loss_hist);

#ifdef __cplusplus
}
#endif

#endif /* ALIEN_HIP_H */
