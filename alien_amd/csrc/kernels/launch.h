// Host-side launchers of the gfx950 kernels (definitions in the .hip files).
// All launches are asynchronous on the given stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "krylov_state.h"

namespace alien_hip {

struct RedWork;
template <typename T> class DeviceArray;

// what the second stage of a reduction does with the totals (reduce.hip)
enum FinOp { kFinStore = 0, kFinRho1 = 1, kFinPQ = 2, kFinXR = 3, kFinXRDiag = 4, kFinRho1B = 5 };
void launch_finalize(hipStream_t, RedWork w, long nblk, int nslot, int op, double* out, CGState* st);
// the same second stage for the BiCGStab state (reduce.hip)
enum BFinOp { kBFinAlpha = 0, kBFinS = 1, kBFinOmega = 2, kBFinXR = 3 };
void launch_finalize_bicg(hipStream_t, RedWork w, long nblk, int nslot, int op, BiCGState* st);

// ---- BLAS-1 (blas1.hip) ----
template <typename T> void launch_copy(hipStream_t, int n, const T* x, T* y);
// y[k] = x[idx[k]], k < n
void launch_gather_f64(hipStream_t, long n, const int* idx, const double* x, double* y);
template <typename T> void launch_axpy(hipStream_t, int n, T a, const T* x, T* y);
template <typename T> void launch_scal(hipStream_t, int n, T a, T* x);
template <typename T> void launch_assign(hipStream_t, int n, T a, T* x);
template <typename T> void launch_pmult(hipStream_t, int n, const T* x, const T* y, T* z);
template <typename T> void launch_dot(hipStream_t, RedWork w, int n, const T* x, const T* y, double* out);
template <typename T> void launch_axpy_dot(hipStream_t, RedWork w, int n, T a, const T* x, T* y, double* out);

// ---- SpMV (spmv.hip) ----
template <typename T>
void launch_spmv_csr(hipStream_t, int n, const int* ro, const int* cols, const T* vals, const T* x, T* y);
template <typename T>
void launch_spmv_dot_csr(hipStream_t, RedWork w, int n, const int* ro, const int* cols, const T* vals,
                         const T* x, T* y, double* out);
template <typename T>
void launch_spmv_rbcsr(hipStream_t, int n, const int* bo, const int* bc, const T* bv, const T* x, T* y);
template <typename T>
void launch_rbcsr_fill(hipStream_t, int n, const int* ro, const int* cols, const double* vals, const int* bo,
                       int* bc, T* bv);
// RB-CSR (doff NULL) or RB-DIA SpMV, y = A x
template <typename T>
void launch_spmv_rb(hipStream_t, int n, RBMat<T> A, const T* x, T* y);
template <typename T>
void launch_rbdia_fill(hipStream_t, int n, const int* ro, const int* cols, const double* vals, const int* bo,
                       const int* doff, T* bv);
// RB-DIA structure of a device CSR built on the device: rbdia_count_device
// returns the slot count and the bundle offsets on the host (-1 when a bundle
// holds more than max_w distinct offsets, a row's columns are not strictly
// ascending or the slots exceed int32 indexing); launch_rbdia_struct writes
// the offsets and lane masks at device bundle offsets bo; rbdia_structure_device
// does both and keeps the structure in the given device arrays.
// pair: the structure of RBMat's pair mode (rows of column pairs (c, c + 1), c even)
long rbdia_count_device(hipStream_t, int n, const int* ro, const int* cols, int max_w, std::vector<int>& bo,
                        bool pair = false);
void launch_rbdia_struct(hipStream_t, int n, const int* ro, const int* cols, const int* bo, int* doff,
                         unsigned long long* dmask, bool pair = false);
// uniform RB-DIA (RBMat::uni): bo[b] = b k, every bundle's slots at the k ascending offsets `off`
// (host array, the union of the matrix's diagonals), lane masks as in RB-DIA
void launch_rbdia_uniform_struct(hipStream_t, int n, const int* ro, const int* cols, int k, const int* off, int* bo,
                                 int* doff, unsigned long long* dmask);
// values of the pair mode (two planes per slot)
template <typename T>
void launch_rbdia2_fill(hipStream_t, int n, const int* ro, const int* cols, const double* vals, const int* bo,
                        const int* doff, T* bv);
long rbdia_structure_device(hipStream_t, int n, const int* ro, const int* cols, int max_w, std::vector<int>& bo,
                            DeviceArray<int>& d_bo, DeviceArray<int>& doff, DeviceArray<unsigned long long>& dmask);
// y = A x on Alien's BEllPack arrays: one wavefront per 64 rows of a slice when the slice
// height is a multiple of 64, else one lane per row
template <typename T>
void launch_spmv_bellpack(hipStream_t, int n, int slice, const int* bro, const int* bc, const T* bv,
                          const T* x, T* y);
// BEllPack arrays (bro[nslices] * slice entries) from a device CSR; any of bc / bv / lmask /
// umask may be NULL; dcol is read only for the masks
template <typename T>
void launch_bellpack_fill(hipStream_t, int n, int slice, const int* ro, const int* cols, const double* vals,
                          const int* dcol, const int* bro, int* bc, T* bv, uint8_t* lmask, uint8_t* umask);
void launch_addmult_bellpack(hipStream_t, int n, int slice, double alpha, const int* bro, const int* bc,
                             const double* bv, const uint8_t* mask, const double* x, double* y);
void launch_addmult_csr(hipStream_t, int n, int upper, double alpha, const int* ro, const int* cols,
                        const int* dcol, const double* vals, const double* x, double* y);
// one fixed-point sweep of FILU0's factorisation (ilu0.hip)
void launch_filu0_sweep(hipStream_t, int n, const int* ro, const int* cols, const int* dcol, const double* a,
                        const double* prev, double* lu);
template <typename T>
void launch_spmv_bsr2(hipStream_t, int nb, const int* ro, const int* cols, const T* vals, const T* x, T* y);
// RB-BSR2: 2x2 block rows bundled by 64 (bundle offsets as RB-CSR's over the block rows)
template <typename T>
void launch_spmv_rbbsr2(hipStream_t, int nbr, const int* bo, const int* bc, const T* bv, const T* x, T* y);
template <typename T>
void launch_rbbsr2_fill(hipStream_t, int nbr, const int* ro, const int* cols, const double* vals, const int* bo,
                        int* bc, T* bv);
// variable-block CSR product, one lane per block row
void launch_spmv_vblock(hipStream_t, int nbr, const int* ro, const int* cols, const int* bcols, const int* bsize,
                        const int* boff, const double* vals, const double* x, double* y);
// diagonal positions of a device CSR (first entry with column >= row)
void launch_dcol(hipStream_t, int n, const int* ro, const int* cols, int* dcol);
void launch_invdiag_csr(hipStream_t, int n, const int* dcol, const double* vals, double* y);
void launch_invdiag_bellpack(hipStream_t, int n, int slice, const int* bro, const int* bc, const double* bv,
                             double* y);

// ---- ILU0 (ilu0.hip) ----
template <typename T>
void launch_ilu0_factor(hipStream_t, int nlevels, const int* ptr_host, const int* rows, const int* ro,
                        const int* cols, const int* dcol, T* lu);
// ILU(0) factor of a colour-permuted matrix split into row-bundled strictly-lower
// and strictly-upper parts (bundle offsets lbo / ubo in 64-entry slots, columns
// -1 = padding) and the diagonal
template <typename T>
struct ILUSplit {
  const int* lbo = nullptr;
  const int* lbc = nullptr;
  const T* lbv = nullptr;
  const int* ubo = nullptr;
  const int* ubc = nullptr;
  const T* ubv = nullptr;
  const T* diag = nullptr;
};
// fills the split parts from LU in CSR order (columns too when lbc / ubc are not null)
template <typename T>
void launch_ilu0_split_fill(hipStream_t, int n, const int* ro, const int* cols, const int* dcol, const double* lu,
                            const int* lbo, const int* ubo, int* lbc, int* ubc, T* lbv, T* ubv, T* diag);
// multicolour ILU(0) apply x = (LU)^-1 y, one launch per colour and sweep
// (cptr_host: colour c = rows [cptr[c], cptr[c+1])); returns early once *done.
// With rdot the upper sweeps also leave the partials of rdot . x in red (their
// count is returned, ilu0_split_partials() beforehand) for a finalize.
int ilu0_split_partials(int ncolours, const int* cptr_host);
// FILU0 solve z = (LU)^-1 y by niter >= 1 Jacobi sweeps per factor part over the
// split factor of the natural-order matrix (F.diag: the inverse diagonal); t1,
// t2: workspace. With rdot the last sweep leaves the partials of rdot . z.
template <typename T>
int launch_filu0_split_apply(hipStream_t, int n, int niter, ILUSplit<T> F, const T* y, T* t1, T* t2, T* z,
                             const int* done, const T* rdot, RedWork red);
template <typename T>
int launch_ilu0_split_apply(hipStream_t, int ncolours, const int* cptr_host, ILUSplit<T> F, const T* y, T* t, T* x,
                            const int* done, const T* rdot, RedWork red);
// One sweep of the natural-order ILU(0) solve in level order: position q of the
// concatenated level sets holds factor row row[q]; its strictly-lower (lower
// sweep) or strictly-upper (upper sweep) entries, in ascending column order,
// sit in slots s = 0..k-1 at [s * n + q] (column -1: no entry); the upper
// sweep also keeps u_ii per position. A level's per-row data is then read
// contiguously; only the solution vector is gathered.
template <typename T>
struct LevelSweep {
  int n = 0, k = 0;          // rows, slots (1..kLevelMaxSlots)
  const int* row = nullptr;  // n
  const int* col = nullptr;  // k * n
  const T* val = nullptr;    // k * n
  const T* diag = nullptr;   // upper sweep: n
  // the sync-free form's work list (nchunks = 0: not built): chunk c holds positions
  // [cq[c], cq[c] + cn[c]) of one level, cn <= 64 * rpl, chunks in level order; each lane works
  // on rpl rows of a chunk; the launch has `blocks` workgroups
  int nchunks = 0;
  const int* cq = nullptr;
  const int* cn = nullptr;
  int rpl = 1, blocks = 64;
};
constexpr int kLevelMaxSlots = 16;
// levels of at most this many rows run in one workgroup (runs of two or more);
// ALIEN_HIP_ILU_WG_ROWS, read when a context is created, overrides it
extern long g_ilu_wg_rows;
// Sync-free sweeps (see ilu0.hip): on / off (ALIEN_HIP_ILU_SYNC_FREE), built for level-set
// schedules whose levels hold at most g_ilu_sf_avg_rows rows on average (ALIEN_HIP_ILU_SF_AVG_ROWS:
// many small levels, where the launches per level dominate); rows per lane and workgroups from
// level_sf_shape unless ALIEN_HIP_ILU_SF_RPL (1, 2 or 4) or ALIEN_HIP_ILU_SF_BLOCKS set them; each
// wait bounded by g_ilu_sf_spins polls (ALIEN_HIP_ILU_SF_SPINS); all read when a context is created.
extern int g_ilu_sync_free;
extern long g_ilu_sf_avg_rows, g_ilu_sf_blocks, g_ilu_sf_rpl, g_ilu_sf_spins;
// level-aligned chunks of at most `rows` positions over the level pointers ptr (nlev levels): cq, cn
void level_chunks(int nlev, const int* ptr, int rows, std::vector<int>& cq, std::vector<int>& cn);
// rows per lane and workgroups of the sync-free sweeps of an n-row schedule of nlev levels
void level_sf_shape(long n, long nlev, int& rpl, int& blocks);
// Slot count of a sweep (the largest strictly-lower or strictly-upper row
// length, 9..16 rounded up to 12 or 16), from the host CSR; 0 when it exceeds
// kLevelMaxSlots.
int level_sweep_slots(int n, const int* ro_host, const int* dcol_host, bool upper);
// Fill one sweep's columns (when col is not null), values and diagonal from the
// factor lu in CSR order (device arrays; row: the level order).
template <typename T>
void launch_level_sweep_fill(hipStream_t, int n, int k, bool upper, const int* row, const int* ro, const int* cols,
                             const int* dcol, const double* lu, int* col, T* val, T* diag);
// ILU(0) solve z = (LU)^-1 y over the level-ordered sweeps (bit-exact with
// launch_ilu0_solve): runs of small levels in one workgroup, larger levels one
// launch each. ptr_host / ptr_dev: the level pointers.
// sf_work: two device ints (ticket, error flag) for the sync-free form, which runs when both
// sweeps carry work lists (and lptr_dev / uptr_dev are given); a wait that runs out of polls
// raises the flag and a one-workgroup kernel redoes that sweep level by level (the result is the
// same either way, bit for bit). The sync-free form makes no host call: it may be captured.
template <typename T>
void launch_ilu0_solve_lev(hipStream_t, int nl, const int* lptr_host, const int* lptr_dev, LevelSweep<T> L, int nu,
                           const int* uptr_host, const int* uptr_dev, LevelSweep<T> U, const T* y, T* t, T* x,
                           const int* done = nullptr, int* sf_work = nullptr);
template <typename T>
void launch_ilu0_solve(hipStream_t, int nl, const int* lptr_host, const int* lrows, int nu,
                       const int* uptr_host, const int* urows, const int* ro, const int* cols, const int* dcol,
                       const T* lu, const T* y, T* t, T* x, const int* done = nullptr,
                       const int* lptr_dev = nullptr, const int* uptr_dev = nullptr);

} // namespace alien_hip
