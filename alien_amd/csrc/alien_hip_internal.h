// Internal declarations shared by the host code and the HIP sources of
// libalien_hip.so. Not part of the public C ABI (that is include/alien_hip.h).
#pragma once

#include "../../include/alien_hip.h"

#include <algorithm>
#include <cstdint>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace alien_hip {

// ---- error reporting (thread-local last error, see alien_hip_last_error) ----
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// ---- host generators (host/generators.cpp) ----
uint64_t splitmix64_next(uint64_t& s);
double splitmix64_u01(uint64_t& s);
void fill_uniform01(uint64_t seed, int64_t n, double* out);
void fill_normal(uint64_t seed, int64_t n, double* out);
int64_t gen_nnz(int kind, int nx, int ny, int nz);
int gen_csr(int kind, int nx, int ny, int nz, const double* params, uint64_t seed, int* ro,
            int* cols, double* vals);
int gen_bsr2(int nx, int ny, int nz, const double* params, uint64_t seed, int* ro, int* cols,
             double* vals);

// ---- host structure builders (host/structure.cpp) ----
void host_dcol(int nrows, const int* ro, const int* cols, int* dcol);
int bellpack_nslices(int nrows, int slice);
void bellpack_bro(int nrows, int slice, const int* ro, int* bro);
void bellpack_fill(int nrows, int slice, const int* ro, const int* cols, const double* vals,
                   const int* bro, int* bcols, double* bvals, uint8_t* lmask, uint8_t* umask);
int host_levels(int nrows, const int* ro, const int* cols, const int* dcol, int upper, int* level);
int host_color(int nrows, const int* ro, const int* cols, int* color);
// rows grouped by level (stable: ascending row inside a level); ptr has nlevels+1 entries
void group_by_level(int nrows, const int* level, int nlevels, int* ptr, int* rows);

constexpr int kBundle = 64; // rows per RB-CSR bundle = one wavefront

// Host array left uninitialised (no zero fill): its pages are first touched by
// whichever threads write it.
template <typename T>
struct HostBuf {
  std::unique_ptr<T[]> p;
  size_t count = 0;
  void alloc(size_t m) {
    p.reset(m ? new T[m] : nullptr);
    count = m;
  }
  T* data() { return p.get(); }
  const T* data() const { return p.get(); }
  size_t size() const { return count; }
  T& operator[](size_t i) { return p[i]; }
  const T& operator[](size_t i) const { return p[i]; }
};

// fn(begin, end) over contiguous ranges of [0, n) on up to 16 host threads
// (one range, the caller's thread, below 64 Ki items)
template <typename F>
void parallel_ranges(long n, F fn) {
  const long hw = std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
  const long nt = std::min(hw, std::max(1L, n / 65536));
  if (nt <= 1) {
    fn(0L, n);
    return;
  }
  std::vector<std::thread> th;
  for (long t = 0; t < nt; ++t) th.emplace_back(fn, n * t / nt, n * (t + 1) / nt);
  for (auto& x : th) x.join();
}
int rbcsr_bundle_offsets(int nrows, const int* ro, int* bundle_off);
// RB-DIA structure of the entries [kb(r), ke(r)) of each row (NULL kb/ke: whole
// rows): bundle offsets (nb + 1), per-slot column offsets and row masks.
// Returns the slot count per lane summed over bundles, or -1 when a bundle has
// more than max_w distinct offsets, a row's columns are not strictly ascending
// or int32 indexing would overflow.
long rbdia_structure(int nrows, const int* ro, const int* cols, const int* kb, const int* ke, int max_w,
                     std::vector<int>& bo, std::vector<int>& doff, std::vector<uint64_t>& dmask);
// ---- MatrixMarket system I/O (host/mtx.cpp) ----
// Variable-block CSR (VBlock: block row i is block_size[i] scalar rows starting at
// block_offset[i]; block j of a row is row-major from vals[block_cols[j]]) -> scalar CSR.
// Scalar row block_offset[i] + k holds, block by block and column by column, the
// entries the variable-block product sums for it, in its order. Returns -1 when
// block_offset is not the running sum of block_size or a block column is out of range.
long vblock_scalar_nnz(int nbrows, const int* ro, const int* cols, const int* block_size);
int vblock_to_csr(int nbrows, const int* ro, const int* cols, const int* block_cols, const int* block_size,
                  const int* block_offset, const double* vals, int* sro, int* scols, double* svals);
int mtx_read_csr(const char* path, int& n, std::vector<int>& ro, std::vector<int>& cols, std::vector<double>& vals,
                 std::string& err);
int mtx_read_vector(const char* path, std::vector<double>& v, std::string& err);
int mtx_write_csr(const char* path, int n, const int* ro, const int* cols, const double* vals, const char* desc,
                  std::string& err);
int mtx_write_vector(const char* path, int n, const double* v, const char* desc, std::string& err);

// RB-DIA is chosen over RB-CSR when it stores at most this many more slots
constexpr double kRbdiaMaxGrowth = 1.25;
constexpr int kRbdiaMaxWidth = 64;

} // namespace alien_hip
