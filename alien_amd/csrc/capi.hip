// C ABI of libalien_hip.so (declarations and reference citations in
// include/alien_hip.h). Each entry point validates what the host can see,
// launches on the context's stream and converts C++ exceptions into status
// codes plus a thread-local message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "alien/hip/HIPBackEnd.h"
#include "alien_hip_internal.h"
#include "kernels/device_common.h"
#include "kernels/launch.h"

namespace alien_hip {

static thread_local std::string t_last_error;

void set_error(const std::string& msg) { t_last_error = msg; }
int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

alien_hip_ctx* make_ctx(void* stream) {
  auto* c = new alien_hip_ctx();
  try {
    if (stream) {
      c->stream = (hipStream_t)stream;
    } else {
      ALIEN_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      c->own_stream = true;
    }
    c->reserve_red(4096);
    c->scalars.resize(alien_hip_ctx::kSlots);
    ALIEN_HIP_CHECK(hipMemsetAsync(c->scalars.data(), 0, sizeof(double) * alien_hip_ctx::kSlots, c->stream));
    c->idle_cg.resize(1);
    ALIEN_HIP_CHECK(hipMemsetAsync(c->idle_cg.data(), 0, sizeof(CGState), c->stream));
    ALIEN_HIP_CHECK(hipHostMalloc((void**)&c->host_scalars, sizeof(double) * alien_hip_ctx::kSlots,
                                  hipHostMallocDefault));
    ALIEN_HIP_CHECK(hipStreamSynchronize(c->stream));
  } catch (...) {
    destroy_ctx(c);
    throw;
  }
  return c;
}

void destroy_ctx(alien_hip_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->partials.reset();
  c->scalars.reset();
  c->idle_cg.reset();
  if (c->host_scalars) (void)hipHostFree(c->host_scalars);
  if (c->pinned_blocks) (void)hipHostFree(c->pinned_blocks);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

alien_hip_ctx* default_ctx() {
  static alien_hip_ctx* c = nullptr;
  static std::once_flag once;
  std::call_once(once, [] { c = make_ctx(nullptr); });
  return c;
}

} // namespace alien_hip

using namespace alien_hip;

#define API_BEGIN try {
#define API_END                                          \
  }                                                      \
  catch (const alien_hip::Error& e) {                    \
    return fail(e.code, e.what());                       \
  }                                                      \
  catch (const std::exception& e) {                      \
    return fail(ALIEN_HIP_EINVAL, e.what());             \
  }                                                      \
  catch (...) {                                          \
    return fail(ALIEN_HIP_EINVAL, "unknown exception");  \
  }                                                      \
  return 0;

#define REQUIRE(cond, msg)                            \
  do {                                                \
    if (!(cond)) return fail(ALIEN_HIP_EINVAL, msg);  \
  } while (0)

#define LAUNCHED(what) check_launch(what)

// uniform RB-DIA operands of the C ABI: k ascending offsets, given by the caller
static bool uni_offsets_ok(int k, const int* off) {
  if (k < 1 || k > kUniMaxSlots || !off) return false;
  for (int j = 1; j < k; ++j)
    if (off[j] <= off[j - 1]) return false;
  return true;
}
template <typename T>
static RBMat<T> uni_mat(int n, int k, const int* off, const T* bv) {
  RBMat<T> A{nullptr, nullptr, bv};
  A.uni = k;
  A.nrows = n;
  for (int j = 0; j < k; ++j) A.uoff[j] = off[j];
  return A;
}


extern "C" {

const char* alien_hip_last_error(void) { return t_last_error.c_str(); }
const char* alien_hip_version(void) { return "alien_hip 0.1.0 (gfx950)"; }

int alien_hip_ctx_create(void* stream, alien_hip_ctx** out) {
  REQUIRE(out, "ctx_create: out is NULL");
  API_BEGIN
  // row-kernel grid caps (device_common.h)
  const char* names[4] = {"ALIEN_HIP_ROW_BLOCKS_F64", "ALIEN_HIP_ROW_BLOCKS_F32", "ALIEN_HIP_ROW_BLOCKS_UNI",
                          "ALIEN_HIP_ROW_BLOCKS_UNI_WIDE"};
  long* caps[4] = {&g_row_blocks_f64, &g_row_blocks_f32, &g_row_blocks_uni, &g_row_blocks_uni_wide};
  for (int i = 0; i < 4; ++i)
    if (const char* e = std::getenv(names[i])) {
      const long v = std::strtol(e, nullptr, 10);
      if (v >= kXcds) *caps[i] = v / kXcds * kXcds;
    }
  if (const char* e = std::getenv("ALIEN_HIP_ILU_WG_ROWS")) g_ilu_wg_rows = std::max(0L, std::strtol(e, nullptr, 10));
  if (const char* e = std::getenv("ALIEN_HIP_ILU_SYNC_FREE")) g_ilu_sync_free = std::atoi(e) != 0;
  if (const char* e = std::getenv("ALIEN_HIP_ILU_SF_AVG_ROWS"))
    g_ilu_sf_avg_rows = std::max(0L, std::strtol(e, nullptr, 10));
  if (const char* e = std::getenv("ALIEN_HIP_ILU_SF_BLOCKS")) // 0: from the level sizes
    g_ilu_sf_blocks = std::min(1024L, std::max(0L, std::strtol(e, nullptr, 10)));
  if (const char* e = std::getenv("ALIEN_HIP_ILU_SF_RPL")) g_ilu_sf_rpl = std::strtol(e, nullptr, 10); // 0: auto
  if (const char* e = std::getenv("ALIEN_HIP_ILU_SF_SPINS"))
    g_ilu_sf_spins = std::max(0L, std::strtol(e, nullptr, 10));
  *out = make_ctx(stream);
  if (const char* e = std::getenv("ALIEN_HIP_TRACE")) (*out)->trace = std::atoi(e) != 0;
  API_END
}

int alien_hip_ctx_destroy(alien_hip_ctx* ctx) {
  API_BEGIN
  destroy_ctx(ctx);
  API_END
}

int alien_hip_ctx_synchronize(alien_hip_ctx* ctx) {
  REQUIRE(ctx, "ctx is NULL");
  API_BEGIN
  ALIEN_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  API_END
}

void* alien_hip_ctx_stream(alien_hip_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int alien_hip_ctx_step_timing(alien_hip_ctx* ctx, int on, double* total_us, long long* launches) {
  REQUIRE(ctx, "ctx is NULL");
  API_BEGIN
  if (total_us) *total_us = ctx->step_us;
  if (launches) *launches = ctx->step_launches;
  if (on) {
    ctx->step_us = 0.0;
    ctx->step_launches = 0;
  }
  ctx->step_timing = on != 0;
  API_END
}

// ---- host generators / builders ------------------------------------------------
int64_t alien_hip_gen_nnz(int kind, int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  (void)kind;
  return gen_nnz(kind, nx, ny, nz);
}

int alien_hip_gen_csr(int kind, int nx, int ny, int nz, const double* params, uint64_t seed, int* ro,
                      int* cols, double* vals) {
  REQUIRE(ro && cols && vals, "gen_csr: NULL output");
  REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "gen_csr: grid sizes must be >= 1");
  REQUIRE(gen_nnz(kind, nx, ny, nz) <= 0x7fffffffLL, "gen_csr: nnz exceeds int32");
  API_BEGIN
  if (gen_csr(kind, nx, ny, nz, params, seed, ro, cols, vals) != 0) return fail(ALIEN_HIP_EINVAL, "gen_csr: bad kind or size");
  API_END
}

int alien_hip_gen_bsr2(int nx, int ny, int nz, const double* params, uint64_t seed, int* ro, int* cols,
                       double* vals) {
  REQUIRE(ro && cols && vals, "gen_bsr2: NULL output");
  REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "gen_bsr2: grid sizes must be >= 1");
  API_BEGIN
  if (gen_bsr2(nx, ny, nz, params, seed, ro, cols, vals) != 0) return fail(ALIEN_HIP_EINVAL, "gen_bsr2: bad size");
  API_END
}

int alien_hip_fill_uniform01(uint64_t seed, int64_t n, double* out) {
  REQUIRE(n >= 0 && (n == 0 || out), "fill_uniform01: bad arguments");
  fill_uniform01(seed, n, out);
  return 0;
}

int alien_hip_host_dcol(int nrows, const int* ro, const int* cols, int* dcol) {
  REQUIRE(nrows >= 0 && ro && (nrows == 0 || (cols && dcol)), "host_dcol: bad arguments");
  host_dcol(nrows, ro, cols, dcol);
  return 0;
}

int alien_hip_host_bellpack_bro(int nrows, int slice, const int* ro, int* bro) {
  REQUIRE(nrows >= 0 && slice > 0 && ro && bro, "host_bellpack_bro: bad arguments");
  bellpack_bro(nrows, slice, ro, bro);
  return 0;
}

int alien_hip_host_bellpack_fill(int nrows, int slice, const int* ro, const int* cols, const double* vals,
                                 const int* bro, int* bcols, double* bvals, uint8_t* lmask, uint8_t* umask) {
  REQUIRE(nrows >= 0 && slice > 0 && ro && bro && bcols, "host_bellpack_fill: bad arguments");
  REQUIRE(!bvals || vals, "host_bellpack_fill: bvals requested without vals");
  bellpack_fill(nrows, slice, ro, cols, vals, bro, bcols, bvals, lmask, umask);
  return 0;
}

int alien_hip_host_levels(int nrows, const int* ro, const int* cols, const int* dcol, int upper, int* level,
                          int* nlevels) {
  REQUIRE(nrows >= 0 && ro && level && nlevels && (nrows == 0 || (cols && dcol)), "host_levels: bad arguments");
  *nlevels = host_levels(nrows, ro, cols, dcol, upper, level);
  return 0;
}

int alien_hip_host_color(int nrows, const int* ro, const int* cols, int* color, int* ncolors) {
  REQUIRE(nrows >= 0 && ro && color && ncolors, "host_color: bad arguments");
  API_BEGIN
  *ncolors = host_color(nrows, ro, cols, color);
  API_END
}

int alien_hip_rbcsr_bundle_offsets(int nrows, const int* ro, int* bo) {
  REQUIRE(nrows >= 0 && ro && bo, "rbcsr_bundle_offsets: bad arguments");
  if (rbcsr_bundle_offsets(nrows, ro, bo) != 0) return fail(ALIEN_HIP_EINVAL, "RB-CSR storage exceeds int32");
  return 0;
}

int64_t alien_hip_rbdia_count(int nrows, const int* ro, const int* cols, int max_w, int* bo) {
  if (!(nrows >= 0 && ro && (cols || nrows == 0) && bo && max_w > 0))
    return fail(ALIEN_HIP_EINVAL, "rbdia_count: bad arguments"), -1;
  try {
    std::vector<int> b, doff;
    std::vector<uint64_t> m;
    const long slots = rbdia_structure(nrows, ro, cols, nullptr, nullptr, max_w, b, doff, m);
    if (slots < 0) return fail(ALIEN_HIP_EINVAL, "rbdia_count: too many offsets in a bundle"), -1;
    std::copy(b.begin(), b.end(), bo);
    return slots;
  } catch (const std::exception& e) {
    return fail(ALIEN_HIP_EINVAL, e.what()), -1;
  }
}

int alien_hip_rbdia_struct(int nrows, const int* ro, const int* cols, int max_w, int* doff_out,
                           uint64_t* dmask_out) {
  REQUIRE(nrows >= 0 && ro && (nrows == 0 || (cols && doff_out && dmask_out)) && max_w > 0,
          "rbdia_struct: bad arguments");
  API_BEGIN
  std::vector<int> b, doff;
  std::vector<uint64_t> m;
  const long slots = rbdia_structure(nrows, ro, cols, nullptr, nullptr, max_w, b, doff, m);
  if (slots < 0) return fail(ALIEN_HIP_EINVAL, "rbdia_struct: too many offsets in a bundle");
  std::copy(doff.begin(), doff.end(), doff_out);
  std::copy(m.begin(), m.end(), dmask_out);
  API_END
}

int64_t alien_hip_rbdia_count_device(alien_hip_ctx* ctx, int nrows, const int* ro, const int* cols, int max_w,
                                     int pair, int* bo) {
  if (!(ctx && nrows >= 0 && (nrows == 0 || (ro && cols)) && bo && max_w > 0))
    return fail(ALIEN_HIP_EINVAL, "rbdia_count_device: bad arguments"), -1;
  try {
    std::vector<int> b;
    const long slots = rbdia_count_device(ctx->stream, nrows, ro, cols, max_w, b, pair != 0);
    if (slots < 0)
      return fail(ALIEN_HIP_EINVAL, "rbdia_count_device: too many offsets in a bundle or columns not ascending"), -1;
    std::copy(b.begin(), b.end(), bo);
    return slots;
  } catch (const Error& e) {
    return fail(e.code, e.what()), -1;
  } catch (const std::exception& e) {
    return fail(ALIEN_HIP_EINVAL, e.what()), -1;
  }
}

int alien_hip_rbdia_struct_device(alien_hip_ctx* ctx, int nrows, const int* ro, const int* cols, int pair,
                                  const int* bo, int* doff, uint64_t* dmask) {
  REQUIRE(ctx, "ctx is NULL");
  REQUIRE(nrows >= 0 && (nrows == 0 || (ro && cols && bo && doff && dmask)), "rbdia_struct_device: bad arguments");
  API_BEGIN
  launch_rbdia_struct(ctx->stream, nrows, ro, cols, bo, doff, reinterpret_cast<unsigned long long*>(dmask),
                      pair != 0);
  LAUNCHED("rbdia_struct_device");
  API_END
}

int alien_hip_mtx_read_csr(const char* path, int* nrows, int64_t* nnz, int* ro, int* cols, double* vals) {
  REQUIRE(path && nrows && nnz, "mtx_read_csr: bad arguments");
  API_BEGIN
  int n = 0;
  std::vector<int> r, c;
  std::vector<double> v;
  std::string err;
  if (mtx_read_csr(path, n, r, c, v, err) != 0) return fail(ALIEN_HIP_EINVAL, "mtx_read_csr: " + err);
  *nrows = n;
  *nnz = (int64_t)c.size();
  if (ro) {
    if (!cols || !vals) return fail(ALIEN_HIP_EINVAL, "mtx_read_csr: NULL cols/vals");
    std::copy(r.begin(), r.end(), ro);
    std::copy(c.begin(), c.end(), cols);
    std::copy(v.begin(), v.end(), vals);
  }
  API_END
}

int alien_hip_mtx_read_vector(const char* path, int* n, double* values) {
  REQUIRE(path && n, "mtx_read_vector: bad arguments");
  API_BEGIN
  std::vector<double> v;
  std::string err;
  if (mtx_read_vector(path, v, err) != 0) return fail(ALIEN_HIP_EINVAL, "mtx_read_vector: " + err);
  *n = (int)v.size();
  if (values) std::copy(v.begin(), v.end(), values);
  API_END
}

int alien_hip_mtx_write_csr(const char* path, int nrows, const int* ro, const int* cols, const double* vals,
                            const char* description) {
  REQUIRE(path && nrows >= 0 && ro && (nrows == 0 || (cols && vals)), "mtx_write_csr: bad arguments");
  API_BEGIN
  std::string err;
  if (mtx_write_csr(path, nrows, ro, cols, vals, description, err) != 0)
    return fail(ALIEN_HIP_EINVAL, "mtx_write_csr: " + err);
  API_END
}

int alien_hip_mtx_write_vector(const char* path, int n, const double* values, const char* description) {
  REQUIRE(path && n >= 0 && (n == 0 || values), "mtx_write_vector: bad arguments");
  API_BEGIN
  std::string err;
  if (mtx_write_vector(path, n, values, description, err) != 0)
    return fail(ALIEN_HIP_EINVAL, "mtx_write_vector: " + err);
  API_END
}

// ---- device kernels -------------------------------------------------------------------
#define CTX_OK(ctx) REQUIRE(ctx, "ctx is NULL")

int alien_hip_spmv_csr_f64(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, const double* vals,
                           const double* x, double* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (ro && cols && vals && x && y)), "spmv_csr_f64: bad arguments");
  API_BEGIN
  launch_spmv_csr<double>(ctx->stream, n, ro, cols, vals, x, y);
  LAUNCHED("spmv_csr_f64");
  API_END
}

int alien_hip_spmv_csr_f32(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, const float* vals,
                           const float* x, float* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (ro && cols && vals && x && y)), "spmv_csr_f32: bad arguments");
  API_BEGIN
  launch_spmv_csr<float>(ctx->stream, n, ro, cols, vals, x, y);
  LAUNCHED("spmv_csr_f32");
  API_END
}

int alien_hip_spmv_bellpack_f64(alien_hip_ctx* ctx, int n, int slice, const int* bro, const int* bc,
                                const double* bv, const double* x, double* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && slice > 0 && (n == 0 || (bro && bc && bv && x && y)), "spmv_bellpack_f64: bad arguments");
  API_BEGIN
  launch_spmv_bellpack<double>(ctx->stream, n, slice, bro, bc, bv, x, y);
  LAUNCHED("spmv_bellpack_f64");
  API_END
}

int alien_hip_spmv_bellpack_f32(alien_hip_ctx* ctx, int n, int slice, const int* bro, const int* bc,
                                const float* bv, const float* x, float* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && slice > 0 && (n == 0 || (bro && bc && bv && x && y)), "spmv_bellpack_f32: bad arguments");
  API_BEGIN
  launch_spmv_bellpack<float>(ctx->stream, n, slice, bro, bc, bv, x, y);
  LAUNCHED("spmv_bellpack_f32");
  API_END
}

int alien_hip_bellpack_fill_f64(alien_hip_ctx* ctx, int n, int slice, const int* ro, const int* cols,
                                const double* vals, const int* dcol, const int* bro, int* bc, double* bv,
                                uint8_t* lmask, uint8_t* umask) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && slice > 0 && ro && bro && (!bv || vals) && (!(lmask || umask) || dcol) && (n == 0 || cols),
          "bellpack_fill_f64: bad arguments");
  API_BEGIN
  launch_bellpack_fill<double>(ctx->stream, n, slice, ro, cols, vals, dcol, bro, bc, bv, lmask, umask);
  LAUNCHED("bellpack_fill_f64");
  API_END
}

int alien_hip_bellpack_fill_f32(alien_hip_ctx* ctx, int n, int slice, const int* ro, const int* cols,
                                const double* vals, const int* bro, int* bc, float* bv) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && slice > 0 && ro && bro && (!bv || vals) && (n == 0 || cols), "bellpack_fill_f32: bad arguments");
  API_BEGIN
  launch_bellpack_fill<float>(ctx->stream, n, slice, ro, cols, vals, nullptr, bro, bc, bv, nullptr, nullptr);
  LAUNCHED("bellpack_fill_f32");
  API_END
}

int alien_hip_addmult_bellpack_f64(alien_hip_ctx* ctx, int upper, double alpha, int n, int slice, const int* bro,
                                   const int* bc, const double* bv, const uint8_t* mask, const double* x,
                                   double* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && slice > 0 && (n == 0 || (bro && bc && bv && mask && x && y)),
          "addmult_bellpack_f64: bad arguments");
  (void)upper; // the mask selects lower or upper
  API_BEGIN
  launch_addmult_bellpack(ctx->stream, n, slice, alpha, bro, bc, bv, mask, x, y);
  LAUNCHED("addmult_bellpack_f64");
  API_END
}

int alien_hip_addmult_csr_f64(alien_hip_ctx* ctx, int upper, double alpha, int n, const int* ro, const int* cols,
                              const int* dcol, const double* vals, const double* x, double* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (ro && cols && dcol && vals && x && y)), "addmult_csr_f64: bad arguments");
  API_BEGIN
  launch_addmult_csr(ctx->stream, n, upper ? 1 : 0, alpha, ro, cols, dcol, vals, x, y);
  LAUNCHED("addmult_csr_f64");
  API_END
}

static bool aligned(const void* p, size_t a) { return ((uintptr_t)p % a) == 0; }

int alien_hip_spmv_bsr2_f64(alien_hip_ctx* ctx, int nb, const int* ro, const int* cols, const double* vals,
                            const double* x, double* y) {
  CTX_OK(ctx);
  REQUIRE(nb >= 0 && (nb == 0 || (ro && cols && vals && x && y)), "spmv_bsr2_f64: bad arguments");
  REQUIRE(aligned(vals, 32) && aligned(x, 16) && aligned(y, 16), "spmv_bsr2_f64: misaligned block arrays");
  API_BEGIN
  launch_spmv_bsr2<double>(ctx->stream, nb, ro, cols, vals, x, y);
  LAUNCHED("spmv_bsr2_f64");
  API_END
}

int alien_hip_rbbsr2_fill_f64(alien_hip_ctx* ctx, int nb, const int* ro, const int* cols, const double* vals,
                              const int* bo, int* bc, double* bv) {
  CTX_OK(ctx);
  REQUIRE(nb >= 0 && (nb == 0 || (ro && cols && vals && bo && bc && bv)), "rbbsr2_fill_f64: bad arguments");
  API_BEGIN
  launch_rbbsr2_fill<double>(ctx->stream, nb, ro, cols, vals, bo, bc, bv);
  LAUNCHED("rbbsr2_fill_f64");
  API_END
}

int alien_hip_rbbsr2_fill_f32(alien_hip_ctx* ctx, int nb, const int* ro, const int* cols, const double* vals,
                              const int* bo, int* bc, float* bv) {
  CTX_OK(ctx);
  REQUIRE(nb >= 0 && (nb == 0 || (ro && cols && vals && bo && bc && bv)), "rbbsr2_fill_f32: bad arguments");
  API_BEGIN
  launch_rbbsr2_fill<float>(ctx->stream, nb, ro, cols, vals, bo, bc, bv);
  LAUNCHED("rbbsr2_fill_f32");
  API_END
}

int alien_hip_spmv_rbbsr2_f64(alien_hip_ctx* ctx, int nb, const int* bo, const int* bc, const double* bv,
                              const double* x, double* y) {
  CTX_OK(ctx);
  REQUIRE(nb >= 0 && (nb == 0 || (bo && bc && bv && x && y)), "spmv_rbbsr2_f64: bad arguments");
  REQUIRE(aligned(x, 16) && aligned(y, 16), "spmv_rbbsr2_f64: misaligned vectors");
  API_BEGIN
  launch_spmv_rbbsr2<double>(ctx->stream, nb, bo, bc, bv, x, y);
  LAUNCHED("spmv_rbbsr2_f64");
  API_END
}

int alien_hip_spmv_rbbsr2_f32(alien_hip_ctx* ctx, int nb, const int* bo, const int* bc, const float* bv,
                              const float* x, float* y) {
  CTX_OK(ctx);
  REQUIRE(nb >= 0 && (nb == 0 || (bo && bc && bv && x && y)), "spmv_rbbsr2_f32: bad arguments");
  REQUIRE(aligned(x, 8) && aligned(y, 8), "spmv_rbbsr2_f32: misaligned vectors");
  API_BEGIN
  launch_spmv_rbbsr2<float>(ctx->stream, nb, bo, bc, bv, x, y);
  LAUNCHED("spmv_rbbsr2_f32");
  API_END
}

int alien_hip_spmv_vblock_f64(alien_hip_ctx* ctx, int nb, const int* ro, const int* cols, const int* bcols,
                              const int* bsize, const int* boff, const double* vals, const double* x, double* y) {
  CTX_OK(ctx);
  REQUIRE(nb >= 0 && (nb == 0 || (ro && cols && bcols && bsize && boff && vals && x && y)),
          "spmv_vblock_f64: bad arguments");
  API_BEGIN
  launch_spmv_vblock(ctx->stream, nb, ro, cols, bcols, bsize, boff, vals, x, y);
  LAUNCHED("spmv_vblock_f64");
  API_END
}

int64_t alien_hip_vblock_scalar_nnz(int nb, const int* ro, const int* cols, const int* bsize) {
  if (!(nb >= 0 && ro && (nb == 0 || (cols && bsize))))
    return fail(ALIEN_HIP_EINVAL, "vblock_scalar_nnz: bad arguments"), -1;
  return vblock_scalar_nnz(nb, ro, cols, bsize);
}

int alien_hip_vblock_to_csr(int nb, const int* ro, const int* cols, const int* bcols, const int* bsize,
                            const int* boff, const double* vals, int* sro, int* scols, double* svals) {
  REQUIRE(nb >= 0 && ro && sro && (nb == 0 || (cols && bcols && bsize && boff && scols)),
          "vblock_to_csr: bad arguments");
  API_BEGIN
  if (vblock_to_csr(nb, ro, cols, bcols, bsize, boff, vals, sro, scols, svals) != 0)
    return fail(ALIEN_HIP_EINVAL, "vblock_to_csr: block_offset is not the running sum of block_size, "
                                  "or a block column is out of range");
  API_END
}

int alien_hip_spmv_bsr2_f32(alien_hip_ctx* ctx, int nb, const int* ro, const int* cols, const float* vals,
                            const float* x, float* y) {
  CTX_OK(ctx);
  REQUIRE(nb >= 0 && (nb == 0 || (ro && cols && vals && x && y)), "spmv_bsr2_f32: bad arguments");
  REQUIRE(aligned(vals, 16) && aligned(x, 8) && aligned(y, 8), "spmv_bsr2_f32: misaligned block arrays");
  API_BEGIN
  launch_spmv_bsr2<float>(ctx->stream, nb, ro, cols, vals, x, y);
  LAUNCHED("spmv_bsr2_f32");
  API_END
}

int alien_hip_invdiag_csr_f64(alien_hip_ctx* ctx, int n, const int* dcol, const double* vals, double* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (dcol && vals && y)), "invdiag_csr_f64: bad arguments");
  API_BEGIN
  launch_invdiag_csr(ctx->stream, n, dcol, vals, y);
  LAUNCHED("invdiag_csr_f64");
  API_END
}

int alien_hip_invdiag_bellpack_f64(alien_hip_ctx* ctx, int n, int slice, const int* bro, const int* bc,
                                   const double* bv, double* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && slice > 0 && (n == 0 || (bro && bc && bv && y)), "invdiag_bellpack_f64: bad arguments");
  API_BEGIN
  launch_invdiag_bellpack(ctx->stream, n, slice, bro, bc, bv, y);
  LAUNCHED("invdiag_bellpack_f64");
  API_END
}

int alien_hip_rbcsr_fill_f64(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, const double* vals,
                             const int* bo, int* bc, double* bv) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (ro && cols && vals && bo && bc && bv)), "rbcsr_fill_f64: bad arguments");
  API_BEGIN
  launch_rbcsr_fill<double>(ctx->stream, n, ro, cols, vals, bo, bc, bv);
  LAUNCHED("rbcsr_fill_f64");
  API_END
}

int alien_hip_rbcsr_fill_f32(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, const double* vals,
                             const int* bo, int* bc, float* bv) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (ro && cols && vals && bo && bc && bv)), "rbcsr_fill_f32: bad arguments");
  API_BEGIN
  launch_rbcsr_fill<float>(ctx->stream, n, ro, cols, vals, bo, bc, bv);
  LAUNCHED("rbcsr_fill_f32");
  API_END
}

int alien_hip_rbdia_fill_f64(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, const double* vals,
                             const int* bo, const int* doff, double* bv) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (ro && cols && vals && bo && doff && bv)), "rbdia_fill_f64: bad arguments");
  API_BEGIN
  launch_rbdia_fill<double>(ctx->stream, n, ro, cols, vals, bo, doff, bv);
  LAUNCHED("rbdia_fill_f64");
  API_END
}

int alien_hip_rbdia_fill_f32(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, const double* vals,
                             const int* bo, const int* doff, float* bv) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (ro && cols && vals && bo && doff && bv)), "rbdia_fill_f32: bad arguments");
  API_BEGIN
  launch_rbdia_fill<float>(ctx->stream, n, ro, cols, vals, bo, doff, bv);
  LAUNCHED("rbdia_fill_f32");
  API_END
}

int alien_hip_rbdia2_fill_f64(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, const double* vals,
                              const int* bo, const int* doff, double* bv) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (ro && cols && vals && bo && doff && bv)), "rbdia2_fill_f64: bad arguments");
  API_BEGIN
  launch_rbdia2_fill<double>(ctx->stream, n, ro, cols, vals, bo, doff, bv);
  LAUNCHED("rbdia2_fill_f64");
  API_END
}

int alien_hip_rbdia2_fill_f32(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, const double* vals,
                              const int* bo, const int* doff, float* bv) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (ro && cols && vals && bo && doff && bv)), "rbdia2_fill_f32: bad arguments");
  API_BEGIN
  launch_rbdia2_fill<float>(ctx->stream, n, ro, cols, vals, bo, doff, bv);
  LAUNCHED("rbdia2_fill_f32");
  API_END
}

int alien_hip_spmv_rbdia2_f64(alien_hip_ctx* ctx, int n, const int* bo, const int* doff, const uint64_t* dmask,
                              const double* bv, const double* x, double* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (bo && doff && dmask && bv && x && y)), "spmv_rbdia2_f64: bad arguments");
  API_BEGIN
  RBMat<double> A{bo, nullptr, bv, doff, reinterpret_cast<const unsigned long long*>(dmask), 1};
  launch_spmv_rb<double>(ctx->stream, n, A, x, y);
  LAUNCHED("spmv_rbdia2_f64");
  API_END
}

int alien_hip_spmv_rbdia2_f32(alien_hip_ctx* ctx, int n, const int* bo, const int* doff, const uint64_t* dmask,
                              const float* bv, const float* x, float* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (bo && doff && dmask && bv && x && y)), "spmv_rbdia2_f32: bad arguments");
  API_BEGIN
  RBMat<float> A{bo, nullptr, bv, doff, reinterpret_cast<const unsigned long long*>(dmask), 1};
  launch_spmv_rb<float>(ctx->stream, n, A, x, y);
  LAUNCHED("spmv_rbdia2_f32");
  API_END
}

int alien_hip_spmv_rbdia_f64(alien_hip_ctx* ctx, int n, const int* bo, const int* doff, const uint64_t* dmask,
                             const double* bv, const double* x, double* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (bo && doff && dmask && bv && x && y)), "spmv_rbdia_f64: bad arguments");
  API_BEGIN
  RBMat<double> A{bo, nullptr, bv, doff, reinterpret_cast<const unsigned long long*>(dmask)};
  launch_spmv_rb<double>(ctx->stream, n, A, x, y);
  LAUNCHED("spmv_rbdia_f64");
  API_END
}

int alien_hip_spmv_rbdia_f32(alien_hip_ctx* ctx, int n, const int* bo, const int* doff, const uint64_t* dmask,
                             const float* bv, const float* x, float* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (bo && doff && dmask && bv && x && y)), "spmv_rbdia_f32: bad arguments");
  API_BEGIN
  RBMat<float> A{bo, nullptr, bv, doff, reinterpret_cast<const unsigned long long*>(dmask)};
  launch_spmv_rb<float>(ctx->stream, n, A, x, y);
  LAUNCHED("spmv_rbdia_f32");
  API_END
}

int alien_hip_spmv_dot_rbdia_f64(alien_hip_ctx* ctx, int n, const int* bo, const int* doff, const uint64_t* dmask,
                                 const double* bv, const double* x, double* y, double* result) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && result && (n == 0 || (bo && doff && dmask && bv && x && y)), "spmv_dot_rbdia_f64: bad arguments");
  API_BEGIN
  ctx->reserve_for(n);
  RBMat<double> A{bo, nullptr, bv, doff, reinterpret_cast<const unsigned long long*>(dmask)};
  launch_spmv_dot_rb<double>(ctx->stream, ctx->red(), n, A, x, y, ctx->idle_cg.data(), result);
  LAUNCHED("spmv_dot_rbdia_f64");
  API_END
}

int alien_hip_spmv_dot_rbdia_f32(alien_hip_ctx* ctx, int n, const int* bo, const int* doff, const uint64_t* dmask,
                                 const float* bv, const float* x, float* y, double* result) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && result && (n == 0 || (bo && doff && dmask && bv && x && y)), "spmv_dot_rbdia_f32: bad arguments");
  API_BEGIN
  ctx->reserve_for(n);
  RBMat<float> A{bo, nullptr, bv, doff, reinterpret_cast<const unsigned long long*>(dmask)};
  launch_spmv_dot_rb<float>(ctx->stream, ctx->red(), n, A, x, y, ctx->idle_cg.data(), result);
  LAUNCHED("spmv_dot_rbdia_f32");
  API_END
}

// ---- uniform RB-DIA -------------------------------------------------------------------
int alien_hip_host_diagonals(int n, const int* ro, const int* cols, int max_k, int* offsets, int* k_out) {
  REQUIRE(n >= 0 && max_k >= 1 && offsets && k_out && (n == 0 || (ro && cols)), "host_diagonals: bad arguments");
  API_BEGIN
  std::vector<int> u;
  int k = 0;
  size_t last = 0; // consecutive entries mostly repeat a recent offset
  for (int r = 0; r < n && k >= 0; ++r)
    for (int e = ro[r]; e < ro[r + 1]; ++e) {
      const long d = (long)cols[e] - r;
      if (last < u.size() && u[last] == (int)d) continue;
      const auto it = std::find(u.begin(), u.end(), (int)d);
      if (it != u.end()) {
        last = (size_t)(it - u.begin());
        continue;
      }
      if ((int)u.size() == max_k) {
        k = -1;
        break;
      }
      u.push_back((int)d);
    }
  std::sort(u.begin(), u.end());
  if (k >= 0) {
    k = (int)u.size();
    for (int j = 0; j < k; ++j) offsets[j] = u[j];
  }
  *k_out = k;
  API_END
}

int alien_hip_rbdia_uniform_struct(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, int k,
                                   const int* offsets, int* bo, int* doff, uint64_t* dmask) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && uni_offsets_ok(k, offsets) && (n == 0 || (ro && cols && bo && doff && dmask)),
          "rbdia_uniform_struct: bad arguments");
  API_BEGIN
  launch_rbdia_uniform_struct(ctx->stream, n, ro, cols, k, offsets, bo, doff,
                              reinterpret_cast<unsigned long long*>(dmask));
  LAUNCHED("rbdia_uniform_struct");
  API_END
}

int alien_hip_spmv_udia_f64(alien_hip_ctx* ctx, int n, int k, const int* offsets, const double* bv, const double* x,
                            double* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && uni_offsets_ok(k, offsets) && (n == 0 || (bv && x && y)), "spmv_udia_f64: bad arguments");
  API_BEGIN
  launch_spmv_rb<double>(ctx->stream, n, uni_mat<double>(n, k, offsets, bv), x, y);
  LAUNCHED("spmv_udia_f64");
  API_END
}

int alien_hip_spmv_udia_f32(alien_hip_ctx* ctx, int n, int k, const int* offsets, const float* bv, const float* x,
                            float* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && uni_offsets_ok(k, offsets) && (n == 0 || (bv && x && y)), "spmv_udia_f32: bad arguments");
  API_BEGIN
  launch_spmv_rb<float>(ctx->stream, n, uni_mat<float>(n, k, offsets, bv), x, y);
  LAUNCHED("spmv_udia_f32");
  API_END
}

int alien_hip_spmv_dot_udia_f64(alien_hip_ctx* ctx, int n, int k, const int* offsets, const double* bv,
                                const double* x, double* y, double* result) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && result && uni_offsets_ok(k, offsets) && (n == 0 || (bv && x && y)),
          "spmv_dot_udia_f64: bad arguments");
  API_BEGIN
  ctx->reserve_for(n);
  launch_spmv_dot_rb<double>(ctx->stream, ctx->red(), n, uni_mat<double>(n, k, offsets, bv), x, y,
                             ctx->idle_cg.data(), result);
  LAUNCHED("spmv_dot_udia_f64");
  API_END
}

int alien_hip_spmv_dot_udia_f32(alien_hip_ctx* ctx, int n, int k, const int* offsets, const float* bv,
                                const float* x, float* y, double* result) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && result && uni_offsets_ok(k, offsets) && (n == 0 || (bv && x && y)),
          "spmv_dot_udia_f32: bad arguments");
  API_BEGIN
  ctx->reserve_for(n);
  launch_spmv_dot_rb<float>(ctx->stream, ctx->red(), n, uni_mat<float>(n, k, offsets, bv), x, y,
                            ctx->idle_cg.data(), result);
  LAUNCHED("spmv_dot_udia_f32");
  API_END
}

int alien_hip_filu0_sweep_f64(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, const int* dcol,
                              const double* vals, const double* prev, double* lu) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (ro && cols && dcol && vals && prev && lu)) && prev != lu,
          "filu0_sweep_f64: bad arguments");
  API_BEGIN
  launch_filu0_sweep(ctx->stream, n, ro, cols, dcol, vals, prev, lu);
  LAUNCHED("filu0_sweep_f64");
  API_END
}

int alien_hip_spmv_rbcsr_f64(alien_hip_ctx* ctx, int n, const int* bo, const int* bc, const double* bv,
                             const double* x, double* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (bo && bc && bv && x && y)), "spmv_rbcsr_f64: bad arguments");
  API_BEGIN
  launch_spmv_rbcsr<double>(ctx->stream, n, bo, bc, bv, x, y);
  LAUNCHED("spmv_rbcsr_f64");
  API_END
}

int alien_hip_spmv_rbcsr_f32(alien_hip_ctx* ctx, int n, const int* bo, const int* bc, const float* bv,
                             const float* x, float* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (bo && bc && bv && x && y)), "spmv_rbcsr_f32: bad arguments");
  API_BEGIN
  launch_spmv_rbcsr<float>(ctx->stream, n, bo, bc, bv, x, y);
  LAUNCHED("spmv_rbcsr_f32");
  API_END
}

// ---- BLAS-1 ----------------------------------------------------------------------------------
int alien_hip_copy_f64(alien_hip_ctx* ctx, int n, const double* x, double* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (x && y)), "copy_f64: bad arguments");
  API_BEGIN
  launch_copy<double>(ctx->stream, n, x, y);
  LAUNCHED("copy_f64");
  API_END
}
int alien_hip_axpy_f64(alien_hip_ctx* ctx, int n, double a, const double* x, double* y) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (x && y)), "axpy_f64: bad arguments");
  API_BEGIN
  launch_axpy<double>(ctx->stream, n, a, x, y);
  LAUNCHED("axpy_f64");
  API_END
}
int alien_hip_scal_f64(alien_hip_ctx* ctx, int n, double a, double* x) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || x), "scal_f64: bad arguments");
  API_BEGIN
  launch_scal<double>(ctx->stream, n, a, x);
  LAUNCHED("scal_f64");
  API_END
}
int alien_hip_assign_f64(alien_hip_ctx* ctx, int n, double a, double* x) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || x), "assign_f64: bad arguments");
  API_BEGIN
  launch_assign<double>(ctx->stream, n, a, x);
  LAUNCHED("assign_f64");
  API_END
}
int alien_hip_pointwise_mult_f64(alien_hip_ctx* ctx, int n, const double* x, const double* y, double* z) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && (n == 0 || (x && y && z)), "pointwise_mult_f64: bad arguments");
  API_BEGIN
  launch_pmult<double>(ctx->stream, n, x, y, z);
  LAUNCHED("pointwise_mult_f64");
  API_END
}
int alien_hip_dot_f64(alien_hip_ctx* ctx, int n, const double* x, const double* y, double* result) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && result && (n == 0 || (x && y)), "dot_f64: bad arguments");
  API_BEGIN
  ctx->reserve_for(n);
  launch_dot<double>(ctx->stream, ctx->red(), n, x, y, result);
  LAUNCHED("dot_f64");
  API_END
}
int alien_hip_dot_f32(alien_hip_ctx* ctx, int n, const float* x, const float* y, double* result) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && result && (n == 0 || (x && y)), "dot_f32: bad arguments");
  API_BEGIN
  ctx->reserve_for(n);
  launch_dot<float>(ctx->stream, ctx->red(), n, x, y, result);
  LAUNCHED("dot_f32");
  API_END
}
int alien_hip_axpy_dot_f64(alien_hip_ctx* ctx, int n, double a, const double* x, double* y, double* result) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && result && (n == 0 || (x && y)), "axpy_dot_f64: bad arguments");
  API_BEGIN
  ctx->reserve_for(n);
  launch_axpy_dot<double>(ctx->stream, ctx->red(), n, a, x, y, result);
  LAUNCHED("axpy_dot_f64");
  API_END
}
int alien_hip_spmv_dot_csr_f64(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, const double* vals,
                               const double* x, double* y, double* result) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && result && (n == 0 || (ro && cols && vals && x && y)), "spmv_dot_csr_f64: bad arguments");
  API_BEGIN
  ctx->reserve_for(n);
  launch_spmv_dot_csr<double>(ctx->stream, ctx->red(), n, ro, cols, vals, x, y, result);
  LAUNCHED("spmv_dot_csr_f64");
  API_END
}
} // extern "C"

// ---- allocation-free fused SpMV + dot ---------------------------------------------------------
// The per-workgroup partials of k_spmv_dot_csr (one double per workgroup of kBlock rows) live in
// a file-scope device array instead of a context's workspace, so a call allocates nothing. Each
// GPU holds its own copy of the array at its own address; the host-side name `scratch` is only a
// stand-in for it, so host code uses the address hipGetSymbolAddress returns for the device
// (taken once per device, on the first call there) and never the name itself.
namespace {
constexpr int kNoallocPartials = 65536;  // 512 KiB per GPU
constexpr int kNoallocMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoallocPartials];

std::mutex g_scratch_mu;
double* g_scratch_ptr[kNoallocMaxDevices] = {};

double* scratch_on_device(int dev) {
  if (dev < 0 || dev >= kNoallocMaxDevices) throw Error(ALIEN_HIP_EINVAL, "spmv_dot_csr_f64_noalloc: device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    ALIEN_HIP_CHECK(hipGetDevice(&cur));
    if (cur != dev) ALIEN_HIP_CHECK(hipSetDevice(dev));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    ALIEN_HIP_CHECK(e);
    if (!ptr) throw Error(ALIEN_HIP_EINVAL, "spmv_dot_csr_f64_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = (double*)ptr;
  }
  return g_scratch_ptr[dev];
}
} // namespace

extern "C" {
int alien_hip_spmv_dot_csr_f64_noalloc(void* stream, int n, const int* ro, const int* cols, const double* vals,
                                       const double* x, double* y, double* result) {
  REQUIRE(n >= 0 && result && (n == 0 || (ro && cols && vals && x && y)),
          "spmv_dot_csr_f64_noalloc: bad arguments");
  REQUIRE(n <= ALIEN_HIP_NOALLOC_MAX_ROWS,
          "spmv_dot_csr_f64_noalloc: nrows exceeds ALIEN_HIP_NOALLOC_MAX_ROWS (16777216)");
  static_assert((long)ALIEN_HIP_NOALLOC_MAX_ROWS <= (long)kNoallocPartials * kBlock,
                "the scratch array holds one partial per workgroup of kBlock rows");
  API_BEGIN
  const hipStream_t st = (hipStream_t)stream;
  int dev = -1;
  if (st) {
    hipDevice_t d = -1;
    ALIEN_HIP_CHECK(hipStreamGetDevice(st, &d));
    dev = (int)d;
  } else {
    ALIEN_HIP_CHECK(hipGetDevice(&dev));
  }
  double* part = scratch_on_device((int)dev);
  launch_spmv_dot_csr<double>(st, RedWork{part, kNoallocPartials}, n, ro, cols, vals, x, y, result);
  LAUNCHED("spmv_dot_csr_f64_noalloc");
  API_END
}

// ---- ILU0 --------------------------------------------------------------------------------------
int alien_hip_ilu0_factor_f64(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, const int* dcol,
                              double* lu, int nlevels, const int* lptr_host, const int* lrows) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && nlevels >= 0 && (n == 0 || (ro && cols && dcol && lu && lptr_host && lrows)),
          "ilu0_factor_f64: bad arguments");
  REQUIRE(nlevels == 0 || (lptr_host[0] == 0 && lptr_host[nlevels] == n), "ilu0_factor_f64: schedule must cover the rows");
  API_BEGIN
  launch_ilu0_factor<double>(ctx->stream, nlevels, lptr_host, lrows, ro, cols, dcol, lu);
  LAUNCHED("ilu0_factor_f64");
  API_END
}

int alien_hip_ilu0_solve_f64(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, const int* dcol,
                             const double* lu, int nl, const int* lptr, const int* lrows, int nu, const int* uptr,
                             const int* urows, const double* y, double* t, double* x) {
  CTX_OK(ctx);
  REQUIRE(n >= 0 && nl >= 0 && nu >= 0 &&
              (n == 0 || (ro && cols && dcol && lu && lptr && lrows && uptr && urows && y && t && x)),
          "ilu0_solve_f64: bad arguments");
  REQUIRE(nl == 0 || (lptr[0] == 0 && lptr[nl] == n), "ilu0_solve_f64: lower schedule must cover the rows");
  REQUIRE(nu == 0 || (uptr[0] == 0 && uptr[nu] == n), "ilu0_solve_f64: upper schedule must cover the rows");
  API_BEGIN
  launch_ilu0_solve<double>(ctx->stream, nl, lptr, lrows, nu, uptr, urows, ro, cols, dcol, lu, y, t, x);
  LAUNCHED("ilu0_solve_f64");
  API_END
}

} // extern "C"
