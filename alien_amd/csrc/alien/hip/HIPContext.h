// HIP execution context of the backend: one stream, the workspace of the
// one-pass reductions, a ring of device-resident scalar slots and a pinned
// host mirror used to read scalars back. Plays the part of the SYCL queue
// holder of Alien's SYCL backend (sycl/algebra/SYCLKernelInternal.h:32-84).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../kernels/device_common.h"
#include "../../kernels/krylov_state.h"

namespace alien_hip {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define ALIEN_HIP_CHECK(expr)                                                                 \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess)                                                                     \
      throw ::alien_hip::Error((int)_e, std::string(#expr) + ": " + hipGetErrorString(_e));   \
  } while (0)

// Check of the most recent launch(es) on this thread.
inline void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error((int)e, std::string(what) + ": " + hipGetErrorString(e));
}

// Freed device blocks of 1 MiB or more, kept by device and size for the next allocation of that
// size: the solvers' work vectors are created and destroyed with each solve, and mapping and
// unmapping device memory each time costs milliseconds. A block is kept only after a device-wide
// synchronisation (what hipFree does), so no queued work can still use it when it is handed out
// again; at most 16 GiB are kept per process, and a failed hipMalloc empties the cache and tries
// once more. The cache is never destroyed (blocks are released with the process).
class BlockCache {
 public:
  static BlockCache& get() {
    static BlockCache* c = new BlockCache();
    return *c;
  }
  void* take(int dev, size_t bytes) {
    std::lock_guard<std::mutex> g(m_);
    auto it = free_.find({dev, bytes});
    if (it == free_.end() || it->second.empty()) return nullptr;
    void* p = it->second.back();
    it->second.pop_back();
    kept_ -= bytes;
    return p;
  }
  // true when the block was kept (else the caller frees it)
  bool give(int dev, void* p, size_t bytes) {
    if (bytes < kMin) return false;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != dev) return false;
    {
      std::lock_guard<std::mutex> g(m_);
      if (kept_ + bytes > kMax) return false;
    }
    if (hipDeviceSynchronize() != hipSuccess) return false;
    std::lock_guard<std::mutex> g(m_);
    if (kept_ + bytes > kMax) return false;
    free_[{dev, bytes}].push_back(p);
    kept_ += bytes;
    return true;
  }
  void release(int dev) {
    std::lock_guard<std::mutex> g(m_);
    for (auto& e : free_)
      if (e.first.first == dev) {
        for (void* p : e.second) {
          (void)hipFree(p);
          kept_ -= e.first.second;
        }
        e.second.clear();
      }
  }

 private:
  static constexpr size_t kMin = size_t(1) << 20, kMax = size_t(16) << 30;
  std::mutex m_;
  std::map<std::pair<int, size_t>, std::vector<void*>> free_;
  size_t kept_ = 0;
};

// Owning device array (its blocks of 1 MiB or more go through BlockCache).
template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n) { resize(n); }
  ~DeviceArray() { reset(); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : m_p(o.m_p), m_n(o.m_n), m_dev(o.m_dev) { o.m_p = nullptr; o.m_n = 0; }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) { reset(); m_p = o.m_p; m_n = o.m_n; m_dev = o.m_dev; o.m_p = nullptr; o.m_n = 0; }
    return *this;
  }
  void resize(size_t n) {
    reset();
    if (n) {
      const size_t bytes = n * sizeof(T);
      ALIEN_HIP_CHECK(hipGetDevice(&m_dev));
      m_p = static_cast<T*>(BlockCache::get().take(m_dev, bytes));
      if (!m_p) {
        hipError_t e = hipMalloc((void**)&m_p, bytes);
        if (e == hipErrorOutOfMemory) {
          (void)hipGetLastError();
          BlockCache::get().release(m_dev);
          e = hipMalloc((void**)&m_p, bytes);
        }
        if (e != hipSuccess) {
          m_p = nullptr;
          ALIEN_HIP_CHECK(e);
        }
      }
    }
    m_n = n;
  }
  void reset() {
    if (m_p && !BlockCache::get().give(m_dev, m_p, m_n * sizeof(T))) (void)hipFree(m_p);
    m_p = nullptr;
    m_n = 0;
  }
  // Synchronous upload: the host buffer may be reused or freed on return.
  void upload(const T* host, size_t n, hipStream_t st) {
    if (n != m_n) resize(n);
    if (n) {
      ALIEN_HIP_CHECK(hipMemcpyAsync(m_p, host, n * sizeof(T), hipMemcpyHostToDevice, st));
      ALIEN_HIP_CHECK(hipStreamSynchronize(st));
    }
  }
  T* data() { return m_p; }
  const T* data() const { return m_p; }
  size_t size() const { return m_n; }
  bool empty() const { return m_n == 0; }

 private:
  T* m_p = nullptr;
  size_t m_n = 0;
  int m_dev = 0;
};

} // namespace alien_hip

// The context object handed through the C ABI.
struct alien_hip_ctx {
  hipStream_t stream = nullptr;
  bool own_stream = false;
  bool trace = false; // ALIEN_HIP_TRACE=1 at creation: set-up phase times and mixed BiCGStab updates on stderr
  // Step-kernel timing (alien_hip_ctx_step_timing): when on, the fused CG solvers bracket each
  // launch of their q = A p kernel with HIP events on `stream` and add up the durations here
  bool step_timing = false;
  double step_us = 0.0;
  long long step_launches = 0;
  // reduction workspace (see RedWork); grows on demand, never while a graph is captured
  alien_hip::DeviceArray<double> partials;
  int red_cap = 0;
  alien_hip::DeviceArray<double> scalars; // device-resident scalar slots
  // zeroed CG state (done = 0, never written): the fused iteration's kernels launched on their own
  alien_hip::DeviceArray<alien_hip::CGState> idle_cg;
  double* host_scalars = nullptr;          // pinned mirror of `scalars`
  // Pinned 512-byte blocks for the fused solvers' host state mirrors, allocated once per context
  // (pinned allocation and release cost far more than a solve's own host work); a solver takes a
  // free block and gives it back, or allocates its own when all are taken.
  static constexpr int kPinnedBlocks = 16, kPinnedBytes = 512;
  char* pinned_blocks = nullptr;
  unsigned pinned_used = 0;
  void* take_pinned() {
    if (!pinned_blocks &&
        hipHostMalloc((void**)&pinned_blocks, (size_t)kPinnedBlocks * kPinnedBytes, hipHostMallocDefault) != hipSuccess) {
      pinned_blocks = nullptr;
      return nullptr;
    }
    for (int i = 0; i < kPinnedBlocks; ++i)
      if (!(pinned_used & (1u << i))) {
        pinned_used |= 1u << i;
        return pinned_blocks + (size_t)i * kPinnedBytes;
      }
    return nullptr;
  }
  // false when p is not one of the blocks (the caller frees it)
  bool give_pinned(void* p) {
    if (!pinned_blocks || (char*)p < pinned_blocks || (char*)p >= pinned_blocks + kPinnedBlocks * kPinnedBytes)
      return false;
    pinned_used &= ~(1u << (int)(((char*)p - pinned_blocks) / kPinnedBytes));
    return true;
  }
  int next_slot = 0;
  static constexpr int kSlots = 64;

  // Make room for reductions launched with up to `nblocks` workgroups.
  void reserve_red(long nblocks) {
    if (nblocks <= red_cap) return;
    long cap = red_cap ? red_cap : 4096;
    while (cap < nblocks) cap *= 2;
    ALIEN_HIP_CHECK(hipStreamSynchronize(stream));
    partials.resize((size_t)alien_hip::kMaxSlots * cap);
    red_cap = (int)cap;
  }
  // Room for every kernel of this library on vectors of n rows.
  void reserve_for(long n) {
    long b = alien_hip::bundle_blocks_max(n);
    long e = alien_hip::ept_blocks(n);
    long r = (n + alien_hip::kBlock - 1) / alien_hip::kBlock;
    reserve_red(std::max(b, std::max(e, r)));
  }
  alien_hip::RedWork red() {
    return alien_hip::RedWork{partials.data(), red_cap};
  }
  // A fresh device scalar slot (ring of kSlots; callers keep at most a few live).
  int take_slot() {
    int s = next_slot;
    next_slot = (next_slot + 1) % kSlots;
    return s;
  }
  double* dslot(int s) { return scalars.data() + s; }
  // Copy one slot to the host (synchronises the stream).
  double read_slot(int s) {
    ALIEN_HIP_CHECK(hipMemcpyAsync(host_scalars + s, scalars.data() + s, sizeof(double),
                                   hipMemcpyDeviceToHost, stream));
    ALIEN_HIP_CHECK(hipStreamSynchronize(stream));
    return host_scalars[s];
  }
};

namespace alien_hip {
alien_hip_ctx* default_ctx();
alien_hip_ctx* make_ctx(void* stream);
void destroy_ctx(alien_hip_ctx* c);
} // namespace alien_hip
