// Device context: caching allocator with budget accounting, events, props.
#pragma once
#include <algorithm>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace sptc {

// Device memory arena: large hipMalloc slabs, best-fit sub-allocation with
// coalescing on release, so repeated calls do not return to the runtime
// allocator.  `live_bytes` counts what callers hold and is what memory
// budgets are checked against -- the analogue of the reference's
// CudaMemoryPool::is_available (cuda_memory.cpp:37-51), which returns a
// status here instead of exit(1) on exhaustion.
class DevicePool {
 public:
  // The calling thread's pool: the process-wide one, unless a Scope made
  // another current (each serving worker allocates from its own pool, so a
  // block freed while its stream still runs is only reused on that stream).
  static DevicePool& get();
  static DevicePool& global();
  class Scope {
   public:
    explicit Scope(DevicePool* p);
    ~Scope();
   private:
    DevicePool* prev_;
  };
  DevicePool() = default;
  ~DevicePool();
  DevicePool(const DevicePool&) = delete;
  DevicePool& operator=(const DevicePool&) = delete;
  void* alloc(size_t bytes);  // throws HipError on failure
  void free(void* p);
  void trim();                // return fully free slabs to the runtime
  size_t live_bytes() const { return live_; }
  size_t cached_bytes() const;
  size_t reserved_bytes() const { return reserved_; }
  uint64_t hip_mallocs() const { return hip_mallocs_; }

 private:
  bool take_locked(size_t want, void** out);
  void insert_free_locked(char* p, size_t sz);
  void erase_size_locked(size_t sz, char* p);
  void release_empty_slabs_locked();
  int slab_of(const char* p) const;
  std::mutex mu_;
  std::vector<std::pair<char*, size_t>> slabs_;
  std::map<char*, size_t> by_addr_;        // free ranges by address
  std::multimap<size_t, char*> by_size_;   // free ranges by size
  std::unordered_map<void*, size_t> sizes_;
  size_t live_ = 0, reserved_ = 0;
  uint64_t hip_mallocs_ = 0;
};

// RAII device buffer from the pool.
template <typename T>
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(uint64_t n) { reset(n); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_), pool_(o.pool_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      pool_ = o.pool_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void reset(uint64_t n) {
    release();
    n_ = n;
    if (n) {
      pool_ = &DevicePool::get();
      p_ = static_cast<T*>(pool_->alloc(n * sizeof(T)));
    }
  }
  void release() {
    if (p_) pool_->free(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  uint64_t size() const { return n_; }
  uint64_t bytes() const { return n_ * sizeof(T); }

 private:
  T* p_ = nullptr;
  uint64_t n_ = 0;
  DevicePool* pool_ = nullptr;
};

struct DeviceInfo {
  int device = 0;
  int num_cus = 0;
  size_t total_bytes = 0;
  double peak_hbm_gbs = 0;
  char arch[64] = {0};
};
const DeviceInfo& device_info();
size_t device_free_bytes();

}  // namespace sptc
