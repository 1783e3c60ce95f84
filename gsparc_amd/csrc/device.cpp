// Device context implementation (host code, built with hipcc).
#include "device.hpp"

#include <cstring>

namespace sptc {

namespace {
constexpr size_t kAlign = 256;
constexpr size_t kSlabMin = 8ull << 30;  // new slabs are at least 8 GiB
size_t align_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
// Size classes for large blocks (>= 64 MiB): round up to 1/8 of the leading
// power of two, so the slightly different sizes that successive calls ask
// for land in the same class and reuse each other's ranges.
size_t size_class(size_t b) {
  b = align_up(b ? b : 1);
  if (b < (64ull << 20)) return b;
  int lg = 63 - __builtin_clzll(b);
  size_t step = 1ull << (lg - 3);
  return (b + step - 1) / step * step;
}
}  // namespace

namespace {
thread_local DevicePool* tl_pool = nullptr;
}

DevicePool& DevicePool::global() {
  static DevicePool* pool = new DevicePool();  // lives until process exit
  return *pool;
}

DevicePool& DevicePool::get() { return tl_pool ? *tl_pool : global(); }

DevicePool::Scope::Scope(DevicePool* p) : prev_(tl_pool) { tl_pool = p; }
DevicePool::Scope::~Scope() { tl_pool = prev_; }

// Best fit over every slab's free ranges; ranges split on allocation and
// coalesce with their neighbours on release.
bool DevicePool::take_locked(size_t want, void** out) {
  auto it = by_size_.lower_bound(want);
  if (it == by_size_.end()) return false;
  const size_t sz = it->first;
  char* base = it->second;
  by_size_.erase(it);
  by_addr_.erase(base);
  if (sz > want) {
    by_addr_[base + want] = sz - want;
    by_size_.emplace(sz - want, base + want);
  }
  sizes_[base] = want;
  live_ += want;
  *out = base;
  return true;
}

void DevicePool::insert_free_locked(char* p, size_t sz) {
  // merge with the following free range
  auto next = by_addr_.find(p + sz);
  if (next != by_addr_.end() && slab_of(p) == slab_of(next->first)) {
    erase_size_locked(next->second, next->first);
    sz += next->second;
    by_addr_.erase(next);
  }
  // merge with the preceding free range
  auto prev = by_addr_.lower_bound(p);
  if (prev != by_addr_.begin()) {
    --prev;
    if (prev->first + prev->second == p && slab_of(prev->first) == slab_of(p)) {
      erase_size_locked(prev->second, prev->first);
      p = prev->first;
      sz += prev->second;
      by_addr_.erase(prev);
    }
  }
  by_addr_[p] = sz;
  by_size_.emplace(sz, p);
}

void DevicePool::erase_size_locked(size_t sz, char* p) {
  auto r = by_size_.equal_range(sz);
  for (auto i = r.first; i != r.second; ++i)
    if (i->second == p) {
      by_size_.erase(i);
      return;
    }
}

int DevicePool::slab_of(const char* p) const {
  for (size_t i = 0; i < slabs_.size(); ++i)
    if (p >= slabs_[i].first && p < slabs_[i].first + slabs_[i].second) return (int)i;
  return -1;
}

void* DevicePool::alloc(size_t bytes) {
  const size_t want = size_class(bytes);
  std::lock_guard<std::mutex> g(mu_);
  void* p = nullptr;
  if (take_locked(want, &p)) return p;
  // New slab: at least kSlabMin, else exactly what is needed.
  size_t fr = 0, tot = 0;
  (void)hipMemGetInfo(&fr, &tot);
  size_t slab = std::max(want, kSlabMin);
  if (slab > fr) slab = want;
  char* base = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&base), slab);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    release_empty_slabs_locked();
    // the process-wide pool may cache the memory this pool needs (a worker
    // pool of the server after large single contractions): release its free
    // slabs too
    if (this != &global()) global().trim();
    e = hipMalloc(reinterpret_cast<void**>(&base), want);
    slab = want;
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw HipError{e, "hipMalloc", __LINE__};
    }
  }
  ++hip_mallocs_;
  slabs_.emplace_back(base, slab);
  reserved_ += slab;
  insert_free_locked(base, slab);
  if (!take_locked(want, &p)) throw HipError{hipErrorOutOfMemory, "pool", __LINE__};
  return p;
}

void DevicePool::free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu_);
  auto it = sizes_.find(p);
  if (it == sizes_.end()) return;
  const size_t sz = it->second;
  sizes_.erase(it);
  live_ -= sz;
  insert_free_locked(static_cast<char*>(p), sz);
}

void DevicePool::release_empty_slabs_locked() {
  for (size_t i = 0; i < slabs_.size();) {
    char* b = slabs_[i].first;
    const size_t s = slabs_[i].second;
    auto f = by_addr_.find(b);
    if (f != by_addr_.end() && f->second == s) {
      erase_size_locked(s, b);
      by_addr_.erase(f);
      (void)hipFree(b);
      reserved_ -= s;
      slabs_.erase(slabs_.begin() + i);
    } else {
      ++i;
    }
  }
}

void DevicePool::trim() {
  std::lock_guard<std::mutex> g(mu_);
  release_empty_slabs_locked();
}

size_t DevicePool::cached_bytes() const { return reserved_ - live_; }

const DeviceInfo& device_info() {
  static DeviceInfo info = [] {
    DeviceInfo d;
    SPTC_HIP(hipGetDevice(&d.device));
    hipDeviceProp_t prop;
    SPTC_HIP(hipGetDeviceProperties(&prop, d.device));
    d.num_cus = prop.multiProcessorCount;
    d.total_bytes = prop.totalGlobalMem;
    // MI355X's published HBM3E peak; the device's clock * bus-width product
    // under-reports HBM3E's data rate, so the published figure is used.
    d.peak_hbm_gbs = 8000.0;
    std::strncpy(d.arch, prop.gcnArchName, sizeof(d.arch) - 1);
    return d;
  }();
  return info;
}

size_t device_free_bytes() {
  size_t fr = 0, tot = 0;
  SPTC_HIP(hipMemGetInfo(&fr, &tot));
  return fr + DevicePool::get().cached_bytes();
}

DevicePool::~DevicePool() {
  for (auto& sl : slabs_) (void)hipFree(sl.first);
}

}  // namespace sptc
