// Device context implementation (host code, built with hipcc).
#include "device.hpp"

#include <cstring>

namespace sptc {

namespace {
constexpr size_t kGranule = 2u << 20;  // round large blocks to 2 MiB
size_t round_size(size_t b) {
  if (b <= (1u << 20)) {
    size_t r = 256;
    while (r < b) r <<= 1;
    return r;
  }
  return (b + kGranule - 1) / kGranule * kGranule;
}
}  // namespace

DevicePool& DevicePool::get() {
  static DevicePool* pool = new DevicePool();  // lives until process exit
  return *pool;
}

void* DevicePool::alloc(size_t bytes) {
  const size_t want = round_size(bytes ? bytes : 1);
  std::lock_guard<std::mutex> g(mu_);
  auto it = cache_.lower_bound(want);
  if (it != cache_.end() && it->first <= want + want / 4 + kGranule) {
    void* p = it->second;
    size_t sz = it->first;
    cache_.erase(it);
    cached_ -= sz;
    live_ += sz;
    sizes_[p] = sz;
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    for (auto& kv : cache_) (void)hipFree(kv.second);
    cache_.clear();
    cached_ = 0;
    e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw HipError{e, "hipMalloc", __LINE__};
    }
  }
  sizes_[p] = want;
  live_ += want;
  return p;
}

void DevicePool::free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu_);
  auto it = sizes_.find(p);
  if (it == sizes_.end()) return;
  size_t sz = it->second;
  sizes_.erase(it);
  live_ -= sz;
  cache_.emplace(sz, p);
  cached_ += sz;
}

void DevicePool::trim() {
  std::lock_guard<std::mutex> g(mu_);
  for (auto& kv : cache_) (void)hipFree(kv.second);
  cache_.clear();
  cached_ = 0;
}

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

}  // namespace sptc
