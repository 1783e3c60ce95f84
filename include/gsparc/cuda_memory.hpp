// gsparc/cuda_memory.hpp -- the reference's per-GPU memory pool
// (include/gsparc/cuda_memory.hpp, src/gsparc/cuda_memory.cpp:17-35) kept
// under its name so callers compile unchanged.  Here it is the device
// memory budget of the engine: the pool size is the free device memory less
// 256 MiB, as in the reference, and TensorManager / SpTC hand it to the
// engine (sptc_opts.mem_budget_bytes), whose own arena allocates within it.
// is_available answers the reference's fit question against that budget.
#ifndef GSPARC_CUDA_MEMORY_HPP_
#define GSPARC_CUDA_MEMORY_HPP_
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../sptc.h"

namespace gsparc {

class CudaMemoryPool {
 public:
  explicit CudaMemoryPool(unsigned id) : id_(id) {
    if (id != 0) throw std::invalid_argument("CudaMemoryPool: this engine runs on GPU 0");
    int32_t cus = 0;
    uint64_t free_b = 0, total_b = 0;
    double peak = 0;
    if (sptc_device_info(&cus, &free_b, &total_b, &peak) != SPTC_OK)
      throw std::runtime_error(std::string("CudaMemoryPool: ") + sptc_last_error());
    const uint64_t margin = 256ull << 20;
    pool_size_ = free_b > margin ? free_b - margin : 0;
  }
  // Bytes the engine may hold for one contraction.
  size_t getPoolSize() const { return pool_size_; }
  size_t getRemainingMemory() const {
    uint64_t live = 0;
    sptc_pool_stats(&live, nullptr, nullptr);
    return live < pool_size_ ? pool_size_ - live : 0;
  }
  bool is_available(size_t size) const { return size <= getRemainingMemory(); }
  // The reference frees every block between iterations (main.cu:88-92); the
  // engine's results and plans own their memory, so this returns cached
  // blocks to the runtime.
  void reset() { sptc_trim_cache(); }
  void printFree() const {
    std::printf("GPU %u: %zu of %zu bytes free\n", id_, getRemainingMemory(), getPoolSize());
  }

 private:
  unsigned id_;
  size_t pool_size_ = 0;
};

}  // namespace gsparc
#endif
