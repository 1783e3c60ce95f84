// Per-launch HIP-event timing of hot kernels, grouped by kernel class.
#pragma once
#include <vector>

#include "common.hpp"
#include "device.hpp"

namespace sptc {

struct KStat {
  uint64_t launches = 0;
  double ms = 0;
  double bytes = 0;  // algorithmic HBM bytes
};

class KTimer {
 public:
  ~KTimer() {
    for (auto e : ev_) (void)hipEventDestroy(e);
  }
  bool on = false;
  // Brackets one launch on stream s; returns a token for end().
  int begin(hipStream_t s) {
    if (!on) return -1;
    const int slot = (int)used_;
    if (ev_.size() < 2 * (used_ + 1)) {
      hipEvent_t a, b;
      SPTC_HIP(hipEventCreate(&a));
      SPTC_HIP(hipEventCreate(&b));
      ev_.push_back(a);
      ev_.push_back(b);
    }
    SPTC_HIP(hipEventRecord(ev_[2 * slot], s));
    ++used_;
    cls_.push_back(-1);
    bytes_.push_back(0);
    return slot;
  }
  void end(int tok, int cls, double bytes, hipStream_t s) {
    if (tok < 0) return;
    SPTC_HIP(hipEventRecord(ev_[2 * tok + 1], s));
    cls_[tok] = cls;
    bytes_[tok] = bytes;
  }
  // A zeroed device counter for the launch of token tok (NULL when tok < 0):
  // the kernel adds units of work it decides on the device (the number of
  // products a sort actually moves, say) and collect() adds per_unit bytes
  // for each to that launch's bytes.  Only for timed, uncaptured launches.
  unsigned long long* counter(int tok, double per_unit, hipStream_t s) {
    if (tok < 0) return nullptr;
    const size_t slot = ctok_.size();
    if (slot / kCountBlock >= cnt_.size()) cnt_.emplace_back(kCountBlock);
    unsigned long long* c = cnt_[slot / kCountBlock].get() + slot % kCountBlock;
    ctok_.push_back(tok);
    cper_.push_back(per_unit);
    SPTC_HIP(hipMemsetAsync(c, 0, sizeof(*c), s));
    return c;
  }
  // Call after the stream has been synchronized.
  void collect(KStat* out, int nclasses) {
    std::vector<unsigned long long> units(kCountBlock);
    for (size_t b = 0; b * kCountBlock < ctok_.size(); ++b) {
      SPTC_HIP(hipMemcpy(units.data(), cnt_[b].get(), kCountBlock * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost));
      for (size_t j = b * kCountBlock; j < ctok_.size() && j < (b + 1) * kCountBlock; ++j)
        bytes_[ctok_[j]] += cper_[j] * (double)units[j - b * kCountBlock];
    }
    for (size_t i = 0; i < used_; ++i) {
      if (cls_[i] < 0 || cls_[i] >= nclasses) continue;
      float ms = 0;
      SPTC_HIP(hipEventElapsedTime(&ms, ev_[2 * i], ev_[2 * i + 1]));
      out[cls_[i]].launches += 1;
      out[cls_[i]].ms += ms;
      out[cls_[i]].bytes += bytes_[i];
    }
    used_ = 0;
    cls_.clear();
    bytes_.clear();
    ctok_.clear();
    cper_.clear();
  }

 private:
  std::vector<hipEvent_t> ev_;
  std::vector<int> cls_;
  std::vector<double> bytes_;
  static constexpr size_t kCountBlock = 1024;
  std::vector<DevBuf<unsigned long long>> cnt_;  // device counters, kCountBlock each
  std::vector<int> ctok_;                        // counter slot -> token
  std::vector<double> cper_;                     // counter slot -> bytes per unit
  size_t used_ = 0;
};

}  // namespace sptc
