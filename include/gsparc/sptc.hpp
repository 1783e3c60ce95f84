// gsparc/sptc.hpp -- C++ entry point with the reference's shape, over the C
// ABI of libsptc_hip.so (include/sptc.h).
//
// The reference's call is
//   gsparc::SpTC<TensorType,SLITOMType>(SX, SY, SZ, pools, gpu_count, copy_timers,
//       idxmatch_timers, dprtn_timers, contraction_timers, dense_accumulator)
// (include/gsparc/sptc.cuh:16-26) after TensorManager::ConvertTensor /
// SortSlitomXY / Partition (main.cu:293-321).  Here one call takes the two
// COO tensors (SparseTensor<uint64_t,float>, sparse_tensor.hpp:12-44) and the
// contracted-mode lists, and returns the output (nnz plus a device-resident
// result that can be decoded to COO); the per-phase timers of main.cu come
// back as seconds.  Errors throw gsparc::Error carrying the sptc_status
// instead of exiting the process.
#ifndef GSPARC_SPTC_HPP_
#define GSPARC_SPTC_HPP_
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../sptc.h"

namespace gsparc {

template <typename IndexType = uint64_t, typename ValueType = float>
struct SparseTensor {
  unsigned short order = 0;
  std::vector<IndexType> dims;                  // per mode (observed max index + 1)
  uint64_t nnz = 0;
  std::vector<std::vector<IndexType>> indices;  // [order][nnz], 0-based
  std::vector<ValueType> values;                // [nnz]
};

struct Error : std::runtime_error {
  sptc_status status;
  Error(sptc_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

// Phase times in seconds, named like main.cu's timers (main.cu:415-432).
struct Timers {
  double slitom = 0;             // linearize + sort of both operands
  double partition = 0;          // static partitioning (none here: 0)
  double idxmatch = 0;
  double contraction = 0;
  double dynamic_partition = 0;  // chunk planning
  double total = 0;
};

class Result {
 public:
  explicit Result(sptc_result* r) : r_(r) { sptc_result_get_info(r_, &info_); }
  Result(Result&& o) noexcept : r_(o.r_), info_(o.info_) { o.r_ = nullptr; }
  Result(const Result&) = delete;
  Result& operator=(const Result&) = delete;
  ~Result() { if (r_) sptc_result_free(r_); }
  uint64_t nnz() const { return info_.nnz; }
  uint64_t ir_nnz() const { return info_.ir_nnz; }
  const sptc_result_info& info() const { return info_; }
  const sptc_result* handle() const { return r_; }
  // Output COO: per-mode 0-based indices (X free modes, then Y free modes)
  // and values.
  SparseTensor<uint32_t, float> decode() const {
    SparseTensor<uint32_t, float> z;
    z.order = (unsigned short)info_.out_order;
    z.nnz = info_.nnz;
    z.indices.assign(z.order, std::vector<uint32_t>(z.nnz));
    z.values.resize(z.nnz);
    std::vector<uint32_t*> p(z.order);
    for (int m = 0; m < z.order; ++m) p[m] = z.indices[m].data();
    check(sptc_result_decode(r_, p.data(), z.values.data()));
    return z;
  }
  static void check(sptc_status s) {
    if (s != SPTC_OK) throw Error(s, sptc_last_error());
  }

 private:
  sptc_result* r_;
  sptc_result_info info_{};
};

// Z = X x_{cmodes_x, cmodes_y} Y.  dense_accumulator follows the reference's
// -d flag (sptc.cuh:201-218): 1 dense, 0 planner's choice, -1 ESC.
inline Result SpTC(const SparseTensor<uint64_t, float>& X, const SparseTensor<uint64_t, float>& Y,
                   int num_cmodes, const int* cmodes_x, const int* cmodes_y,
                   int dense_accumulator = 0, Timers* timers = nullptr,
                   uint64_t mem_budget_bytes = 0) {
  auto as_coo = [](const SparseTensor<uint64_t, float>& t, std::vector<const void*>& ptrs) {
    sptc_coo c{};
    c.order = t.order;
    c.nnz = t.nnz;
    c.dims = t.dims.data();
    ptrs.resize(t.order);
    for (int m = 0; m < t.order; ++m) ptrs[m] = t.indices[m].data();
    c.idx = ptrs.data();
    c.idx_bytes = 8;
    c.val = t.values.data();
    c.on_device = 0;
    return c;
  };
  std::vector<const void*> px, py;
  const sptc_coo cx = as_coo(X, px), cy = as_coo(Y, py);
  std::vector<int32_t> mx(cmodes_x, cmodes_x + num_cmodes), my(cmodes_y, cmodes_y + num_cmodes);
  sptc_opts o{};
  o.dense_accumulator = dense_accumulator;
  o.mem_budget_bytes = mem_budget_bytes;
  sptc_result* r = nullptr;
  Result::check(sptc_contract(&cx, &cy, num_cmodes, mx.data(), my.data(), &o, nullptr, &r));
  Result res(r);
  if (timers) {
    const double* ms = res.info().phase_ms;
    timers->slitom = (ms[0] + ms[1]) * 1e-3;
    timers->partition = 0;
    timers->idxmatch = ms[2] * 1e-3;
    timers->dynamic_partition = ms[3] * 1e-3;
    timers->contraction = ms[4] * 1e-3;
    timers->total = (ms[0] + ms[5]) * 1e-3;
  }
  return res;
}

// ---- the reference's call sequence (main.cu:293-376) ----------------------
//
//   gsparc::TensorManager tm(1);
//   tm.FindPartitionNum(X.nnz, Y.nnz, pool_bytes, nbits);
//   gsparc::SLITOM SX(1), SY(2), SZ;
//   tm.ConvertTensor(&X, &SX, nc, cmodes_x);
//   tm.ConvertTensor(&Y, &SY, nc, cmodes_y);
//   tm.SortSlitomXY(&SX, &SY);
//   tm.Partition(&SX); tm.Partition(&SY);
//   gsparc::SpTC(&SX, &SY, &SZ, dense_accumulator);
//   result_nnz += SZ.nnz;
//
// SLITOM (slitom.hpp:27-92) is here an operand handle: after ConvertTensor it
// names a COO tensor and its contracted modes; SortSlitomXY builds the pair's
// plan -- both operands linearized and sorted in device memory, shared by SX
// and SY (sptc_plan); SpTC executes it (sptc_execute) and fills SZ.  There is
// no static partitioning: the accumulation chunks are planned per execution
// from the product counts, so FindPartitionNum and Partition only keep the
// call sequence.
struct SLITOM {
  int input_order = 0;  // 1 = X, 2 = Y, 0 = output
  const SparseTensor<uint64_t, float>* tensor = nullptr;
  std::vector<int32_t> cmodes;
  uint64_t nnz = 0;
  int nprtn = 1;
  std::shared_ptr<sptc_plan_t> plan;  // set on SX and SY by SortSlitomXY
  std::shared_ptr<Result> result;     // set on SZ by SpTC
  Timers timers;                      // SZ: phase times of the SpTC call (s)
  double slitom_seconds = 0;          // SX, SY: linearize + sort of the pair
  SLITOM() = default;
  explicit SLITOM(int order) : input_order(order) {}
};

class TensorManager {
 public:
  explicit TensorManager(int gpu_count = 1) {
    if (gpu_count != 1) throw Error(SPTC_ERR_BAD_ARGS, "this engine runs on one GPU");
  }
  // tensor_manager.tpp:32-61 decides host vs GPU sorting and the static
  // partition count; operands are always sorted on the GPU here.
  bool FindPartitionNum(uint64_t, uint64_t, uint64_t, int) const { return true; }
  // tensor_manager.tpp:212-298: records the tensor and its contracted modes
  // (linearization happens in SortSlitomXY, on the device).
  void ConvertTensor(const SparseTensor<uint64_t, float>* t, SLITOM* s, int num_cmodes,
                     const int* cmodes) const {
    s->tensor = t;
    s->cmodes.assign(cmodes, cmodes + num_cmodes);
    s->nnz = t->nnz;
  }
  // tensor_manager.tpp:487-522: lays out, linearizes and sorts both operands.
  void SortSlitomXY(SLITOM* SX, SLITOM* SY, uint64_t mem_budget_bytes = 0) const {
    if (!SX->tensor || !SY->tensor || SX->cmodes.size() != SY->cmodes.size())
      throw Error(SPTC_ERR_BAD_ARGS, "SortSlitomXY: ConvertTensor both operands first");
    std::vector<const void*> px, py;
    const sptc_coo cx = as_coo(*SX->tensor, px), cy = as_coo(*SY->tensor, py);
    sptc_opts o{};
    o.mem_budget_bytes = mem_budget_bytes;
    sptc_plan_t* p = nullptr;
    Result::check(sptc_plan(&cx, &cy, (int32_t)SX->cmodes.size(), SX->cmodes.data(),
                            SY->cmodes.data(), &o, nullptr, &p));
    std::shared_ptr<sptc_plan_t> plan(p, sptc_plan_destroy);
    sptc_plan_info pi{};
    Result::check(sptc_plan_get_info(p, &pi));
    SX->plan = SY->plan = plan;
    SX->slitom_seconds = SY->slitom_seconds = pi.prepare_ms * 1e-3;
  }
  // tensor_manager.tpp:633-749: nothing to do (chunks are planned per call).
  void Partition(SLITOM*) const {}

  static sptc_coo as_coo(const SparseTensor<uint64_t, float>& t, std::vector<const void*>& ptrs) {
    sptc_coo c{};
    c.order = t.order;
    c.nnz = t.nnz;
    c.dims = t.dims.data();
    ptrs.resize(t.order);
    for (int m = 0; m < t.order; ++m) ptrs[m] = t.indices[m].data();
    c.idx = ptrs.data();
    c.idx_bytes = 8;
    c.val = t.values.data();
    c.on_device = 0;
    return c;
  }
};

// sptc.cuh:16-26 on a SortSlitomXY'd pair: SZ->nnz and SZ->result (the
// reference's pools, GPU count and per-GPU timers are replaced by the
// engine's budget and the phase times in SZ->timers).
inline void SpTC(SLITOM* SX, SLITOM* SY, SLITOM* SZ, int dense_accumulator = 0) {
  if (!SX->plan || SX->plan != SY->plan)
    throw Error(SPTC_ERR_BAD_ARGS, "SpTC: SortSlitomXY(SX, SY) first");
  sptc_opts o{};
  o.dense_accumulator = dense_accumulator;
  sptc_result* r = nullptr;
  Result::check(sptc_execute(SX->plan.get(), &o, nullptr, &r));
  SZ->result = std::make_shared<Result>(r);
  SZ->nnz = SZ->result->nnz();
  const double* ms = SZ->result->info().phase_ms;
  SZ->timers = Timers{};
  SZ->timers.slitom = SX->slitom_seconds;
  SZ->timers.idxmatch = ms[2] * 1e-3;
  SZ->timers.dynamic_partition = ms[3] * 1e-3;
  SZ->timers.contraction = ms[4] * 1e-3;
  SZ->timers.total = SX->slitom_seconds + ms[5] * 1e-3;
}

}  // namespace gsparc
#endif
