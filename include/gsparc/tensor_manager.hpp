// gsparc/tensor_manager.hpp -- the reference's operand preparation
// (include/gsparc/tensor_manager.hpp, tensor_manager.tpp) over the engine's
// C ABI.  The call sequence of main.cu:293-321 works unchanged:
//
//   FindPartitionNum  -> sptc_find_partition_num (part counts for the sizes)
//   ConvertTensor     -> layout fields of the SLITOM (sptc_layout); the
//                        linearization itself runs on the device later
//   SortSlitomXY      -> sptc_plan: both operands uploaded, linearized and
//                        sorted on the device (or split into static parts
//                        when they do not fit half the pool)
//   Partition         -> the plan's static parts into nprtn / prtn_idx /
//                        prtn_coord (sptc_plan_get_partitions)
//   convert_and_sort_64 -> nothing to do: each Y part is re-keyed to the Y
//                        layout and sorted on the device when SpTC reaches
//                        it (the engine's partition-pair loop)
//
// The _extra / _128 forms (keys above 64 bits, tensor_manager.tpp:300-406,
// 524-631) are the same calls: the engine lays out wide keys itself.
#ifndef GSPARC_TENSOR_MANAGER_HPP_
#define GSPARC_TENSOR_MANAGER_HPP_
#include <algorithm>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../sptc.h"
#include "cuda_memory.hpp"
#include "slitom.hpp"
#include "sparse_tensor.hpp"
#include "timer.hpp"

namespace gsparc {

struct Error : std::runtime_error {
  sptc_status status;
  Error(sptc_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check(sptc_status s) {
  if (s != SPTC_OK) throw Error(s, sptc_last_error());
}

namespace detail {

// The output of a contraction as an operand: trailing `nc` modes
// contracted (the Y free modes of the previous step, main.cu:345-363).
template <typename S>
void output_as_operand(S* s, int nc, void* stream) {
  if (!s->result) throw Error(SPTC_ERR_BAD_ARGS, "SortSlitomXY: the operand holds no tensor");
  check(sptc_result_as_coo(s->result.get(), stream, &s->coo));
  s->cmodes.clear();
  for (int i = 0; i < nc; ++i) s->cmodes.push_back(s->coo.order - nc + i);
}

}  // namespace detail

template <typename TensorType, typename SLITOMType>
class TensorManager {
 public:
  using sptensor_t = TensorType;
  using slitom_t = SLITOMType;
  using index_t = typename sptensor_t::index_t;
  using ulindex_t = typename slitom_t::ulindex_t;
  using lindex_t = typename slitom_t::lindex_t;
  using value_t = typename sptensor_t::value_t;
  using memorypool_t = CudaMemoryPool;
  using mask_t = typename slitom_t::mask_t;

  explicit TensorManager(int gpu_count) : gpu_count(gpu_count) {
    if (gpu_count != 1) throw Error(SPTC_ERR_BAD_ARGS, "this engine runs on one GPU (-g 1)");
  }

  // tensor_manager.tpp:32-61: part counts for the sizes under the pool
  // (order 4 assumed for the staged indices; SortSlitomXY decides with the
  // real orders).  Returns whether the operands stay resident whole.
  bool FindPartitionNum(uint64_t nnz_count_x, uint64_t nnz_count_y, size_t poolSize,
                        int /*nbits*/) {
    int32_t px = 1, py = 1;
    check(sptc_find_partition_num(nnz_count_x, 4, nnz_count_y, 4, (int32_t)sizeof(index_t),
                                  poolSize, &px, &py));
    prtn_num = std::max(px, py);
    return prtn_num == 1;
  }

  // tensor_manager.tpp:212-298: the operand's layout (its own dims, as in
  // the reference) and a view of its arrays.
  void ConvertTensor(sptensor_t* sptensor, slitom_t* slitom, int cnmodes, const int* cpos,
                     Timer* timer) {
    if (timer) timer->start();
    const int order = sptensor->order;
    slitom->coo_dims.assign(order, 0);
    for (int m = 0; m < order; ++m) slitom->coo_dims[m] = (uint64_t)sptensor->dims[m];
    slitom->coo_ptr.assign(order, nullptr);
    for (int m = 0; m < order; ++m) slitom->coo_ptr[m] = sptensor->indices[m];
    sptc_coo& c = slitom->coo;
    c.order = order;
    c.nnz = sptensor->nnz;
    c.dims = slitom->coo_dims.data();
    c.idx = slitom->coo_ptr.data();
    c.idx_bytes = (int32_t)sizeof(index_t);
    c.val = reinterpret_cast<const float*>(sptensor->values);
    c.on_device = 0;
    slitom->cmodes.assign(cpos, cpos + cnmodes);
    fill_layout(slitom, order, slitom->coo_dims.data(), cnmodes, cpos,
                slitom->input_order == 1 ? 1 : 2);
    slitom->nnz = sptensor->nnz;
    if (timer) timer->stop();
  }
  void ConvertTensor_extra(sptensor_t* sptensor, slitom_t* slitom, int cnmodes, const int* cpos,
                           Timer* timer) {
    ConvertTensor(sptensor, slitom, cnmodes, cpos, timer);
  }

  // tensor_manager.tpp:487-522 (+ the GPU sorts of sort.cuh:39-159): builds
  // the pair's plan under pools[0]'s size.  An SX holding the output of a
  // previous contraction is used from device memory, its trailing modes
  // contracted with SY's.
  void SortSlitomXY(slitom_t* SX, slitom_t* SY, memorypool_t** pools, Timer* timer) {
    static_assert(sizeof(value_t) == 4, "values are FP32");
    if (timer) timer->start();
    if (!SX->coo.nnz && SX->result) detail::output_as_operand(SX, (int)SY->cmodes.size(), nullptr);
    if (SX->cmodes.size() != SY->cmodes.size())
      throw Error(SPTC_ERR_BAD_ARGS, "SortSlitomXY: contracted mode counts differ");
    sptc_opts o{};
    o.mem_budget_bytes = pools && pools[0] ? pools[0]->getPoolSize() : 0;
    sptc_plan_t* p = nullptr;
    check(sptc_plan(&SX->coo, &SY->coo, (int32_t)SX->cmodes.size(), SX->cmodes.data(),
                    SY->cmodes.data(), &o, nullptr, &p));
    std::shared_ptr<sptc_plan_t> plan(p, sptc_plan_destroy);
    sptc_plan_info pi{};
    check(sptc_plan_get_info(p, &pi));
    SX->plan = SY->plan = plan;
    SX->plan_role = 1;
    SY->plan_role = 2;
    SX->peer = SY;
    SY->peer = SX;
    SX->sort_gpu = pi.x_parts == 1;
    SY->sort_gpu = pi.y_parts == 1;
    SX->nfbits = pi.x_nfbits;
    SY->nfbits = pi.y_nfbits;
    SX->ncbits = SY->ncbits = pi.ncbits;
    SX->nbits = SX->nfbits + SX->ncbits;
    SY->nbits = SY->nfbits + SY->ncbits;
    if (timer) timer->stop();
  }

  // tensor_manager.tpp:633-749: the plan's static parts of this operand.
  void Partition(slitom_t* slitom) {
    if (!slitom->plan) throw Error(SPTC_ERR_BAD_ARGS, "Partition: SortSlitomXY first");
    int32_t n = 0;
    check(sptc_plan_get_partitions(slitom->plan.get(), slitom->plan_role, 0, nullptr, nullptr,
                                   &n));
    std::vector<uint64_t> nb(n + 1), kb(n + 1);
    check(sptc_plan_get_partitions(slitom->plan.get(), slitom->plan_role, n, nb.data(),
                                   kb.data(), &n));
    slitom->nprtn = n;
    deallocate(slitom->prtn_idx);
    deallocate(slitom->prtn_coord);
    slitom->prtn_idx = allocate<uint64_t>(n + 1);
    slitom->prtn_coord = allocate<lindex_t>(n + 1);
    slitom->max_prtn_size = 0;
    for (int k = 0; k <= n; ++k) {
      slitom->prtn_idx[k] = nb[k];
      slitom->prtn_coord[k] = (lindex_t)kb[k];
      if (k) slitom->max_prtn_size = std::max<uint64_t>(slitom->max_prtn_size, nb[k] - nb[k - 1]);
    }
  }
  void Partition_128(slitom_t* slitom) { Partition(slitom); }

  // tensor_manager.tpp:751-798: the engine re-keys and sorts each Y part on
  // the device in the partition-pair loop; nothing is left to do here.
  void convert_and_sort_64(slitom_t* slitom, CudaMemoryPool** /*pool*/) {
    if (!slitom->plan) throw Error(SPTC_ERR_BAD_ARGS, "convert_and_sort_64: SortSlitomXY first");
  }
  void convert_and_sort_128(slitom_t* slitom, CudaMemoryPool** pools) {
    convert_and_sort_64(slitom, pools);
  }

  int prtn_num = 1;
  int gpu_count = 1;
  int max_block_size = 0;

 private:
  static void fill_layout(slitom_t* s, int order, const uint64_t* dims, int nc, const int* cpos,
                          int role) {
    std::vector<int32_t> cm(cpos, cpos + nc), sh(order), wd(order);
    int32_t nf = 0, ncb = 0;
    const sptc_status st =
        sptc_layout(order, dims, nc, cm.data(), nullptr, role, sh.data(), wd.data(), &nf, &ncb);
    if (st != SPTC_OK && st != SPTC_ERR_KEY_OVERFLOW) check(st);
    s->nmodes = (unsigned short)order;
    s->cnmodes = (unsigned short)nc;
    s->fnmodes = (unsigned short)(order - nc);
    s->nfbits = nf;
    s->ncbits = ncb;
    s->nbits = nf + ncb;
    deallocate(s->dims);
    deallocate(s->cpos);
    deallocate(s->fpos);
    s->dims = allocate<lindex_t>(order);
    for (int m = 0; m < order; ++m) s->dims[m] = (lindex_t)dims[m];
    s->cpos = allocate<int>(nc);
    for (int i = 0; i < nc; ++i) s->cpos[i] = cpos[i];
    s->fpos = allocate<int>(order - nc);
    int k = 0;
    for (int m = 0; m < order; ++m)
      if (std::find(cpos, cpos + nc, m) == cpos + nc) s->fpos[k++] = m;
  }
};

}  // namespace gsparc
#endif
