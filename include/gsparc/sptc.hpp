// gsparc/sptc.hpp -- the reference's SpTC entry points over the C ABI of
// libsptc_hip.so (include/sptc.h), with the reference's signatures:
//
//   template <class TensorType, class SLITOMType>
//   void gsparc::SpTC(SX, SY, SZ, CudaMemoryPool** pools, int gpu_count,
//                     Timer** copy, Timer** idxmatch, Timer** dprtn,
//                     Timer** contraction, int dense_accumulator = 0)
//                                                    (include/gsparc/sptc.cuh:16-26)
//   gsparc::multi_SpTC<...>(same, without dense_accumulator) (sptc.cuh:362-372)
//   gsparc::extra::SpTC<...>(same as SpTC)                   (sptc.cuh:702-712)
//
// so main.cu's sequence (main.cu:93-381: TensorManager calls, then SpTC /
// multi_SpTC / extra::SpTC) compiles unchanged against these headers.  SX
// and SY must have been paired by SortSlitomXY; the contraction runs on one
// GPU (gpu_count 1) with pools[0]'s size as the device budget, and the
// per-GPU timer arrays receive the engine's phase times (HIP events): copy
// = uploads and per-part preparation, idxmatch = index match, dprtn =
// dynamic partitioning, contraction = accumulation.  SZ receives nnz and the
// device-resident result (SZ->result).  Errors throw gsparc::Error with the
// sptc_status instead of exiting the process.
#ifndef GSPARC_SPTC_HPP_
#define GSPARC_SPTC_HPP_
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../sptc.h"
#include "cmdline_opts.hpp"
#include "cuda_memory.hpp"
#include "io_manager.hpp"
#include "slitom.hpp"
#include "sparse_tensor.hpp"
#include "tensor_manager.hpp"
#include "timer.hpp"

namespace gsparc {

// Decoded COO output of a result: per-mode 0-based indices (X free modes
// ascending, then Y free modes) and values.
inline SparseTensor<uint32_t, float>* decode(const sptc_result* r) {
  sptc_result_info info{};
  check(sptc_result_get_info(r, &info));
  std::vector<uint32_t> dims(info.out_order, 0);
  auto* z = new SparseTensor<uint32_t, float>((unsigned short)info.out_order, dims.data(),
                                              info.nnz, 0);
  try {
    check(sptc_result_decode(r, z->indices, z->values));
  } catch (...) {
    delete z;
    throw;
  }
  for (int m = 0; m < z->order; ++m)
    for (uint64_t i = 0; i < z->nnz; ++i)
      if (z->indices[m][i] + 1 > z->dims[m]) z->dims[m] = z->indices[m][i] + 1;
  return z;
}

namespace detail {

template <typename S>
sptc_result_info execute(S* SX, S* SZ, CudaMemoryPool** pools, int gpu_count, Timer** copy,
                         Timer** idxmatch, Timer** dprtn, Timer** contraction, int dense) {
  if (gpu_count != 1) throw Error(SPTC_ERR_BAD_ARGS, "this engine runs on one GPU (-g 1)");
  if (!SX->plan) throw Error(SPTC_ERR_BAD_ARGS, "SpTC: SortSlitomXY(SX, SY) first");
  sptc_opts o{};
  o.dense_accumulator = dense;
  o.mem_budget_bytes = pools && pools[0] ? pools[0]->getPoolSize() : 0;
  sptc_result* r = nullptr;
  check(sptc_execute(SX->plan.get(), &o, nullptr, &r));
  SZ->result.reset(r, sptc_result_free);
  sptc_result_info info{};
  check(sptc_result_get_info(r, &info));
  SZ->nnz = info.nnz;
  const double* ms = info.phase_ms;
  if (copy && copy[0]) copy[0]->add((ms[0] + ms[1]) * 1e-3);
  if (idxmatch && idxmatch[0]) idxmatch[0]->add(ms[2] * 1e-3);
  if (dprtn && dprtn[0]) dprtn[0]->add(ms[3] * 1e-3);
  if (contraction && contraction[0]) contraction[0]->add(ms[4] * 1e-3);
  return info;
}

}  // namespace detail

// sptc.cuh:16-360: Z = X x Y on the pair's plan.
template <typename TensorType, typename SLITOMType>
void SpTC(SLITOMType* SX, SLITOMType* SY, SLITOMType* SZ, CudaMemoryPool** memory_pools,
          int gpu_count, Timer** copy_timers, Timer** idxmatch_timers, Timer** dprtn_timers,
          Timer** contraction_timers, int dense_accumulator = 0) {
  if (SX->peer != SY && SX->plan != SY->plan)
    throw Error(SPTC_ERR_BAD_ARGS, "SpTC: SX and SY were not paired by SortSlitomXY");
  detail::execute(SX, SZ, memory_pools, gpu_count, copy_timers, idxmatch_timers, dprtn_timers,
                  contraction_timers, dense_accumulator);
}

// sptc.cuh:362-698: a contraction whose output feeds the next one.  SZ
// keeps the device result, which SortSlitomXY(SZ, W) then uses as the X
// operand (its trailing modes -- the Y free modes -- contracted with W's),
// and the reference's host fields: packed keys (X free << Y nfbits | Y
// free; left zero when they need more than 64 bits) and values in
// SZ->indices / SZ->values, dims, nfbits = Y's, ncbits = X's.  The
// operand pair is the one SortSlitomXY made for SX: for the second step of
// main.cu:345-363 that is (SZ, W), so the SY argument is not used.
template <typename TensorType, typename SLITOMType>
void multi_SpTC(SLITOMType* SX, SLITOMType* SY, SLITOMType* SZ, CudaMemoryPool** memory_pools,
                int gpu_count, Timer** copy_timers, Timer** idxmatch_timers,
                Timer** dprtn_timers, Timer** contraction_timers) {
  (void)SY;
  using lindex_t = typename SLITOMType::lindex_t;
  using value_t = typename SLITOMType::value_t;
  const sptc_result_info info =
      detail::execute(SX, SZ, memory_pools, gpu_count, copy_timers, idxmatch_timers,
                      dprtn_timers, contraction_timers, 0);
  SLITOMType* Y = SX->peer;
  sptc_coo as{};
  check(sptc_result_as_coo(SZ->result.get(), nullptr, &as));
  const int order = as.order;
  const int fx = info.out_order - (Y ? Y->fnmodes : 0);
  deallocate(SZ->dims);
  deallocate(SZ->cpos);
  deallocate(SZ->fpos);
  deallocate(SZ->indices);
  deallocate(SZ->values);
  deallocate(SZ->prtn_idx);
  SZ->dims = allocate<lindex_t>(order);
  for (int m = 0; m < order; ++m) SZ->dims[m] = (lindex_t)as.dims[m];
  SZ->nmodes = (unsigned short)order;
  SZ->fnmodes = (unsigned short)fx;
  SZ->cnmodes = (unsigned short)(order - fx);
  SZ->fpos = allocate<int>(fx);
  for (int m = 0; m < fx; ++m) SZ->fpos[m] = m;
  SZ->cpos = allocate<int>(order - fx);
  for (int m = fx; m < order; ++m) SZ->cpos[m - fx] = m;
  SZ->nfbits = info.y_nfbits;
  SZ->ncbits = info.x_nfbits;
  SZ->nbits = info.x_nfbits + info.y_nfbits;
  SZ->input_order = 1;
  SZ->indices = allocate<lindex_t>(info.nnz);
  SZ->values = allocate<value_t>(info.nnz);
  if (info.out_bits <= 64 && sizeof(lindex_t) == 8) {
    check(sptc_result_copy_to_host(SZ->result.get(), reinterpret_cast<uint64_t*>(SZ->indices),
                                   nullptr, nullptr, reinterpret_cast<float*>(SZ->values)));
  } else {
    std::memset(SZ->indices, 0, sizeof(lindex_t) * (info.nnz ? info.nnz : 1));
    check(sptc_result_copy_to_host(SZ->result.get(), nullptr, nullptr, nullptr,
                                   reinterpret_cast<float*>(SZ->values)));
  }
  SZ->nprtn = 1;
  SZ->prtn_idx = allocate<uint64_t>(2);
  SZ->prtn_idx[0] = 0;
  SZ->prtn_idx[1] = info.nnz;
  SZ->coo = sptc_coo{};  // SortSlitomXY takes the device result
  SZ->cmodes.clear();
}

namespace extra {
// sptc.cuh:702-1048: keys above 64 bits; the engine lays them out itself.
template <typename TensorType, typename SLITOMType>
void SpTC(SLITOMType* SX, SLITOMType* SY, SLITOMType* SZ, CudaMemoryPool** memory_pools,
          int gpu_count, Timer** copy_timers, Timer** idxmatch_timers, Timer** dprtn_timers,
          Timer** contraction_timers, int dense_accumulator = 0) {
  gsparc::SpTC<TensorType, SLITOMType>(SX, SY, SZ, memory_pools, gpu_count, copy_timers,
                                       idxmatch_timers, dprtn_timers, contraction_timers,
                                       dense_accumulator);
}
}  // namespace extra

}  // namespace gsparc
#endif
