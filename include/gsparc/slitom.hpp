// gsparc/slitom.hpp -- the reference's linearized operand descriptor
// (include/gsparc/slitom.hpp:27-92) with its public fields, so code that
// reads them compiles unchanged.  Here the key arrays live on the device,
// inside the engine: ConvertTensor records the tensor and fills the layout
// fields, SortSlitomXY builds the pair's plan (sptc_plan) that both
// operands share, Partition copies the plan's static parts into
// nprtn / prtn_idx / prtn_coord, and SpTC leaves its output in the output
// SLITOM (`result`, nnz, and for chained contractions the host keys).
#ifndef GSPARC_SLITOM_HPP_
#define GSPARC_SLITOM_HPP_
#include <cstdint>
#include <memory>
#include <vector>

#include "../sptc.h"
#include "sparse_tensor.hpp"

namespace gsparc {

typedef enum PackOrder_ { LSB_FIRST, MSB_FIRST } PackOrder;
typedef enum ModeOrder_ { SHORT_FIRST, LONG_FIRST, NATURAL } ModeOrder;
struct MPair {
  int mode;
  int bits;
};

template <typename MaskType, typename ULIndexType, typename LIndexType, typename ValueType>
class SLITOM {
 public:
  using mask_t = MaskType;
  using ulindex_t = ULIndexType;
  using lindex_t = LIndexType;
  using value_t = ValueType;

  SLITOM() : SLITOM(0, true) {}
  explicit SLITOM(int input_order) : SLITOM(input_order, true) {}
  SLITOM(int input_order, bool sort_gpu) : input_order(input_order), sort_gpu(sort_gpu) {}
  SLITOM(const SLITOM&) = delete;
  SLITOM& operator=(const SLITOM&) = delete;
  ~SLITOM() {
    deallocate(dims);
    deallocate(cpos);
    deallocate(fpos);
    deallocate(indices);
    deallocate(values);
    deallocate(prtn_idx);
    deallocate(prtn_coord);
  }

  int input_order = 0;  // 1 = X, 2 = Y, 0 = output
  MPair* mode_bits = nullptr;
  lindex_t* dims = nullptr;
  uint64_t nnz = 0;
  unsigned short nmodes = 0, fnmodes = 0, cnmodes = 0;
  int nfbits = 0, ncbits = 0, nbits = 0;
  int* cpos = nullptr;  // contracted modes, in the caller's order
  int* fpos = nullptr;  // free modes, ascending
  mask_t fmask = 0, cmask = 0, slitom_mask = 0;
  // Host keys and values: filled for the output of a chained contraction
  // (multi_SpTC), whose packed keys the reference keeps on the host; the
  // operands' keys stay on the device.
  ulindex_t* uindices = nullptr;
  lindex_t* indices = nullptr;
  value_t* values = nullptr;
  int nprtn = 1;
  bool sort_gpu = true;
  uint64_t* prtn_idx = nullptr;     // nnz before each static part [nprtn + 1]
  lindex_t* prtn_coord = nullptr;   // first free key of each part [nprtn + 1]
  uint64_t max_prtn_size = 0;

  // ---- engine side ---------------------------------------------------
  sptc_coo coo{};                      // the operand (ConvertTensor)
  std::vector<const void*> coo_ptr;
  std::vector<uint64_t> coo_dims;
  std::vector<int32_t> cmodes;         // contracted modes (cpos)
  std::shared_ptr<sptc_plan_t> plan;   // shared by a SortSlitomXY'd pair
  int plan_role = 0;                   // 1: the plan's X, 2: its Y
  SLITOM* peer = nullptr;              // the other operand of the plan
  std::shared_ptr<sptc_result> result; // output of SpTC / multi_SpTC
};

}  // namespace gsparc
#endif
