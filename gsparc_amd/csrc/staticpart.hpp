// Static operand partitioning: operands whose device footprint exceeds half
// the memory budget are split into free-key ranges on the host and
// contracted pair by pair (the reference's FindPartitionNum, Partition and
// convert_and_sort_64 with the X x Y partition-pair loop,
// tensor_manager.tpp:32-61,633-798, sptc.cuh:42-81).
#pragma once
#include <vector>

#include "engine.hpp"

namespace sptc {

// Free-key range split of one operand: part k holds the nonzeros whose free
// key (the X key's high field, the Y key's low field) lies in
// [key_lo[k], key_lo[k + 1]); nnz_lo[k] is the number of nonzeros in the
// parts before k (so nnz_lo[parts] = nnz).  Ranges follow free keys, so no
// free row (X) or output column (Y) is split across parts.
struct HostPart {
  uint64_t nnz = 0;
  int idx_bytes = 4;
  std::vector<std::vector<uint8_t>> idx;  // per mode, nnz * idx_bytes bytes
  std::vector<float> val;
  std::vector<const void*> ptr;
  sptc_coo coo(int order, const uint64_t* dims);
};

struct OperandSplit {
  int parts = 1;
  std::vector<uint64_t> key_lo;  // parts + 1 entries (last = 2^nfbits)
  std::vector<uint64_t> nnz_lo;  // parts + 1 entries
  std::vector<HostPart> part;    // empty when parts == 1
};

// Device bytes an operand of n nonzeros holds through the contraction when
// resident (sorted keys and values, match tables or the Y contract and rank
// tables) and the peak while it is prepared (staged indices, sort
// workspace).
uint64_t resident_bytes(uint64_t n, int role);
uint64_t prepare_bytes(uint64_t n, int order, int idx_bytes);

// FindPartitionNum: numbers of X and Y parts so that one X part, one Y part
// and the preparation of the larger of them fit in half the budget (the
// other half is left to the products and outputs).  Both 1 when the
// operands fit whole.  Returns false when 64 parts per operand do not fit.
bool find_partition_num(uint64_t nx, int xorder, uint64_t ny, int yorder, int idx_bytes,
                        uint64_t budget, int* px, int* py);

// Partition: splits a host COO into `parts` free-key ranges of about equal
// nonzero counts (a histogram of the top free-key bits chooses the cuts).
// Entries keep their input order within a part.  Device inputs are copied
// to the host first.  With parts == 1 the operand is copied into part[0]
// only when `own` is set (a plan keeps its operands), else part is empty.
void split_operand(const sptc_coo& t, const Layout& lay, int parts, bool own,
                   OperandSplit* out);

// The X x Y partition-pair loop: for each X part, every Y part is staged,
// linearized and sorted on the device (the Y re-key and sort of
// convert_and_sort_64), contracted with the X part, and the pair results
// are merged row by row into one chunk of the output (Y parts hold disjoint,
// ascending column ranges, so a row's entries are its pieces in part
// order).  Results are bit-identical to the unpartitioned contraction.
sptc_status contract_partitioned(Engine& e, const sptc_coo& X, const sptc_coo& Y,
                                 const Layout& xl, const Layout& yl, OperandSplit& xs,
                                 OperandSplit& ys, sptc_result_s* res);

}  // namespace sptc
