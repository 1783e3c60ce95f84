// Host launchers of the result-side kernels (result.hip).
#pragma once
#include <algorithm>

#include "common.hpp"

namespace sptc {

// Per-workgroup partial of the verification pass.
struct VerifyPart {
  double sum;    // FP64 sum of the values the workgroup read
  uint64_t bad;  // entries out of order or outside the chunk's rows
};

constexpr uint32_t kVerifyMaxBlocks = 4096;
uint32_t verify_grid(uint64_t n);  // workgroups (and partials) for n entries

// Checks one chunk: r0 <= row[i] < r1 and strictly ascending (row, col);
// part receives verify_grid(n) partials.
void launch_verify(const uint32_t* row, const uint64_t* col, const float* val, uint64_t n,
                   uint64_t r0, uint64_t r1, VerifyPart* part, hipStream_t s);

constexpr int kMaxOutModes = 2 * SPTC_MAX_ORDER;

// Output field layout for decoding: modes [0, nx) come from the X free key,
// [nx, nx + ny) from the Y free key; index = (key >> shift) & mask.
struct DecodeArgs {
  int nx, ny;
  int shift[kMaxOutModes];
  uint64_t mask[kMaxOutModes];
  uint32_t* out[kMaxOutModes];  // device arrays, entry i of the chunk at out0 + i
  float* vals;                  // may be NULL
};

void launch_decode(const uint32_t* row, const uint64_t* col, const float* val, uint64_t n,
                   const uint64_t* row_key, const DecodeArgs& a, uint64_t out0, hipStream_t s);

// ---- merge of pair results (static operand partitioning) -----------------
// The results of one X part against each Y part share the X part's row
// table; Y parts hold ascending, disjoint column ranges, so the merged
// output lists each row's entries part by part.

constexpr int kMaxMergeParts = 64;

// Entry range of each row of one output chunk whose rows ascend:
// start[r] = base + first entry of row r, end[r] = base + last + 1.  Rows
// without entries keep what the caller filled in (zeros: an empty range).
void launch_row_extent(const uint32_t* row, uint64_t n, uint64_t base, uint64_t* start,
                       uint64_t* end, hipStream_t s);

struct ExtentSet {
  int Q;
  const uint64_t* start[kMaxMergeParts];
  const uint64_t* end[kMaxMergeParts];
};
// cnt[r * Q + q] = entries of row r in part q's result (r < R).
void launch_interleave_counts(const ExtentSet& e, uint64_t R, uint64_t* cnt, hipStream_t s);

// Moves entries [0, n) of one chunk of part q's result (its entries base,
// base + 1, ... of the part's result) to dst_base[r * Q + q] + (base + i -
// start[r]), adding row_off to the row index.
void launch_merge_scatter(const uint32_t* row, const uint64_t* col, const float* val,
                          uint64_t n, uint64_t base, const uint64_t* start,
                          const uint64_t* dst_base, int Q, int q, uint32_t row_off,
                          uint32_t* orow, uint64_t* ocol, float* oval, hipStream_t s);

}  // namespace sptc
