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

}  // namespace sptc
