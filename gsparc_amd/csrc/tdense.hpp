// Tiled dense accumulator (tdense.hip): per-row dense accumulation in LDS
// over column tiles, for column spaces wider than one LDS row.
#pragma once
#include "kernels.hpp"

namespace sptc {

// Tile widths: 2^15 column ranks per tile in the symbolic pass (bitmap
// only), 2^12 in the numeric pass (FP64 slots).
constexpr int kTDenseCountBits = 15;
constexpr int kTDenseNumBits = 12;

struct TDenseArgs {
  const uint64_t* act_prefix;  // [A + 1] product prefix of the active nonzeros
  const uint64_t* act_ypos;    // [A] first matching Y entry
  const float* act_xval;       // [A]
  const uint64_t* row_act;     // [R + 1] first active nonzero of each X row
  uint64_t r0;                 // first row of the launch (block b = row r0 + b)
  const uint32_t* ycol;        // [ny] column rank of each Y entry
  const float* yvals;          // [ny]
  const uint64_t* ytab;        // rank -> Y free key (NULL: the rank is the key)
  uint64_t* cur;               // [A] cursors (scratch, written by the kernels)
  uint32_t* row_cnt;           // symbolic: outputs of each row of the launch
  const uint64_t* out_off;     // numeric: first output of each row of the launch
  uint64_t out0;               // numeric: subtracted from out_off (chunk start)
  uint32_t* out_row;
  uint64_t* out_col;
  float* out_val;
  double* vsum;                // numeric: += sum of the products
};

void launch_tdense_count(const TDenseArgs& a, uint64_t rows, hipStream_t s);
void launch_tdense_numeric(const TDenseArgs& a, uint64_t rows, hipStream_t s);

}  // namespace sptc
