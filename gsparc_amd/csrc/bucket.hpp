// Big-row reduce of the high-digit form (RowSegPlan::msd).
//
// After one scatter pass on the top kBucketRB column bits, each big row's
// products lie in 2^kBucketRB buckets, bucket d holding the products whose
// columns are d << low .. (d + 1) << low - 1, in expansion order (the pass is
// stable).  A bucket's columns fit an LDS row, so it is reduced there instead
// of being sorted on its low bits: each product is added to its column's
// slot with the compensated FP32 step in product order (products of one
// column in the same 64-product window are applied lowest lane first), the
// touched slots are marked in a bitmap and emitted in column order.  The sums
// are those of the fully sorted path (same values, same order, same steps),
// so the outputs agree bit for bit.
//
// An option (SPTC_BIG_TAIL=buckets): on the NIPS workload it replaces the
// second scatter pass, its count and the narrow reduce of the big rows
// (about 28 ms per 2.6e9-product chunk) with a count and an emit that take
// about 30 and 96 ms -- per 64-product window the column matching, the
// ordered slot updates and the per-bucket compaction cost several hundred
// wave64 VALU instructions, against about 90 for a scatter pass.
//
// A count pass first gives each bucket's distinct columns; a per-row scan
// turns them into the buckets' output offsets, and each row's total joins
// the narrow reduce's wave-tile counts at the tile after the row's first
// product (that tile lies inside the row, which holds more than half a sort
// tile of products), so one scan of the tile counts places small-row and
// big-row outputs in row order.
#pragma once
#include "segsort.hpp"

namespace sptc {

struct BucketArgs {
  const SegDesc* segs;     // the plan's big rows (bigseg), in row order
  const SegTile* tiles;    // their sort tiles
  uint64_t nsegs;
  const uint32_t* starts;  // count matrix: a row's first-tile row holds its digit starts
  uint32_t* boff;          // offset matrix: first-tile rows get bucket counts, then offsets
  const uint32_t* cols;    // the pass's output (column ranks) and values
  const float* vals;
  int low;                 // column bits below the digit (col_bits - kBucketRB)
  int gshift;              // log2 of the buckets each wave takes (<= 6)
  const uint32_t* bigrow;  // chunk-local row of each big row (RowSegPlan::bigrow)
  U2* tile_sums;           // the narrow reduce's wave-tile counts / offsets
  uint64_t r0;             // first row of the chunk
  uint32_t* out_row;
  uint64_t* out_col;
  float* out_val;
  const uint64_t* ytab;    // rank -> Y free key (NULL: the rank is the key)
  const uint32_t* ytab32;  // the same as 32-bit words (NULL: use ytab)
};

// Buckets per wave for big rows holding nprod products: about 2048 products
// per wave, one to 64 buckets.
int bucket_group_shift(uint64_t nprod, uint64_t nsegs);
// Count pass: per-bucket distinct columns, per-row offsets, row totals
// added to a.tile_sums (after the narrow count wrote them, before their
// scan).  Launches only.
void run_bucket_count(const BucketArgs& a, hipStream_t s);
// Emit: outputs at the scanned tile offsets plus the bucket offsets.
void run_bucket_emit(const BucketArgs& a, hipStream_t s);

}  // namespace sptc
