// Segmented LSD radix sort: every segment of a key array is sorted on its
// own, with tiles that never straddle a segment.  Used for the operand sorts
// of a batch (one segment per operand) and for product sorts (one segment per
// X row, so only the column bits need sorting: the rows are already in
// order after the expansion).
#pragma once
#include <vector>

#include "kernels.hpp"

namespace sptc {

constexpr int kSegRB = 8;     // digit bits
constexpr int kSegIPT = 12;   // keys per thread
constexpr int kSegNT = 512;   // threads per tile
constexpr uint64_t kSegTile = (uint64_t)kSegIPT * kSegNT;
// Big-row passes of the row-segmented product sort move 32-bit column keys,
// so more of them fit a tile's LDS staging: kNarrowIPT keys per thread (12,
// 16 or 24; SPTC_SEG_IPT overrides the default).  16 keeps two 512-thread
// workgroups per CU (112 VGPRs); 24 drops to one.
constexpr int kNarrowIPT = 16;
// High-digit form of the big-row sort: one pass on the top kBucketRB column
// bits, leaving buckets of at most 2^kBucketMaxLow columns.
constexpr int kBucketRB = 10;
constexpr int kBucketMaxLow = 10;

// One tile of a segmented sort (never straddles a segment).  Its digit
// counts are row `tile` of a tile-major count matrix; the scan leaves in
// the same row of the offset matrix each digit's keys in the segment's
// earlier tiles, and in row `first` of the count matrix the segment's digit
// starts: a key's destination is their sum plus its rank (both rows read
// coalesced).
struct SegTile {
  uint64_t start;  // first key (global index)
  uint64_t tile;   // global tile index
  uint32_t count;  // keys in the tile
  uint32_t first;  // global index of the segment's first tile
};
// One segment: its tiles are first_tile .. first_tile + ntiles - 1; one
// workgroup per segment scans their counts (see SegTile).
struct SegDesc {
  uint64_t start;
  uint32_t first_tile;
  uint32_t ntiles;
};

// A segment small enough to be sorted inside one workgroup's LDS: the
// products of consecutive small rows from chunk-local row row0 on.
struct LocalSeg {
  uint64_t start;
  uint32_t count;  // <= kSegTile
  uint32_t row0;
};

// Segment of global index i: the s with off[s] <= i < off[s+1] (off is
// non-decreasing with nseg+1 entries; empty segments are skipped).
__device__ __forceinline__ uint32_t find_seg(const uint64_t* off, uint32_t nseg, uint64_t i) {
  uint32_t lo = 0, hi = nseg;  // last s with off[s] <= i
  while (lo + 1 < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

// Tile table of a segmented sort: each segment [seg_off[s], seg_off[s+1])
// is cut into its own tiles.  Built once per batch shape (one host->device
// copy).
struct SegSortPlan {
  DevBuf<SegTile> tiles;
  DevBuf<SegDesc> segs;
  uint64_t ntiles = 0, nsegs = 0;
  uint64_t n = 0;
  void build(const std::vector<uint64_t>& seg_off, hipStream_t s);
};
// Sorts each segment of (k0, v0) on bits [0, nbits) with 8-bit reduce-then-
// scan passes (count per tile, per-segment scan, scatter), ping-ponging with
// (k1, v1); ws supplies the count matrices (segmented_sort_reserve).
// Launches only (no allocation or host synchronization), so it can be
// replayed and captured.  Returns true when the result ended in (k1, v1).
void segmented_sort_reserve(SortWorkspace& ws, uint64_t ntiles, hipStream_t s,
                            int rb = kSegRB);
void segmented_sort_reserve(SortWorkspace& ws, const SegSortPlan& plan, hipStream_t s);
bool segmented_sort(SortWorkspace& ws, const SegTile* tiles, uint64_t ntiles,
                    const SegDesc* segs, uint64_t nsegs, uint64_t n, uint64_t* k0, float* v0,
                    uint64_t* k1, float* v1, int nbits, hipStream_t s, KTimer* kt = nullptr,
                    int cls = SPTC_K_SORT_INPUT);
bool segmented_sort(SortWorkspace& ws, const SegSortPlan& plan, uint64_t* k0, float* v0,
                    uint64_t* k1, float* v1, int nbits, hipStream_t s, KTimer* kt = nullptr,
                    int cls = SPTC_K_SORT_INPUT);
int segmented_sort_passes(int nbits);

// Row-segmented tiles over a chunk's products.  The chunk's rows come in
// ranges of consecutive global rows (one range for a plain chunk; one per
// piece of a batched chunk); range q covers global rows [ra, ra + rows) and
// its first product sits at chunk offset out0.  Products of global row r
// number row_prod[r+1] - row_prod[r].
struct RowRange {
  uint64_t ra;
  uint64_t rows;
  uint64_t lrow0;  // chunk-local row of ra (prefix of rows)
  uint64_t out0;   // chunk offset of ra's first product
};
// Row-segmented product sort of a chunk.  Products are grouped by row and
// travel narrow -- a 32-bit column (col_bits <= 32) and a value, the row
// implied by the row starts -- and are sorted by (row, column) into
// (col, val).  Rows holding more than half a 6144-product tile are segments
// of their own, sorted on the column bits by the multi-tile passes above
// (digits of 8 bits, or 10 where that saves a pass; tiles of ipt * kSegNT
// columns) from their narrow target (narrow_target) into (col, val) --
// except rows with a single owner (with row_act given), which are already
// in column order and are cut into group items of at most 6144 products
// instead.  Runs of smaller rows are cut into groups of at most 6144
// products (a group starts where the running small-row total crosses a
// multiple of half a tile); each group item is sorted on its (row in group,
// column) bits entirely in one workgroup's LDS, in place in (col, val),
// unless it is already in that order.  Every big-row product is read and
// written once per digit pass (at least one), 8 bytes each way; a group
// item reads its columns, and when it needs sorting its values, and writes
// both back.
struct RowSegPlan {
  // nsmall: products in group items (small-row groups and one-owner rows)
  uint64_t nlocal = 0, nbig_tiles = 0, nbig_rows = 0, nsmall = 0;
  uint64_t lrows = 0;
  int col_bits = 0;
  int rb = kSegRB;  // digit bits of the big-row passes (8 or 10)
  int ipt = kNarrowIPT;  // keys per thread of their tiles
  // high-digit form: the big rows take one pass on their top rb column bits
  // (10 < col_bits <= 20), which leaves each row's products in 2^rb buckets
  // of 2^(col_bits - rb) columns, still in expansion order; the buckets are
  // reduced in LDS (bucket.hpp) instead of being sorted on the low bits
  bool msd = false;
  DevBuf<LocalSeg> local;
  DevBuf<SegTile> big;
  DevBuf<SegDesc> bigseg;
  DevBuf<uint32_t> bigrow;   // high-digit form: chunk-local row of each big segment
  DevBuf<uint64_t> scratch;  // 10 x (rows + 1) per-row values and scans
  // per chunk-local row: nonzero when the row is a big segment
  const uint64_t* row_big() const { return scratch.get() + (lrows + 1); }
  // per chunk-local row: its first product; [lrows] = the product count
  const uint64_t* row_start() const { return scratch.get() + 8 * (lrows + 1); }
};
// Builds the plan (one host synchronization for the counts); tmp holds
// scan_tmp_count(lrows) + 1 entries.  Reserves ws for the big-row passes.
// msd: use the high-digit form where the column bits allow it.
// row_act (global row -> its first active X nonzero, [rows + 1]; may be
// NULL): a row above half a tile with a single owner is already in column
// order (its products are one Y segment), so it skips the big-row passes
// and becomes group items of at most kSegTile products each, which the
// group sort finds in order.
void build_row_segments(const RowRange* ranges, uint32_t nranges, uint64_t lrows,
                        const uint64_t* row_prod, int col_bits, RowSegPlan* plan,
                        SortWorkspace& ws, U2* tmp, hipStream_t s, bool msd = false,
                        const uint64_t* row_act = nullptr);
// Where the expansion writes big-row products of an n-product chunk sorted
// into (col, v) with spare arrays (k1, v1) of n 64-bit keys and n values: a
// 32-bit column array and a value array, chosen so that the big-row passes
// end in (col, v).
struct NarrowTarget {
  uint32_t* col;
  float* val;
};
NarrowTarget narrow_target(const RowSegPlan& plan, uint64_t n, float* v, uint64_t* k1,
                           float* v1);
// Sorts with a plan: big rows from their narrow target, groups in place;
// row_start as RowSegPlan::row_start (kept by the caller when the plan's
// scratch is released).  Launches only (replayable).
// counted: the first pass's digit counts are already in ws (launch_expand_items).
// v and v1 NULL: the columns are sorted alone (count-only runs).
void sort_row_segments(SortWorkspace& ws, const RowSegPlan& plan, uint64_t n, uint32_t* col,
                       float* v, uint64_t* k1, float* v1, const uint64_t* row_start,
                       hipStream_t s, KTimer* kt = nullptr, int cls = SPTC_K_SORT_PROD,
                       bool counted = false);
// The chunk's expansion organised by the plan's work items (big-row tiles,
// then small-row groups): a.big_col/big_val from narrow_target, a.out_col32/
// out_vals the sorted arrays, a.ycol required.  Also writes the big tiles'
// first-pass digit counts into ws, so sort_row_segments(..., counted=true)
// skips that pass's upsweep.  NULL value targets (a.big_val, a.out_vals):
// columns only, the value sum still taken.
void launch_expand_items(const ExpandArgs& a, const RowSegPlan& plan, SortWorkspace& ws,
                         hipStream_t s);

}  // namespace sptc
