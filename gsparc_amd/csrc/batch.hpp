// Batched (multi-contraction) kernels and their host launchers.
//
// A batch concatenates the operands of B contractions into shared arrays
// with per-instance offset tables; every phase is one launch over all
// instances.  X rows are numbered globally across the batch, so the product
// keys of an ESC chunk (chunk-local row << K | column rank) sort the
// products of every instance in the chunk with one unsegmented radix sort;
// operand sorts are segmented (one segment per operand, tiles never
// straddle segments).
#pragma once
#include <vector>

#include "kernels.hpp"
#include "segsort.hpp"

namespace sptc {

constexpr int kBatchMaxInstances = 4096;

// Per-operand linearize arguments (device table, one per segment).  With
// skeys set the segment is an already linearized and sorted operand and is
// copied instead.
struct BLinArgs {
  const void* idx[SPTC_MAX_ORDER];
  int shift[SPTC_MAX_ORDER];
  int width[SPTC_MAX_ORDER];
  int order;
  int idx_bytes;
  const float* val;
  const uint64_t* skeys;
  const float* svals;
};

// Per-instance parameters used by the match/accumulate kernels.
struct BInst {
  uint64_t cmask;    // X contract field mask
  int ncbits;        // X contract field width
  int kyf;           // Y column (rank) field width
  uint64_t yt_off;   // first entry of this instance's rank -> free-key table
};

void launch_blinearize(const BLinArgs* args, const uint64_t* seg_off, uint32_t nseg,
                       uint64_t n, uint64_t* keys, float* vals, hipStream_t s);

// Y free keys (key & fmask of the segment's instance) for rank compression.
void launch_bfree_extract(const uint64_t* ykeys, const uint64_t* yoff, uint32_t B,
                          const uint64_t* fmask, uint64_t n, uint64_t* out, hipStream_t s);
// Distinct values per segment of a segmented-sorted array: table[] gets
// them, seg_first[s] the table index of segment s's first distinct value.
void run_bunique(const uint64_t* sorted, const uint64_t* seg_off, uint32_t nseg, uint64_t n,
                 uint64_t* table, uint64_t* seg_first, U2* tmp, hipStream_t s);
// Y key (c << ynf | f) -> (c << kyf | rank of f in its instance's table).
// ycol: the rank (the key's column field) of every Y key as 32 bits.
void launch_brank(const uint64_t* ykeys, const uint64_t* yoff, uint32_t B, const int* ynf,
                  const int* kyf, const uint64_t* ytab, const uint64_t* yt_off, uint64_t n,
                  uint64_t* out, uint32_t* ycol, hipStream_t s);
// Distinct contract values per Y instance: ukey/ustart (global Y index),
// u_off[s] = first table entry of instance s; ustart[total] = n.
void run_bycontract(const uint64_t* ykeys, const uint64_t* yoff, uint32_t B, const int* kyf,
                    uint64_t n, uint64_t* ukey, uint64_t* ustart, uint64_t* u_off, U2* tmp,
                    hipStream_t s);

struct BMatchOut {
  uint64_t* act_prefix;
  uint64_t* act_ypos;
  float* act_xval;
  uint32_t* act_row;
  uint64_t* row_key;
  uint64_t* row_prod;
  uint64_t* row_act;
  uint64_t* inst_P;  // [B] first product / active entry / row of instance s
  uint64_t* inst_A;
  uint64_t* inst_R;
  const uint32_t* bad;  // set: operands changed since planning; write nothing
};
void run_bmatch(const uint64_t* xkeys, const float* xvals, const uint64_t* xoff, uint32_t B,
                const BInst* inst, const uint64_t* ukey, const uint64_t* ustart,
                const uint64_t* u_off, uint64_t nx, const BMatchOut& o, U2* tmp, hipStream_t s);

// An ESC chunk is a list of pieces, each a contiguous global row range of
// ESC instances; dense instances' rows between pieces are skipped.  Chunk-
// local row = lrow0 + (row - ra); product key = local row << K | rank.
struct BPiece {
  uint64_t ra;     // first global row
  uint64_t lrow0;  // chunk-local row of ra
  uint64_t A0, A1; // active X entries of the piece
};
// Expand tile: products [p0, p0 + cnt) of piece `piece`, written at chunk
// offset out0.  Tiles never straddle pieces.
struct BETile {
  uint64_t p0;
  uint64_t out0;
  uint32_t cnt;
  uint32_t piece;
};
constexpr uint32_t kBExpandTileItems = 256 * 8;  // products per expand tile

struct BExpandArgs {
  const uint64_t* act_prefix;
  const uint64_t* act_ypos;
  const float* act_xval;
  const uint32_t* act_row;
  const uint32_t* ycol;   // Y region: every key's column field (rank) as 32 bits
  const float* yvals;
  const BPiece* pieces;
  const BETile* tiles;
  uint32_t ntiles;
  int K;
  uint64_t* out_keys;
  float* out_vals;
  // narrow form (row_big non-null): products go out as 32-bit columns, to
  // (big_col, big_val) for big rows and (out_col32, out_vals) for the others
  const uint64_t* row_big;  // chunk-local row -> big segment
  uint32_t* big_col;
  float* big_val;
  uint32_t* out_col32;
  const uint32_t* bad;     // nonzero: operands changed since planning, skip
  uint32_t remap;          // XCD-contiguous tile order (set by launch_bexpand)
};
void launch_bexpand(const BExpandArgs& a, hipStream_t s);

// Reduce-by-key emit with per-instance column tables.  Outputs are written with instance-local rows; bounds[s] (s <= B, filled
// with the output count beforehand) receives the first output of instance s.
struct BReduceArgs {
  const uint32_t* map_row;   // chunk-local row -> global row
  const uint32_t* map_inst;  // chunk-local row -> instance
  const uint64_t* inst_R;
  const BInst* inst;
  const uint64_t* ytab;
  const uint32_t* ytab32;  // ytab as 32-bit words when every Y free key fits (or NULL)
  uint32_t B;
  uint64_t* bounds;
  const uint32_t* bad;
};
// Chunk-local row -> (global row, instance) for the rows of a chunk's pieces.
void launch_chunk_row_maps(const BPiece* pieces, uint32_t npieces, uint64_t lrows,
                           const uint64_t* inst_R, uint32_t B, uint32_t* map_row,
                           uint32_t* map_inst, hipStream_t s);
// Per-instance output starts of a dense class buffer (instances not in the
// class get the next class instance's start, or nnz).
void launch_dense_bounds(const uint64_t* grp_rows, const uint32_t* grp_inst, uint32_t G,
                         const uint64_t* out_off, uint64_t nnz, uint32_t B, uint64_t* bounds,
                         const uint32_t* bad, hipStream_t s);
void run_breduce_emit(const uint64_t* keys, const float* vals, uint64_t n, uint32_t* out_row,
                      uint64_t* out_col, float* out_val, int K, const BReduceArgs& b,
                      const U2* tmp, hipStream_t s);
// The same over a chunk's narrow form (keys: row << 32 | column).
void run_breduce_emit_narrow(const NarrowRows& nr, const float* vals, uint64_t n,
                             uint32_t* out_row, uint64_t* out_col, float* out_val,
                             const BReduceArgs& b, const U2* tmp, hipStream_t s);

// Dense accumulation over a group of instances sharing one slot width.
struct BDenseArgs {
  DenseArgs d;                 // d.r0 unused; d.fmask/d.ytab unused
  const uint64_t* grp_rows;    // [G+1] prefix of the group's row counts
  const uint64_t* grp_row0;    // [G] first global row of each group instance
  const uint32_t* grp_inst;    // [G] instance id
  const BInst* inst;
  const uint64_t* ytab;
  uint32_t G;
  const uint32_t* bad;
};
void launch_bdense_count(const BDenseArgs& a, int eb, uint64_t rows, hipStream_t s);
void launch_bdense_numeric(const BDenseArgs& a, int eb, uint64_t rows, hipStream_t s);

// arr[s] of every empty segment s (seg_off[s] == seg_off[s+1]) takes the
// value of the next segment, so per-segment starts are monotone with
// arr[nseg] = total.
void launch_fix_empty(uint64_t* arr, const uint64_t* seg_off, uint32_t nseg, hipStream_t s);
// flags[s] = 1 when segment s of a sorted key array holds equal neighbours.
void launch_bany_dup(const uint64_t* keys, const uint64_t* seg_off, uint32_t nseg, uint64_t n,
                     uint32_t* flags, hipStream_t s);
// Position-dependent 64-bit digest of n keys, added into *out (cleared by
// the caller); launch_check_digest sets *bad when *got != *want.
void launch_digest(const uint64_t* keys, uint64_t n, uint64_t* out, hipStream_t s);
void launch_check_digest(const uint64_t* got, const uint64_t* want, uint32_t* bad,
                         hipStream_t s);
// *out += sum of vals[0, n) in FP64.
void launch_sum_f64(const float* vals, uint64_t n, double* out, hipStream_t s);

}  // namespace sptc
