// Host launchers and argument structs of the SpTC phase kernels.
#pragma once
#include "common.hpp"
#include "scan.hpp"
#include "sort.hpp"

namespace sptc {

constexpr int kMaxFusedPasses = 8;  // producers fuse histograms up to this many passes
constexpr int kDirectBits = 22;     // direct-addressed Y segment table up to 2^22 keys

struct LinArgs {
  const void* idx[SPTC_MAX_ORDER];
  int shift[SPTC_MAX_ORDER];
  int width[SPTC_MAX_ORDER];
  int order;
};

// Packed key of a group of contracted coordinates, above the previous
// group's rank: key = prev << prev_shift | OR_j (idx_j & (2^w_j-1)) << s_j.
struct TupleArgs {
  const void* idx[SPTC_MAX_ORDER];
  int shift[SPTC_MAX_ORDER];
  int width[SPTC_MAX_ORDER];
  int n;
  int bytes;               // 4 or 8 (index arrays)
  const uint64_t* prev;    // NULL for the first group
  int prev_shift;
};

struct YTable {
  int direct;
  const uint64_t* lo;      // direct: [2^ncbits]
  const uint64_t* hi;
  const uint64_t* ukey;    // search: [U] distinct contract values
  const uint64_t* ustart;  // [U+1]
  uint64_t U;
};

struct MatchOut {
  // engine form (compacted active X nonzeros and X rows)
  uint64_t* act_prefix;
  uint64_t* act_ypos;
  float* act_xval;
  uint32_t* act_row;
  uint64_t* row_key;
  uint64_t* row_prod;
  uint64_t* row_act;
  // per-phase API form
  uint64_t* mpos;
  uint64_t* mcnt;
  uint64_t* prefix;
};

struct ExpandArgs {
  const uint64_t* act_prefix;
  const uint64_t* act_ypos;
  const float* act_xval;
  const uint32_t* act_row;
  uint64_t A0, A1, P0, P1;
  uint32_t r0;
  const uint64_t* ykeys;
  const float* yvals;
  int ynf;
  uint64_t fmask;
  uint64_t* out_keys;
  float* out_vals;
  // row-segmented sort (launch_expand_items, narrow form): products go out
  // as 32-bit columns, to (big_col, big_val) for big-row tiles and to
  // (out_col32, out_vals) for small-row groups; launch_expand writes 64-bit
  // keys to out_keys
  uint32_t* big_col;
  float* big_val;
  uint32_t* out_col32;
  // column field of every Y key as 32 bits (NULL: taken from ykeys & fmask)
  const uint32_t* ycol;
  uint64_t* hist;
  int passes;
  double* vsum;
  uint32_t remap;  // XCD-contiguous tile order (set by launch_expand)
  // big-row tiles count the first pass's digits at this column shift
  // (set by launch_expand_items from the plan)
  int dshift = 0;
};

struct DenseArgs {
  const uint64_t* act_prefix;  // [A+1], act_prefix[A] = ir
  const uint64_t* act_ypos;
  const float* act_xval;
  const uint64_t* row_act;     // [R+1]
  uint64_t r0;                 // first row of the chunk
  const uint64_t* ykeys;
  const float* yvals;
  uint64_t fmask;              // column field mask
  const uint64_t* ytab;        // rank -> Y free key (NULL: field is the key)
  uint32_t* row_cnt;           // [rows] symbolic counts
  const uint64_t* out_off;     // [rows] output offsets
  uint32_t* out_row;
  uint64_t* out_col;
  float* out_val;
  double* vsum;
};

constexpr int kDenseMaxBits = 13;  // dense LDS rows up to 8192 slots
// The dense kernels spread one active X nonzero's Y segment over a 256-thread
// workgroup at a time; below this many products per active nonzero most
// threads idle and ESC is faster (measured: ~3e9 products/s at 3-4 per
// nonzero, against ~1e10 for ESC on the same instances).
constexpr uint64_t kDenseMinRun = 16;
// The ordered dense pass applies a round's owners (active X nonzeros) one
// after another, a barrier between them, and a row is one workgroup: rows
// with many owners serialise on those barriers.  On the batched workload's
// 57 dense-eligible instances dense won (0.74-0.99x of ESC's time) up to 5.1
// owners per row and lost from 5.5 on (1.03x, up to 2.9x at 4000 per row).
constexpr uint64_t kDenseMaxOwnersPerRow = 5;
// The automatic choice (single and batched planners alike): dense when its
// rows are dense enough for the slots (ir * 32 >= rows * slots), the Y
// segments long enough and the owners per row few enough.
inline bool dense_pays(uint64_t ir, uint64_t rows, uint64_t active, uint64_t slots) {
  return ir * 32 >= rows * slots && ir >= kDenseMinRun * active &&
         active <= kDenseMaxOwnersPerRow * rows;
}

// Dense numeric passes sum in owner order (bit-identical to ESC) unless the
// FP64 LDS-atomic form is selected (SPTC_DENSE_ORDER=atomic).
bool dense_ordered();
void launch_dense_count(const DenseArgs& a, int eb, uint64_t rows, hipStream_t s);
void launch_dense_numeric(const DenseArgs& a, int eb, uint64_t rows, hipStream_t s);
void launch_any_dup(const uint64_t* keys, uint64_t n, uint32_t* flag, hipStream_t s);
void run_row_count_scan(const uint32_t* cnt, uint64_t* off, uint64_t n, U2* tmp, hipStream_t s);

void launch_linearize(const LinArgs& a, int idx_bytes, uint64_t n, uint64_t* keys,
                      uint64_t* hist, int passes, int rb, hipStream_t s);
void launch_ytable_direct(const uint64_t* ykeys, uint64_t ny, int ynf, uint64_t K,
                          uint64_t* lo, uint64_t* hi, hipStream_t s);
void run_yunique(const uint64_t* ykeys, uint64_t ny, int ynf, uint64_t* ukey,
                 uint64_t* ustart, U2* tmp, hipStream_t s);
void run_match(const uint64_t* xkeys, const float* xvals, uint64_t nx, int ncbits,
               const YTable& t, const MatchOut& o, int api, U2* tmp, hipStream_t s);
// Reduce half of run_match: tmp[scan_tmp_count(nx)] = {ir, A << 32 | R}.
void run_match_count(const uint64_t* xkeys, uint64_t nx, int ncbits, const YTable& t, U2* tmp,
                     hipStream_t s);
void run_exclusive_scan(const uint64_t* in, uint64_t* out, uint64_t n, U2* tmp,
                        hipStream_t s);
void launch_chunk_end(const uint64_t* row_prod, const uint64_t* row_act, uint64_t R,
                      uint64_t r0, uint64_t rmax, uint64_t cap, uint64_t* out,
                      hipStream_t s);
void launch_max_row(const uint64_t* row_prod, uint64_t R, uint64_t* out, hipStream_t s);
void launch_expand(const ExpandArgs& a, int rb, hipStream_t s);
// Reduce counts: tmp holds reduce_tmp_count(n) entries (tile sums, total,
// block partials); the scanned tile sums are the emit's tile offsets.
void run_reduce_count(const uint64_t* keys, uint64_t n, U2* tmp, hipStream_t s);
void run_reduce_emit(const uint64_t* keys, const float* vals, uint64_t n, uint32_t* out_row,
                     uint64_t* out_col, float* out_val, uint64_t r0, int ynf,
                     uint64_t fmask, const uint64_t* ytab, const U2* tmp, hipStream_t s);
// NarrowRows::tile_row for a chunk of rstart[nrows] products: 2 entries per
// wave tile, 2 * (narrow_tile_count(n) + 1) in all.
void launch_tile_rows(const uint64_t* rstart, uint64_t nrows, uint64_t n, uint32_t* tile_row,
                      hipStream_t s);
// Narrow-form count and emit (outputs' rows are r0 + chunk-local row): tmp
// holds narrow_tmp_count(n) entries, the total at tmp[narrow_tile_count(n)];
// columns and values 16-byte aligned.
// scan = false leaves the per-tile counts in tmp unscanned (more counts are
// added to them before scan_narrow_tiles).
void run_reduce_count_narrow(const NarrowRows& nr, uint64_t n, U2* tmp, hipStream_t s,
                             bool scan = true);
void scan_narrow_tiles(uint64_t n, U2* tmp, hipStream_t s);
// ytab32 (may be NULL): the same table as 32-bit words, used when set.
void run_reduce_emit_narrow(const NarrowRows& nr, const float* vals, uint64_t n,
                            uint32_t* out_row, uint64_t* out_col, float* out_val, uint64_t r0,
                            const uint64_t* ytab, const uint32_t* ytab32, const U2* tmp,
                            hipStream_t s);
void launch_yfree_extract(const uint64_t* ykeys, uint64_t ny, uint64_t fmask, uint64_t* out,
                          hipStream_t s);
void run_unique(const uint64_t* sorted, uint64_t n, uint64_t* out, U2* tmp, hipStream_t s);
void launch_tuple_key(const TupleArgs& a, uint64_t n, uint64_t* out, hipStream_t s);
// *out = max(*out, largest of the n indices); *out cleared by the caller.
void launch_mode_max(const void* idx, int bytes, uint64_t n, uint64_t* out, hipStream_t s);
// ycol[i] = ykeys[i] & fmask (a column field of at most 32 bits).
void launch_ycol(const uint64_t* ykeys, uint64_t ny, uint64_t fmask, uint32_t* ycol,
                 hipStream_t s);
void launch_yrank(const uint64_t* ykeys, uint64_t ny, int ynf, uint64_t fmask,
                  const uint64_t* ytab, uint64_t U, int yrb, uint64_t* out, hipStream_t s);
void launch_pack_keys(const uint32_t* row, const uint64_t* col, const uint64_t* row_key,
                      uint64_t n, int ynf, uint64_t* out, hipStream_t s);
void launch_gather_rowkey(const uint32_t* row, const uint64_t* row_key, uint64_t n,
                          uint64_t* out, hipStream_t s);

}  // namespace sptc
