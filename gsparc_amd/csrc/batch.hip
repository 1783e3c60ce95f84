// Batched (multi-contraction) kernels for gfx950; see batch.hpp.
#include "batch.hpp"
#include "dense_core.hpp"
#include "reduce_core.hpp"

#include <algorithm>
#include <stdexcept>

namespace sptc {

namespace {

// ---- batched linearize -------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_blinearize(const BLinArgs* __restrict__ args,
                                                       const uint64_t* __restrict__ seg_off,
                                                       uint32_t nseg, uint64_t n,
                                                       uint64_t* __restrict__ keys,
                                                       float* __restrict__ vals) {
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; i < n; i += stride) {
    const uint32_t sgm = find_seg(seg_off, nseg, i);
    const BLinArgs& a = args[sgm];
    const uint64_t li = i - seg_off[sgm];
    if (a.skeys) {
      keys[i] = a.skeys[li];
      vals[i] = a.svals[li];
      continue;
    }
    uint64_t k = 0;
    for (int m = 0; m < a.order; ++m) {
      const uint64_t x = a.idx_bytes == 4 ? (uint64_t) static_cast<const uint32_t*>(a.idx[m])[li]
                                          : static_cast<const uint64_t*>(a.idx[m])[li];
      k |= shl64(x & low_mask(a.width[m]), a.shift[m]);
    }
    keys[i] = k;
    vals[i] = a.val[li];
  }
}

// ---- Y free-key rank compression --------------------------------------------
__global__ __launch_bounds__(kBlock) void k_bfree_extract(const uint64_t* __restrict__ ykeys,
                                                          const uint64_t* __restrict__ yoff,
                                                          uint32_t B,
                                                          const uint64_t* __restrict__ fmask,
                                                          uint64_t n, uint64_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; i < n; i += stride) out[i] = ykeys[i] & fmask[find_seg(yoff, B, i)];
}

struct BUniqueOp {
  const uint64_t* keys;
  const uint64_t* seg_off;
  uint32_t nseg;
  uint64_t* table;
  uint64_t* seg_first;
  __device__ U2 value(uint64_t i) const {
    const uint32_t s = find_seg(seg_off, nseg, i);
    return mk2(i == seg_off[s] || keys[i] != keys[i - 1], 0);
  }
  __device__ void emit(uint64_t i, U2 excl, U2 v) const {
    const uint32_t s = find_seg(seg_off, nseg, i);
    if (i == seg_off[s]) seg_first[s] = excl.a;
    if (v.a) table[excl.a] = keys[i];
    if (i == seg_off[nseg] - 1) seg_first[nseg] = excl.a + v.a;
  }
};

__global__ __launch_bounds__(kBlock) void k_brank(const uint64_t* __restrict__ ykeys,
                                                  const uint64_t* __restrict__ yoff, uint32_t B,
                                                  const int* __restrict__ ynf,
                                                  const int* __restrict__ kyf,
                                                  const uint64_t* __restrict__ ytab,
                                                  const uint64_t* __restrict__ yt_off, uint64_t n,
                                                  uint64_t* __restrict__ out,
                                                  uint32_t* __restrict__ ycol) {
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; i < n; i += stride) {
    const uint32_t s = find_seg(yoff, B, i);
    const uint64_t k = ykeys[i];
    const uint64_t f = k & low_mask(ynf[s]);
    uint64_t lo = yt_off[s], hi = yt_off[s + 1];
    const uint64_t base = lo;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (ytab[mid] < f) lo = mid + 1; else hi = mid;
    }
    out[i] = shl64(shr64(k, ynf[s]), kyf[s]) | (lo - base);
    ycol[i] = (uint32_t)(lo - base);  // the column field (ranks stay below 2^32)
  }
}

struct BContractOp {
  const uint64_t* ykeys;
  const uint64_t* yoff;
  uint32_t B;
  const int* kyf;
  uint64_t n;
  uint64_t* ukey;
  uint64_t* ustart;
  uint64_t* u_off;
  __device__ U2 value(uint64_t i) const {
    const uint32_t s = find_seg(yoff, B, i);
    return mk2(i == yoff[s] || shr64(ykeys[i], kyf[s]) != shr64(ykeys[i - 1], kyf[s]), 0);
  }
  __device__ void emit(uint64_t i, U2 excl, U2 v) const {
    const uint32_t s = find_seg(yoff, B, i);
    if (i == yoff[s]) u_off[s] = excl.a;
    if (v.a) {
      ukey[excl.a] = shr64(ykeys[i], kyf[s]);
      ustart[excl.a] = i;
    }
    if (i == n - 1) {
      ustart[excl.a + v.a] = n;
      u_off[B] = excl.a + v.a;
    }
  }
};

// value: a = match count; b = (active << 32) | row_head
struct BMatchOp {
  const uint64_t* xkeys;
  const float* xvals;
  const uint64_t* xoff;
  uint32_t B;
  const BInst* inst;
  const uint64_t* ukey;
  const uint64_t* ustart;
  const uint64_t* u_off;  // [B+1]
  BMatchOut o;

  __device__ void lookup(uint32_t s, uint64_t c, uint64_t& pos, uint64_t& cnt) const {
    uint64_t lo = u_off[s], hi = u_off[s + 1];
    const uint64_t end = hi;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (ukey[mid] < c) lo = mid + 1; else hi = mid;
    }
    if (lo < end && ukey[lo] == c) {
      pos = ustart[lo];
      cnt = ustart[lo + 1] - pos;
    } else {
      pos = 0;
      cnt = 0;
    }
  }
  // the item's instance and Y segment start, from value() to emit() in
  // registers (both are searches)
  struct Aux {
    uint64_t pos;
    uint32_t s;
  };
  __device__ U2 value(uint64_t i, Aux& x) const {
    const uint32_t s = find_seg(xoff, B, i);
    const BInst in = inst[s];
    const uint64_t k = xkeys[i];
    uint64_t pos, cnt;
    lookup(s, k & in.cmask, pos, cnt);
    x.pos = pos;
    x.s = s;
    const uint64_t head =
        (i == xoff[s]) || (shr64(xkeys[i - 1], in.ncbits) != shr64(k, in.ncbits));
    return mk2(cnt, ((uint64_t)(cnt != 0) << 32) | head);
  }
  __device__ void emit(uint64_t i, U2 excl, U2 v, const Aux& x) const {
    if (*o.bad) return;  // the row table is part of the planned results
    const uint32_t s = x.s;
    const uint64_t pos = x.pos, cnt = v.a;
    const uint64_t head = v.b & 1;
    const uint64_t row = (excl.b & 0xFFFFFFFFull) + head - 1;
    if (i == xoff[s]) {
      o.inst_P[s] = excl.a;
      o.inst_A[s] = excl.b >> 32;
      o.inst_R[s] = excl.b & 0xFFFFFFFFull;
    }
    if (head) {
      o.row_key[row] = shr64(xkeys[i], inst[s].ncbits);
      o.row_prod[row] = excl.a;
      o.row_act[row] = excl.b >> 32;
    }
    if (cnt) {
      const uint64_t j = excl.b >> 32;
      o.act_prefix[j] = excl.a;
      o.act_ypos[j] = pos;
      o.act_xval[j] = xvals[i];
      o.act_row[j] = (uint32_t)row;
    }
    if (i == xoff[B] - 1) {  // totals close every table
      const uint64_t P = excl.a + cnt;
      const uint64_t A = (excl.b >> 32) + (cnt != 0);
      const uint64_t R = (excl.b & 0xFFFFFFFFull) + head;
      o.inst_P[B] = P;
      o.inst_A[B] = A;
      o.inst_R[B] = R;
      o.act_prefix[A] = P;
      o.row_prod[R] = P;
      o.row_act[R] = A;
    }
  }
};

// ---- batched ESC expand ----------------------------------------------------
// One block per tile of the chunk's tile table (grid-stride).  The tile's
// owners are found by one binary search over act_prefix inside its piece and
// their prefix run staged in LDS; products are lane-striped (coalesced Y
// gathers and stores), each finding its owner by an LDS binary search.
constexpr int kBExpandItems = 8;
static_assert(kBExpandTileItems == kBlock * kBExpandItems, "expand tile");

__global__ __launch_bounds__(kBlock) void k_bexpand(BExpandArgs a) {
  constexpr uint32_t TILE = kBExpandTileItems;
  __shared__ uint64_t s_pref[TILE + 1];
  __shared__ uint64_t s_ja, s_jl;
  if (*a.bad) return;
  uint64_t t0, t1, step;
  xcd_tiles(a.ntiles, a.remap, &t0, &t1, &step);
  for (uint64_t t = t0; t < t1; t += step) {
    const BETile tl = a.tiles[t];
    const BPiece pc = a.pieces[tl.piece];
    const uint64_t pb = tl.p0;
    const uint32_t cnt = tl.cnt;
    __syncthreads();
    // owners of the tile's first and last products (last j in the piece
    // with prefix <= p; waves 0 and 1): only their prefixes are staged
    if (threadIdx.x < 2 * kWave) {
      const uint64_t j =
          wave_last_le(a.act_prefix, pc.A0, pc.A1, threadIdx.x < kWave ? pb : pb + cnt - 1);
      if (threadIdx.x == 0) s_ja = j;
      if (threadIdx.x == kWave) s_jl = j;
    }
    __syncthreads();
    const uint64_t ja = s_ja;
    const uint64_t need = s_jl - ja + 1;
    const uint32_t m = (uint32_t)(need < TILE + 1 ? need : TILE + 1);
    for (uint32_t k = threadIdx.x; k < m; k += kBlock) s_pref[k] = a.act_prefix[ja + k];
    __syncthreads();
    uint32_t klo = 0, kc = ~0u;
    uint64_t rbits = 0, ybase = 0;
    float xv = 0.f;
    bool big = false;
#pragma unroll
    for (int e = 0; e < kBExpandItems; ++e) {
      const uint32_t q = (uint32_t)e * kBlock + threadIdx.x;
      if (q < cnt) {
        const uint64_t p = pb + q;
        uint32_t k = klo;  // same owner as the previous product: one LDS read
        if (k + 1 < m && s_pref[k + 1] <= p) {
          uint32_t lo = k + 1, hi = m;  // last k with s_pref[k] <= p
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_pref[mid] <= p) lo = mid + 1; else hi = mid;
          }
          k = lo - 1;
        }
        klo = k;
        const uint64_t j = ja + k;
        if (k != kc) {  // new owner: its row's key bits, Y segment and X value
          kc = k;
          const uint32_t row = a.act_row[j];
          const uint64_t lr = pc.lrow0 + (row - pc.ra);
          rbits = shl64(lr, a.K);
          big = a.row_big && a.row_big[lr] != 0;
          ybase = a.act_ypos[j] - s_pref[k];
          xv = a.act_xval[j];
        }
        const uint64_t y = ybase + p;
        const uint32_t col = a.ycol[y];
        const float v = xv * a.yvals[y];
        // streaming stores: the gathered Y arrays stay cached
        if (big) {
          __builtin_nontemporal_store(col, &a.big_col[tl.out0 + q]);
          __builtin_nontemporal_store(v, &a.big_val[tl.out0 + q]);
        } else if (a.row_big) {
          __builtin_nontemporal_store(col, &a.out_col32[tl.out0 + q]);
          __builtin_nontemporal_store(v, &a.out_vals[tl.out0 + q]);
        } else {
          __builtin_nontemporal_store(rbits | col, &a.out_keys[tl.out0 + q]);
          __builtin_nontemporal_store(v, &a.out_vals[tl.out0 + q]);
        }
      }
    }
  }
}

// ---- batched reduce emit ---------------------------------------------------
struct BReduceOp {
  const uint64_t* keys;
  const float* vals;
  uint64_t n;
  uint32_t* out_row;
  uint64_t* out_col;
  float* out_val;
  int K;
  BReduceArgs b;
  NarrowRows nr;  // narrow form (keys unused)
};

// Same run detection and compensated summation as the plain reduce; the
// chunk-local row maps back to a global row through the piece table.
__global__ __launch_bounds__(kBlock) void k_breduce_emit(BReduceOp op, const U2* tile_offs) {
  __shared__ ReduceSmem sm;
  if (*op.b.bad) return;
  const uint32_t tile = blockIdx.x;
  const uint64_t base = (uint64_t)tile * kScanTile;
  uint64_t prev_key;
  const uint32_t nout = reduce_tile(op.keys, op.vals, op.n, tile, sm, &prev_key);
  const uint64_t obase = tile_offs[tile].a;
  const uint64_t cm = low_mask(op.K);
  // the thread's gathers (row and instance maps, then instance and column
  // tables) are issued a stage at a time for all its outputs, stores last
  uint64_t key[kRedRunsPerThread], col[kRedRunsPerThread];
  uint32_t row[kRedRunsPerThread], si[kRedRunsPerThread];
#pragma unroll
  for (int q = 0; q < kRedRunsPerThread; ++q) {
    const uint32_t i = threadIdx.x + (uint32_t)q * kBlock;
    key[q] = i < nout ? sm.k[i] : 0;
    const uint64_t lrow = shr64(key[q], op.K);
    row[q] = i < nout ? op.b.map_row[lrow] : 0;
    si[q] = i < nout ? op.b.map_inst[lrow] : 0;
  }
#pragma unroll
  for (int q = 0; q < kRedRunsPerThread; ++q) {
    const uint32_t i = threadIdx.x + (uint32_t)q * kBlock;
    col[q] = 0;
    if (i < nout) {
      col[q] = op.b.ytab[op.b.inst[si[q]].yt_off + (key[q] & cm)];
      row[q] -= (uint32_t)op.b.inst_R[si[q]];
    }
  }
#pragma unroll
  for (int q = 0; q < kRedRunsPerThread; ++q) {
    const uint32_t i = threadIdx.x + (uint32_t)q * kBlock;
    if (i >= nout) continue;
    op.out_row[obase + i] = row[q];
    op.out_col[obase + i] = col[q];
    op.out_val[obase + i] = sm.v[i];
    if (!op.b.bounds) continue;
    // first output of every instance in (previous output's instance, s]
    int64_t sp = -1;
    if (i > 0) sp = op.b.map_inst[shr64(sm.k[i - 1], op.K)];
    else if (base > 0) sp = op.b.map_inst[shr64(prev_key, op.K)];
    for (int64_t t = sp + 1; t <= (int64_t)si[q]; ++t) op.b.bounds[t] = obase + i;
  }
}

// The narrow form, one wave per wave tile (the count pass's tiling).
__global__ __launch_bounds__(kBlock) void k_breduce_emit_narrow(BReduceOp op,
                                                                const U2* tile_offs,
                                                                uint64_t ntiles) {
  __shared__ WaveSmem sm[kBlock / kWave];
  if (*op.b.bad) return;
  const uint64_t wt = (uint64_t)blockIdx.x * (kBlock / kWave) + wave_id();
  if (wt >= ntiles) return;
  WaveSmem& w = sm[wave_id()];
  uint64_t r0, rp;
  bool rows;
  const uint32_t T = wave_tile_reduce(op.nr, op.vals, op.n, wt, w, &r0, &rp, &rows);
  const uint64_t obase = tile_offs[wt].a;
  const uint32_t lane = lane_id();
  uint64_t col[kScanItems];
  uint32_t row[kScanItems], si[kScanItems];
  if (!rows) {
    // one row for the whole tile: its global row, instance and column table
    // once for the wave instead of once per output
    const uint32_t s0 = op.b.map_inst[r0];
    const uint64_t yt = op.b.inst[s0].yt_off;
    const uint32_t rr = op.b.map_row[r0] - (uint32_t)op.b.inst_R[s0];
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
      const uint32_t i = lane + (uint32_t)q * kWave;
      row[q] = rr;
      si[q] = s0;
      col[q] = 0;
      if (i < T) {
        const uint64_t t = yt + w.c[i];
        col[q] = op.b.ytab32 ? (uint64_t)op.b.ytab32[t] : op.b.ytab[t];  // 32-bit: half the gather
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
      const uint32_t i = lane + (uint32_t)q * kWave;
      const uint64_t lrow = i < T ? r0 + w.m[i] : 0;
      row[q] = i < T ? op.b.map_row[lrow] : 0;
      si[q] = i < T ? op.b.map_inst[lrow] : 0;
    }
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
      const uint32_t i = lane + (uint32_t)q * kWave;
      col[q] = 0;
      if (i < T) {
        const uint64_t t = op.b.inst[si[q]].yt_off + w.c[i];
        col[q] = op.b.ytab32 ? (uint64_t)op.b.ytab32[t] : op.b.ytab[t];
        row[q] -= (uint32_t)op.b.inst_R[si[q]];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    const uint32_t i = lane + (uint32_t)q * kWave;
    if (i >= T) continue;
    op.out_row[obase + i] = row[q];
    op.out_col[obase + i] = col[q];
    op.out_val[obase + i] = w.v[i];
    if (!op.b.bounds) continue;
    if (!rows && i > 0) continue;  // the previous output has the same row and instance
    // first output of every instance in (previous output's instance, s]
    int64_t sp = -1;
    if (i > 0) sp = op.b.map_inst[r0 + (rows ? w.m[i - 1] : 0)];
    else if (wt > 0) sp = op.b.map_inst[rp];
    for (int64_t t = sp + 1; t <= (int64_t)si[q]; ++t) op.b.bounds[t] = obase + i;
  }
}

// ---- batched dense ---------------------------------------------------------
// Block b handles the b-th row of the group: group entry g (binary search in
// grp_rows), global row grp_row0[g] + (b - grp_rows[g]).
__device__ __forceinline__ void bdense_row(const BDenseArgs& a, uint64_t b, uint64_t& row,
                                           BInst& in) {
  const uint32_t g = find_seg(a.grp_rows, a.G, b);
  row = a.grp_row0[g] + (b - a.grp_rows[g]);
  in = a.inst[a.grp_inst[g]];
}

template <int EB>
__global__ __launch_bounds__(kBlock) void k_bdense_count(BDenseArgs a) {
  __shared__ DenseSmem<EB> sm;
  __shared__ uint64_t s_ja;
  if (*a.bad) return;
  uint64_t r;
  BInst in;
  bdense_row(a, blockIdx.x, r, in);
  const uint32_t t = dense_row_count<EB>(a.d, a.d.row_act[r], a.d.row_act[r + 1],
                                         low_mask(in.kyf), sm, &s_ja);
  if (threadIdx.x == 0) a.d.row_cnt[blockIdx.x] = t;
}

template <int EB, bool ORDERED>
__global__ __launch_bounds__(kBlock) void k_bdense_numeric(BDenseArgs a) {
  constexpr int E = 1 << EB;
  __shared__ double s_acc[E];
  __shared__ DenseSmem<EB> sm;
  __shared__ uint64_t s_ja;
  if (*a.bad) return;
  uint64_t r;
  BInst in;
  bdense_row(a, blockIdx.x, r, in);
  const uint32_t g = find_seg(a.grp_rows, a.G, blockIdx.x);
  // instance-local row id
  dense_row_numeric<EB, ORDERED>(a.d, a.d.row_act[r], a.d.row_act[r + 1], low_mask(in.kyf),
                        a.ytab + in.yt_off, (uint32_t)(blockIdx.x - a.grp_rows[g]),
                        a.d.out_off[blockIdx.x], s_acc, sm, &s_ja);
}

__global__ __launch_bounds__(kBlock) void k_chunk_row_maps(const BPiece* pieces,
                                                           uint32_t npieces, uint64_t lrows,
                                                           const uint64_t* inst_R, uint32_t B,
                                                           uint32_t* map_row,
                                                           uint32_t* map_inst) {
  uint64_t l = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; l < lrows; l += stride) {
    uint32_t lo = 0, hi = npieces;  // last piece with lrow0 <= l
    while (lo + 1 < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (pieces[mid].lrow0 <= l) lo = mid; else hi = mid;
    }
    const uint64_t g = pieces[lo].ra + (l - pieces[lo].lrow0);
    map_row[l] = (uint32_t)g;
    map_inst[l] = find_seg(inst_R, B, g);
  }
}

__global__ void k_dense_bounds(const uint64_t* grp_rows, const uint32_t* grp_inst, uint32_t G,
                               const uint64_t* out_off, uint64_t nnz, uint32_t B,
                               uint64_t* bounds, const uint32_t* bad) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > B || *bad) return;
  uint32_t lo = 0, hi = G;  // first group instance >= t
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (grp_inst[mid] < t) lo = mid + 1; else hi = mid;
  }
  bounds[t] = lo < G ? out_off[grp_rows[lo]] : nnz;
}

__global__ void k_fix_empty(uint64_t* arr, const uint64_t* seg_off, uint32_t nseg) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int64_t g = (int64_t)nseg - 1; g >= 0; --g)
    if (seg_off[g] == seg_off[g + 1]) arr[g] = arr[g + 1];
}

__global__ __launch_bounds__(kBlock) void k_bany_dup(const uint64_t* keys,
                                                     const uint64_t* seg_off, uint32_t nseg,
                                                     uint64_t n, uint32_t* flags) {
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x + 1;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; i < n; i += stride) {
    if (keys[i] != keys[i - 1]) continue;
    const uint32_t sg = find_seg(seg_off, nseg, i);
    if (i > seg_off[sg]) flags[sg] = 1;
  }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void k_digest(const uint64_t* keys, uint64_t n,
                                                   unsigned long long* out) {
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  uint64_t h = 0;
  for (; i < n; i += stride) h += mix64(keys[i] ^ (i * 0x9E3779B97F4A7C15ull));
  for (int off = kWave / 2; off > 0; off >>= 1) h += __shfl_xor(h, off, kWave);
  if (lane_id() == 0 && h) atomicAdd(out, (unsigned long long)h);
}

__global__ void k_check_digest(const uint64_t* got, const uint64_t* want, uint32_t* bad) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && *got != *want) *bad = 1;
}

__global__ __launch_bounds__(kBlock) void k_sum_f64(const float* v, uint64_t n, double* out) {
  __shared__ double s_red[kBlock / kWave];
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  double t = 0.0;
  for (; i < n; i += stride) t += (double)v[i];
  for (int off = kWave / 2; off > 0; off >>= 1) t += __shfl_xor(t, off, kWave);
  if (lane_id() == 0) s_red[threadIdx.x / kWave] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    double u = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) u += s_red[w];
    if (u != 0.0) atomicAdd(out, u);
  }
}

uint32_t grid_for(uint64_t n, uint32_t cap = 4096) {
  return (uint32_t)std::min<uint64_t>(ceil_div(n ? n : 1, kBlock), cap);
}

}  // namespace

// ---- host launchers --------------------------------------------------------

void launch_blinearize(const BLinArgs* args, const uint64_t* seg_off, uint32_t nseg,
                       uint64_t n, uint64_t* keys, float* vals, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_blinearize, dim3(grid_for(n)), dim3(kBlock), 0, s, args, seg_off, nseg,
                     n, keys, vals);
}

void launch_bfree_extract(const uint64_t* ykeys, const uint64_t* yoff, uint32_t B,
                          const uint64_t* fmask, uint64_t n, uint64_t* out, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_bfree_extract, dim3(grid_for(n)), dim3(kBlock), 0, s, ykeys, yoff, B,
                     fmask, n, out);
}

void run_bunique(const uint64_t* sorted, const uint64_t* seg_off, uint32_t nseg, uint64_t n,
                 uint64_t* table, uint64_t* seg_first, U2* tmp, hipStream_t s) {
  run_scan(BUniqueOp{sorted, seg_off, nseg, table, seg_first}, n, tmp, s);
}

void launch_brank(const uint64_t* ykeys, const uint64_t* yoff, uint32_t B, const int* ynf,
                  const int* kyf, const uint64_t* ytab, const uint64_t* yt_off, uint64_t n,
                  uint64_t* out, uint32_t* ycol, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_brank, dim3(grid_for(n)), dim3(kBlock), 0, s, ykeys, yoff, B, ynf, kyf,
                     ytab, yt_off, n, out, ycol);
}

void run_bycontract(const uint64_t* ykeys, const uint64_t* yoff, uint32_t B, const int* kyf,
                    uint64_t n, uint64_t* ukey, uint64_t* ustart, uint64_t* u_off, U2* tmp,
                    hipStream_t s) {
  run_scan(BContractOp{ykeys, yoff, B, kyf, n, ukey, ustart, u_off}, n, tmp, s);
}

void run_bmatch(const uint64_t* xkeys, const float* xvals, const uint64_t* xoff, uint32_t B,
                const BInst* inst, const uint64_t* ukey, const uint64_t* ustart,
                const uint64_t* u_off, uint64_t nx, const BMatchOut& o, U2* tmp, hipStream_t s) {
  run_scan(BMatchOp{xkeys, xvals, xoff, B, inst, ukey, ustart, u_off, o}, nx, tmp, s);
}

void launch_bexpand(const BExpandArgs& a, hipStream_t s) {
  if (!a.ntiles) return;
  const uint32_t grid = std::min<uint32_t>(a.ntiles, 4096);
  BExpandArgs b = a;
  b.remap = xcd_remap_enabled();
  hipLaunchKernelGGL(k_bexpand, dim3(grid), dim3(kBlock), 0, s, b);
}

void run_breduce_emit(const uint64_t* keys, const float* vals, uint64_t n, uint32_t* out_row,
                      uint64_t* out_col, float* out_val, int K, const BReduceArgs& b,
                      const U2* tmp, hipStream_t s) {
  if (!n) return;
  BReduceOp op{keys, vals, n, out_row, out_col, out_val, K, b, NarrowRows{}};
  hipLaunchKernelGGL(k_breduce_emit, dim3((uint32_t)scan_tmp_count(n)), dim3(kBlock), 0, s, op,
                     tmp);
}

void run_breduce_emit_narrow(const NarrowRows& nr, const float* vals, uint64_t n,
                             uint32_t* out_row, uint64_t* out_col, float* out_val,
                             const BReduceArgs& b, const U2* tmp, hipStream_t s) {
  if (!n) return;
  if ((reinterpret_cast<uintptr_t>(nr.cols) | reinterpret_cast<uintptr_t>(vals)) & 15)
    throw std::logic_error("narrow products must be 16-byte aligned");
  BReduceOp op{nullptr, vals, n, out_row, out_col, out_val, 32, b, nr};
  const uint64_t nt = narrow_tile_count(n);
  hipLaunchKernelGGL(k_breduce_emit_narrow, dim3((uint32_t)ceil_div(nt, kBlock / kWave)),
                     dim3(kBlock), 0, s, op, tmp, nt);
}

void launch_bdense_count(const BDenseArgs& a, int eb, uint64_t rows, hipStream_t s) {
  if (!rows) return;
  switch (eb) {
    case 8: hipLaunchKernelGGL(k_bdense_count<8>, dim3((uint32_t)rows), dim3(kBlock), 0, s, a); break;
    case 10: hipLaunchKernelGGL(k_bdense_count<10>, dim3((uint32_t)rows), dim3(kBlock), 0, s, a); break;
    default: hipLaunchKernelGGL(k_bdense_count<13>, dim3((uint32_t)rows), dim3(kBlock), 0, s, a); break;
  }
}

void launch_bdense_numeric(const BDenseArgs& a, int eb, uint64_t rows, hipStream_t s) {
  if (!rows) return;
  const dim3 g((uint32_t)rows), b(kBlock);
  if (dense_ordered()) {
    switch (eb) {
      case 8: hipLaunchKernelGGL((k_bdense_numeric<8, true>), g, b, 0, s, a); break;
      case 10: hipLaunchKernelGGL((k_bdense_numeric<10, true>), g, b, 0, s, a); break;
      default: hipLaunchKernelGGL((k_bdense_numeric<13, true>), g, b, 0, s, a); break;
    }
  } else {
    switch (eb) {
      case 8: hipLaunchKernelGGL((k_bdense_numeric<8, false>), g, b, 0, s, a); break;
      case 10: hipLaunchKernelGGL((k_bdense_numeric<10, false>), g, b, 0, s, a); break;
      default: hipLaunchKernelGGL((k_bdense_numeric<13, false>), g, b, 0, s, a); break;
    }
  }
}

void launch_chunk_row_maps(const BPiece* pieces, uint32_t npieces, uint64_t lrows,
                           const uint64_t* inst_R, uint32_t B, uint32_t* map_row,
                           uint32_t* map_inst, hipStream_t s) {
  if (!lrows) return;
  hipLaunchKernelGGL(k_chunk_row_maps, dim3(grid_for(lrows)), dim3(kBlock), 0, s, pieces,
                     npieces, lrows, inst_R, B, map_row, map_inst);
}

void launch_dense_bounds(const uint64_t* grp_rows, const uint32_t* grp_inst, uint32_t G,
                         const uint64_t* out_off, uint64_t nnz, uint32_t B, uint64_t* bounds,
                         const uint32_t* bad, hipStream_t s) {
  hipLaunchKernelGGL(k_dense_bounds, dim3(ceil_div(B + 1, 256)), dim3(256), 0, s, grp_rows,
                     grp_inst, G, out_off, nnz, B, bounds, bad);
}

void launch_fix_empty(uint64_t* arr, const uint64_t* seg_off, uint32_t nseg, hipStream_t s) {
  hipLaunchKernelGGL(k_fix_empty, dim3(1), dim3(64), 0, s, arr, seg_off, nseg);
}

void launch_bany_dup(const uint64_t* keys, const uint64_t* seg_off, uint32_t nseg, uint64_t n,
                     uint32_t* flags, hipStream_t s) {
  fill_u32(flags, 0, nseg, s);
  if (n < 2) return;
  hipLaunchKernelGGL(k_bany_dup, dim3(grid_for(n)), dim3(kBlock), 0, s, keys, seg_off, nseg, n,
                     flags);
}

void launch_digest(const uint64_t* keys, uint64_t n, uint64_t* out, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_digest, dim3(grid_for(n, 2048)), dim3(kBlock), 0, s, keys, n,
                     reinterpret_cast<unsigned long long*>(out));
}

void launch_check_digest(const uint64_t* got, const uint64_t* want, uint32_t* bad,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_check_digest, dim3(1), dim3(64), 0, s, got, want, bad);
}

void launch_sum_f64(const float* vals, uint64_t n, double* out, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_sum_f64, dim3(grid_for(n, 1024)), dim3(kBlock), 0, s, vals, n, out);
}

}  // namespace sptc
