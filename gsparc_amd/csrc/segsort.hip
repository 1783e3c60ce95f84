// Segmented LSD radix sort kernels (gfx950); see segsort.hpp.
#include "segsort.hpp"
#include "sort_core.hpp"

#include <algorithm>
#include <cstdlib>
#include <stdexcept>

namespace sptc {

namespace {

// ---- segmented reduce-then-scan radix pass (8-bit digits, 512 x 12 tiles) --

template <typename KT, int RB, int IPT>
__global__ __launch_bounds__(kSegNT) void k_seg_upsweep(const KT* __restrict__ kin, int shift,
                                                        const SegTile* __restrict__ tiles,
                                                        uint32_t* __restrict__ cnt) {
  constexpr int R = 1 << RB;
  __shared__ uint32_t sh[R];
  for (int i = threadIdx.x; i < R; i += kSegNT) sh[i] = 0;
  __syncthreads();
  const SegTile t = tiles[blockIdx.x];
  const KT* __restrict__ kb = kin + t.start;  // uniform base, 32-bit offsets
  // all the thread's loads first (IPT in flight), then one LDS atomic per key
  KT k[IPT];
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const uint32_t j = (uint32_t)i * kSegNT + threadIdx.x;
    k[i] = j < t.count ? kb[j] : 0;
  }
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const uint32_t j = (uint32_t)i * kSegNT + threadIdx.x;
    if (j < t.count) atomicAdd(&sh[(uint32_t)(k[i] >> shift) & (R - 1)], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < R; d += kSegNT) cnt[(uint64_t)blockIdx.x * R + d] = sh[d];
}

// KI -> KO: 64 -> 64 (operand sorts), 32 -> 32 (column passes of big rows).
// NT threads x IPT keys per tile (NT * IPT = the plan's tile size).
template <typename KI, typename KO, int RB, int IPT, int NT = kSegNT>
__global__ __launch_bounds__(NT) void k_seg_downsweep(
    const KI* __restrict__ kin, const float* __restrict__ vin, KO* __restrict__ kout,
    float* __restrict__ vout, int shift, const SegTile* __restrict__ tiles,
    const uint32_t* __restrict__ offs, const uint32_t* __restrict__ dstarts, uint32_t remap) {
  constexpr int R = 1 << RB;
  __shared__ TileSmem<RB, IPT, NT, KI> sm;
  const uint32_t tid = threadIdx.x;
  const SegTile t = tiles[remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x];
  for (int i = tid; i < (NT / kWave) * R; i += NT) (&sm.cnt[0][0])[i] = 0;
  for (int d = tid; d < R; d += NT)
    sm.gbase[d] = offs[t.tile * R + d] + dstarts[(uint64_t)t.first * R + d];
  __syncthreads();
  const bool full = t.count == NT * IPT;
  KI k[IPT];
  float v[IPT];
  uint32_t pos[IPT];
  const float* vt = vin ? vin + t.start : nullptr;  // NULL values: keys only
  if (full)
    tile_load_rank<true, RB, IPT, NT>(sm, kin + t.start, vt, t.count, shift, k, v, pos);
  else
    tile_load_rank<false, RB, IPT, NT>(sm, kin + t.start, vt, t.count, shift, k, v, pos);
  __syncthreads();
  tile_prefix_fold<RB, IPT, NT>(sm);
  if (full)
    tile_stage_scatter<true, RB, IPT, NT, KI, KI, KO, true>(sm, t.count, shift, k, v, pos, kout,
                                                           vout);
  else
    tile_stage_scatter<false, RB, IPT, NT, KI, KI, KO, true>(sm, t.count, shift, k, v, pos, kout,
                                                            vout);
}

// Per segment (one workgroup, thread d = digit d): off[t][d] = digit d's
// keys in the segment's tiles before t (one coalesced read and write per
// tile row), then cnt[first][d] = start + keys of smaller digits in the
// segment (the first tile's counts are consumed by then).
template <int RB>
__global__ __launch_bounds__(1 << RB) void k_seg_scan(const SegDesc* __restrict__ segs,
                                                     uint32_t* __restrict__ cnt,
                                                     uint32_t* __restrict__ off) {
  constexpr int R = 1 << RB;
  __shared__ uint32_t scratch[R / kWave];
  const SegDesc g = segs[blockIdx.x];
  const uint32_t d = threadIdx.x;
  const uint64_t t0 = g.first_tile, nt = g.ntiles;
  uint32_t run = 0;
  uint64_t t = 0;
  for (; t + 8 <= nt; t += 8) {  // eight loads in flight
    uint32_t c[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) c[e] = cnt[(t0 + t + e) * R + d];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      off[(t0 + t + e) * R + d] = run;
      run += c[e];
    }
  }
  for (; t < nt; ++t) {
    const uint32_t c = cnt[(t0 + t) * R + d];
    off[(t0 + t) * R + d] = run;
    run += c;
  }
  uint32_t all;
  const uint32_t base = (uint32_t)g.start + block_scan_u32<R>(run, scratch, &all);
  cnt[t0 * R + d] = base;
}

// ---- row segments: small-row groups (LDS) and big rows (multi-tile) --------
constexpr uint64_t kHalfTile = kSegTile / 2;

__device__ __forceinline__ uint64_t rowseg_locate(const RowRange* ranges, uint32_t nranges,
                                                  uint64_t l, const uint64_t* row_prod,
                                                  uint64_t* first, uint64_t* grow = nullptr) {
  uint32_t lo = 0, hi = nranges;  // last range with lrow0 <= l
  while (lo + 1 < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ranges[mid].lrow0 <= l) lo = mid; else hi = mid;
  }
  const RowRange rr = ranges[lo];
  const uint64_t g = rr.ra + (l - rr.lrow0);
  *first = rr.out0 + (row_prod[g] - row_prod[rr.ra]);
  if (grow) *grow = g;
  return row_prod[g + 1] - row_prod[g];
}

// Per row: products in small-row groups (0 for big rows), big-row tile
// counts, and for a row above half a tile with one owner -- its products are
// that owner's Y segment, already in column order -- the number of
// kSegTile-product pieces it is cut into (group items the group sort finds
// in order; row_act NULL: no such rows).
__global__ __launch_bounds__(kBlock) void k_rowseg_sizes(const RowRange* ranges, uint32_t nr,
                                                         uint64_t lrows, const uint64_t* row_prod,
                                                         const uint64_t* row_act, uint64_t tile,
                                                         uint64_t* small_n, uint64_t* big_t,
                                                         uint64_t* big_f, uint64_t* rstart,
                                                         uint64_t* pre_p) {
  uint64_t l = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; l < lrows; l += stride) {
    uint64_t first, g;
    const uint64_t n = rowseg_locate(ranges, nr, l, row_prod, &first, &g);
    const bool one = row_act && row_act[g + 1] - row_act[g] == 1;
    const bool pre = n > kHalfTile && one;
    const bool big = n > kHalfTile && !one;
    small_n[l] = big ? 0 : n;
    big_t[l] = big ? ceil_div(n, tile) : 0;
    big_f[l] = big ? 1 : 0;
    pre_p[l] = pre ? ceil_div(n, kSegTile) : 0;
    rstart[l] = first;
    if (l + 1 == lrows) rstart[lrows] = first + n;
  }
}

// Per row: the group items it starts -- 1 when a small row starts a group
// (after a big or pieced row, or where the running small-row total enters a
// new half-tile window), a pieced row's piece count.
__global__ __launch_bounds__(kBlock) void k_rowseg_heads(uint64_t lrows, const uint64_t* small_n,
                                                         const uint64_t* big_t,
                                                         const uint64_t* pre_p,
                                                         const uint64_t* ps, uint64_t* head) {
  uint64_t l = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; l < lrows; l += stride) {
    uint64_t h = 0;
    if (pre_p[l])
      h = pre_p[l];
    else if (!big_t[l])
      h = l == 0 || big_t[l - 1] != 0 || pre_p[l - 1] != 0 ||
          ps[l] / kHalfTile != ps[l - 1] / kHalfTile;
    head[l] = h;
  }
}

__global__ __launch_bounds__(kBlock) void k_rowseg_emit(
    const RowRange* ranges, uint32_t nr, uint64_t lrows, const uint64_t* row_prod, int rb,
    uint64_t tile,
    const uint64_t* big_t, const uint64_t* bt, const uint64_t* bs, const uint64_t* head,
    const uint64_t* hs, const uint64_t* pre_p, LocalSeg* local, SegTile* big, SegDesc* bigseg,
    uint32_t* bigrow) {
  uint64_t l = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; l < lrows; l += stride) {
    uint64_t first;
    const uint64_t n = rowseg_locate(ranges, nr, l, row_prod, &first);
    const uint64_t nt = big_t[l];
    if (nt) {
      const uint64_t o = bt[l];
      SegDesc g;
      g.start = first;
      g.first_tile = (uint32_t)o;
      g.ntiles = (uint32_t)nt;
      bigseg[bs[l]] = g;
      if (bigrow) bigrow[bs[l]] = (uint32_t)l;
      for (uint64_t t = 0; t < nt; ++t) {
        SegTile d;
        d.start = first + t * tile;
        d.count = (uint32_t)(n - t * tile < tile ? n - t * tile : tile);
        d.tile = o + t;
        d.first = (uint32_t)o;
        big[o + t] = d;
      }
    } else if (pre_p[l]) {
      for (uint64_t t = 0; t < pre_p[l]; ++t) {  // pieces of a row already in order
        LocalSeg q;
        q.start = first + t * kSegTile;
        q.count = (uint32_t)(n - t * kSegTile < kSegTile ? n - t * kSegTile : kSegTile);
        q.row0 = (uint32_t)l;
        local[hs[l] + t] = q;
      }
    } else {
      // the group holding row l: its head is the last head at or before l
      const uint64_t gidx = hs[l] + head[l] - 1;
      if (head[l]) {
        local[gidx].start = first;
        local[gidx].row0 = (uint32_t)l;
      }
      if (n) atomicAdd(reinterpret_cast<unsigned int*>(&local[gidx].count), (unsigned int)n);
    }
  }
}

__global__ void k_zero_local(LocalSeg* local, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    local[i].count = 0;
    local[i].row0 = 0;
  }
}

// One workgroup per small group of (32-bit column, value) products of the
// consecutive rows row0..: each product's row rank within the group comes
// from the row starts (marks at the starts, then a prefix count over the
// wave-striped layout), the composite key (row rank << b | column, b the
// bits the group's columns use)
// is sorted on the bits that vary inside the group entirely in LDS (stable
// LSD passes), and columns and values are stored back in place.
// KT/RB: 32-bit keys with 10-bit digits when (row in group, column) fits 32
// bits (fewer passes, half the key staging), else 64-bit keys, 8-bit digits.
template <typename KT, int RB>
// moved (may be NULL): gets the products of the groups that are sorted.
__global__ __launch_bounds__(kSegNT) void k_local_sort(uint32_t* cols, float* vals,
                                                       const LocalSeg* __restrict__ segs,
                                                       const uint64_t* __restrict__ rstart,
                                                       uint64_t lrows,
                                                       unsigned long long* moved) {
  constexpr int R = 1 << RB;
  __shared__ TileSmem<RB, kSegIPT, kSegNT, KT> sm;
  __shared__ float vbuf[kSegTile];
  __shared__ uint64_t s_or[kSegNT / kWave];
  __shared__ uint32_t s_wmax[kSegNT / kWave];  // row starts per wave
  __shared__ uint32_t s_cor[kSegNT / kWave];   // OR of the wave's columns
  __shared__ uint64_t s_k0;
  __shared__ KT s_last[kSegNT / kWave];  // each wave's last key
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t wave = tid / kWave;
  const LocalSeg sg = segs[blockIdx.x];
  const uint64_t n = sg.count;
  const uint64_t wbase = (uint64_t)wave * kWave * kSegIPT;  // tile-local
  // mark[p] = 1 where a row starts inside the group (p > 0); a product's
  // row key is the number of such starts at or before it -- the rank of
  // its row among the group's non-empty rows (< 6144, 13 bits), which
  // orders like the row itself
  uint32_t* mark = reinterpret_cast<uint32_t*>(sm.buf);
  for (uint32_t i = tid; i < n; i += kSegNT) mark[i] = 0;
  __syncthreads();
  for (uint64_t r = (uint64_t)sg.row0 + 1 + tid; r < lrows; r += kSegNT) {
    const uint64_t st = rstart[r];
    if (st >= sg.start + n) break;
    if (st > sg.start) mark[st - sg.start] = 1;  // empty rows share the next row's start
  }
  __syncthreads();
  KT k[kSegIPT];
  float v[kSegIPT];
  uint32_t pos[kSegIPT];
  uint32_t carry = 0;
#pragma unroll
  for (int i = 0; i < kSegIPT; ++i) {
    const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    pos[i] = carry + wave_incl_scan32(idx < n ? mark[idx] : 0u);  // starts so far (this wave)
    carry = __shfl(pos[i], kWave - 1, kWave);
  }
  // the group's columns, and the bits they use: the row rank goes right
  // above those (a chunk's column field is as wide as its widest instance)
  uint32_t cv[kSegIPT];
  uint32_t cor = 0;
#pragma unroll
  for (int i = 0; i < kSegIPT; ++i) {
    const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    cv[i] = idx < n ? cols[sg.start + idx] : 0u;
    cor |= cv[i];
  }
  for (int off = kWave / 2; off > 0; off >>= 1) cor |= __shfl_xor(cor, off, kWave);
  if (lane == 0) {
    s_wmax[wave] = carry;
    s_cor[wave] = cor;
  }
  __syncthreads();
  uint32_t before = 0;
  cor = 0;
  for (uint32_t w = 0; w < kSegNT / kWave; ++w) {
    if (w < wave) before += s_wmax[w];
    cor |= s_cor[w];
  }
  const int cb = cor ? 32 - __builtin_clz(cor) : 0;  // at most the chunk's column bits
  const uint64_t cmask = low_mask(cb);
#pragma unroll
  for (int i = 0; i < kSegIPT; ++i) {
    const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    const bool valid = idx < n;
    const uint64_t row = pos[i] + before;
    k[i] = valid ? (KT)(shl64(row, cb) | cv[i]) : (KT)0;
  }
  if (lane == kWave - 1) s_last[wave] = k[kSegIPT - 1];
  if (tid == 0) s_k0 = (uint64_t)k[0];
  __syncthreads();  // marks read, sm.buf free for the sort
  // a group already in (row, column) order -- a row whose products are one
  // owner's Y segment is -- stays as it is: no sort, no value loads, no
  // stores.  The key before item i of lane 0 is lane 63's of item i - 1
  // (or the previous wave's last key, that wave being full).
  uint32_t desc = 0;
#pragma unroll
  for (int i = 0; i < kSegIPT; ++i) {
    const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    const KT up = __shfl_up(k[i], 1, kWave);
    const KT p63 = i > 0 ? (KT)__shfl(k[i > 0 ? i - 1 : 0], kWave - 1, kWave)
                         : (wave > 0 ? s_last[wave > 0 ? wave - 1 : 0] : (KT)0);
    const KT prev = lane ? up : p63;
    if (idx < n && idx > 0 && k[i] < prev) desc = 1;
  }
  if (!__syncthreads_or(desc)) return;
  if (moved && tid == 0) atomicAdd(moved, (unsigned long long)n);
#pragma unroll
  for (int i = 0; i < kSegIPT; ++i) {
    const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    v[i] = idx < n && vals ? vals[sg.start + idx] : 0.f;  // vals NULL: columns only
  }
  const uint64_t k0 = s_k0;
  uint64_t vary = 0;
#pragma unroll
  for (int i = 0; i < kSegIPT; ++i) {
    const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    if (idx < n) vary |= (uint64_t)k[i] ^ k0;
  }
  for (int off = kWave / 2; off > 0; off >>= 1) vary |= __shfl_xor(vary, off, kWave);
  if (lane == 0) s_or[wave] = vary;
  __syncthreads();
  vary = 0;
#pragma unroll
  for (int w = 0; w < kSegNT / kWave; ++w) vary |= s_or[w];
  const int passes = vary ? (64 - __builtin_clzll(vary) + RB - 1) / RB : 0;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * RB;
    for (int i = tid; i < (kSegNT / kWave) * R; i += kSegNT) (&sm.cnt[0][0])[i] = 0;
    __syncthreads();
    tile_rank<false, RB, kSegIPT, kSegNT>(sm, (uint32_t)n, shift, k, pos);
    __syncthreads();
    tile_prefix_fold<RB, kSegIPT, kSegNT>(sm);  // cnt[w][d] = run start + wave prefix
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSegIPT; ++i) {
      const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
      if (idx < n) {
        const uint32_t d = (uint32_t)(k[i] >> shift) & (R - 1);
        const uint32_t q = pos[i] + (uint32_t)sm.cnt[wave][d];
        sm.buf[q] = k[i];
        vbuf[q] = v[i];
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSegIPT; ++i) {
      const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
      if (idx < n) {
        k[i] = sm.buf[idx];
        v[i] = vbuf[idx];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < kSegIPT; ++i) {
    const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    if (idx < n) {
      cols[sg.start + idx] = (uint32_t)(k[i] & cmask);
      if (vals) vals[sg.start + idx] = v[i];
    }
  }
}

// Expansion by sort work item (row-segmented sort): items are the big-row
// tiles, then the small-row groups; an item's products are consecutive and
// are taken in sub-tiles of kExpandSub like the plain expansion (one global
// search for the sub-tile's first owner, owner prefixes staged in LDS, a
// product's owner found from its previous one).  Products are written
// narrow to the big rows' first-pass set or to (out_col32, out_vals); a big
// tile also counts its first-pass digits into its count row, which takes
// the place of that pass's upsweep.
constexpr uint32_t kExpandSub = kBlock * 8;

template <int RB>
__global__ __launch_bounds__(kBlock) void k_expand_items(ExpandArgs a,
                                                         const SegTile* __restrict__ big,
                                                         uint32_t nbig,
                                                         const LocalSeg* __restrict__ local,
                                                         uint32_t nlocal,
                                                         uint32_t* __restrict__ cnt) {
  constexpr int R = 1 << RB;
  __shared__ uint64_t s_pref[kExpandSub + 1];
  __shared__ uint32_t sh[R];
  __shared__ uint64_t s_ja, s_jl;
  __shared__ double s_sum[kBlock / kWave];
  double vsum = 0.0;
  uint64_t t0, t1, step;
  xcd_tiles((uint64_t)nbig + nlocal, a.remap, &t0, &t1, &step);
  for (uint64_t it = t0; it < t1; it += step) {
    const bool isbig = it < nbig;
    uint64_t start;
    uint32_t count;
    if (isbig) {
      start = big[it].start;
      count = big[it].count;
    } else {
      start = local[it - nbig].start;
      count = local[it - nbig].count;
    }
    uint32_t* ocol = isbig ? a.big_col : a.out_col32;
    float* oval = isbig ? a.big_val : a.out_vals;
    if (isbig)
      for (int i = threadIdx.x; i < R; i += kBlock) sh[i] = 0;
    uint64_t jlo = a.A0;  // owners only move forward within an item
    // windows start at multiples of 64 products, so each wave's stores are
    // whole 256-byte runs; the first window's products before the item are
    // masked off
    const uint64_t iend = start + count;
    for (uint64_t sub = count ? start & ~63ull : iend; sub < iend; sub += kExpandSub) {
      const uint64_t pb = a.P0 + sub;                          // window start
      const uint64_t p0 = a.P0 + (sub > start ? sub : start);  // its first product
      const uint64_t pe = a.P0 + (iend - sub < kExpandSub ? iend : sub + kExpandSub);
      __syncthreads();
      // owners of the first and the last product (waves 0 and 1): only
      // their prefixes are staged
      if (threadIdx.x < 2 * kWave) {
        const uint64_t j = wave_last_le(a.act_prefix, jlo, a.A1, threadIdx.x < kWave ? p0 : pe - 1);
        if (threadIdx.x == 0) s_ja = j;
        if (threadIdx.x == kWave) s_jl = j;
      }
      __syncthreads();
      const uint64_t ja = s_ja;
      jlo = ja;
      const uint64_t need = s_jl - ja + 1;
      const uint32_t m = (uint32_t)(need < kExpandSub + 1 ? need : kExpandSub + 1);
      for (uint32_t k = threadIdx.x; k < m; k += kBlock) s_pref[k] = a.act_prefix[ja + k];
      __syncthreads();
      uint32_t klo = 0, kc = ~0u;
      uint64_t ybase = 0;
      float xv = 0.f;
#pragma unroll
      for (int e = 0; e < (int)(kExpandSub / kBlock); ++e) {
        const uint64_t p = pb + (uint64_t)e * kBlock + threadIdx.x;
        const bool valid = p >= p0 && p < pe;
        uint32_t col = 0;
        if (valid) {
          uint32_t k = klo;  // same owner as the previous product: one LDS read
          if (k + 1 < m && s_pref[k + 1] <= p) {
            uint32_t lo = k + 1, hi = m;
            while (lo < hi) {
              const uint32_t mid = (lo + hi) >> 1;
              if (s_pref[mid] <= p) lo = mid + 1; else hi = mid;
            }
            k = lo - 1;
          }
          klo = k;
          if (k != kc) {
            kc = k;
            ybase = a.act_ypos[ja + k] - s_pref[k];
            xv = a.act_xval[ja + k];
          }
          const uint64_t y = ybase + p;
          const float v = xv * a.yvals[y];
          col = a.ycol[y];
          __builtin_nontemporal_store(col, &ocol[p - a.P0]);
          if (oval) __builtin_nontemporal_store(v, &oval[p - a.P0]);  // NULL: count-only
          vsum += (double)v;
        }
        if (isbig) {
          if (!a.dshift) {
            // low column bits are spread: plain LDS atomics
            if (valid) atomicAdd(&sh[col & (R - 1)], 1u);
          } else {
            // high digits come in runs (each owner's Y segment is sorted by
            // column): the first lane of each run adds the run's length
            const uint32_t lane = lane_id();
            const uint32_t d = (col >> a.dshift) & (R - 1);
            const uint32_t pd = __shfl_up(d, 1, kWave);
            const bool pv = __shfl_up((uint32_t)valid, 1, kWave) != 0u;
            const bool lead = valid && (lane == 0 || !pv || pd != d);
            const uint64_t lm = __ballot(lead);
            const uint64_t vm = __ballot(valid);  // one run of lanes
            if (lead) {
              const uint32_t hi = (uint32_t)kWave - (uint32_t)__builtin_clzll(vm);
              const uint64_t above = lane + 1 < (uint32_t)kWave ? lm >> (lane + 1) : 0ull;
              const uint32_t end = above ? lane + 1 + (uint32_t)__builtin_ctzll(above) : hi;
              atomicAdd(&sh[d], end - lane);
            }
          }
        }
      }
    }
    if (isbig) {
      __syncthreads();
      for (int d = threadIdx.x; d < R; d += kBlock) cnt[it * R + d] = sh[d];
    }
  }
  for (int off = kWave / 2; off > 0; off >>= 1) vsum += __shfl_xor(vsum, off, kWave);
  if (lane_id() == 0) s_sum[threadIdx.x / kWave] = vsum;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < kBlock / kWave; ++w) t += s_sum[w];
    atomicAdd(a.vsum, t);
  }
}

}  // namespace

static uint32_t xcd_remap() { return xcd_remap_enabled(); }

void SegSortPlan::build(const std::vector<uint64_t>& seg_off, hipStream_t s) {
  std::vector<SegTile> tv;
  std::vector<SegDesc> sv;
  for (size_t g = 0; g + 1 < seg_off.size(); ++g) {
    const uint64_t len = seg_off[g + 1] - seg_off[g];
    const uint64_t nt = ceil_div(len, kSegTile);
    if (!nt) continue;
    const uint64_t t0 = tv.size();
    sv.push_back(SegDesc{seg_off[g], (uint32_t)t0, (uint32_t)nt});
    for (uint64_t t = 0; t < nt; ++t) {
      SegTile d;
      d.start = seg_off[g] + t * kSegTile;
      d.count = (uint32_t)std::min<uint64_t>(kSegTile, len - t * kSegTile);
      d.tile = t0 + t;
      d.first = (uint32_t)t0;
      tv.push_back(d);
    }
  }
  n = seg_off.empty() ? 0 : seg_off.back();
  if (n >= (1ull << 32)) throw std::length_error("segmented sort of 2^32 or more keys");
  ntiles = tv.size();
  nsegs = sv.size();
  tiles.reset(ntiles ? ntiles : 1);
  segs.reset(nsegs ? nsegs : 1);
  if (ntiles) {
    SPTC_HIP(hipMemcpyAsync(tiles.get(), tv.data(), ntiles * sizeof(SegTile),
                            hipMemcpyHostToDevice, s));
    SPTC_HIP(hipMemcpyAsync(segs.get(), sv.data(), nsegs * sizeof(SegDesc),
                            hipMemcpyHostToDevice, s));
  }
  SPTC_HIP(hipStreamSynchronize(s));  // tv, sv are released on return
}

void segmented_sort_reserve(SortWorkspace& ws, const SegSortPlan& plan, hipStream_t s) {
  segmented_sort_reserve(ws, plan.ntiles, s);
}

void segmented_sort_reserve(SortWorkspace& ws, uint64_t ntiles, hipStream_t s, int rb) {
  const uint64_t m = ntiles << rb;
  ws.ensure(1, kSegRB, s);
  if (ws.rts_cnt.size() < m) {
    ws.rts_cnt.reset(m);
    ws.rts_off.reset(m);
  }
}

int segmented_sort_passes(int nbits) { return radix_passes(nbits, kSegRB); }

bool segmented_sort(SortWorkspace& ws, const SegTile* tiles, uint64_t ntiles,
                    const SegDesc* segs, uint64_t nsegs, uint64_t n, uint64_t* k0, float* v0,
                    uint64_t* k1, float* v1, int nbits, hipStream_t s, KTimer* kt, int cls) {
  const int passes = radix_passes(nbits, kSegRB);
  if (!n || !passes || !ntiles) return false;
  const uint64_t m = ntiles << kSegRB;
  if (ws.rts_cnt.size() < m) throw std::logic_error("segmented sort workspace not reserved");
  if (ntiles >= (1ull << 31) || nsegs >= (1ull << 31))
    throw std::length_error("segmented sort tile count");
  uint64_t* ka = k0;
  float* va = v0;
  uint64_t* kb = k1;
  float* vb = v1;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * kSegRB;
    int tok = kt ? kt->begin(s) : -1;
    hipLaunchKernelGGL((k_seg_upsweep<uint64_t, kSegRB, kSegIPT>), dim3((uint32_t)ntiles),
                       dim3(kSegNT), 0, s, ka, shift, tiles, ws.rts_cnt.get());
    hipLaunchKernelGGL(k_seg_scan<kSegRB>, dim3((uint32_t)nsegs), dim3(1 << kSegRB), 0, s, segs,
                       ws.rts_cnt.get(), ws.rts_off.get());
    if (kt) kt->end(tok, SPTC_K_SORT_COUNT, 8.0 * (double)n + 16.0 * (double)m, s);
    tok = kt ? kt->begin(s) : -1;
    hipLaunchKernelGGL((k_seg_downsweep<uint64_t, uint64_t, kSegRB, kSegIPT>),
                       dim3((uint32_t)ntiles), dim3(kSegNT), 0, s, ka, va, kb, vb, shift, tiles,
                       ws.rts_off.get(), ws.rts_cnt.get(), xcd_remap());
    if (kt) kt->end(tok, cls, 24.0 * (double)n, s);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  return (passes & 1) != 0;
}

bool segmented_sort(SortWorkspace& ws, const SegSortPlan& plan, uint64_t* k0, float* v0,
                    uint64_t* k1, float* v1, int nbits, hipStream_t s, KTimer* kt, int cls) {
  return segmented_sort(ws, plan.tiles.get(), plan.ntiles, plan.segs.get(), plan.nsegs, plan.n,
                        k0, v0, k1, v1, nbits, s, kt, cls);
}

// Digit width of the big-row passes: 10 bits where that saves a pass over
// 8-bit digits (column bits 9-10, 17-20), else 8 (longer runs per digit).
static int narrow_rb(int col_bits) {
  static const int forced = [] {
    const char* e = getenv("SPTC_SEG_RB");
    return e ? atoi(e) : 0;
  }();
  if (forced == 8) return 8;
  const int p8 = std::max(1, radix_passes(col_bits, 8));
  const int p10 = std::max(1, radix_passes(col_bits, 10));
  return p10 < p8 ? 10 : 8;
}
// SPTC_SEG_WIDE=1: big-row scatter passes on 1024-thread workgroups
static bool seg_wide() {
  static const bool w = [] {
    const char* e = getenv("SPTC_SEG_WIDE");
    return e && atoi(e) == 1;
  }();
  return w;
}
static int narrow_ipt() {
  static const int ipt = [] {
    const char* e = getenv("SPTC_SEG_IPT");
    const int v = e ? atoi(e) : kNarrowIPT;
    return v == 12 || v == 16 || v == 24 ? v : kNarrowIPT;
  }();
  return ipt;
}
static int narrow_passes(int col_bits) {
  return std::max(1, radix_passes(col_bits, narrow_rb(col_bits)));
}
static int plan_passes(const RowSegPlan& plan) {
  return plan.msd ? 1 : narrow_passes(plan.col_bits);
}

void build_row_segments(const RowRange* ranges, uint32_t nranges, uint64_t lrows,
                        const uint64_t* row_prod, int col_bits, RowSegPlan* plan,
                        SortWorkspace& ws, U2* tmp, hipStream_t s, bool msd,
                        const uint64_t* row_act) {
  if (col_bits > 32) throw std::logic_error("row-segmented sort of more than 32 column bits");
  plan->nlocal = plan->nbig_tiles = plan->nbig_rows = plan->nsmall = 0;
  plan->lrows = lrows;
  plan->col_bits = col_bits;
  plan->msd = msd && col_bits > kBucketRB && col_bits <= kBucketRB + kBucketMaxLow;
  plan->rb = plan->msd ? kBucketRB : narrow_rb(col_bits);
  plan->ipt = narrow_ipt();
  const uint64_t tile = (uint64_t)plan->ipt * kSegNT;
  if (!lrows) return;
  const uint64_t m = lrows + 1;
  if (plan->scratch.size() < 10 * m) plan->scratch.reset(10 * m);
  uint64_t* small_n = plan->scratch.get();
  uint64_t* big_t = small_n + m;
  uint64_t* ps = big_t + m;
  uint64_t* bt = ps + m;
  uint64_t* head = bt + m;
  uint64_t* hs = head + m;
  uint64_t* big_f = hs + m;
  uint64_t* bs = big_f + m;
  uint64_t* rstart = bs + m;
  uint64_t* pre_p = rstart + m;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(lrows, kBlock), 4096);
  hipLaunchKernelGGL(k_rowseg_sizes, dim3(grid), dim3(kBlock), 0, s, ranges, nranges, lrows,
                     row_prod, row_act, tile, small_n, big_t, big_f, rstart, pre_p);
  run_exclusive_scan(small_n, ps, lrows, tmp, s);
  run_exclusive_scan(big_t, bt, lrows, tmp, s);
  run_exclusive_scan(big_f, bs, lrows, tmp, s);
  hipLaunchKernelGGL(k_rowseg_heads, dim3(grid), dim3(kBlock), 0, s, lrows, small_n, big_t,
                     pre_p, ps, head);
  run_exclusive_scan(head, hs, lrows, tmp, s);
  uint64_t tot[4];
  SPTC_HIP(hipMemcpyAsync(&tot[0], hs + lrows, 8, hipMemcpyDeviceToHost, s));
  SPTC_HIP(hipMemcpyAsync(&tot[1], bt + lrows, 8, hipMemcpyDeviceToHost, s));
  SPTC_HIP(hipMemcpyAsync(&tot[2], ps + lrows, 8, hipMemcpyDeviceToHost, s));
  SPTC_HIP(hipMemcpyAsync(&tot[3], bs + lrows, 8, hipMemcpyDeviceToHost, s));
  SPTC_HIP(hipStreamSynchronize(s));
  plan->nlocal = tot[0];
  plan->nbig_tiles = tot[1];
  plan->nsmall = tot[2];
  plan->nbig_rows = tot[3];
  if (plan->nlocal >= (1ull << 31) || plan->nbig_tiles >= (1ull << 31))
    throw std::length_error("row segments");
  if (plan->local.size() < plan->nlocal) plan->local.reset(plan->nlocal);
  if (plan->big.size() < plan->nbig_tiles) plan->big.reset(plan->nbig_tiles);
  if (plan->bigseg.size() < plan->nbig_rows) plan->bigseg.reset(plan->nbig_rows);
  if (plan->msd && plan->bigrow.size() < plan->nbig_rows) plan->bigrow.reset(plan->nbig_rows);
  if (plan->nlocal)
    hipLaunchKernelGGL(k_zero_local, dim3((uint32_t)ceil_div(plan->nlocal, 256)), dim3(256), 0,
                       s, plan->local.get(), plan->nlocal);
  hipLaunchKernelGGL(k_rowseg_emit, dim3(grid), dim3(kBlock), 0, s, ranges, nranges, lrows,
                     row_prod, plan->rb, tile, big_t, bt, bs, head, hs, pre_p,
                     plan->local.get(), plan->big.get(), plan->bigseg.get(),
                     plan->msd ? plan->bigrow.get() : nullptr);
  if (plan->nbig_tiles) segmented_sort_reserve(ws, plan->nbig_tiles, s, plan->rb);
}

// Column-array sets of the big-row passes: set 0 = (k1 as 32-bit words
// [0, n), v1), set 1 = (k1 words [n, 2n), v).  The passes alternate sets and
// the last one reads set 0 and writes (col, v), so the expansion starts in
// set 0 for an odd pass count and in set 1 for an even one.  Set 1 shares v
// with the groups, at the big rows' positions only.
NarrowTarget narrow_target(const RowSegPlan& plan, uint64_t n, float* v, uint64_t* k1,
                           float* v1) {
  uint32_t* c = reinterpret_cast<uint32_t*>(k1);
  if (plan_passes(plan) & 1) return NarrowTarget{c, v1};
  return NarrowTarget{c + n, v};
}

template <int RB, int IPT>
static void narrow_sort(SortWorkspace& ws, const RowSegPlan& plan, uint64_t n, uint32_t* fcol,
                        float* v, uint64_t* k1, float* v1, bool counted, hipStream_t s,
                        KTimer* kt, int cls) {
  const uint64_t ntiles = plan.nbig_tiles;
  const uint64_t m = ntiles << RB;
  if (ws.rts_cnt.size() < m) throw std::logic_error("segmented sort workspace not reserved");
  const uint64_t nb = n - plan.nsmall;  // big-row products
  uint32_t* col[2] = {reinterpret_cast<uint32_t*>(k1), reinterpret_cast<uint32_t*>(k1) + n};
  float* val[2] = {v1, v};
  const int passes = plan_passes(plan);
  int cur = (passes & 1) ? 0 : 1;
  const SegTile* tiles = plan.big.get();
  for (int p = 0; p < passes; ++p) {
    const int shift = plan.msd ? plan.col_bits - RB : p * RB;
    int tok = kt ? kt->begin(s) : -1;
    if (p > 0 || !counted)  // else the expansion counted the first pass
      hipLaunchKernelGGL((k_seg_upsweep<uint32_t, RB, IPT>), dim3((uint32_t)ntiles),
                         dim3(kSegNT), 0, s, col[cur], shift, tiles, ws.rts_cnt.get());
    hipLaunchKernelGGL(k_seg_scan<RB>, dim3((uint32_t)plan.nbig_rows), dim3(1 << RB), 0, s,
                       plan.bigseg.get(), ws.rts_cnt.get(), ws.rts_off.get());
    const double up = (p > 0 || !counted) ? 4.0 * (double)nb : 0.0;
    if (kt) kt->end(tok, SPTC_K_SORT_COUNT, up + 16.0 * (double)m, s);
    tok = kt ? kt->begin(s) : -1;
    const bool last = p + 1 == passes;
    if (IPT == 16 && seg_wide())  // same 8192-key tiles, 1024 threads x 8 keys
      hipLaunchKernelGGL((k_seg_downsweep<uint32_t, uint32_t, RB, 8, 1024>),
                         dim3((uint32_t)ntiles), dim3(1024), 0, s, col[cur], val[cur],
                         last ? fcol : col[1 - cur], last ? v : val[1 - cur], shift, tiles,
                         ws.rts_off.get(), ws.rts_cnt.get(), xcd_remap());
    else
      hipLaunchKernelGGL((k_seg_downsweep<uint32_t, uint32_t, RB, IPT>), dim3((uint32_t)ntiles),
                         dim3(kSegNT), 0, s, col[cur], val[cur], last ? fcol : col[1 - cur],
                         last ? v : val[1 - cur], shift, tiles, ws.rts_off.get(),
                         ws.rts_cnt.get(), xcd_remap());
    if (kt) kt->end(tok, cls, (v ? 16.0 : 8.0) * (double)nb, s);  // column (+ value) moved
    cur = 1 - cur;
  }
}

void launch_expand_items(const ExpandArgs& a, const RowSegPlan& plan, SortWorkspace& ws,
                         hipStream_t s) {
  const uint64_t items = plan.nbig_tiles + plan.nlocal;
  if (!items) return;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(items, 8192);
  ExpandArgs b = a;
  b.remap = xcd_remap_enabled();
  b.dshift = plan.msd ? plan.col_bits - plan.rb : 0;
  if (plan.rb == 10)
    hipLaunchKernelGGL(k_expand_items<10>, dim3(grid), dim3(kBlock), 0, s, b, plan.big.get(),
                       (uint32_t)plan.nbig_tiles, plan.local.get(), (uint32_t)plan.nlocal,
                       ws.rts_cnt.get());
  else
    hipLaunchKernelGGL(k_expand_items<8>, dim3(grid), dim3(kBlock), 0, s, b, plan.big.get(),
                       (uint32_t)plan.nbig_tiles, plan.local.get(), (uint32_t)plan.nlocal,
                       ws.rts_cnt.get());
}

void sort_row_segments(SortWorkspace& ws, const RowSegPlan& plan, uint64_t n, uint32_t* col,
                       float* v, uint64_t* k1, float* v1, const uint64_t* row_start,
                       hipStream_t s, KTimer* kt, int cls, bool counted) {
  if (plan.nbig_tiles) {
    const bool c = counted;
    switch (plan.rb * 100 + plan.ipt) {
      case 812: narrow_sort<8, 12>(ws, plan, n, col, v, k1, v1, c, s, kt, cls); break;
      case 816: narrow_sort<8, 16>(ws, plan, n, col, v, k1, v1, c, s, kt, cls); break;
      case 824: narrow_sort<8, 24>(ws, plan, n, col, v, k1, v1, c, s, kt, cls); break;
      case 1012: narrow_sort<10, 12>(ws, plan, n, col, v, k1, v1, c, s, kt, cls); break;
      case 1016: narrow_sort<10, 16>(ws, plan, n, col, v, k1, v1, c, s, kt, cls); break;
      case 1024: narrow_sort<10, 24>(ws, plan, n, col, v, k1, v1, c, s, kt, cls); break;
      default: throw std::logic_error("row-segmented sort tile shape");
    }
  }
  if (plan.nlocal) {
    int tok = kt ? kt->begin(s) : -1;
    // every group's columns are read (4 B a product); a group not already in
    // order also reads its values and writes both back (12 B, or 4 B for
    // columns alone), counted on the device
    unsigned long long* moved = kt ? kt->counter(tok, v ? 12.0 : 4.0, s) : nullptr;
    // a group's row offset takes at most 13 bits (6144 products)
    if (plan.col_bits + 13 <= 32)
      hipLaunchKernelGGL((k_local_sort<uint32_t, 10>), dim3((uint32_t)plan.nlocal), dim3(kSegNT),
                         0, s, col, v, plan.local.get(), row_start, plan.lrows, moved);
    else
      hipLaunchKernelGGL((k_local_sort<uint64_t, kSegRB>), dim3((uint32_t)plan.nlocal),
                         dim3(kSegNT), 0, s, col, v, plan.local.get(), row_start, plan.lrows,
                         moved);
    if (kt) kt->end(tok, SPTC_K_SORT_LOCAL, 4.0 * (double)plan.nsmall, s);
  }
}

}  // namespace sptc
