// SpTC phase kernels for gfx950: linearize, index match, ESC expand and
// reduce.  Host launchers are declared in kernels.hpp.
#include "dense_core.hpp"
#include "kernels.hpp"

#include <string>
#include "reduce_core.hpp"

#include <cstdlib>
#include <stdexcept>

namespace sptc {

// --------------------------------------------------------------------------
// Tile-sum scan (one workgroup, loops over the tile sums with a carry).
__global__ __launch_bounds__(kBlock) void k_scan_tiles(U2* sums, uint64_t nt, U2* total) {
  __shared__ U2 scratch[kBlock / kWave];
  U2 carry = mk2(0, 0);
  for (uint64_t base = 0; base < nt; base += kScanTile) {
    const uint64_t first = base + (uint64_t)threadIdx.x * kScanItems;
    U2 v[kScanItems];
    U2 s = mk2(0, 0);
#pragma unroll
    for (int e = 0; e < kScanItems; ++e) {
      v[e] = first + e < nt ? sums[first + e] : mk2(0, 0);
      s = s + v[e];
    }
    U2 tot;
    U2 ex = block_excl_scan_u2(s, scratch, &tot) + carry;
#pragma unroll
    for (int e = 0; e < kScanItems; ++e) {
      if (first + e < nt) sums[first + e] = ex;
      ex = ex + v[e];
    }
    carry = carry + tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

// Many-workgroup tile-sum scan: block sums, one scan of the block sums,
// then each block rescans its sums from its offset (1024 threads x 16 sums).
constexpr int kPartNT = 1024, kPartItems = (int)(kScanSumsPerBlock / kPartNT);

__device__ __forceinline__ U2 block_scan_part(U2 t, U2* scratch, U2* tot) {
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  U2 inc = t;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint64_t oa = __shfl_up(inc.a, d, kWave), ob = __shfl_up(inc.b, d, kWave);
    if (lane >= (uint32_t)d) {
      inc.a += oa;
      inc.b += ob;
    }
  }
  if (lane == kWave - 1) scratch[wave] = inc;
  __syncthreads();
  U2 off = mk2(0, 0), all = mk2(0, 0);
  for (int w = 0; w < kPartNT / kWave; ++w) {
    const U2 x = scratch[w];
    if ((uint32_t)w < wave) off = off + x;
    all = all + x;
  }
  __syncthreads();
  *tot = all;
  return off + inc - t;
}

__global__ __launch_bounds__(kPartNT) void k_part_sums(const U2* __restrict__ sums, uint64_t nt,
                                                       U2* __restrict__ parts) {
  __shared__ U2 scratch[kPartNT / kWave];
  const uint64_t first = (uint64_t)blockIdx.x * kScanSumsPerBlock + threadIdx.x;
  U2 t = mk2(0, 0);
#pragma unroll
  for (int e = 0; e < kPartItems; ++e) {
    const uint64_t i = first + (uint64_t)e * kPartNT;
    if (i < nt) t = t + sums[i];
  }
  U2 tot;
  block_scan_part(t, scratch, &tot);
  if (threadIdx.x == 0) parts[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kPartNT) void k_part_apply(U2* __restrict__ sums, uint64_t nt,
                                                        const U2* __restrict__ parts) {
  __shared__ U2 scratch[kPartNT / kWave];
  const uint64_t first = (uint64_t)blockIdx.x * kScanSumsPerBlock +
                         (uint64_t)threadIdx.x * kPartItems;  // blocked: own run of sums
  U2 v[kPartItems];
  U2 t = mk2(0, 0);
#pragma unroll
  for (int e = 0; e < kPartItems; ++e) {
    v[e] = first + e < nt ? sums[first + e] : mk2(0, 0);
    t = t + v[e];
  }
  U2 tot;
  U2 ex = block_scan_part(t, scratch, &tot) + parts[blockIdx.x];
#pragma unroll
  for (int e = 0; e < kPartItems; ++e) {
    if (first + e < nt) sums[first + e] = ex;
    ex = ex + v[e];
  }
}

void scan_tiles_parts(U2* sums, uint64_t nt, U2* total, U2* parts, hipStream_t s) {
  if (nt <= 2 * kScanSumsPerBlock) {
    scan_tiles(sums, nt, total, s);
    return;
  }
  const uint32_t nb = (uint32_t)ceil_div(nt, kScanSumsPerBlock);
  hipLaunchKernelGGL(k_part_sums, dim3(nb), dim3(kPartNT), 0, s, sums, nt, parts);
  scan_tiles(parts, nb, total, s);  // block offsets in place, grand total
  hipLaunchKernelGGL(k_part_apply, dim3(nb), dim3(kPartNT), 0, s, sums, nt, parts);
}

void scan_tiles(U2* sums, uint64_t nt, U2* total, hipStream_t s) {
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kBlock), 0, s, sums, nt, total);
}

// --------------------------------------------------------------------------
// (a) Linearize: key = OR_m (idx[m] & (2^w[m]-1)) << shift[m].  The fields
// are contiguous, so the reference's pdep with mask (2^w-1) << s is a masked
// shift (pdep likewise keeps only the low w bits of an index).  Fused: the
// digit histograms of every radix pass of the following sort.
template <typename IdxT, int RB>
__global__ __launch_bounds__(kBlock) void k_linearize(LinArgs a, uint64_t n,
                                                      uint64_t* __restrict__ keys,
                                                      uint64_t* __restrict__ hist,
                                                      int passes) {
  constexpr int R = 1 << RB;
  __shared__ uint32_t sh[kMaxFusedPasses * R];
  for (int i = threadIdx.x; i < passes * R; i += kBlock) sh[i] = 0;
  __syncthreads();
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; i < n; i += stride) {
    uint64_t k = 0;
#pragma unroll
    for (int m = 0; m < SPTC_MAX_ORDER; ++m)
      if (m < a.order)
        k |= shl64((uint64_t)static_cast<const IdxT*>(a.idx[m])[i] & low_mask(a.width[m]),
                   a.shift[m]);
    keys[i] = k;
    if (passes) hist_add<RB>(sh, k, passes);
  }
  if (passes) {
    __syncthreads();
    hist_flush<RB>(sh, hist, passes);
  }
}

void launch_linearize(const LinArgs& a, int idx_bytes, uint64_t n, uint64_t* keys,
                      uint64_t* hist, int passes, int rb, hipStream_t s) {
  if (!n) return;
  uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(n, kBlock * 4), 2048);
  if (passes > kMaxFusedPasses) passes = 0;  // caller falls back to histogram()
#define SPTC_LIN(T, RBV)                                                                   \
  hipLaunchKernelGGL((k_linearize<T, RBV>), dim3(grid), dim3(kBlock), 0, s, a, n, keys, hist, \
                     passes)
  if (idx_bytes == 4) {
    switch (rb) {
      case 9: SPTC_LIN(uint32_t, 9); break;
      case 10: SPTC_LIN(uint32_t, 10); break;
      case 11: SPTC_LIN(uint32_t, 11); break;
      default: SPTC_LIN(uint32_t, 8); break;
    }
  } else {
    switch (rb) {
      case 9: SPTC_LIN(uint64_t, 9); break;
      case 10: SPTC_LIN(uint64_t, 10); break;
      case 11: SPTC_LIN(uint64_t, 11); break;
      default: SPTC_LIN(uint64_t, 8); break;
    }
  }
#undef SPTC_LIN
}

// --------------------------------------------------------------------------
// (c) Index match.
//
// Y is sorted by key = c << y_nfbits | f, so every contract value c owns one
// contiguous segment.  With a small contract field (ncbits <= kDirectBits)
// the segment table is direct-addressed: lo[c], hi[c]; otherwise it is the
// compacted list of distinct c with their start offsets, searched per X
// nonzero.

__global__ __launch_bounds__(kBlock) void k_ytable_direct(const uint64_t* __restrict__ ykeys,
                                                          uint64_t ny, int ynf, uint64_t K,
                                                          uint64_t* __restrict__ lo,
                                                          uint64_t* __restrict__ hi) {
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; i < ny; i += stride) {
    uint64_t c = shr64(ykeys[i], ynf);
    if (c >= K) continue;  // key outside the contract field: matches nothing
    if (i == 0 || shr64(ykeys[i - 1], ynf) != c) lo[c] = i;
    if (i == ny - 1 || shr64(ykeys[i + 1], ynf) != c) hi[c] = i + 1;
  }
}

void launch_ytable_direct(const uint64_t* ykeys, uint64_t ny, int ynf, uint64_t K,
                          uint64_t* lo, uint64_t* hi, hipStream_t s) {
  if (!ny) return;
  uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(ny, kBlock), 4096);
  hipLaunchKernelGGL(k_ytable_direct, dim3(grid), dim3(kBlock), 0, s, ykeys, ny, ynf, K, lo,
                     hi);
}

struct YUniqueOp {
  const uint64_t* ykeys;
  uint64_t ny;
  int ynf;
  uint64_t* ukey;
  uint64_t* ustart;
  __device__ U2 value(uint64_t i) const {
    return mk2(i == 0 || shr64(ykeys[i - 1], ynf) != shr64(ykeys[i], ynf), 0);
  }
  __device__ void emit(uint64_t i, U2 excl, U2 v) const {
    if (v.a) {
      ukey[excl.a] = shr64(ykeys[i], ynf);
      ustart[excl.a] = i;
    }
    if (i == ny - 1) ustart[excl.a + v.a] = ny;
  }
};

void run_yunique(const uint64_t* ykeys, uint64_t ny, int ynf, uint64_t* ukey,
                 uint64_t* ustart, U2* tmp, hipStream_t s) {
  YUniqueOp op{ykeys, ny, ynf, ukey, ustart};
  run_scan(op, ny, tmp, s);
}

__device__ __forceinline__ void ytable_lookup(const YTable& t, uint64_t c,
                                              uint64_t& pos, uint64_t& cnt) {
  if (t.direct) {
    uint64_t l = t.lo[c], h = t.hi[c];
    pos = l;
    cnt = h - l;
  } else {
    uint64_t lo = 0, hi = t.U;
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (t.ukey[mid] < c) lo = mid + 1; else hi = mid;
    }
    if (lo < t.U && t.ukey[lo] == c) {
      pos = t.ustart[lo];
      cnt = t.ustart[lo + 1] - pos;
    } else {
      pos = 0;
      cnt = 0;
    }
  }
}

// value: a = match count; b = (active << 32) | row_head
struct MatchOp {
  const uint64_t* xkeys;
  const float* xvals;
  uint64_t nx;
  uint64_t cmask;
  int ncbits;
  YTable t;
  MatchOut o;
  int api;  // 1: write mpos/mcnt/prefix (per-phase entry point)

  // the X key's Y segment start, from value() to emit() in registers
  struct Aux {
    uint64_t pos;
  };
  __device__ U2 value(uint64_t i, Aux& x) const {
    const uint64_t k = xkeys[i];
    uint64_t pos, cnt;
    ytable_lookup(t, k & cmask, pos, cnt);
    x.pos = pos;
    const uint64_t head = (i == 0) || (shr64(xkeys[i - 1], ncbits) != shr64(k, ncbits));
    return mk2(cnt, ((uint64_t)(cnt != 0) << 32) | head);
  }
  __device__ void emit(uint64_t i, U2 excl, U2 v, const Aux& x) const {
    const uint64_t pos = x.pos, cnt = v.a;
    if (api) {
      o.mpos[i] = cnt ? pos : 0;
      o.mcnt[i] = cnt;
      o.prefix[i] = excl.a;
      if (i == nx - 1) o.prefix[nx] = excl.a + v.a;
      return;
    }
    const uint64_t head = v.b & 1;
    const uint64_t row = (excl.b & 0xFFFFFFFFull) + head - 1;
    if (head) {
      o.row_key[row] = shr64(xkeys[i], ncbits);
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
  }
};

void run_match(const uint64_t* xkeys, const float* xvals, uint64_t nx, int ncbits,
               const YTable& t, const MatchOut& o, int api, U2* tmp, hipStream_t s) {
  MatchOp op{xkeys, xvals, nx, low_mask(ncbits), ncbits, t, o, api};
  run_scan(op, nx, tmp, s);
}

void run_match_count(const uint64_t* xkeys, uint64_t nx, int ncbits, const YTable& t, U2* tmp,
                     hipStream_t s) {
  MatchOut o{};
  MatchOp op{xkeys, nullptr, nx, low_mask(ncbits), ncbits, t, o, 0};
  run_scan(op, nx, tmp, s, /*emit=*/false);
}

struct ExclScanOp {
  const uint64_t* in;
  uint64_t* out;
  uint64_t n;
  __device__ U2 value(uint64_t i) const { return mk2(in[i], 0); }
  __device__ void emit(uint64_t i, U2 excl, U2 v) const {
    out[i] = excl.a;
    if (i == n - 1) out[n] = excl.a + v.a;
  }
};

void run_exclusive_scan(const uint64_t* in, uint64_t* out, uint64_t n, U2* tmp,
                        hipStream_t s) {
  ExclScanOp op{in, out, n};
  run_scan(op, n, tmp, s);
}

// --------------------------------------------------------------------------
// Y free-key rank compression.  Y's free field (ynf bits) usually takes far
// fewer distinct values than 2^ynf; replacing it by its rank among Y's
// distinct free keys shortens every product key (fewer radix passes).  The
// map is monotone, so Y stays sorted by (c, rank).

__global__ __launch_bounds__(kBlock) void k_yfree_extract(const uint64_t* __restrict__ ykeys,
                                                          uint64_t ny, uint64_t fmask,
                                                          uint64_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; i < ny; i += stride) out[i] = ykeys[i] & fmask;
}

struct UniqueOp {
  const uint64_t* keys;  // sorted
  uint64_t* out;
  __device__ U2 value(uint64_t i) const { return mk2(i == 0 || keys[i] != keys[i - 1], 0); }
  __device__ void emit(uint64_t i, U2 excl, U2 v) const {
    if (v.a) out[excl.a] = keys[i];
  }
};

__global__ __launch_bounds__(kBlock) void k_yrank(const uint64_t* __restrict__ ykeys, uint64_t ny,
                                                  int ynf, uint64_t fmask,
                                                  const uint64_t* __restrict__ ytab, uint64_t U,
                                                  int yrb, uint64_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; i < ny; i += stride) {
    const uint64_t k = ykeys[i];
    const uint64_t f = k & fmask;
    uint64_t lo = 0, hi = U;
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (ytab[mid] < f) lo = mid + 1; else hi = mid;
    }
    out[i] = shl64(shr64(k, ynf), yrb) | lo;
  }
}

__global__ __launch_bounds__(kBlock) void k_tuple_key(TupleArgs a, uint64_t n,
                                                      uint64_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; i < n; i += stride) {
    uint64_t k = a.prev ? shl64(a.prev[i], a.prev_shift) : 0;
    for (int m = 0; m < a.n; ++m) {
      const uint64_t v = a.bytes == 4 ? (uint64_t) static_cast<const uint32_t*>(a.idx[m])[i]
                                      : static_cast<const uint64_t*>(a.idx[m])[i];
      k |= shl64(v & low_mask(a.width[m]), a.shift[m]);
    }
    out[i] = k;
  }
}

// Largest index of one mode: a grid-stride max, a wave reduction, one
// atomicMax per wave into *out (cleared by the caller).
__global__ __launch_bounds__(kBlock) void k_mode_max(const void* __restrict__ idx, int bytes,
                                                     uint64_t n,
                                                     unsigned long long* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  unsigned long long m = 0;
  for (; i < n; i += stride) {
    const unsigned long long v = bytes == 4 ? (unsigned long long)static_cast<const uint32_t*>(idx)[i]
                                            : static_cast<const uint64_t*>(idx)[i];
    m = v > m ? v : m;
  }
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const unsigned long long o = __shfl_down(m, d, kWave);
    m = o > m ? o : m;
  }
  if (lane_id() == 0) atomicMax(out, m);
}

void launch_mode_max(const void* idx, int bytes, uint64_t n, uint64_t* out, hipStream_t s) {
  if (!n) return;
  uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(n, kBlock * 8), 2048);
  hipLaunchKernelGGL(k_mode_max, dim3(grid), dim3(kBlock), 0, s, idx, bytes, n,
                     reinterpret_cast<unsigned long long*>(out));
}

void launch_tuple_key(const TupleArgs& a, uint64_t n, uint64_t* out, hipStream_t s) {
  if (!n) return;
  uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(n, kBlock), 4096);
  hipLaunchKernelGGL(k_tuple_key, dim3(grid), dim3(kBlock), 0, s, a, n, out);
}

void launch_yfree_extract(const uint64_t* ykeys, uint64_t ny, uint64_t fmask, uint64_t* out,
                          hipStream_t s) {
  if (!ny) return;
  uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(ny, kBlock), 4096);
  hipLaunchKernelGGL(k_yfree_extract, dim3(grid), dim3(kBlock), 0, s, ykeys, ny, fmask, out);
}

void run_unique(const uint64_t* sorted, uint64_t n, uint64_t* out, U2* tmp, hipStream_t s) {
  UniqueOp op{sorted, out};
  run_scan(op, n, tmp, s);
}

void launch_yrank(const uint64_t* ykeys, uint64_t ny, int ynf, uint64_t fmask,
                  const uint64_t* ytab, uint64_t U, int yrb, uint64_t* out, hipStream_t s) {
  if (!ny) return;
  uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(ny, kBlock), 4096);
  hipLaunchKernelGGL(k_yrank, dim3(grid), dim3(kBlock), 0, s, ykeys, ny, ynf, fmask, ytab, U,
                     yrb, out);
}

// --------------------------------------------------------------------------
// Partition: end row of the next chunk.  Largest r1 in (r0, min(R, r0+rmax)]
// with row_prod[r1] - row_prod[r0] <= cap; r1 = r0 when even one row does
// not fit.  out = {r1, row_prod[r1], row_act[r1]}.
__global__ void k_chunk_end(const uint64_t* row_prod, const uint64_t* row_act,
                            uint64_t R, uint64_t r0, uint64_t rmax, uint64_t cap,
                            uint64_t* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint64_t p0 = row_prod[r0];
  uint64_t hi_r = r0 + rmax < R ? r0 + rmax : R;
  uint64_t lo = r0, hi = hi_r;  // answer in [r0, hi_r]
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo + 1) / 2;
    if (row_prod[mid] - p0 <= cap) lo = mid; else hi = mid - 1;
  }
  out[0] = lo;
  out[1] = row_prod[lo];
  out[2] = row_act[lo];
}

void launch_chunk_end(const uint64_t* row_prod, const uint64_t* row_act, uint64_t R,
                      uint64_t r0, uint64_t rmax, uint64_t cap, uint64_t* out,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_chunk_end, dim3(1), dim3(64), 0, s, row_prod, row_act, R, r0,
                     rmax, cap, out);
}

// Largest product count of any row: max over r of row_prod[r+1]-row_prod[r].
__global__ __launch_bounds__(kBlock) void k_max_row(const uint64_t* row_prod, uint64_t R,
                                                    unsigned long long* out) {
  uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  uint64_t m = 0;
  for (; r < R; r += stride) {
    uint64_t d = row_prod[r + 1] - row_prod[r];
    m = d > m ? d : m;
  }
  for (int off = kWave / 2; off > 0; off >>= 1) {
    uint64_t o = __shfl_xor(m, off, kWave);
    m = o > m ? o : m;
  }
  if (lane_id() == 0 && m) atomicMax(out, (unsigned long long)m);
}

void launch_max_row(const uint64_t* row_prod, uint64_t R, uint64_t* out, hipStream_t s) {
  fill_u64(out, 0, 1, s);
  if (!R) return;
  uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(R, kBlock), 1024);
  hipLaunchKernelGGL(k_max_row, dim3(grid), dim3(kBlock), 0, s, row_prod, R,
                     (unsigned long long*)out);
}

// --------------------------------------------------------------------------
// (d) ESC expand with load-balanced search.  The chunk's products
// [P0, P1) are cut into tiles of kExpandTile; each active X nonzero owns at
// least one product, so a tile spans at most kExpandTile+1 of them.  One
// global binary search per tile finds the first; their product offsets are
// staged in LDS and each product finds its owner with an LDS search.
// Product key = (row - r0) << ynf | (ykey & fmask); value = vx * vy (FP32,
// contraction.cuh:132-136).  Fused: digit histograms for the product sort
// and an FP64 running sum of products.
constexpr int kExpandItems = 8;
constexpr uint32_t kExpandTile = kBlock * kExpandItems;

// Products are lane-striped (product pb + e*kBlock + t goes to thread t), so
// the Y gathers and the product stores of a wave are contiguous runs; each
// product finds its owner by a binary search over the staged prefixes,
// starting from its previous product's owner.  HIST adds the fused digit
// histograms for the look-back sort (its LDS is only reserved then).
template <int RB, bool HIST>
__global__ __launch_bounds__(kBlock) void k_expand(ExpandArgs a) {
  constexpr int R = 1 << RB;
  __shared__ uint64_t s_pref[kExpandTile + 1];
  __shared__ uint32_t sh[HIST ? kMaxFusedPasses * R : 1];
  __shared__ uint64_t s_ja, s_jl;
  __shared__ double s_sum[kBlock / kWave];
  if (HIST)
    for (int i = threadIdx.x; i < a.passes * R; i += kBlock) sh[i] = 0;
  double vsum = 0.0;
  const uint64_t nprod = a.P1 - a.P0;
  const uint64_t ntiles = ceil_div(nprod, kExpandTile);
  uint64_t t0, t1, step;
  xcd_tiles(ntiles, a.remap, &t0, &t1, &step);
  for (uint64_t tile = t0; tile < t1; tile += step) {
    const uint64_t pb = a.P0 + tile * kExpandTile;
    const uint64_t pe = pb + kExpandTile < a.P1 ? pb + kExpandTile : a.P1;
    __syncthreads();
    // owners of the first and the last product (last j in [A0, A1) with
    // act_prefix[j] <= p; waves 0 and 1): only their prefixes are staged
    if (threadIdx.x < 2 * kWave) {
      const uint64_t j = wave_last_le(a.act_prefix, a.A0, a.A1, threadIdx.x < kWave ? pb : pe - 1);
      if (threadIdx.x == 0) s_ja = j;
      if (threadIdx.x == kWave) s_jl = j;
    }
    __syncthreads();
    const uint64_t ja = s_ja;
    const uint64_t need = s_jl - ja + 1;
    const uint32_t m = (uint32_t)(need < kExpandTile + 1 ? need : kExpandTile + 1);
    for (uint32_t k = threadIdx.x; k < m; k += kBlock) s_pref[k] = a.act_prefix[ja + k];
    __syncthreads();
    const uint64_t obase = pb - a.P0;
    uint32_t klo = 0, kc = ~0u;
    uint64_t ybase = 0, rbits = 0;
    float xv = 0.f;
#pragma unroll
    for (int e = 0; e < kExpandItems; ++e) {
      const uint64_t p = pb + (uint64_t)e * kBlock + threadIdx.x;
      const bool valid = p < pe;
      uint64_t key = 0;
      if (valid) {
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
        if (k != kc) {  // new owner
          kc = k;
          const uint64_t j = ja + k;
          ybase = a.act_ypos[j] - s_pref[k];
          xv = a.act_xval[j];
          rbits = shl64((uint64_t)(a.act_row[j] - a.r0), a.ynf);
        }
        const uint64_t y = ybase + p;
        const float v = xv * a.yvals[y];
        const uint64_t col = a.ycol ? (uint64_t)a.ycol[y] : (a.ykeys[y] & a.fmask);
        key = rbits | col;
        a.out_keys[obase + (p - pb)] = key;
        a.out_vals[obase + (p - pb)] = v;
        vsum += (double)v;
      }
      if (HIST) hist_add_wave<RB>(sh, key, valid, a.passes);
    }
  }
  __syncthreads();
  if (HIST) hist_flush<RB>(sh, a.hist, a.passes);
  for (int off = kWave / 2; off > 0; off >>= 1) vsum += __shfl_xor(vsum, off, kWave);
  if (lane_id() == 0) s_sum[threadIdx.x / kWave] = vsum;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < kBlock / kWave; ++w) t += s_sum[w];
    atomicAdd(a.vsum, t);
  }
}

__global__ void k_ycol(const uint64_t* __restrict__ ykeys, uint64_t ny, uint64_t fmask,
                       uint32_t* __restrict__ ycol) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < ny; i += stride) ycol[i] = (uint32_t)(ykeys[i] & fmask);
}

void launch_ycol(const uint64_t* ykeys, uint64_t ny, uint64_t fmask, uint32_t* ycol,
                 hipStream_t s) {
  if (!ny) return;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(ny, kBlock), 4096);
  hipLaunchKernelGGL(k_ycol, dim3(grid), dim3(kBlock), 0, s, ykeys, ny, fmask, ycol);
}

uint32_t xcd_remap_enabled() {
  static const uint32_t on = [] {
    const char* e = getenv("SPTC_XCD_REMAP");
    return (uint32_t)!(e && std::atoi(e) == 0);
  }();
  return on;
}

void launch_expand(const ExpandArgs& a, int rb, hipStream_t s) {
  const uint64_t nprod = a.P1 - a.P0;
  if (!nprod) return;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(nprod, kExpandTile), 8192);
  ExpandArgs b = a;
  b.remap = xcd_remap_enabled();
  if (!a.passes) {
    hipLaunchKernelGGL((k_expand<8, false>), dim3(grid), dim3(kBlock), 0, s, b);
    return;
  }
  switch (rb) {
    case 9: hipLaunchKernelGGL((k_expand<9, true>), dim3(grid), dim3(kBlock), 0, s, b); break;
    case 10: hipLaunchKernelGGL((k_expand<10, true>), dim3(grid), dim3(kBlock), 0, s, b); break;
    case 11: hipLaunchKernelGGL((k_expand<11, true>), dim3(grid), dim3(kBlock), 0, s, b); break;
    default: hipLaunchKernelGGL((k_expand<8, true>), dim3(grid), dim3(kBlock), 0, s, b); break;
  }
}

// --------------------------------------------------------------------------
// (d) Reduce by key over sorted products: one output per run of equal keys;
// the run's values are summed in sorted (stable) order with compensated FP32
// arithmetic (cub::DeviceReduce::ReduceByKey(Sum), contraction.cuh:512-537).
struct ReduceOp {
  const uint64_t* keys;
  const float* vals;
  uint64_t n;
  uint32_t* out_row;
  uint64_t* out_col;
  float* out_val;
  uint64_t r0;
  int ynf;             // width of the key's column field
  uint64_t fmask;      // its mask
  const uint64_t* ytab;  // rank -> Y free key (NULL: the field is the key)
  NarrowRows nr;         // narrow form (keys unused): columns + row starts
  const uint32_t* ytab32 = nullptr;  // ytab as 32-bit words (narrow emit; keys < 2^32)
};

// Per tile of kScanTile sorted keys: the number of run heads (key differs
// from its predecessor) in tile_sums[tile].a, the tiling of the emit below.
// Items are lane-striped (coalesced loads); the predecessor comes from the
// neighbouring lane.
__global__ __launch_bounds__(kBlock) void k_reduce_heads(const uint64_t* __restrict__ keys,
                                                         uint64_t n, U2* __restrict__ tile_sums) {
  __shared__ U2 scratch[kBlock / kWave];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  uint64_t c = 0;
#pragma unroll
  for (int e = 0; e < kScanItems; ++e) {
    const uint64_t i = base + (uint64_t)e * kBlock + threadIdx.x;
    const uint64_t k = i < n ? keys[i] : 0;
    uint64_t pk = __shfl_up(k, 1, kWave);
    if (lane_id() == 0) pk = i > 0 && i < n ? keys[i - 1] : 0;
    c += i < n && (i == 0 || k != pk);
  }
  U2 tot;
  block_excl_scan_u2(mk2(c, 0), scratch, &tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

void run_reduce_count(const uint64_t* keys, uint64_t n, U2* tmp, hipStream_t s) {
  const uint64_t nt = scan_tmp_count(n);
  if (n)
    hipLaunchKernelGGL(k_reduce_heads, dim3((uint32_t)nt), dim3(kBlock), 0, s, keys, n, tmp);
  scan_tiles_parts(tmp, n ? nt : 0, tmp + nt, tmp + nt + 1, s);
}

// Emit for a tile of kScanTile sorted products (same tiling as the count
// pass, whose scanned tile sums give each tile's first output slot); the
// tile's outputs are stored coalesced.
__global__ __launch_bounds__(kBlock) void k_reduce_emit(ReduceOp op, const U2* tile_offs) {
  __shared__ ReduceSmem sm;
  uint64_t prev;
  const uint32_t tile = blockIdx.x;
  const uint32_t nout = reduce_tile(op.keys, op.vals, op.n, tile, sm, &prev);
  const uint64_t obase = tile_offs[tile].a;
  // the thread's column-table gathers are all issued before its stores
  const uint64_t* __restrict__ ytab = op.ytab;
  uint32_t* __restrict__ orow = op.out_row + obase;
  uint64_t* __restrict__ ocol = op.out_col + obase;
  float* __restrict__ oval = op.out_val + obase;
  uint64_t key[kRedRunsPerThread], col[kRedRunsPerThread];
#pragma unroll
  for (int q = 0; q < kRedRunsPerThread; ++q) {
    const uint32_t i = threadIdx.x + (uint32_t)q * kBlock;
    key[q] = i < nout ? sm.k[i] : 0;
    col[q] = key[q] & op.fmask;
    if (ytab && i < nout) col[q] = ytab[col[q]];
  }
#pragma unroll
  for (int q = 0; q < kRedRunsPerThread; ++q) {
    const uint32_t i = threadIdx.x + (uint32_t)q * kBlock;
    if (i < nout) {  // streaming stores
      __builtin_nontemporal_store((uint32_t)(op.r0 + shr64(key[q], op.ynf)), &orow[i]);
      __builtin_nontemporal_store(col[q], &ocol[i]);
      __builtin_nontemporal_store(sm.v[i], &oval[i]);
    }
  }
}

void run_reduce_emit(const uint64_t* keys, const float* vals, uint64_t n, uint32_t* out_row,
                     uint64_t* out_col, float* out_val, uint64_t r0, int ynf,
                     uint64_t fmask, const uint64_t* ytab, const U2* tmp, hipStream_t s) {
  if (!n) return;
  ReduceOp op{keys, vals, n, out_row, out_col, out_val, r0, ynf, fmask, ytab, NarrowRows{}};
  hipLaunchKernelGGL(k_reduce_emit, dim3((uint32_t)scan_tmp_count(n)), dim3(kBlock), 0, s, op,
                     tmp);
}

// Narrow-form emit, one wave per wave tile (the count pass's tiling).  The
// workgroup's four wave tiles are consecutive, so their outputs are one
// contiguous run: after a barrier the workgroup stores that run together,
// column-table gathers issued first, with streaming stores whose lanes are
// counted from the 32-output boundary at or below the run's start, so every
// store instruction covers whole 128-byte lines except at the run's ends
// (streaming stores do not merge partial lines in L2).
__global__ __launch_bounds__(kBlock) void k_reduce_emit_narrow(ReduceOp op, const U2* tile_offs,
                                                               uint64_t ntiles) {
  constexpr uint32_t NW = kBlock / kWave;
  __shared__ WaveSmem sm[NW];
  __shared__ uint32_t s_T[NW], s_rb[NW], s_rows[NW];
  const uint32_t wave = wave_id();
  const uint64_t wt0 = (uint64_t)blockIdx.x * NW;
  const uint64_t wt = wt0 + wave;
  uint32_t T = 0, rb = 0;
  bool rows = false;
  if (wt < ntiles) {
    const uint64_t b = wt * kWaveTile;
    const uint32_t cnt = (uint32_t)(op.n - b < kWaveTile ? op.n - b : kWaveTile);
    // a tile inside one big row has no outputs here
    if (!wave_tile_all_big(op.nr, b, cnt, op.nr.tile_row[2 * wt])) {
      uint64_t r0, rp;
      T = wave_tile_reduce(op.nr, op.vals, op.n, wt, sm[wave], &r0, &rp, &rows);
      rb = (uint32_t)(op.r0 + r0);
    }
  }
  if (lane_id() == 0) {
    s_T[wave] = T;
    s_rb[wave] = rb;
    s_rows[wave] = rows;
  }
  __syncthreads();
  uint32_t pre[NW + 1];
  pre[0] = 0;
#pragma unroll
  for (uint32_t k = 0; k < NW; ++k) pre[k + 1] = pre[k] + s_T[k];
  const uint32_t total = pre[NW];
  const uint64_t obase = tile_offs[wt0].a;  // the first wave tile's outputs start the run
  const uint64_t abase = obase & ~31ull;     // a 128-byte line of rows and values
  const uint32_t lead = (uint32_t)(obase - abase);
  const uint64_t* __restrict__ ytab = op.ytab;
  const uint32_t* __restrict__ ytab32 = op.ytab32;
  uint32_t* __restrict__ orow = op.out_row + abase;  // uniform bases
  uint64_t* __restrict__ ocol = op.out_col + abase;
  float* __restrict__ oval = op.out_val + abase;
  // slots [lead, lead + total) of at most 31 + 4 * 512
  constexpr int kPer = (int)((NW * kWaveTile + 31 + kBlock - 1) / kBlock);
  uint64_t col[kPer];
  uint32_t row[kPer];
  float val[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const uint32_t jj = threadIdx.x + (uint32_t)q * kBlock;
    const uint32_t j = jj - lead;  // wraps below lead: fails j < total
    col[q] = 0;
    row[q] = 0;
    val[q] = 0.f;
    if (jj >= lead && j < total) {
      uint32_t k = 0;
#pragma unroll
      for (uint32_t e = 1; e < NW; ++e) k += j >= pre[e] ? 1u : 0u;
      const uint32_t i = j - pre[k];
      const WaveSmem& w = sm[k];
      row[q] = s_rows[k] ? s_rb[k] + w.m[i] : s_rb[k];
      val[q] = w.v[i];
      const uint32_t c = w.c[i];
      col[q] = ytab32 ? (uint64_t)ytab32[c] : ytab ? ytab[c] : (uint64_t)c;  // half the gathered table
    }
  }
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const uint32_t jj = threadIdx.x + (uint32_t)q * kBlock;
    if (jj >= lead && jj - lead < total) {
      __builtin_nontemporal_store(row[q], &orow[jj]);
      __builtin_nontemporal_store(col[q], &ocol[jj]);
      __builtin_nontemporal_store(val[q], &oval[jj]);
    }
  }
}

static void check_narrow(const NarrowRows& nr, const float* vals) {
  if ((reinterpret_cast<uintptr_t>(nr.cols) | reinterpret_cast<uintptr_t>(vals)) & 15)
    throw std::logic_error("narrow products must be 16-byte aligned");
}

void run_reduce_emit_narrow(const NarrowRows& nr, const float* vals, uint64_t n,
                            uint32_t* out_row, uint64_t* out_col, float* out_val, uint64_t r0,
                            const uint64_t* ytab, const uint32_t* ytab32, const U2* tmp,
                            hipStream_t s) {
  if (!n) return;
  check_narrow(nr, vals);
  ReduceOp op{nullptr, vals, n, out_row, out_col, out_val, r0, 32, 0xFFFFFFFFull, ytab, nr,
              ytab32};
  const uint64_t nt = narrow_tile_count(n);
  hipLaunchKernelGGL(k_reduce_emit_narrow, dim3((uint32_t)ceil_div(nt, kBlock / kWave)),
                     dim3(kBlock), 0, s, op, tmp, nt);
}

// Narrow-form count: per wave tile, the number of run heads in
// tile_sums[wt].a (wave_tile_heads).
__global__ __launch_bounds__(kBlock) void k_reduce_heads_narrow(NarrowRows nr, uint64_t n,
                                                                uint64_t ntiles,
                                                                U2* __restrict__ tile_sums) {
  __shared__ __align__(16) uint32_t mark[kBlock / kWave][kWaveTile];
  const uint64_t wt = (uint64_t)blockIdx.x * (kBlock / kWave) + wave_id();
  if (wt >= ntiles) return;
  const uint64_t b = wt * kWaveTile;
  const uint32_t cnt = (uint32_t)(n - b < kWaveTile ? n - b : kWaveTile);
  const uint64_t r0 = nr.tile_row[2 * wt];
  if (wave_tile_all_big(nr, b, cnt, r0)) {
    if (lane_id() == 0) tile_sums[wt] = mk2(0, 0);
    return;
  }
  uint32_t col[kScanItems], mk[kScanItems];
  float unused[kScanItems];
  wave_tile_load<false>(nr.cols + b, nullptr, cnt, kScanItems * lane_id(), col, unused);
  bool rowsin;
  uint32_t hm = wave_tile_heads(nr, b, cnt, r0, col, mark[wave_id()], mk, &rowsin);
  if (nr.big) {
    uint32_t ro[kScanItems];
    wave_tile_rows(mk, rowsin, ro);
    hm = wave_emit_mask(nr, r0, rowsin, ro, hm);
  }
  const uint32_t tot = __builtin_amdgcn_readlane(wave_incl_scan32((uint32_t)__popc(hm)),
                                                 kWave - 1);
  if (lane_id() == 0) tile_sums[wt] = mk2(tot, 0);
}

// One thread per wave tile t: [2t] = row of its first product b = 512 t,
// [2t + 1] = row of product b - 1 (t > 0); the row of product p is the last
// row r with rstart[r] <= p (empty rows share their start with the next row).
__device__ __forceinline__ uint32_t row_of(const uint64_t* __restrict__ rstart, uint64_t nrows,
                                           uint64_t p) {
  uint64_t lo = 0, hi = nrows;  // first r with rstart[r] > p, minus one
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (rstart[mid] <= p) lo = mid + 1; else hi = mid;
  }
  return (uint32_t)(lo - 1);
}

__global__ __launch_bounds__(kBlock) void k_tile_rows(const uint64_t* __restrict__ rstart,
                                                      uint64_t nrows, uint64_t ntiles,
                                                      uint32_t* __restrict__ tile_row) {
  const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= ntiles) return;
  const uint64_t b = t * kWaveTile;
  tile_row[2 * t] = row_of(rstart, nrows, b);
  if (t) tile_row[2 * t + 1] = row_of(rstart, nrows, b - 1);
}

void launch_tile_rows(const uint64_t* rstart, uint64_t nrows, uint64_t n, uint32_t* tile_row,
                      hipStream_t s) {
  if (!nrows || !n) return;
  const uint64_t nt = narrow_tile_count(n);
  hipLaunchKernelGGL(k_tile_rows, dim3((uint32_t)ceil_div(nt, kBlock)), dim3(kBlock), 0, s,
                     rstart, nrows, nt, tile_row);
}

void run_reduce_count_narrow(const NarrowRows& nr, uint64_t n, U2* tmp, hipStream_t s,
                             bool scan) {
  const uint64_t nt = narrow_tile_count(n);
  if (n) {
    check_narrow(nr, nullptr);
    hipLaunchKernelGGL(k_reduce_heads_narrow, dim3((uint32_t)ceil_div(nt, kBlock / kWave)),
                       dim3(kBlock), 0, s, nr, n, nt, tmp);
  }
  if (scan) scan_narrow_tiles(n, tmp, s);
}

void scan_narrow_tiles(uint64_t n, U2* tmp, hipStream_t s) {
  const uint64_t nt = narrow_tile_count(n);
  scan_tiles_parts(tmp, n ? nt : 0, tmp + nt, tmp + nt + 1, s);
}

// --------------------------------------------------------------------------
// (d) Dense per-row accumulation for small column spaces (E = 2^EB slots,
// the rank space of Y's free keys).  One workgroup per X row.  The row's
// active X nonzeros are taken in order; each one's Y segment is spread over
// the threads and added into the dense LDS row with compensated FP32
// arithmetic, a barrier separating segments, so every slot sums its products
// in the same order as the ESC path's stable sort (Y free keys are distinct
// within a segment when Y has no duplicate coordinates, which the caller
// checks).  A first, symbolic kernel counts each row's touched slots so the
// numeric kernel writes its outputs, in column order, straight to their
// final place.


template <int EB>
__global__ __launch_bounds__(kBlock) void k_dense_count(DenseArgs a) {
  __shared__ DenseSmem<EB> sm;
  __shared__ uint64_t s_ja;
  const uint64_t r = a.r0 + blockIdx.x;
  const uint32_t t = dense_row_count<EB>(a, a.row_act[r], a.row_act[r + 1], a.fmask, sm, &s_ja);
  if (threadIdx.x == 0) a.row_cnt[blockIdx.x] = t;
}

template <int EB, bool ORDERED>
__global__ __launch_bounds__(kBlock) void k_dense_numeric(DenseArgs a) {
  constexpr int E = 1 << EB;
  __shared__ double s_acc[E];
  __shared__ DenseSmem<EB> sm;
  __shared__ uint64_t s_ja;
  const uint64_t r = a.r0 + blockIdx.x;
  dense_row_numeric<EB, ORDERED>(a, a.row_act[r], a.row_act[r + 1], a.fmask, a.ytab,
                                 (uint32_t)r, a.out_off[blockIdx.x], s_acc, sm, &s_ja);
}

// Ordered (bit-identical to ESC) unless SPTC_DENSE_ORDER=atomic selects the
// FP64 LDS-atomic accumulation.
bool dense_ordered() {
  static const bool ord = [] {
    const char* e = getenv("SPTC_DENSE_ORDER");
    return !(e && std::string(e) == "atomic");
  }();
  return ord;
}

void launch_dense_count(const DenseArgs& a, int eb, uint64_t rows, hipStream_t s) {
  if (!rows) return;
  switch (eb) {
    case 8: hipLaunchKernelGGL(k_dense_count<8>, dim3((uint32_t)rows), dim3(kBlock), 0, s, a); break;
    case 10: hipLaunchKernelGGL(k_dense_count<10>, dim3((uint32_t)rows), dim3(kBlock), 0, s, a); break;
    default: hipLaunchKernelGGL(k_dense_count<13>, dim3((uint32_t)rows), dim3(kBlock), 0, s, a); break;
  }
}

void launch_dense_numeric(const DenseArgs& a, int eb, uint64_t rows, hipStream_t s) {
  if (!rows) return;
  const dim3 g((uint32_t)rows), b(kBlock);
  if (dense_ordered()) {
    switch (eb) {
      case 8: hipLaunchKernelGGL((k_dense_numeric<8, true>), g, b, 0, s, a); break;
      case 10: hipLaunchKernelGGL((k_dense_numeric<10, true>), g, b, 0, s, a); break;
      default: hipLaunchKernelGGL((k_dense_numeric<13, true>), g, b, 0, s, a); break;
    }
  } else {
    switch (eb) {
      case 8: hipLaunchKernelGGL((k_dense_numeric<8, false>), g, b, 0, s, a); break;
      case 10: hipLaunchKernelGGL((k_dense_numeric<10, false>), g, b, 0, s, a); break;
      default: hipLaunchKernelGGL((k_dense_numeric<13, false>), g, b, 0, s, a); break;
    }
  }
}

// Any two equal adjacent keys in a sorted array (duplicate coordinates).
__global__ __launch_bounds__(kBlock) void k_any_dup(const uint64_t* __restrict__ keys, uint64_t n,
                                                    uint32_t* flag) {
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x + 1;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  bool dup = false;
  for (; i < n; i += stride) dup |= keys[i] == keys[i - 1];
  if (__ballot(dup) && lane_id() == 0) atomicOr(flag, 1u);
}

void launch_any_dup(const uint64_t* keys, uint64_t n, uint32_t* flag, hipStream_t s) {
  fill_u32(flag, 0, 1, s);
  if (n < 2) return;
  uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(n, kBlock), 2048);
  hipLaunchKernelGGL(k_any_dup, dim3(grid), dim3(kBlock), 0, s, keys, n, flag);
}

// Exclusive scan of per-row uint32 counts into uint64 offsets (+ total).
struct RowCountScanOp {
  const uint32_t* in;
  uint64_t* out;
  __device__ U2 value(uint64_t i) const { return mk2(in[i], 0); }
  __device__ void emit(uint64_t i, U2 excl, U2) const { out[i] = excl.a; }
};

void run_row_count_scan(const uint32_t* cnt, uint64_t* off, uint64_t n, U2* tmp, hipStream_t s) {
  run_scan(RowCountScanOp{cnt, off}, n, tmp, s);
}

// --------------------------------------------------------------------------
// Result decode helpers.
__global__ __launch_bounds__(kBlock) void k_pack_keys(const uint32_t* row, const uint64_t* col,
                                                      const uint64_t* row_key, uint64_t n,
                                                      int ynf, uint64_t* out) {
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; i < n; i += stride) out[i] = shl64(row_key[row[i]], ynf) | col[i];
}

__global__ __launch_bounds__(kBlock) void k_gather_rowkey(const uint32_t* row,
                                                          const uint64_t* row_key,
                                                          uint64_t n, uint64_t* out) {
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; i < n; i += stride) out[i] = row_key[row[i]];
}

void launch_pack_keys(const uint32_t* row, const uint64_t* col, const uint64_t* row_key,
                      uint64_t n, int ynf, uint64_t* out, hipStream_t s) {
  if (!n) return;
  uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(n, kBlock), 4096);
  hipLaunchKernelGGL(k_pack_keys, dim3(grid), dim3(kBlock), 0, s, row, col, row_key, n, ynf, out);
}
void launch_gather_rowkey(const uint32_t* row, const uint64_t* row_key, uint64_t n,
                          uint64_t* out, hipStream_t s) {
  if (!n) return;
  uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(n, kBlock), 4096);
  hipLaunchKernelGGL(k_gather_rowkey, dim3(grid), dim3(kBlock), 0, s, row, row_key, n, out);
}

}  // namespace sptc
