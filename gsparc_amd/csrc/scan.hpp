// Generic device-wide scan with a fused producer and consumer.
//
// An Op supplies, for item i of [0, n):
//   __device__ V value(uint64_t i) const;            // the item's contribution
//   __device__ void emit(uint64_t i, V excl, V val);   // consumer (optional)
// where V is U2 (two uint64 lanes, added lane-wise).  An Op that declares a
// type Aux instead supplies value(i, Aux&) and emit(i, excl, val, const
// Aux&): what value() looked up reaches emit() in registers, not loaded
// again.  Three kernels:
//   k_scan_reduce : per tile of kBlock*kScanItems items, the tile sum
//   k_scan_tiles  : exclusive scan of tile sums (single workgroup), total
//   k_scan_emit   : recompute values, block scan + tile offset, emit()
// Values are computed and emitted tile-striped (item e*kBlock + t of a tile
// to thread t, so each of the kScanItems loads and stores of a workgroup is
// one contiguous run); the emit's scan goes through LDS, where thread t
// takes items [t*kScanItems, t*kScanItems + kScanItems) in order.
#pragma once
#include "common.hpp"

#include <type_traits>

namespace sptc {

constexpr int kScanItems = 8;
constexpr uint64_t kScanTile = (uint64_t)kBlock * kScanItems;

struct U2 {
  uint64_t a, b;
};
__host__ __device__ __forceinline__ U2 mk2(uint64_t a, uint64_t b) {
  U2 r;
  r.a = a;
  r.b = b;
  return r;
}
__host__ __device__ __forceinline__ U2 operator+(const U2& x, const U2& y) {
  return mk2(x.a + y.a, x.b + y.b);
}
__host__ __device__ __forceinline__ U2 operator-(const U2& x, const U2& y) {
  return mk2(x.a - y.a, x.b - y.b);
}

__device__ __forceinline__ U2 block_excl_scan_u2(U2 v, U2* scratch, U2* total) {
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  U2 inc = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint64_t oa = __shfl_up(inc.a, d, kWave);
    uint64_t ob = __shfl_up(inc.b, d, kWave);
    if (lane >= (uint32_t)d) { inc.a += oa; inc.b += ob; }
  }
  if (lane == kWave - 1) scratch[wave] = inc;
  __syncthreads();
  U2 off = mk2(0, 0), tot = mk2(0, 0);
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) {
    U2 s = scratch[w];
    if ((uint32_t)w < wave) off = off + s;
    tot = tot + s;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// Op::Aux when the Op declares one (see above).
template <class Op, class = void>
struct ScanAux {
  struct type {};
  static constexpr bool on = false;
};
template <class Op>
struct ScanAux<Op, std::void_t<typename Op::Aux>> {
  using type = typename Op::Aux;
  static constexpr bool on = true;
};
template <class Op>
__device__ __forceinline__ U2 scan_value(const Op& op, uint64_t i,
                                         typename ScanAux<Op>::type& aux) {
  if constexpr (ScanAux<Op>::on) return op.value(i, aux);
  else return op.value(i);
}

template <class Op>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(Op op, uint64_t n, U2* tile_sums) {
  __shared__ U2 scratch[kBlock / kWave];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x;
  U2 s = mk2(0, 0);
#pragma unroll
  for (int e = 0; e < kScanItems; ++e) {
    const uint64_t i = base + (uint64_t)e * kBlock;
    typename ScanAux<Op>::type aux;
    if (i < n) s = s + scan_value(op, i, aux);
  }
  U2 tot;
  block_excl_scan_u2(s, scratch, &tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

template <class Op>
__global__ __launch_bounds__(kBlock) void k_scan_emit(Op op, uint64_t n, const U2* tile_offs) {
  __shared__ U2 scratch[kBlock / kWave];
  __shared__ U2 sv[kScanTile];  // values, then exclusive prefixes, by tile index
  const uint32_t tid = threadIdx.x;
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  U2 vals[kScanItems];
  typename ScanAux<Op>::type aux[kScanItems];
#pragma unroll
  for (int e = 0; e < kScanItems; ++e) {
    const uint64_t i = base + (uint64_t)e * kBlock + tid;
    vals[e] = i < n ? scan_value(op, i, aux[e]) : mk2(0, 0);
    sv[e * kBlock + tid] = vals[e];
  }
  __syncthreads();
  U2 mine[kScanItems];
  U2 s = mk2(0, 0);
#pragma unroll
  for (int e = 0; e < kScanItems; ++e) {
    mine[e] = sv[tid * kScanItems + e];
    s = s + mine[e];
  }
  U2 tot;  // (barriers inside: every thread's reads of sv are done after it)
  U2 excl = block_excl_scan_u2(s, scratch, &tot) + tile_offs[blockIdx.x];
#pragma unroll
  for (int e = 0; e < kScanItems; ++e) {
    sv[tid * kScanItems + e] = excl;
    excl = excl + mine[e];
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < kScanItems; ++e) {
    const uint64_t i = base + (uint64_t)e * kBlock + tid;
    if (i < n) {
      if constexpr (ScanAux<Op>::on) op.emit(i, sv[e * kBlock + tid], vals[e], aux[e]);
      else op.emit(i, sv[e * kBlock + tid], vals[e]);
    }
  }
}

// In-place exclusive scan of nt tile sums by one workgroup; total -> *total
// (defined in kernels.hip).
void scan_tiles(U2* sums, uint64_t nt, U2* total, hipStream_t s);

// Runs reduce -> tile scan -> emit.  `tmp` must hold scan_tmp_count(n)+1 U2
// entries; the grand total lands in tmp[scan_tmp_count(n)].
inline uint64_t scan_tmp_count(uint64_t n) { return ceil_div(n ? n : 1, kScanTile); }

// Long tile-sum arrays (one sum per 2048 products of a reduce count) are
// scanned by many workgroups, kScanSumsPerBlock sums each, with the block
// partials kept after the total: reduce_tmp_count(n) entries of tmp.
constexpr uint64_t kScanSumsPerBlock = 16384;
inline uint64_t reduce_tmp_count(uint64_t n) {
  const uint64_t nt = scan_tmp_count(n);
  return nt + 1 + ceil_div(nt, kScanSumsPerBlock) + 1;
}
// The narrow form's reduce (count and emit) works in wave tiles of
// kWaveTile products, kScanItems consecutive ones per lane: one tile sum
// per wave tile, the grand total at tmp[narrow_tile_count(n)], and
// narrow_tmp_count(n) entries of tmp in all.
constexpr uint64_t kWaveTile = (uint64_t)kWave * kScanItems;
inline uint64_t narrow_tile_count(uint64_t n) { return ceil_div(n ? n : 1, kWaveTile); }
inline uint64_t narrow_tmp_count(uint64_t n) {
  const uint64_t nt = narrow_tile_count(n);
  return nt + 1 + ceil_div(nt, kScanSumsPerBlock) + 1;
}
// Exclusive scan of nt tile sums in place, total -> *total; parts holds
// ceil(nt / kScanSumsPerBlock) + 1 entries (defined in kernels.hip).
void scan_tiles_parts(U2* sums, uint64_t nt, U2* total, U2* parts, hipStream_t s);

template <class Op>
void run_scan(const Op& op, uint64_t n, U2* tmp, hipStream_t s, bool emit = true) {
  const uint64_t nt = scan_tmp_count(n);
  U2* total = tmp + nt;
  if (n == 0) {
    scan_tiles(tmp, 0, total, s);
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce<Op>, dim3((uint32_t)nt), dim3(kBlock), 0, s, op, n, tmp);
  scan_tiles(tmp, nt, total, s);
  if (emit)
    hipLaunchKernelGGL(k_scan_emit<Op>, dim3((uint32_t)nt), dim3(kBlock), 0, s, op, n,
                       (const U2*)tmp);
}

}  // namespace sptc
