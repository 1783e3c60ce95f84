// Reduce-by-key tile core shared by the single-contraction and the batched
// emit kernels (cub::DeviceReduce::ReduceByKey(Sum) semantics,
// contraction.cuh:512-537): one output per run of equal sorted keys, the
// run's values summed in sorted order with compensated FP32 arithmetic.
#pragma once
#include "scan.hpp"

namespace sptc {

constexpr int kRedRunsPerThread = (int)(kScanTile / kBlock);  // 8

struct ReduceSmem {
  uint64_t k[kScanTile];  // the tile's keys, then its outputs' keys
  float v[kScanTile];     // the tile's values, then its outputs' sums
  uint8_t head[kScanTile];
  uint16_t rs[kScanTile + 1];  // tile-local start of each run, rs[nruns] = count
  U2 scratch[kBlock / kWave];
  uint64_t tail_end;
};

// Tile `tile` of kScanTile sorted products.  Runs are found from head
// flags (key differs from its predecessor), so each run's sum walks a known
// range: no key compares on the dependent chain, loads issued ahead of the
// adds.  Runs go to threads round-robin; the run that crosses the tile end
// continues in global memory up to an end the workgroup finds with
// coalesced probes.  Leaves the tile's outputs -- one per run starting in
// it, in key order -- in (sm.k, sm.v)[0 .. nout-1] and returns nout; *prev
// is the key before the tile (0 for the first tile).
__device__ __forceinline__ uint32_t reduce_tile(const uint64_t* __restrict__ keys,
                                                const float* __restrict__ vals, uint64_t n,
                                                uint32_t tile, ReduceSmem& sm, uint64_t* prev) {
  const uint32_t tid = threadIdx.x;
  const uint64_t base = (uint64_t)tile * kScanTile;
  const uint32_t cnt = (uint32_t)(n - base < kScanTile ? n - base : kScanTile);
  for (uint32_t i = tid; i < cnt; i += kBlock) {
    sm.k[i] = keys[base + i];
    sm.v[i] = vals[base + i];
  }
  const uint64_t prev_key = base ? keys[base - 1] : 0;
  *prev = prev_key;
  __syncthreads();
  for (uint32_t i = tid; i < kScanTile; i += kBlock) {
    uint8_t h = 0;
    if (i < cnt) h = (base + i == 0) || sm.k[i] != (i ? sm.k[i - 1] : prev_key);
    sm.head[i] = h;
  }
  __syncthreads();
  // thread t's flags are bytes [8t, 8t+8): run indices by a block scan
  const uint64_t f = reinterpret_cast<const uint64_t*>(sm.head)[tid];
  U2 tot;
  uint32_t r = (uint32_t)block_excl_scan_u2(mk2((uint64_t)__popcll(f), 0), sm.scratch, &tot).a;
  const uint32_t nruns = (uint32_t)tot.a;
  for (uint64_t m = f; m; m &= m - 1) sm.rs[r++] = (uint16_t)(tid * 8 + (__builtin_ctzll(m) >> 3));
  if (tid == 0) {
    sm.rs[nruns] = (uint16_t)cnt;
    sm.tail_end = base + cnt;
  }
  __syncthreads();
  // the last run may continue past the tile: first index with another key
  if (nruns && base + cnt < n && tid < kWave) {
    const uint64_t last = sm.k[sm.rs[nruns - 1]];
    for (uint64_t g0 = base + cnt;; g0 += kWave) {
      const uint64_t g = g0 + tid;
      const bool stop = g >= n || keys[g] != last;
      const uint64_t m = __ballot(stop);
      if (m) {
        if (tid == 0) sm.tail_end = g0 + __builtin_ctzll(m);
        break;
      }
    }
  }
  __syncthreads();
  const uint64_t tail_end = sm.tail_end;
  uint64_t ok[kRedRunsPerThread];
  float ov[kRedRunsPerThread];
#pragma unroll
  for (int q = 0; q < kRedRunsPerThread; ++q) {
    const uint32_t run = tid + (uint32_t)q * kBlock;
    ok[q] = 0;
    ov[q] = 0.f;
    if (run >= nruns) continue;
    const uint32_t j0 = sm.rs[run], j1 = sm.rs[run + 1];
    KahanF32 acc;
    uint32_t j = j0;
    for (; j + 4 <= j1; j += 4) {
      const float a = sm.v[j], b = sm.v[j + 1], c = sm.v[j + 2], d = sm.v[j + 3];
      acc.add(a);
      acc.add(b);
      acc.add(c);
      acc.add(d);
    }
    for (; j < j1; ++j) acc.add(sm.v[j]);
    if (run + 1 == nruns) {
      for (uint64_t g = base + cnt; g < tail_end; g += 8) {
        float t[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) t[e] = g + e < tail_end ? vals[g + e] : 0.f;
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (g + e < tail_end) acc.add(t[e]);
      }
    }
    ok[q] = sm.k[j0];
    ov[q] = acc.value();
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kRedRunsPerThread; ++q) {
    const uint32_t run = tid + (uint32_t)q * kBlock;
    if (run < nruns) {
      sm.k[run] = ok[q];
      sm.v[run] = ov[q];
    }
  }
  __syncthreads();
  return nruns;
}

// ---- narrow form: wave tiles ----------------------------------------------
// The narrow form is reduced in wave tiles of kWaveTile sorted products,
// lane l holding products [8l, 8l + 8) of its tile in registers; every wave
// works alone (no workgroup barriers).  A product starts a run (a head)
// when it is the first product, a row starts at it or its column differs
// from its predecessor's, as the 64-bit keys (row << 32 | column) would
// give.  Each run is summed in sorted order with compensated FP32
// arithmetic by the lane holding its head; a run that continues past the
// lane reads the later lanes' values from LDS and, past the tile, global
// memory up to another column or its row's end.
static_assert(kScanItems == 8, "wave tiles hold 8 products per lane");
// (may_alias: LDS words are written as vectors and read one by one)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4), __may_alias__));
typedef float f32x4 __attribute__((ext_vector_type(4), __may_alias__));

struct alignas(16) WaveSmem {
  float v[kWaveTile];     // the tile's values, then its outputs' sums
  uint32_t c[kWaveTile];  // the outputs' columns
  uint32_t m[kWaveTile];  // row-start marks, then the outputs' rows (- r0)
};

// The lane's 8 products q0 .. q0 + 7 of the tile (cnt products at cols,
// vals; uniform bases, so the addresses are 32-bit offsets): streaming
// vector loads when all are in the tile (cols and vals 16-byte aligned, q0
// a multiple of 8), else one by one.
template <bool VALS>
__device__ __forceinline__ void wave_tile_load(const uint32_t* __restrict__ cols,
                                               const float* __restrict__ vals, uint32_t cnt,
                                               uint32_t q0, uint32_t (&col)[kScanItems],
                                               float (&val)[kScanItems]) {
  if (q0 + kScanItems <= cnt) {
    const u32x4* cp = reinterpret_cast<const u32x4*>(cols + q0);
    const u32x4 c0 = __builtin_nontemporal_load(cp), c1 = __builtin_nontemporal_load(cp + 1);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      col[j] = c0[j];
      col[4 + j] = c1[j];
    }
    if (VALS) {
      const f32x4* vp = reinterpret_cast<const f32x4*>(vals + q0);
      const f32x4 v0 = __builtin_nontemporal_load(vp), v1 = __builtin_nontemporal_load(vp + 1);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        val[j] = v0[j];
        val[4 + j] = v1[j];
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
      col[j] = q0 + j < cnt ? cols[q0 + j] : 0;
      if (VALS) val[j] = q0 + j < cnt ? vals[q0 + j] : 0.f;
    }
  }
}

// Head mask of the lane's products of the wave tile at b (cnt products,
// the first in row r0): bit j set when product 8 * lane + j starts a run.
// When rows start after the tile's first product (*rowsin), mk[j] = row - r0
// of the last row starting at the lane's product j (0: none), and `mark`
// (the wave's kWaveTile words of LDS) holds them for the whole tile.
__device__ __forceinline__ uint32_t wave_tile_heads(const NarrowRows& nr, uint64_t b,
                                                    uint32_t cnt, uint64_t r0,
                                                    const uint32_t (&col)[kScanItems],
                                                    uint32_t* __restrict__ mark,
                                                    uint32_t (&mk)[kScanItems], bool* rowsin) {
  const uint32_t lane = lane_id();
  const uint32_t q0 = kScanItems * lane;
  const uint32_t pc0 = b ? nr.cols[b - 1] : 0;  // same address in every lane
  const bool at0 = nr.rstart[r0] == b;
  // lane l's candidate: row r0 + 1 + l (later rows start after product b)
  const uint64_t r1 = r0 + 1 + lane;
  const uint64_t st1 = r1 < nr.nrows ? nr.rstart[r1] : ~0ull;
  const bool any = __ballot(st1 < b + cnt) != 0;
  *rowsin = any;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) mk[j] = 0;
  if (any) {
    u32x4* mp = reinterpret_cast<u32x4*>(mark + q0);
    mp[0] = u32x4{0u, 0u, 0u, 0u};
    mp[1] = u32x4{0u, 0u, 0u, 0u};
    wave_sync();
    for (uint64_t r = r1; r < nr.nrows; r += kWave) {
      const uint64_t st = r == r1 ? st1 : nr.rstart[r];
      if (st >= b + cnt) break;
      if (st > b) atomicMax(&mark[st - b], (uint32_t)(r - r0));  // empty rows: the last wins
    }
    wave_sync();
    const u32x4 m0 = mp[0], m1 = mp[1];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      mk[j] = m0[j];
      mk[4 + j] = m1[j];
    }
  }
  uint32_t prev = __shfl_up(col[kScanItems - 1], 1, kWave);
  if (lane == 0) prev = pc0;
  // the first product of all (b = 0, lane 0, j = 0) and a row starting at
  // the tile's first product count as row starts
  const bool first = lane == 0 && (b == 0 || at0);
  uint32_t hm = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const uint32_t pk = j ? col[j - 1] : prev;
    const bool rs = mk[j] != 0 || (j == 0 && first);
    hm |= (uint32_t)(q0 + j < cnt && (rs || col[j] != pk)) << j;
  }
  return hm;
}

// Row of each of the lane's products as an offset from the tile's first row
// r0: a prefix maximum of the marks mk (lane, then wave); all 0 when no row
// starts inside the tile.
__device__ __forceinline__ void wave_tile_rows(const uint32_t (&mk)[kScanItems], bool rowsin,
                                               uint32_t (&ro)[kScanItems]) {
  if (!rowsin) {
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) ro[j] = 0;
    return;
  }
  const uint32_t lane = lane_id();
  uint32_t pm = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    pm = mk[j] > pm ? mk[j] : pm;
    ro[j] = pm;
  }
  uint32_t inc = pm;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d, kWave);
    if (lane >= (uint32_t)d) inc = o > inc ? o : inc;
  }
  uint32_t ex = __shfl_up(inc, 1, kWave);
  if (lane == 0) ex = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) ro[j] = ro[j] > ex ? ro[j] : ex;
}

// The heads of hm that produce outputs here: all of them unless rows are
// marked big (nr.big), whose heads only end the run before them.
__device__ __forceinline__ uint32_t wave_emit_mask(const NarrowRows& nr, uint64_t r0,
                                                   bool rowsin,
                                                   const uint32_t (&ro)[kScanItems],
                                                   uint32_t hm) {
  if (!nr.big || !hm) return hm;
  if (!rowsin) return nr.big[r0] ? 0u : hm;
  uint32_t em = hm;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j)
    if (((hm >> j) & 1u) && nr.big[r0 + ro[j]]) em &= ~(1u << j);
  return em;
}

// True when all cnt products of the wave tile at b lie in row r0 and that
// row is big (uniform across the wave).
__device__ __forceinline__ bool wave_tile_all_big(const NarrowRows& nr, uint64_t b, uint32_t cnt,
                                                  uint64_t r0) {
  return nr.big && nr.big[r0] && (r0 + 1 >= nr.nrows || nr.rstart[r0 + 1] >= b + cnt);
}

// Compensated sums of the lane's products: res[j] = the running sum after
// product j of the run it belongs to (runs restart at the heads); acc is
// left holding the lane's last run.  FULL: all 8 products are in the tile.
template <bool FULL>
__device__ __forceinline__ void wave_lane_sums(uint32_t hm, uint32_t q0, uint32_t cnt,
                                               const float (&val)[kScanItems], KahanF32& acc,
                                               float (&res)[kScanItems]) {
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    if ((hm >> j) & 1u) acc = KahanF32();
    if (FULL || q0 + j < cnt) acc.add(val[j]);
    res[j] = acc.value();
  }
}

// Reduces wave tile wt: leaves its outputs -- one per head in the tile, in
// key order -- in (w.c, w.m, w.v)[0 .. T-1] (column, row - r0, sum) and
// returns T; *r0 = row of the tile's first product, *rp = row of the
// product before it.  When no row starts after the tile's first product
// (*rows false, most tiles of long rows) every output is in row r0 and the
// row offsets are not staged.
__device__ __forceinline__ uint32_t wave_tile_reduce(const NarrowRows& nr,
                                                     const float* __restrict__ vals, uint64_t n,
                                                     uint64_t wt, WaveSmem& w, uint64_t* r0p,
                                                     uint64_t* rpp, bool* rows) {
  const uint32_t lane = lane_id();
  const uint64_t b = wt * kWaveTile;
  const uint32_t cnt = (uint32_t)(n - b < kWaveTile ? n - b : kWaveTile);
  const uint64_t r0 = nr.tile_row[2 * wt];
  *r0p = r0;
  *rpp = b ? nr.tile_row[2 * wt + 1] : 0;
  const uint32_t q0 = kScanItems * lane;
  uint32_t col[kScanItems], mk[kScanItems];
  float val[kScanItems];
  wave_tile_load<true>(nr.cols + b, vals + b, cnt, q0, col, val);
  bool rowsin;
  const uint32_t hm = wave_tile_heads(nr, b, cnt, r0, col, w.m, mk, &rowsin);
  // row of each product: prefix maximum of the marks (lane, then wave)
  uint32_t ro[kScanItems];
  wave_tile_rows(mk, rowsin, ro);
  // output slots: exclusive scan of the lanes' emitting head counts
  const uint32_t em = wave_emit_mask(nr, r0, rowsin, ro, hm);
  const uint32_t h = (uint32_t)__popc(em);
  const uint32_t hinc = wave_incl_scan32(h);
  const uint32_t o = hinc - h;
  const uint32_t T = __builtin_amdgcn_readlane(hinc, kWave - 1);
  // first head after the lane's products (kWaveTile: none), suffix minimum
  uint32_t fs = hm ? q0 + (uint32_t)__builtin_ctz(hm) : (uint32_t)kWaveTile;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t x = __shfl_down(fs, d, kWave);
    if (lane + (uint32_t)d < kWave) fs = x < fs ? x : fs;
  }
  uint32_t nxt = __shfl_down(fs, 1, kWave);
  if (lane == kWave - 1) nxt = (uint32_t)kWaveTile;
  // the tile's last run may continue past it: up to another column or its
  // row's end (a full tile here, so its last product is lane 63's item 7)
  uint64_t tail_end = b + cnt;
  if (T && b + cnt < n) {
    const uint32_t lc = __builtin_amdgcn_readlane(col[kScanItems - 1], kWave - 1);
    const uint32_t lro = __builtin_amdgcn_readlane(ro[kScanItems - 1], kWave - 1);
    const uint64_t rend = nr.rstart[r0 + lro + 1];
    for (uint64_t g0 = b + cnt;; g0 += kWave) {
      const uint64_t g = g0 + lane;
      const uint32_t c = g < n ? nr.cols[g] : 0;
      const uint64_t m = __ballot(g >= n || c != lc || g >= rend);
      if (m) {
        tail_end = g0 + (uint64_t)__builtin_ctzll(m);
        break;
      }
    }
  }
  // values to LDS for the runs that continue into later lanes
  {
    f32x4* vp = reinterpret_cast<f32x4*>(w.v + q0);
    vp[0] = f32x4{val[0], val[1], val[2], val[3]};
    vp[1] = f32x4{val[4], val[5], val[6], val[7]};
  }
  wave_sync();
  KahanF32 acc;
  float res[kScanItems];
  if (cnt == kWaveTile) wave_lane_sums<true>(hm, q0, cnt, val, acc, res);
  else wave_lane_sums<false>(hm, q0, cnt, val, acc, res);
  float last = 0.f;
  if (hm) {  // the lane's last run: later lanes' products, then past the tile
    const uint32_t stop = nxt < cnt ? nxt : cnt;
    for (uint32_t q = q0 + kScanItems; q < stop; ++q) acc.add(w.v[q]);
    if (nxt >= cnt) {
      for (uint64_t g = b + cnt; g < tail_end; g += 8) {
        float t[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) t[e] = g + e < tail_end ? vals[g + e] : 0.f;
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (g + e < tail_end) acc.add(t[e]);
      }
    }
    last = acc.value();
  }
  wave_sync();  // every lane's LDS reads of values done
  uint32_t k = o;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    if ((em >> j) & 1u) {
      w.c[k] = col[j];
      if (rowsin) w.m[k] = ro[j];
      ++k;
    }
  }
  // a run ends before each later head of the lane; the lane's last run
  // ends with `last`; only emitting heads' runs take an output slot
  k = o;
  bool emits = false;  // the current run (started in this lane) emits
#pragma unroll
  for (int j = 0; j + 1 < kScanItems; ++j) {
    if ((hm >> j) & 1u) emits = (em >> j) & 1u;
    if (((hm >> (j + 1)) & 1u) && emits) w.v[k++] = res[j];
  }
  if (hm && ((em >> (31 - __builtin_clz(hm))) & 1u)) w.v[o + h - 1] = last;
  wave_sync();
  *rows = rowsin;
  return T;
}

}  // namespace sptc
