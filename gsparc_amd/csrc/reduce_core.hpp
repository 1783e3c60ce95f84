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

// reduce_tile for the narrow form: the tile's packed keys are rebuilt in LDS
// from the row starts (a mark at each row's first product, then a prefix
// maximum over the tile), the run crossing the tile end stops at its row's
// end as well.  Same outputs, keys (row << 32 | column).
__device__ __forceinline__ uint32_t reduce_tile_narrow(const NarrowRows& nr,
                                                       const float* __restrict__ vals,
                                                       uint64_t n, uint32_t tile,
                                                       ReduceSmem& sm, uint64_t* prev) {
  uint32_t* mark = reinterpret_cast<uint32_t*>(sm.v);  // aliases v until the keys are built
  const uint32_t tid = threadIdx.x;
  const uint64_t base = (uint64_t)tile * kScanTile;
  const uint32_t cnt = (uint32_t)(n - base < kScanTile ? n - base : kScanTile);
  // every load the tile needs first is issued at once: row table entries
  // (same address for all threads), columns, values (held in registers
  // while v holds the marks), the first candidate row start
  const uint64_t r0 = nr.tile_row[2 * (uint64_t)tile];
  const uint64_t rp = nr.tile_row[2 * (uint64_t)tile + 1];
  const uint64_t pcol = base ? nr.cols[base - 1] : 0;
  float vreg[kScanItems];
  const uint64_t r1 = r0 + 1 + tid;
  uint64_t st1 = r1 < nr.nrows ? nr.rstart[r1] : ~0ull;
#pragma unroll
  for (int e = 0; e < kScanItems; ++e) {
    const uint32_t i = (uint32_t)e * kBlock + tid;
    // streaming loads (non-temporal): they should not evict the column
    // table the emit gathers from
    if (i < cnt) sm.k[i] = r0 << 32 | __builtin_nontemporal_load(&nr.cols[base + i]);
    vreg[e] = i < cnt ? __builtin_nontemporal_load(&vals[base + i]) : 0.f;
  }
  // most tiles lie inside one row (thread 0's candidate is the next row):
  // their keys are complete; otherwise rows come from marks at the row
  // starts and a prefix maximum
  if (__syncthreads_or(st1 < base + cnt)) {
#pragma unroll
    for (int e = 0; e < kScanItems; ++e) mark[(uint32_t)e * kBlock + tid] = 0;
    __syncthreads();
    for (uint64_t r = r1; r < nr.nrows; r += kBlock) {
      const uint64_t st = r == r1 ? st1 : nr.rstart[r];
      if (st >= base + cnt) break;
      atomicMax(&mark[st - base], (uint32_t)(r - r0));
    }
    __syncthreads();
    // Wave w walks positions [w * 512, w * 512 + 512) in 8 lane-striped
    // steps (conflict-free LDS access), carrying the maximum; the waves
    // before it contribute theirs.
    constexpr uint32_t kSpan = (uint32_t)kScanTile / (kBlock / kWave);
    const uint32_t lane = lane_id(), wv = tid / kWave;
    uint32_t loc[kSpan / kWave];
    uint32_t carry = 0;
#pragma unroll
    for (int e = 0; e < (int)(kSpan / kWave); ++e) {
      uint32_t m = mark[wv * kSpan + e * kWave + lane];
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t o = __shfl_up(m, d, kWave);
        if (lane >= (uint32_t)d) m = m > o ? m : o;
      }
      m = m > carry ? m : carry;
      loc[e] = m;
      carry = __shfl(m, kWave - 1, kWave);
    }
    __shared__ uint32_t s_wave[kBlock / kWave];
    if (lane == 0) s_wave[wv] = carry;
    __syncthreads();
    uint32_t ex = 0;
    for (uint32_t w = 0; w < wv; ++w) ex = s_wave[w] > ex ? s_wave[w] : ex;
#pragma unroll
    for (int e = 0; e < (int)(kSpan / kWave); ++e) {
      const uint32_t i = wv * kSpan + e * kWave + lane;
      if (i < cnt) {
        const uint32_t rr = loc[e] > ex ? loc[e] : ex;
        sm.k[i] = (r0 + rr) << 32 | (sm.k[i] & 0xFFFFFFFFull);
      }
    }
  }
  __syncthreads();  // marks consumed
#pragma unroll
  for (int e = 0; e < kScanItems; ++e) sm.v[(uint32_t)e * kBlock + tid] = vreg[e];
  const uint64_t prev_key = base ? (rp << 32 | pcol) : 0;
  *prev = prev_key;
  __syncthreads();
  for (uint32_t i = tid; i < kScanTile; i += kBlock) {
    uint8_t h = 0;
    if (i < cnt) h = (base + i == 0) || sm.k[i] != (i ? sm.k[i - 1] : prev_key);
    sm.head[i] = h;
  }
  __syncthreads();
  const uint64_t f = reinterpret_cast<const uint64_t*>(sm.head)[tid];
  U2 tot;
  uint32_t rix = (uint32_t)block_excl_scan_u2(mk2((uint64_t)__popcll(f), 0), sm.scratch, &tot).a;
  const uint32_t nruns = (uint32_t)tot.a;
  for (uint64_t m = f; m; m &= m - 1)
    sm.rs[rix++] = (uint16_t)(tid * 8 + (__builtin_ctzll(m) >> 3));
  if (tid == 0) {
    sm.rs[nruns] = (uint16_t)cnt;
    sm.tail_end = base + cnt;
  }
  __syncthreads();
  // the last run may continue past the tile, up to another column or its row's end
  if (nruns && base + cnt < n && tid < kWave) {
    const uint64_t last = sm.k[sm.rs[nruns - 1]];
    const uint32_t lcol = (uint32_t)last;
    const uint64_t rend = nr.rstart[(last >> 32) + 1];
    for (uint64_t g0 = base + cnt;; g0 += kWave) {
      const uint64_t g = g0 + tid;
      const uint32_t c = g < n ? nr.cols[g] : 0;  // issued beside the row end
      const bool stop = g >= n || c != lcol || g >= rend;
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
    const uint32_t rn = tid + (uint32_t)q * kBlock;
    ok[q] = 0;
    ov[q] = 0.f;
    if (rn >= nruns) continue;
    const uint32_t j0 = sm.rs[rn], j1 = sm.rs[rn + 1];
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
    if (rn + 1 == nruns) {
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
    const uint32_t rn = tid + (uint32_t)q * kBlock;
    if (rn < nruns) {
      sm.k[rn] = ok[q];
      sm.v[rn] = ov[q];
    }
  }
  __syncthreads();
  return nruns;
}

}  // namespace sptc
