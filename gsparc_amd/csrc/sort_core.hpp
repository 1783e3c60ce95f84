// Tile core of the radix sort: load + stable wave64 ranking, LDS staging and
// coalesced scatter.  Shared by the sort passes and the batched (segmented)
// passes.
#pragma once
#include "common.hpp"

namespace sptc {

// Block-wide exclusive scan of uint32 for NT threads.
template <int NT>
__device__ __forceinline__ uint32_t block_scan_u32(uint32_t v, uint32_t* scratch,
                                                   uint32_t* total) {
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  uint32_t inc = wave_incl_scan32(v);
  if (lane == kWave - 1) scratch[wave] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NT / kWave; ++w) {
    uint32_t s = scratch[w];
    if ((uint32_t)w < wave) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// ST: staging element (keys, then values, pass through buf): 64-bit for
// full keys, 32-bit for column keys.
template <int RB, int IPT, int NT, typename ST = uint64_t>
struct TileSmem {
  static constexpr int R = 1 << RB;
  static constexpr int TILE = NT * IPT;
  static constexpr int NW = NT / kWave;
  uint16_t cnt[NW][R];     // per-wave running counts, then each wave's staging start
  uint32_t gbase[R];       // global start of this tile's run of each digit (then minus
                           // the run's tile-local start)
  uint16_t dstart[R];      // tile totals / run starts (paths without tile_prefix_fold)
  ST buf[TILE];            // staged keys, then staged values
  uint32_t scan[NW];
  uint32_t tile;
};

// Lanes of the wave holding the same RB-bit digit d as this lane, among
// the lanes in `valid`.  Per digit bit b: m = -(bit b of d) (0 or all ones)
// and the ballot of that bit; (ballot ^ m) marks the lanes whose bit b
// differs from this lane's, accumulated per 32-lane half.
template <int RB>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) {
  uint32_t nlo = 0, nhi = 0;
#pragma unroll
  for (int b = 0; b < RB; ++b) {
    uint32_t m;  // one v_bfe_i32 (the compiler splits a sign extract into two shifts)
    asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(m) : "v"(d), "i"(b));
    const uint64_t bb = __ballot(m != 0);
    nlo |= (uint32_t)bb ^ m;
    nhi |= (uint32_t)(bb >> 32) ^ m;
  }
  return valid & ~((uint64_t)nhi << 32 | nlo);
}

// Tile-local index of the thread's key i: wave w owns keys
// [w*64*IPT, (w+1)*64*IPT) of the tile, lane-striped.
template <int IPT>
__device__ __forceinline__ uint32_t tile_local(int i) {
  return wave_id() * (uint32_t)(kWave * IPT) + (uint32_t)i * kWave + lane_id();
}

// Ranks the keys a thread holds (tile_local layout; valid when below cnt,
// FULL: cnt is the whole tile) within its wave's run of equal digits.
// sm.cnt must be zero; afterwards sm.cnt[w][d] = number of keys of digit d
// in wave w.  Keys are 64-bit, or 32-bit column keys of the row-segmented
// product sort.
template <bool FULL, int RB, int IPT, int NT, typename ST, typename KT>
__device__ __forceinline__ void tile_rank(TileSmem<RB, IPT, NT, ST>& sm, uint32_t cnt, int shift,
                                          const KT (&k)[IPT], uint32_t (&pos)[IPT]) {
  constexpr int R = 1 << RB;
  const uint32_t lane = lane_id();
  const uint32_t wave = wave_id();
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    // a wave's item past the tile's end on all lanes (partly filled tiles,
    // small-row groups) is skipped whole
    if (!FULL && wave * (uint32_t)(kWave * IPT) + (uint32_t)i * kWave >= cnt) {
      pos[i] = 0;
      continue;
    }
    const bool valid = FULL || tile_local<IPT>(i) < cnt;
    const uint32_t d = (uint32_t)(k[i] >> shift) & (R - 1);
    const uint64_t peers = match_digit<RB>(d, FULL ? ~0ull : ballot(valid));
    uint32_t before = 0;
    if (valid) before = sm.cnt[wave][d];
    __builtin_amdgcn_wave_barrier();
    pos[i] = before + popc_below(peers);  // rank within the wave's digit run
    if (valid && (peers >> lane) == 1ull)  // highest lane of the group advances
      sm.cnt[wave][d] = (uint16_t)(before + (uint32_t)__popcll(peers));
    __builtin_amdgcn_wave_barrier();
  }
}

// Loads the tile (cnt keys at kb, values at vb: the tile's own bases, so
// the addresses are 32-bit offsets) and ranks it.
template <bool FULL, int RB, int IPT, int NT, typename ST, typename KT>
__device__ __forceinline__ void tile_load_rank(TileSmem<RB, IPT, NT, ST>& sm,
                                               const KT* __restrict__ kb,
                                               const float* __restrict__ vb, uint32_t cnt,
                                               int shift, KT (&k)[IPT], float (&v)[IPT],
                                               uint32_t (&pos)[IPT]) {
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const uint32_t li = tile_local<IPT>(i);
    const bool valid = FULL || li < cnt;
    k[i] = valid ? kb[li] : 0;
    v[i] = valid && vb ? vb[li] : 0.f;  // vb NULL: keys only
  }
  tile_rank<FULL, RB, IPT, NT>(sm, cnt, shift, k, pos);
}

// sm.dstart holds each digit's tile total: turn it into the tile-local start
// of each digit's run (block-wide exclusive scan).
template <int RB, int IPT, int NT, typename ST>
__device__ __forceinline__ void tile_digit_starts(TileSmem<RB, IPT, NT, ST>& sm) {
  constexpr int R = 1 << RB;
  constexpr int PER = (R + NT - 1) / NT;
  const uint32_t tid = threadIdx.x;
  uint32_t loc[PER];
  uint32_t sum = 0;
  for (int j = 0; j < PER; ++j) {
    const int d = tid * PER + j;
    loc[j] = d < R ? sm.dstart[d] : 0;
    sum += loc[j];
  }
  uint32_t tot;
  uint32_t ex = block_scan_u32<NT>(sum, sm.scan, &tot);
  for (int j = 0; j < PER; ++j) {
    const int d = tid * PER + j;
    if (d < R) sm.dstart[d] = (uint16_t)ex;
    ex += loc[j];
  }
}

// After ranking (sm.cnt = per-wave digit counts) in one sweep: each thread
// takes PER consecutive digits, turns their per-wave counts into wave
// prefixes in registers, the block scans the digit totals, and the thread
// writes back cnt[w][d] = (tile-local start of d) + (wave prefix) and
// gbase[d] -= start (modulo 2^32) -- what tile_digit_starts and the fold in
// tile_stage_scatter do in three sweeps.
template <int RB, int IPT, int NT, typename ST>
__device__ __forceinline__ void tile_prefix_fold(TileSmem<RB, IPT, NT, ST>& sm) {
  constexpr int R = 1 << RB;
  constexpr int PER = (R + NT - 1) / NT;
  constexpr int NW = NT / kWave;
  const uint32_t tid = threadIdx.x;
  uint32_t pre[PER][NW], tot[PER];
  uint32_t sum = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int d = tid * PER + j;
    uint32_t acc = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      pre[j][w] = acc;
      acc += d < R ? (uint32_t)sm.cnt[w][d] : 0u;
    }
    tot[j] = acc;
    sum += acc;
  }
  uint32_t all;
  uint32_t ex = block_scan_u32<NT>(sum, sm.scan, &all);  // barriers inside
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int d = tid * PER + j;
    if (d < R) {
#pragma unroll
      for (int w = 0; w < NW; ++w) sm.cnt[w][d] = (uint16_t)(ex + pre[j][w]);
      sm.gbase[d] -= ex;
    }
    ex += tot[j];
  }
}

// vout NULL: keys only (no value staging or stores).
// Given sm.gbase and sm.dstart (tile totals per digit on entry), scans the
// totals into tile-local run starts, stages keys in digit order, writes them
// in coalesced runs, then does the same for the values.  The run starts are
// folded into the per-wave prefixes (cnt[w][d] += dstart[d]) and the global
// bases (gbase[d] -= dstart[d], modulo 2^32) first, so a key's staging slot
// and destination take one LDS read each.
// FOLDED: tile_prefix_fold has already run (sm.dstart unused).
template <bool FULL, int RB, int IPT, int NT, typename ST, typename KT, typename KO,
          bool FOLDED = false>
__device__ __forceinline__ void tile_stage_scatter(TileSmem<RB, IPT, NT, ST>& sm, uint32_t cnt,
                                                   int shift, const KT (&k)[IPT],
                                                   const float (&v)[IPT], uint32_t (&pos)[IPT],
                                                   KO* __restrict__ kout,
                                                   float* __restrict__ vout) {
  constexpr int R = 1 << RB;
  const uint32_t tid = threadIdx.x;
  const uint32_t wave = wave_id();
  static_assert(sizeof(KT) <= sizeof(ST), "staging element narrower than the key");
  KT* kbuf = reinterpret_cast<KT*>(sm.buf);
  if constexpr (!FOLDED) {
    tile_digit_starts<RB, IPT, NT>(sm);
    __syncthreads();
    for (int d = tid; d < R; d += NT) {
      const uint32_t ds = sm.dstart[d];
#pragma unroll
      for (int w = 0; w < NT / kWave; ++w) sm.cnt[w][d] = (uint16_t)(sm.cnt[w][d] + ds);
      sm.gbase[d] -= ds;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const uint32_t d = (uint32_t)(k[i] >> shift) & (R - 1);
    pos[i] += (uint32_t)sm.cnt[wave][d];
    if (FULL || tile_local<IPT>(i) < cnt) kbuf[pos[i]] = k[i];
  }
  __syncthreads();
  uint32_t dst[IPT];
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const uint32_t j = tid + (uint32_t)i * NT;
    if (FULL || j < cnt) {
      const KT key = kbuf[j];
      const uint32_t d = (uint32_t)(key >> shift) & (R - 1);
      dst[i] = sm.gbase[d] + j;
      kout[dst[i]] = (KO)key;
    }
  }
  if (!vout) return;  // keys only
  __syncthreads();
  float* vbuf = reinterpret_cast<float*>(sm.buf);
#pragma unroll
  for (int i = 0; i < IPT; ++i)
    if (FULL || tile_local<IPT>(i) < cnt) vbuf[pos[i]] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const uint32_t j = tid + (uint32_t)i * NT;
    if (FULL || j < cnt) vout[dst[i]] = vbuf[j];
  }
}

// Per digit: turn per-wave counts into wave prefixes; returns the tile total.
template <int RB, int IPT, int NT, typename ST>
__device__ __forceinline__ uint32_t wave_prefixes(TileSmem<RB, IPT, NT, ST>& sm, int d) {
  uint32_t acc = 0;
#pragma unroll
  for (int w = 0; w < NT / kWave; ++w) {
    const uint32_t c = sm.cnt[w][d];
    sm.cnt[w][d] = (uint16_t)acc;
    acc += c;
  }
  return acc;
}

}  // namespace sptc
