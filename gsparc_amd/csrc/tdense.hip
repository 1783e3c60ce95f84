// Tiled dense accumulator: per-row dense accumulation in LDS over column
// tiles, for column spaces too wide for one LDS row.
//
// One workgroup owns one X row.  The row's products are the matched Y
// segments of its active nonzeros ("owners"), each segment sorted by column
// rank (Y keys are sorted by contract field, then free field).  The column
// space is walked in tiles of 2^TWB ranks: every owner keeps a cursor into
// its segment, and for the current tile each owner's entries below the
// tile's end are its products there (a binary search from the cursor).  The
// tile's products are spread over the workgroup (owner by LDS search over
// the owners' product prefix), accumulated into the tile's LDS slots and
// marked in an LDS bitmap; then the marked slots are emitted in column
// order and cleared.  Tiles without products are skipped: the next tile is
// the one holding the smallest column any cursor points at.  Products never
// reach HBM: the row's Y segments are read from cache and each output is
// written once.
//
// Symbolic pass (NUM = false): bitmap only, wide tiles, per-row output
// counts.  Numeric pass (NUM = true): FP64 slots (LDS atomic add of the
// FP32 products), outputs at the row's offset.  Sums in FP64 rounded once
// to FP32: the result is within FP32 rounding of the exact sum whatever the
// order of the adds (the order of concurrent LDS adds is not fixed, so an
// FP64 sum lying at an FP32 rounding boundary could round either way).
//
// This is the reference's dense accumulator (contraction.cuh:247-362: an
// FP32 slab indexed by column with atomicAdd, compacted on the host) with
// the slab tiled through LDS instead of held in HBM, and kept exact zeros
// (the reference drops outputs whose sum is 0).
#include "tdense.hpp"

namespace sptc {

namespace {

__device__ __forceinline__ uint32_t block_min_u32(uint32_t v, uint32_t* red) {
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const uint32_t o = __shfl_xor(v, d, kWave);
    v = o < v ? o : v;
  }
  __syncthreads();
  if (lane_id() == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  uint32_t m = red[0];
#pragma unroll
  for (int w = 1; w < kBlock / kWave; ++w) m = red[w] < m ? red[w] : m;
  return m;
}

// 64-bit OR over the wave.
__device__ __forceinline__ uint64_t wave_or64(uint64_t v) {
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const uint32_t lo = __shfl_xor((uint32_t)v, d, kWave);
    const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), d, kWave);
    v |= ((uint64_t)hi << 32) | lo;
  }
  return v;
}

template <int TWB, bool NUM>
__global__ __launch_bounds__(kBlock) void k_tdense(TDenseArgs a) {
  constexpr uint32_t TW = 1u << TWB;
  constexpr uint32_t NWORD = TW / 32;
  constexpr uint32_t S = TW / kBlock;  // slots per thread in the compaction
  constexpr uint32_t NWAVE = kBlock / kWave;
  static_assert(!NUM || (S >= 1 && S <= 32 && 32 % S == 0), "numeric tile width");
  static_assert(NUM || NWORD % kBlock == 0, "symbolic tile width");
  __shared__ uint32_t bm[NWORD];
  __shared__ double slot[NUM ? TW : 1];
  // the batch's owners with products in the tile, compacted: first product
  // (exclusive prefix), first Y entry, X value; cpref[nk] = the total
  __shared__ uint32_t cpref[kBlock + 1];
  __shared__ uint64_t cbase[kBlock];
  __shared__ float cxv[kBlock];
  __shared__ uint64_t red64[NWAVE];
  __shared__ uint32_t red[NWAVE];
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t wave = wave_id();
  const uint64_t r = a.r0 + blockIdx.x;
  const uint64_t j0 = a.row_act[r], j1 = a.row_act[r + 1];
  for (uint32_t i = tid; i < NWORD; i += kBlock) bm[i] = 0;
  if constexpr (NUM)
    for (uint32_t i = tid; i < TW; i += kBlock) slot[i] = 0.0;
  // cursors at the segments' starts; first tile = that of the smallest column
  uint32_t mn = ~0u;
  for (uint64_t j = j0 + tid; j < j1; j += kBlock) {
    const uint64_t y = a.act_ypos[j];
    a.cur[j] = y;
    const uint32_t c = a.ycol[y];
    mn = c < mn ? c : mn;
  }
  mn = block_min_u32(mn, red);
  const uint64_t obase = NUM ? a.out_off[blockIdx.x] - a.out0 : 0;
  uint64_t ocount = 0;
  double vs = 0.0;
  while (mn != ~0u) {
    const uint32_t t = mn >> TWB;
    const uint64_t t0 = (uint64_t)t << TWB;
    const uint64_t tend = t0 + TW;
    uint32_t nmin = ~0u;
    for (uint64_t jb = j0; jb < j1; jb += kBlock) {
      const uint64_t j = jb + tid;
      uint32_t cnt = 0;
      uint64_t s = 0;
      if (j < j1) {
        s = a.cur[j];
        const uint64_t e = a.act_ypos[j] + (a.act_prefix[j + 1] - a.act_prefix[j]);
        uint64_t lo = s, hi = e;  // first entry at or past the tile's end
        while (lo < hi) {
          const uint64_t mid = (lo + hi) >> 1;
          if (a.ycol[mid] < tend) lo = mid + 1; else hi = mid;
        }
        cnt = (uint32_t)(lo - s);
        a.cur[j] = lo;
        if (lo < e) {
          const uint32_t c = a.ycol[lo];
          nmin = c < nmin ? c : nmin;
        }
      }
      // one scan gives both the product prefix and the compacted position
      uint64_t tot2;
      const uint64_t ex =
          block_excl_scan<uint64_t>(((uint64_t)cnt << 32) | (cnt ? 1u : 0u), red64, &tot2);
      const uint32_t tot = (uint32_t)(tot2 >> 32), nk = (uint32_t)tot2;
      if (cnt) {
        const uint32_t k = (uint32_t)ex;
        cpref[k] = (uint32_t)(ex >> 32);
        cbase[k] = s;
        if constexpr (NUM) cxv[k] = a.act_xval[j];
      }
      if (tid == 0) cpref[nk] = tot;
      __syncthreads();
      // wave w takes products [p0, p1), 64 consecutive ones per step; the
      // owner of a lane's product is k0 plus the owners that start in the
      // window at or below it (a mask built from the window's boundaries)
      const uint32_t per = (uint32_t)ceil_div(ceil_div(tot, NWAVE), kWave) * kWave;
      const uint32_t p0 = wave * per < tot ? wave * per : tot;
      const uint32_t p1 = p0 + per < tot ? p0 + per : tot;
      uint32_t k0 = 0;
      if (p0 < p1) {  // last owner with cpref <= p0: sampled search over nk
        uint32_t lo = 0, hi = nk;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (cpref[mid] <= p0) lo = mid; else hi = mid;
        }
        k0 = lo;
      }
      for (uint32_t pw = p0; pw < p1; pw += kWave) {
        const uint32_t p = pw + lane;
        const uint32_t kb = k0 + 1 + lane;
        const uint32_t b = kb < nk ? cpref[kb] : ~0u;
        // boundaries inside the window mark their offset (distinct: every
        // compacted owner has products)
        const uint64_t mark = (b >= pw && b < pw + kWave) ? (1ull << (b - pw)) : 0ull;
        const uint64_t m = wave_or64(mark);
        const uint64_t below = lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1);
        const uint32_t k = k0 + (uint32_t)__popcll(m & below);
        if (p < p1) {
          const uint64_t y = cbase[k] + (p - cpref[k]);
          const uint32_t sl = (uint32_t)(a.ycol[y] - t0);
          atomicOr(&bm[sl >> 5], 1u << (sl & 31));
          if constexpr (NUM) {
            const float v = cxv[k] * a.yvals[y];
            vs += (double)v;
            atomicAdd(&slot[sl], (double)v);
          }
        }
        k0 += (uint32_t)__popcll(m);  // owner of the next window's first product
      }
      __syncthreads();
    }
    // the tile's touched slots in column order (numeric), or their count
    uint32_t tot;
    if constexpr (NUM) {
      const uint32_t first = tid * S;
      uint32_t bits = bm[first >> 5];
      if constexpr (S < 32) bits = (bits >> (first & 31)) & ((1u << S) - 1);
      uint32_t o = block_excl_scan<uint32_t>((uint32_t)__popc(bits), red, &tot);
      while (bits) {
        const uint32_t e = __builtin_ctz(bits);
        bits &= bits - 1;
        const uint32_t sl = first + e;
        const uint64_t rk = t0 + sl;
        const uint64_t d = obase + ocount + o;
        a.out_row[d] = (uint32_t)r;
        a.out_col[d] = a.ytab ? a.ytab[rk] : rk;
        a.out_val[d] = (float)slot[sl];
        slot[sl] = 0.0;
        ++o;
      }
    } else {
      uint32_t c = 0;
      for (uint32_t i = tid; i < NWORD; i += kBlock) c += __popc(bm[i]);
      block_excl_scan<uint32_t>(c, red, &tot);
    }
    ocount += tot;
    __syncthreads();
    for (uint32_t i = tid; i < NWORD; i += kBlock) bm[i] = 0;
    mn = block_min_u32(nmin, red);
  }
  if constexpr (NUM) {
    for (int d = kWave / 2; d > 0; d >>= 1) vs += __shfl_xor(vs, d, kWave);
    if (lane == 0 && vs != 0.0) atomicAdd(a.vsum, vs);
  } else {
    if (tid == 0) a.row_cnt[blockIdx.x] = (uint32_t)ocount;
  }
}

}  // namespace

void launch_tdense_count(const TDenseArgs& a, uint64_t rows, hipStream_t s) {
  if (!rows) return;
  hipLaunchKernelGGL((k_tdense<kTDenseCountBits, false>), dim3((uint32_t)rows), dim3(kBlock), 0,
                     s, a);
  SPTC_HIP(hipGetLastError());
}

void launch_tdense_numeric(const TDenseArgs& a, uint64_t rows, hipStream_t s) {
  if (!rows) return;
  hipLaunchKernelGGL((k_tdense<kTDenseNumBits, true>), dim3((uint32_t)rows), dim3(kBlock), 0, s,
                     a);
  SPTC_HIP(hipGetLastError());
}

}  // namespace sptc
