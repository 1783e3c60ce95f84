// Dense per-row accumulation core shared by the single-contraction and the
// batched dense kernels.
//
// One workgroup owns one X row.  The row's products -- the concatenation, in
// active-nonzero order j, of each nonzero's matched Y segment -- are taken in
// rounds of kBlock * G products: every thread loads its G products of the
// round at once (strided, so loads coalesce and G of them are in flight per
// thread), finds their owners j by binary search over the round's
// act_prefix entries staged in LDS, and then the round's owners are applied
// one after another with a barrier between owners.  Within one owner the
// column ranks are distinct (no duplicate Y coordinates), so each rank's
// products are added in j order -- the order of the ESC path's stable sort --
// with the same Neumaier-compensated FP32 steps, which makes the two
// accumulators bit-identical.
#pragma once
#include "kernels.hpp"

namespace sptc {

constexpr int kDenseG = 4;                               // products per thread per round
constexpr uint32_t kDenseRound = kBlock * kDenseG;       // products per round

template <int EB>
struct DenseSmem {
  static constexpr int E = 1 << EB;
  static constexpr int W = E / 32;
  uint32_t bm[W];
  uint64_t pref[kDenseRound + 1];
  uint32_t red[kBlock / kWave];
  uint64_t jl;  // owner of the round's last product
};

// Stages the round's owner prefixes: owners of products [pb, pl] are
// entries [ja, ja + m) of [j0, j1).  Returns ja through *ja.
template <typename SM>
__device__ __forceinline__ uint32_t dense_stage_owners(const uint64_t* __restrict__ act_prefix,
                                                       uint64_t j0, uint64_t j1, uint64_t pb,
                                                       uint64_t pl, SM& sm, uint64_t* s_ja,
                                                       uint64_t* ja) {
  __syncthreads();
  // last j with act_prefix[j] <= pb (wave 0), <= pl (wave 1)
  if (threadIdx.x < 2 * kWave) {
    const uint64_t j = wave_last_le(act_prefix, j0, j1, threadIdx.x < kWave ? pb : pl);
    if (threadIdx.x == 0) *s_ja = j;
    if (threadIdx.x == kWave) sm.jl = j;
  }
  __syncthreads();
  *ja = *s_ja;
  uint64_t* pref = sm.pref;
  const uint64_t need = sm.jl - *ja + 1;
  const uint32_t m = (uint32_t)(need < kDenseRound ? need : kDenseRound);
  for (uint32_t k = threadIdx.x; k <= m; k += kBlock) pref[k] = act_prefix[*ja + k];
  __syncthreads();
  return m;
}

// Owner (index into the staged run) of product p: last k with pref[k] <= p.
__device__ __forceinline__ uint32_t dense_owner(const uint64_t* pref, uint32_t m, uint64_t p) {
  uint32_t lo = 0, hi = m;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pref[mid] <= p) lo = mid + 1; else hi = mid;
  }
  return lo - 1;
}

// Symbolic pass: bitmap of the row's touched column ranks; returns the count
// (valid in thread 0).
template <int EB>
__device__ __forceinline__ uint32_t dense_row_count(const DenseArgs& a, uint64_t j0, uint64_t j1,
                                                    uint64_t cmask, DenseSmem<EB>& sm,
                                                    uint64_t* s_ja) {
  for (int i = threadIdx.x; i < DenseSmem<EB>::W; i += kBlock) sm.bm[i] = 0;
  const uint64_t p0 = j0 < j1 ? a.act_prefix[j0] : 0;
  const uint64_t p1 = j0 < j1 ? a.act_prefix[j1] : 0;
  for (uint64_t pb = p0; pb < p1; pb += kDenseRound) {
    uint64_t ja;
    const uint64_t pl = (pb + kDenseRound < p1 ? pb + kDenseRound : p1) - 1;
    const uint32_t m = dense_stage_owners(a.act_prefix, j0, j1, pb, pl, sm, s_ja, &ja);
#pragma unroll
    for (int g = 0; g < kDenseG; ++g) {
      const uint64_t p = pb + (uint64_t)g * kBlock + threadIdx.x;
      if (p < p1) {
        const uint32_t k = dense_owner(sm.pref, m, p);
        const uint64_t y = a.act_ypos[ja + k] + (p - sm.pref[k]);
        const uint32_t rank = (uint32_t)(a.ykeys[y] & cmask);
        atomicOr(&sm.bm[rank >> 5], 1u << (rank & 31));
      }
    }
  }
  __syncthreads();
  uint32_t cnt = 0;
  for (int i = threadIdx.x; i < DenseSmem<EB>::W; i += kBlock) cnt += __popc(sm.bm[i]);
  for (int off = kWave / 2; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, kWave);
  if (lane_id() == 0) sm.red[threadIdx.x / kWave] = cnt;
  __syncthreads();
  uint32_t t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / kWave; ++w) t += sm.red[w];
  return t;
}

// Numeric pass, then the touched slots in rank order to out_* at obase (row
// id `row`, column ytab[rank] or rank).  `acc` holds E doubles of LDS.
//   ORDERED: per slot a Neumaier-compensated FP32 sum in owner order (the
//            round's owners applied one after another, a barrier between
//            owners) -- bit-identical to the ESC path;
//   else:    per slot an FP64 sum by LDS atomic add (no barriers), rounded
//            once to FP32 -- within FP32 rounding of the exact sum, the
//            order of concurrent adds not fixed.
template <int EB, bool ORDERED>
__device__ __forceinline__ void dense_row_numeric(const DenseArgs& a, uint64_t j0, uint64_t j1,
                                                  uint64_t cmask, const uint64_t* ytab,
                                                  uint32_t row, uint64_t obase, double* acc,
                                                  DenseSmem<EB>& sm, uint64_t* s_ja) {
  constexpr int E = DenseSmem<EB>::E;
  constexpr int S = E / kBlock;  // slots per thread in the compaction (1..32)
  static_assert(E >= kBlock && E % kBlock == 0, "dense row width");
  float* s_sum = reinterpret_cast<float*>(acc);
  float* s_cmp = s_sum + E;
  for (int i = threadIdx.x; i < E; i += kBlock) acc[i] = 0.0;
  for (int i = threadIdx.x; i < DenseSmem<EB>::W; i += kBlock) sm.bm[i] = 0;
  const uint64_t p0 = j0 < j1 ? a.act_prefix[j0] : 0;
  const uint64_t p1 = j0 < j1 ? a.act_prefix[j1] : 0;
  double vsum = 0.0;
  for (uint64_t pb = p0; pb < p1; pb += kDenseRound) {
    uint64_t ja;
    const uint64_t pl = (pb + kDenseRound < p1 ? pb + kDenseRound : p1) - 1;
    const uint32_t m = dense_stage_owners(a.act_prefix, j0, j1, pb, pl, sm, s_ja, &ja);
    uint32_t own[kDenseG], rank[kDenseG];
    float val[kDenseG];
#pragma unroll
    for (int g = 0; g < kDenseG; ++g) {
      const uint64_t p = pb + (uint64_t)g * kBlock + threadIdx.x;
      own[g] = ~0u;
      rank[g] = 0;
      val[g] = 0.f;
      if (p < p1) {
        const uint32_t k = dense_owner(sm.pref, m, p);
        const uint64_t y = a.act_ypos[ja + k] + (p - sm.pref[k]);
        own[g] = k;
        rank[g] = (uint32_t)(a.ykeys[y] & cmask);
        val[g] = a.act_xval[ja + k] * a.yvals[y];
      }
    }
    if constexpr (!ORDERED) {
#pragma unroll
      for (int g = 0; g < kDenseG; ++g) {
        if (own[g] != ~0u) {
          vsum += (double)val[g];
          atomicAdd(&acc[rank[g]], (double)val[g]);
          atomicOr(&sm.bm[rank[g] >> 5], 1u << (rank[g] & 31));
        }
      }
      continue;  // the next round's staging starts with a barrier
    }
    // owners of this round in order: the first is own of product pb, the
    // last that of pl = min(pb + round, p1) - 1
    const uint32_t klast = dense_owner(sm.pref, m, pl);
    for (uint32_t k = 0; k <= klast; ++k) {
#pragma unroll
      for (int g = 0; g < kDenseG; ++g) {
        if (own[g] == k) {
          const float v = val[g];
          const uint32_t r = rank[g];
          vsum += (double)v;
          const float smv = s_sum[r];
          const float t = smv + v;
          if (fabsf(smv) >= fabsf(v)) s_cmp[r] += (smv - t) + v;
          else s_cmp[r] += (v - t) + smv;
          s_sum[r] = t;
          atomicOr(&sm.bm[r >> 5], 1u << (r & 31));
        }
      }
      __syncthreads();
    }
  }
  __syncthreads();
  // Compaction in column order: thread t owns slots [t*S, t*S+S).
  const uint32_t first = threadIdx.x * S;
  const uint32_t word = sm.bm[first >> 5];
  uint32_t bits;
  if constexpr (S == 32) bits = word;
  else bits = (word >> (first & 31)) & ((1u << S) - 1);
  uint32_t tot;
  uint32_t o = block_excl_scan<uint32_t>((uint32_t)__popc(bits), sm.red, &tot);
  for (int e = 0; e < S; ++e) {
    if ((bits >> e) & 1u) {
      const uint32_t rk = first + e;
      a.out_row[obase + o] = row;
      a.out_col[obase + o] = ytab ? ytab[rk] : rk;
      a.out_val[obase + o] = ORDERED ? s_sum[rk] + s_cmp[rk] : (float)acc[rk];
      ++o;
    }
  }
  for (int off = kWave / 2; off > 0; off >>= 1) vsum += __shfl_xor(vsum, off, kWave);
  if (lane_id() == 0 && vsum != 0.0) atomicAdd(a.vsum, vsum);
}

}  // namespace sptc
