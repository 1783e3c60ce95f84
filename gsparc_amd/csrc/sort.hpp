// Radix sort workspace and host launchers (see sort.hip).
#pragma once
#include "common.hpp"
#include "device.hpp"
#include "ktimer.hpp"
#include "scan.hpp"

namespace sptc {

constexpr int kMaxPasses = 16;

struct SortWorkspace {
  int rb = 8;                   // radix bits per pass
  uint64_t cap = 0;             // max keys
  uint64_t max_tiles = 0;
  DevBuf<uint64_t> alt_keys;
  DevBuf<float> alt_vals;
  DevBuf<uint64_t> status;      // [max_tiles][R] look-back words
  DevBuf<uint64_t> hist;        // [kMaxPasses][R] digit counts
  DevBuf<uint64_t> base;        // [kMaxPasses][R] digit starts
  DevBuf<uint32_t> counters;    // [kMaxPasses] tile counters, [32] error flag
  DevBuf<uint32_t> rts_cnt;     // [R][tiles] per-tile digit counts
  DevBuf<uint32_t> rts_off;     // [R][tiles] their exclusive scan
  DevBuf<U2> rts_tmp;           // scan scratch
  uint32_t epoch = 0;

  // Grows the workspace to hold `n` keys at `rb` radix bits.
  void ensure(uint64_t n, int rb, hipStream_t s);
  // Clears histograms and tile counters before a producer accumulates.
  void prepare(int passes, hipStream_t s);
  uint32_t* err() const { return counters.get() + 32; }
  static uint64_t bytes_per_key(int rb);
};

int radix_passes(int nbits, int rb);
uint64_t sort_tile_items(int rb);
bool radix_bits_supported(int rb);
// True when the configured pass algorithm needs up-front digit histograms
// (onesweep); reduce-then-scan counts per tile itself.
bool sort_needs_histogram();
// Digit width for a sort over nbits key bits (fewest passes, digits <= 10).
int choose_radix_bits(int nbits);
void histogram(const SortWorkspace& ws, const uint64_t* keys, uint64_t n,
               int passes, hipStream_t s);
// Sorts (keys, vals) on bits [0, nbits), ws.hist holding the histograms of
// all passes.  Returns true when the sorted data ended in (ws.alt_keys,
// ws.alt_vals) (odd pass count).
// kt (optional) times every pass under kernel class `cls`.
bool sort_after_histogram(SortWorkspace& ws, uint64_t* keys, float* vals,
                          uint64_t n, int nbits, hipStream_t s, KTimer* kt = nullptr,
                          int cls = 0);

// Device-side histogram accumulation for producers: call once per key with
// an LDS array of passes*R uint32 cleared by the block, then flush.
template <int RB>
__device__ __forceinline__ void hist_add(uint32_t* sh, uint64_t key, int passes) {
  constexpr uint32_t R = 1u << RB;
  for (int p = 0; p < passes; ++p)
    atomicAdd(&sh[p * R + ((uint32_t)(key >> (p * RB)) & (R - 1))], 1u);
}
// Wave-aggregated variant: when every valid lane of the wave has the same
// digit (typical for the high, row-index digits of product keys) one lane
// adds the whole count; otherwise each lane adds one.  All lanes of the wave
// must call it together.
template <int RB>
__device__ __forceinline__ void hist_add_wave(uint32_t* sh, uint64_t key, bool valid,
                                             int passes) {
  constexpr uint32_t R = 1u << RB;
  const uint64_t vmask = __ballot(valid);
  if (!vmask) return;
  const int first = __builtin_ctzll(vmask);
  for (int p = 0; p < passes; ++p) {
    const uint32_t d = (uint32_t)(key >> (p * RB)) & (R - 1);
    const uint32_t d0 = __shfl(d, first, 64);
    const uint64_t same = __ballot(valid && d == d0);
    if (same == vmask) {
      if ((int)__lane_id() == first) atomicAdd(&sh[p * R + d0], (uint32_t)__popcll(vmask));
    } else if (valid) {
      atomicAdd(&sh[p * R + d], 1u);
    }
  }
}

template <int RB>
__device__ __forceinline__ void hist_flush(const uint32_t* sh, uint64_t* hist,
                                           int passes) {
  constexpr int R = 1 << RB;
  for (int i = threadIdx.x; i < passes * R; i += blockDim.x)
    if (sh[i]) atomicAdd((unsigned long long*)&hist[i], (unsigned long long)sh[i]);
}

}  // namespace sptc
