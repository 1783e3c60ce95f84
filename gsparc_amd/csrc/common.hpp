// Shared definitions for the gfx950 SpTC kernels and their host launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#include "../../include/sptc.h"

namespace sptc {

constexpr int kBlock = 256;  // 4 wave64s per workgroup
constexpr int kWave = 64;

// ---- host-side error plumbing -------------------------------------------
struct HipError {
  hipError_t err;
  const char* what;
  int line;
};

#define SPTC_HIP(call)                                                  \
  do {                                                                  \
    hipError_t e_ = (call);                                             \
    if (e_ != hipSuccess) throw ::sptc::HipError{e_, #call, __LINE__}; \
  } while (0)

__host__ __device__ inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
// Shifts that are defined for every count in [0, 64].
__host__ __device__ __forceinline__ uint64_t shr64(uint64_t k, int s) { return s >= 64 ? 0 : k >> s; }
__host__ __device__ __forceinline__ uint64_t shl64(uint64_t k, int s) { return s >= 64 ? 0 : k << s; }
__host__ __device__ __forceinline__ uint64_t low_mask(int bits) {
  return bits >= 64 ? ~0ull : ((1ull << bits) - 1);
}

// ---- device helpers ------------------------------------------------------
// Workgroups are dispatched to the 8 XCDs round-robin by blockIdx, each XCD
// with its own L2.  xcd_tile maps block b of n to a tile so that XCD x runs a
// contiguous range of tiles in order: neighbouring tiles, whose writes meet
// in the same cache lines, then share one L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
  const uint32_t per = n / kXcds, rem = n % kXcds;
  const uint32_t x = b % kXcds, i = b / kXcds;
  return x * per + (x < rem ? x : rem) + i;
}

// Persistent-loop form: the tiles of XCD x are one contiguous range that its
// workgroups walk in turn (tile t0, t0 + step, ... below t1).  With fewer
// workgroups than XCDs the plain grid-stride order is kept.
__device__ __forceinline__ void xcd_tiles(uint64_t ntiles, bool remap, uint64_t* t0,
                                          uint64_t* t1, uint64_t* step) {
  const uint32_t G = gridDim.x;
  if (!remap || G < kXcds) {
    *t0 = blockIdx.x;
    *t1 = ntiles;
    *step = G;
    return;
  }
  const uint32_t x = blockIdx.x % kXcds, i = blockIdx.x / kXcds;
  const uint64_t per = ntiles / kXcds, rem = ntiles % kXcds;
  const uint64_t lo = x * per + (x < rem ? x : rem);
  *t0 = lo + i;
  *t1 = lo + per + (x < rem ? 1 : 0);
  *step = (G + kXcds - 1 - x) / kXcds;  // workgroups on XCD x
}

// The wave's index in its workgroup, in a scalar register (the compiler
// does not know threadIdx.x / 64 is the same across a wave).
__device__ __forceinline__ uint32_t wave_id() {
  return __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
}
__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
// Number of set bits of m in lanes strictly below this lane.
__device__ __forceinline__ uint32_t popc_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Orders one wave's LDS accesses across its lanes: the fences keep the
// compiler from moving memory operations across the point (the wave
// barrier alone does not), LDS executes a wave's accesses in order.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Agent-scope (whole device, across XCD L2s) relaxed atomics for look-back.
__device__ __forceinline__ uint64_t ld_agent(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-level inclusive scan of uint64 via LDS-free shuffles.
__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint64_t o = __shfl_up(v, d, kWave);
    if (lane >= (uint32_t)d) v += o;
  }
  return v;
}
__device__ __forceinline__ uint32_t wave_incl_scan32(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint32_t o = __shfl_up(v, d, kWave);
    if (lane >= (uint32_t)d) v += o;
  }
  return v;
}

// Last j in [lo, hi) with pre[j] <= p, for non-decreasing pre with
// pre[lo] <= p; called by a whole wave, the result is uniform.  The first
// probe reads the 64 entries from lo (an owner search that starts from the
// previous owner usually ends there); then each round samples 64 evenly
// spaced entries and keeps the interval between the last two that bracket
// p, so a range of N entries takes about log64(N) dependent loads.
__device__ __forceinline__ uint64_t wave_last_le(const uint64_t* __restrict__ pre, uint64_t lo,
                                                 uint64_t hi, uint64_t p) {
  const uint32_t lane = lane_id();
  {
    const uint64_t j = lo + lane;
    const uint64_t m = __ballot(j < hi && pre[j] <= p);  // a prefix of the lanes
    if (!(m >> (kWave - 1))) return lo + (uint64_t)__popcll(m) - 1;
    lo += kWave - 1;
  }
  while (hi - lo > (uint64_t)kWave) {
    const uint64_t step = (hi - lo + kWave - 1) / kWave;
    const uint64_t j = lo + lane * step;
    const uint64_t m = __ballot(j < hi && pre[j] <= p);  // lane 0 holds (j = lo)
    lo += (uint64_t)(kWave - 1 - __builtin_clzll(m)) * step;
    hi = hi < lo + step ? hi : lo + step;
  }
  const uint64_t j = lo + lane;
  const uint64_t m = __ballot(j < hi && pre[j] <= p);
  return lo + (uint64_t)__popcll(m) - 1;
}

// Block-wide exclusive scan (blockDim == kBlock), returns exclusive prefix
// and the block total through *total.  `scratch` holds kBlock/kWave values.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* scratch, T* total) {
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  T inc;
  if constexpr (sizeof(T) == 8) inc = wave_incl_scan(v);
  else inc = wave_incl_scan32(v);
  if (lane == kWave - 1) scratch[wave] = inc;
  __syncthreads();
  T wave_off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) {
    T s = scratch[w];
    if ((uint32_t)w < wave) wave_off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return wave_off + inc - v;
}

// One Neumaier-compensated FP32 step: sum s, compensation c, the sum's
// value is s + c (error ~2u independent of the count).  Every accumulator
// takes this one step, so sums of the same values in the same order agree
// bit for bit whichever kernel adds them.
__device__ __forceinline__ void kahan_step(float& s, float& c, float x) {
  const float t = s + x;
  if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
  else c += (x - t) + s;
  s = t;
}
struct KahanF32 {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) { kahan_step(s, c, x); }
  __device__ __forceinline__ float value() const { return s + c; }
};

// XCD-contiguous tile order on (SPTC_XCD_REMAP=0 turns it off); host side.
uint32_t xcd_remap_enabled();

// Narrow form of a chunk's sorted products: 32-bit columns (+ values)
// grouped by row, rows given by their first products rstart[0..nrows]
// (non-decreasing, rstart[nrows] = product count).  A product's key is
// (row << 32 | column); run heads are where the column changes or a row
// starts.
struct NarrowRows {
  const uint32_t* cols;
  const uint64_t* rstart;
  uint64_t nrows;
  // per wave tile t (512 products from t * 512): [2t] = row of its first
  // product, [2t + 1] = row of the product before it
  const uint32_t* tile_row;
  // per row: nonzero when the row's outputs are produced elsewhere (big
  // rows reduced by high-digit buckets): its products delimit runs but are
  // neither counted nor emitted.  NULL: every row is reduced here.
  const uint64_t* big = nullptr;
};

// ---- launchers (defined in the .hip files) --------------------------------
struct SortWorkspace;

// fill / copy helpers (graph-safe: kernels, not hipMemsetAsync)
void fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s);
void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);

}  // namespace sptc
