// Big-row reduce of the high-digit form: per-bucket counts, per-row
// offsets, and the LDS slot reduce that emits each bucket's outputs in
// column order (the method is described in bucket.hpp).
//
// Each wave takes 2^gshift consecutive buckets of one big row.  Their
// products are one contiguous range (the pass sorted the row by digit) and
// a product's bucket is its column's digit, so both kernels stream the
// range in 64-product windows, the next window's loads in flight while the
// current one is applied.
#include "bucket.hpp"

#include <algorithm>
#include <stdexcept>

#include "sort_core.hpp"

namespace sptc {

namespace {

constexpr int kR = 1 << kBucketRB;
constexpr int kBW = kBlock / kWave;  // waves per workgroup
constexpr int kMaxGroup = 6;         // at most 64 buckets per wave

struct BucketWave {
  uint64_t seg;
  uint64_t ft;     // the row's first sort tile (its rows of the matrices)
  uint64_t start;  // the row's first product
  uint32_t d0, nb;
  uint32_t s0, s1;  // products of buckets d0 .. d0 + nb - 1
  bool live;
};

// Work item gw of the persistent waves: big row gw >> (kBucketRB - gshift),
// its bucket group gw's low bits.
__device__ __forceinline__ uint64_t bucket_items(const BucketArgs& a) {
  return a.nsegs << (kBucketRB - a.gshift);
}
__device__ __forceinline__ BucketWave bucket_wave(const BucketArgs& a, uint64_t gw) {
  BucketWave w;
  w.seg = gw >> (kBucketRB - a.gshift);
  w.live = w.seg < a.nsegs;
  w.nb = 1u << a.gshift;
  w.d0 = (uint32_t)(gw & (((uint64_t)kR >> a.gshift) - 1)) << a.gshift;
  w.ft = w.start = 0;
  w.s0 = w.s1 = 0;
  if (!w.live) return w;
  const SegDesc g = a.segs[w.seg];
  w.ft = g.first_tile;
  w.start = g.start;
  w.s0 = a.starts[w.ft * kR + w.d0];
  if (w.d0 + w.nb < (uint32_t)kR) {
    w.s1 = a.starts[w.ft * kR + w.d0 + w.nb];
  } else {
    const SegTile lt = a.tiles[w.ft + g.ntiles - 1];
    w.s1 = (uint32_t)(lt.start + lt.count);
  }
  return w;
}

// Marks of the touched columns of one bucket, a byte per column (plain
// stores: lanes marking one column write the same byte), read back as
// words: lane l takes columns [l * S, l * S + S); returns their bit mask
// and clears them.
template <int LM>
__device__ __forceinline__ uint32_t take_marks(uint8_t* seen, int L, uint32_t lane,
                                               uint32_t* first) {
  const uint32_t nslot = 1u << L;
  uint32_t bits = 0;
  if (L >= 8) {  // S = 4 .. 16 bytes: whole words
    const uint32_t S = nslot >> 6;
    *first = lane * S;
    uint32_t* wp = reinterpret_cast<uint32_t*>(seen + *first);
    for (uint32_t k = 0; k < S / 4; ++k) {
      const uint32_t x = wp[k];
      bits |= (((x & 1u) | ((x >> 7) & 2u) | ((x >> 14) & 4u) | ((x >> 21) & 8u))) << (4 * k);
      wp[k] = 0;
    }
  } else if (L >= 6) {  // S = 1 or 2 bytes
    const uint32_t S = nslot >> 6;
    *first = lane * S;
    for (uint32_t e = 0; e < S; ++e) {
      bits |= (uint32_t)(seen[*first + e] != 0) << e;
      seen[*first + e] = 0;
    }
  } else {
    *first = lane;
    if (lane < nslot) {
      bits = seen[lane] != 0;
      seen[lane] = 0;
    }
  }
  return bits;
}

// Distinct columns per bucket: the wave's buckets one after another, each
// column marked in LDS; the count of a bucket is taken when the stream
// reaches a later bucket (empty buckets count 0).
template <int LM>
__global__ __launch_bounds__(kBlock) void k_bucket_count(BucketArgs a) {
  __shared__ __align__(16) uint8_t seen_all[kBW][1 << LM];
  const uint32_t lane = lane_id();
  uint8_t* seen = seen_all[wave_id()];
  const int L = a.low;
  const uint32_t nslot = 1u << L;
  const uint32_t lmask = nslot - 1;
  for (uint32_t i = lane; i < nslot; i += kWave) seen[i] = 0;  // left clear by every close
  wave_sync();
  const uint64_t items = bucket_items(a), stride = (uint64_t)gridDim.x * kBW;
  for (uint64_t gw = (uint64_t)blockIdx.x * kBW + wave_id(); gw < items; gw += stride) {
    const BucketWave w = bucket_wave(a, gw);
    uint32_t mine_cnt = 0;  // lane i: bucket d0 + i
    auto close = [&](uint32_t cur) {
      wave_sync();
      uint32_t first;
      const uint32_t bits = take_marks<LM>(seen, L, lane, &first);
      const uint32_t t = __builtin_amdgcn_readlane(wave_incl_scan32((uint32_t)__popc(bits)),
                                                   kWave - 1);
      if (lane == cur - w.d0) mine_cnt = t;
      wave_sync();
    };
    uint32_t cur = w.d0;
    uint32_t c = w.s0 + lane < w.s1 ? a.cols[w.s0 + lane] : 0u;
    for (uint32_t p = w.s0; p < w.s1; p += kWave) {
      const uint32_t pn = p + kWave;
      const uint32_t c2 = pn + lane < w.s1 ? a.cols[pn + lane] : 0u;
      const uint32_t d = c >> L;
      uint64_t pending = __ballot(p + lane < w.s1);
      while (pending) {
        const uint64_t mine = pending & __ballot(d == cur);
        if ((mine >> lane) & 1ull) seen[c & lmask] = 1;
        pending &= ~mine;
        if (pending) {
          close(cur);
          cur = __shfl(d, (int)__builtin_ctzll(pending), kWave);
        }
      }
      c = c2;
    }
    close(cur);
    if (lane < w.nb) a.boff[w.ft * kR + w.d0 + lane] = mine_cnt;
  }
}

// One workgroup per big row: bucket counts -> offsets within the row; the
// row's total joins the wave-tile count of the tile after its first product.
__global__ __launch_bounds__(kR) void k_bucket_scan(BucketArgs a) {
  __shared__ uint32_t scratch[kR / kWave];
  const SegDesc g = a.segs[blockIdx.x];
  const uint64_t i = (uint64_t)g.first_tile * kR + threadIdx.x;
  uint32_t tot;
  const uint32_t ex = block_scan_u32<kR>(a.boff[i], scratch, &tot);
  a.boff[i] = ex;
  if (threadIdx.x == 0 && tot)
    atomicAdd(reinterpret_cast<unsigned long long*>(&a.tile_sums[g.start / kWaveTile + 1].a),
              (unsigned long long)tot);
}

template <int LM>
struct EmitSmem {
  static constexpr int E = 1 << LM;
  float2 slot[E];       // (sum, compensation) per column of the bucket
  uint32_t srank[E];    // staged outputs: column rank, value
  float sval[E];
  alignas(16) uint8_t seen[E];  // touched columns
};

// Adds the products of lanes `mine` (one bucket) to their columns' slots:
// each column's products in lane (= product) order, one compensated step
// each.  The lowest lane of each column (peers found by one ballot per
// column bit) reads the slot once, takes the others' values in lane order
// and writes it back, so no two lanes touch one slot.
template <int LM>
__device__ __forceinline__ void bucket_accumulate(EmitSmem<LM>& sm, uint64_t mine, uint32_t cl,
                                                  float v, int L, uint32_t lane) {
  uint64_t peers = mine;
  for (int b = 0; b < L; ++b) {
    const bool bit = (cl >> b) & 1u;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  const bool lead = ((mine >> lane) & 1ull) && popc_below(peers) == 0;
  float s = 0.f, cc = 0.f;
  uint64_t rest = 0;
  if (lead) {
    const float2 q = sm.slot[cl];
    s = q.x;
    cc = q.y;
    kahan_step(s, cc, v);
    rest = peers & (peers - 1);
  }
  while (__ballot(rest != 0)) {
    const uint32_t src = rest ? (uint32_t)__builtin_ctzll(rest) : lane;
    const float x = __shfl(v, (int)src, kWave);
    if (rest) {
      kahan_step(s, cc, x);
      rest &= rest - 1;
    }
  }
  if (lead) {
    sm.slot[cl] = make_float2(s, cc);
    sm.seen[cl] = 1;
  }
  wave_sync();
}

// Stores the staged outputs at q (ranks resolved through the Y table),
// four 64-output rows of gathers in flight.
template <int LM>
__device__ __forceinline__ void bucket_flush(const BucketArgs& a, EmitSmem<LM>& sm,
                                             uint32_t staged, uint64_t q, uint32_t row,
                                             uint32_t lane) {
  for (uint32_t k = 0; k < staged; k += 4 * kWave) {
    uint32_t rk[4];
    float val[4];
    uint64_t col[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t i = k + (uint32_t)u * kWave + lane;
      rk[u] = i < staged ? sm.srank[i] : 0u;
      val[u] = i < staged ? sm.sval[i] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t i = k + (uint32_t)u * kWave + lane;
      col[u] = rk[u];
      if (i < staged) {
        if (a.ytab32) col[u] = a.ytab32[rk[u]];
        else if (a.ytab) col[u] = a.ytab[rk[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint64_t i = k + (uint32_t)u * kWave + lane;
      if (i < staged) {
        __builtin_nontemporal_store(row, &a.out_row[q + i]);
        __builtin_nontemporal_store(col[u], &a.out_col[q + i]);
        __builtin_nontemporal_store(val[u], &a.out_val[q + i]);
      }
    }
  }
}

template <int LM>
__global__ __launch_bounds__(kBlock) void k_bucket_emit(BucketArgs a) {
  __shared__ EmitSmem<LM> sm_all[kBW];
  constexpr uint32_t E = EmitSmem<LM>::E;
  const uint32_t lane = lane_id();
  EmitSmem<LM>& sm = sm_all[wave_id()];
  const int L = a.low;
  const uint32_t nslot = 1u << L;
  const uint32_t lmask = nslot - 1;
  // slots and marks start clear; every compaction clears what it takes
  for (uint32_t i = lane; i < nslot; i += kWave) {
    sm.slot[i] = make_float2(0.f, 0.f);
    sm.seen[i] = 0;
  }
  wave_sync();
  const uint64_t items = bucket_items(a), stride = (uint64_t)gridDim.x * kBW;
  for (uint64_t gw = (uint64_t)blockIdx.x * kBW + wave_id(); gw < items; gw += stride) {
    const BucketWave w = bucket_wave(a, gw);
    const uint32_t row = (uint32_t)a.r0 + a.bigrow[w.seg];
    // the wave's outputs are contiguous: the row's first output plus the
    // first bucket's offset within the row
    uint64_t q = a.tile_sums[w.start / kWaveTile + 1].a + a.boff[w.ft * kR + w.d0];
    uint32_t staged = 0;
    // the touched slots of bucket `cur` to the stage in column order (lane l
    // takes slots [l * S, l * S + S)), then cleared
    auto compact = [&](uint32_t cur) {
      uint32_t first;
      uint32_t bits = take_marks<LM>(sm.seen, L, lane, &first);
      const uint32_t h = (uint32_t)__popc(bits);
      const uint32_t inc = wave_incl_scan32(h);
      const uint32_t T = __builtin_amdgcn_readlane(inc, kWave - 1);
      if (staged + T > E) {
        bucket_flush(a, sm, staged, q, row, lane);
        q += staged;
        staged = 0;
        wave_sync();
      }
      uint32_t o = staged + inc - h;
      while (bits) {
        const uint32_t sl = first + (uint32_t)__builtin_ctz(bits);
        bits &= bits - 1;
        const float2 acc = sm.slot[sl];
        sm.slot[sl] = make_float2(0.f, 0.f);
        sm.srank[o] = (cur << L) | sl;
        sm.sval[o] = acc.x + acc.y;
        ++o;
      }
      staged += T;
      wave_sync();
    };
    uint32_t cur = w.d0;
    uint32_t c = 0;
    float v = 0.f;
    if (w.s0 + lane < w.s1) {
      c = a.cols[w.s0 + lane];
      v = a.vals[w.s0 + lane];
    }
    for (uint32_t p = w.s0; p < w.s1; p += kWave) {
      const uint32_t pn = p + kWave;
      uint32_t c2 = 0;
      float v2 = 0.f;
      if (pn + lane < w.s1) {
        c2 = a.cols[pn + lane];
        v2 = a.vals[pn + lane];
      }
      const uint32_t d = c >> L;
      const uint32_t cl = c & lmask;
      uint64_t pending = __ballot(p + lane < w.s1);
      while (pending) {
        const uint64_t mine = pending & __ballot(d == cur);
        if (mine) bucket_accumulate(sm, mine, cl, v, L, lane);
        pending &= ~mine;
        if (pending) {  // later buckets in this window: close the current one
          compact(cur);
          cur = __shfl(d, (int)__builtin_ctzll(pending), kWave);
        }
      }
      c = c2;
      v = v2;
    }
    compact(cur);
    bucket_flush(a, sm, staged, q, row, lane);
    wave_sync();
  }
}

// Persistent grid: about as many workgroups as stay resident (per CU,
// `per_cu` of them), never more than there are items.
uint32_t bucket_blocks(const BucketArgs& a, uint32_t per_cu) {
  static const int cus = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      n = 256;
    return n > 0 ? n : 256;
  }();
  const uint64_t waves = a.nsegs << (kBucketRB - a.gshift);
  return (uint32_t)std::min<uint64_t>(ceil_div(waves, kBW), (uint64_t)cus * per_cu);
}

void check_shape(const BucketArgs& a) {
  if (a.low < 1 || a.low > kBucketMaxLow || a.gshift < 0 || a.gshift > kMaxGroup)
    throw std::logic_error("bucket reduce shape");
}

}  // namespace

int bucket_group_shift(uint64_t nprod, uint64_t nsegs) {
  const double avg = nsegs ? (double)nprod / ((double)nsegs * kR) : 1.0;
  int g = 0;
  while (g < kMaxGroup && avg * (double)(2u << g) <= 2048.0) ++g;
  return g;
}

void run_bucket_count(const BucketArgs& a, hipStream_t s) {
  if (!a.nsegs) return;
  check_shape(a);
  const dim3 g(bucket_blocks(a, 8)), b(kBlock);
  if (a.low <= 6) hipLaunchKernelGGL(k_bucket_count<6>, g, b, 0, s, a);
  else if (a.low <= 8) hipLaunchKernelGGL(k_bucket_count<8>, g, b, 0, s, a);
  else hipLaunchKernelGGL(k_bucket_count<10>, g, b, 0, s, a);
  SPTC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_bucket_scan, dim3((uint32_t)a.nsegs), dim3(kR), 0, s, a);
  SPTC_HIP(hipGetLastError());
}

void run_bucket_emit(const BucketArgs& a, hipStream_t s) {
  if (!a.nsegs) return;
  check_shape(a);
  // resident workgroups per CU as the LDS allows (160 KB per CU)
  const uint32_t per_cu = a.low <= 8 ? 8 : a.low == 9 ? 4 : 2;
  const dim3 g(bucket_blocks(a, per_cu)), b(kBlock);
  switch (a.low <= 6 ? 6 : a.low) {
    case 6: hipLaunchKernelGGL(k_bucket_emit<6>, g, b, 0, s, a); break;
    case 7: hipLaunchKernelGGL(k_bucket_emit<7>, g, b, 0, s, a); break;
    case 8: hipLaunchKernelGGL(k_bucket_emit<8>, g, b, 0, s, a); break;
    case 9: hipLaunchKernelGGL(k_bucket_emit<9>, g, b, 0, s, a); break;
    default: hipLaunchKernelGGL(k_bucket_emit<10>, g, b, 0, s, a); break;
  }
  SPTC_HIP(hipGetLastError());
}

}  // namespace sptc
