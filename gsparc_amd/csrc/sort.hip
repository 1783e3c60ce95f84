// Stable LSD radix sort of (uint64 key, float value) pairs for gfx950.
//
// Replaces cub::DeviceRadixSort::SortPairs in the reference (sort.cuh:90-108,
// contraction.cuh:506-530, tensor_manager.tpp:779-786).  Design:
//   * Only the significant key bits are sorted: ceil(nbits / RB) passes.
//   * One up-front pass computes every digit histogram at once; producers
//     (linearize, ESC expand) fuse it, so the sort never re-reads keys for
//     counting.  digit_scan turns them into per-pass digit bases.
//   * Each pass is a single kernel ("onesweep"): a workgroup takes the next
//     tile id from an atomic counter, ranks its 256*IPT keys with wave64
//     ballot matching (one ballot per digit bit, stable within the wave),
//     publishes per-digit tile counts, resolves its global offsets by
//     decoupled look-back over predecessor tiles, stages the tile in LDS in
//     digit order and writes it out in coalesced digit runs.
//   * Look-back status words carry a pass epoch so the status array is never
//     cleared between passes: [63:62] flag (1 aggregate, 2 inclusive prefix),
//     [61:48] epoch, [47:0] count.
#include "common.hpp"
#include "sort.hpp"
#include "scan.hpp"

#include <cstdlib>
#include <string>

namespace sptc {

namespace {

constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPre = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 48) - 1;
constexpr uint32_t kSpinLimit = 1u << 28;

__device__ __forceinline__ uint64_t pack_status(uint64_t flag, uint32_t epoch,
                                                uint64_t v) {
  return flag | ((uint64_t)epoch << 48) | (v & kValMask);
}

__global__ void k_fill_u64(uint64_t* p, uint64_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Standalone histogram of all passes (used when no producer fused it).
template <int RB>
__global__ __launch_bounds__(kBlock) void k_histogram(const uint64_t* __restrict__ keys,
                                                      uint64_t n, int passes,
                                                      uint64_t* __restrict__ hist) {
  constexpr int R = 1 << RB;
  __shared__ uint32_t sh[kMaxPasses * R];
  for (int i = threadIdx.x; i < passes * R; i += kBlock) sh[i] = 0;
  __syncthreads();
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; i < n; i += stride) {
    uint64_t k = keys[i];
    for (int p = 0; p < passes; ++p)
      atomicAdd(&sh[p * R + ((k >> (p * RB)) & (R - 1))], 1u);
  }
  __syncthreads();
  for (int i2 = threadIdx.x; i2 < passes * R; i2 += kBlock)
    if (sh[i2]) atomicAdd((unsigned long long*)&hist[i2], (unsigned long long)sh[i2]);
}

// hist [passes][R] counts -> base [passes][R] exclusive starts.  One block per
// pass.
template <int RB>
__global__ __launch_bounds__(kBlock) void k_digit_scan(const uint64_t* __restrict__ hist,
                                                       uint64_t* __restrict__ base) {
  constexpr int R = 1 << RB;
  constexpr int PER = (R + kBlock - 1) / kBlock;
  __shared__ uint64_t scratch[kBlock / kWave];
  const int p = blockIdx.x;
  uint64_t local[PER];
  uint64_t sum = 0;
  // thread t owns digits [t*PER, t*PER+PER)
  for (int j = 0; j < PER; ++j) {
    int d = threadIdx.x * PER + j;
    local[j] = d < R ? hist[p * R + d] : 0;
    sum += local[j];
  }
  uint64_t tot;
  uint64_t ex = block_excl_scan<uint64_t>(sum, scratch, &tot);
  for (int j = 0; j < PER; ++j) {
    int d = threadIdx.x * PER + j;
    if (d < R) base[p * R + d] = ex;
    ex += local[j];
  }
}

template <int RB, int IPT>
__global__ __launch_bounds__(kBlock) void k_onesweep(
    const uint64_t* __restrict__ kin, const float* __restrict__ vin,
    uint64_t* __restrict__ kout, float* __restrict__ vout, uint64_t n, int shift,
    const uint64_t* __restrict__ digit_base, uint64_t* status,
    uint32_t* tile_counter, uint32_t epoch, uint32_t* err) {
  constexpr int R = 1 << RB;
  constexpr int TILE = kBlock * IPT;
  constexpr int NW = kBlock / kWave;
  static_assert(TILE < 65536, "16-bit tile offsets");
  __shared__ uint16_t s_cnt[NW][R];     // per-wave running counts, then prefixes
  __shared__ uint64_t s_gbase[R];       // global start of this tile's digit run
  __shared__ uint16_t s_dstart[R];      // tile-local start of each digit run
  __shared__ uint64_t s_buf[TILE];      // staged keys, then staged values
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_scan[NW];

  const uint32_t tid = threadIdx.x;
  const uint32_t wave = tid / kWave;
  const uint32_t lane = lane_id();
  if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
  for (int i = tid; i < NW * R; i += kBlock) (&s_cnt[0][0])[i] = 0;
  __syncthreads();
  const uint64_t tile = s_tile;
  const uint64_t tile_begin = tile * TILE;
  const uint64_t wbase = tile_begin + (uint64_t)wave * kWave * IPT;

  uint64_t k[IPT];
  float v[IPT];
  uint32_t pos[IPT];
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    bool valid = idx < n;
    k[i] = valid ? kin[idx] : 0;
    v[i] = valid ? vin[idx] : 0.f;
  }
  // Rank: rounds in order, lanes in order => stable within the wave's run.
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    bool valid = idx < n;
    uint32_t d = (uint32_t)(k[i] >> shift) & (R - 1);
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      bool bit = (d >> b) & 1u;
      uint64_t bb = ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    uint32_t before = 0;
    if (valid) before = s_cnt[wave][d];
    __builtin_amdgcn_wave_barrier();
    pos[i] = before + popc_below(peers);  // rank within the wave's digit run
    if (valid && (peers >> lane) == 1ull)  // highest lane of the group advances
      s_cnt[wave][d] = (uint16_t)(before + (uint32_t)__popcll(peers));
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();

  // Per digit: wave prefixes, tile total, publish + look-back.
  for (int d = tid; d < R; d += kBlock) {
    uint32_t acc = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      uint32_t c = s_cnt[w][d];
      s_cnt[w][d] = (uint16_t)acc;
      acc += c;
    }
    const uint64_t total = acc;
    uint64_t excl = 0;
    uint64_t* my = status + tile * R + d;
    if (tile == 0) {
      st_agent(my, pack_status(kFlagPre, epoch, total));
    } else {
      st_agent(my, pack_status(kFlagAgg, epoch, total));
      int64_t j = (int64_t)tile - 1;
      uint32_t spins = 0;
      while (j >= 0) {
        uint64_t st = ld_agent(status + (uint64_t)j * R + d);
        uint32_t se = (uint32_t)((st >> 48) & 0x3FFF);
        uint64_t fl = st & (3ull << 62);
        if (se != epoch || fl == 0) {
          if (++spins > kSpinLimit) { atomicOr(err, 1u); break; }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        excl += st & kValMask;
        if (fl == kFlagPre) break;
        --j;
      }
      st_agent(my, pack_status(kFlagPre, epoch, excl + total));
    }
    s_gbase[d] = digit_base[d] + excl;
    s_dstart[d] = (uint16_t)acc;  // temporarily the tile total
  }
  __syncthreads();
  // Exclusive scan of tile totals over digits -> s_dstart.
  {
    constexpr int PER = (R + kBlock - 1) / kBlock;
    uint32_t loc[PER];
    uint32_t sum = 0;
    for (int j = 0; j < PER; ++j) {
      int d = tid * PER + j;
      loc[j] = d < R ? s_dstart[d] : 0;
      sum += loc[j];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan<uint32_t>(sum, s_scan, &tot);
    for (int j = 0; j < PER; ++j) {
      int d = tid * PER + j;
      if (d < R) s_dstart[d] = (uint16_t)ex;
      ex += loc[j];
    }
  }
  __syncthreads();
  // Phase 1: stage keys in digit order, write them out in coalesced runs.
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    uint32_t d = (uint32_t)(k[i] >> shift) & (R - 1);
    pos[i] += (uint32_t)s_dstart[d] + (uint32_t)s_cnt[wave][d];
    if (idx < n) s_buf[pos[i]] = k[i];
  }
  __syncthreads();
  const uint32_t cnt = (uint32_t)((n - tile_begin) < (uint64_t)TILE ? (n - tile_begin) : TILE);
  uint64_t dst[IPT];
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const uint32_t j = tid + (uint32_t)i * kBlock;
    if (j < cnt) {
      const uint64_t key = s_buf[j];
      const uint32_t d = (uint32_t)(key >> shift) & (R - 1);
      dst[i] = s_gbase[d] + (j - s_dstart[d]);
      kout[dst[i]] = key;
    }
  }
  __syncthreads();
  // Phase 2: the values through the same buffer, same destinations.
  float* s_vbuf = reinterpret_cast<float*>(s_buf);
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    if (idx < n) s_vbuf[pos[i]] = v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const uint32_t j = tid + (uint32_t)i * kBlock;
    if (j < cnt) vout[dst[i]] = s_vbuf[j];
  }
}

// ---- reduce-then-scan pass (no inter-tile serialization) -----------------
// upsweep: per-tile digit counts into a digit-major matrix cnt[d][tile];
// an exclusive scan of that matrix, read in memory order, is exactly each
// tile's global start for each digit; downsweep ranks, stages and scatters
// like the onesweep kernel but takes its offsets from the scanned matrix.
template <int RB, int IPT>
__global__ __launch_bounds__(kBlock) void k_rts_upsweep(const uint64_t* __restrict__ kin,
                                                        uint64_t n, int shift,
                                                        uint32_t* __restrict__ cnt,
                                                        uint64_t ntiles) {
  constexpr int R = 1 << RB;
  constexpr int TILE = kBlock * IPT;
  __shared__ uint32_t sh[R];
  for (int i = threadIdx.x; i < R; i += kBlock) sh[i] = 0;
  __syncthreads();
  const uint64_t tile = blockIdx.x;
  const uint64_t base = tile * TILE;
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
    const bool valid = idx < n;
    const uint64_t k = valid ? kin[idx] : 0;
    hist_add_wave<RB>(sh, k >> shift, valid, 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < R; d += kBlock) cnt[(uint64_t)d * ntiles + tile] = sh[d];
}

template <int RB, int IPT, bool STAGE>
__global__ __launch_bounds__(kBlock) void k_rts_downsweep(
    const uint64_t* __restrict__ kin, const float* __restrict__ vin,
    uint64_t* __restrict__ kout, float* __restrict__ vout, uint64_t n, int shift,
    const uint32_t* __restrict__ offs, uint64_t ntiles) {
  constexpr int R = 1 << RB;
  constexpr int TILE = kBlock * IPT;
  constexpr int NW = kBlock / kWave;
  __shared__ uint16_t s_cnt[NW][R];
  __shared__ uint64_t s_gbase[R];
  __shared__ uint16_t s_dstart[R];
  __shared__ uint64_t s_buf[TILE];
  __shared__ uint32_t s_scan[NW];
  const uint32_t tid = threadIdx.x;
  const uint32_t wave = tid / kWave;
  const uint32_t lane = lane_id();
  for (int i = tid; i < NW * R; i += kBlock) (&s_cnt[0][0])[i] = 0;
  const uint64_t tile = blockIdx.x;
  for (int d = tid; d < R; d += kBlock) s_gbase[d] = offs[(uint64_t)d * ntiles + tile];
  __syncthreads();
  const uint64_t tile_begin = tile * TILE;
  const uint64_t wbase = tile_begin + (uint64_t)wave * kWave * IPT;
  uint64_t k[IPT];
  float v[IPT];
  uint32_t pos[IPT];
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    bool valid = idx < n;
    k[i] = valid ? kin[idx] : 0;
    v[i] = valid ? vin[idx] : 0.f;
  }
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    bool valid = idx < n;
    uint32_t d = (uint32_t)(k[i] >> shift) & (R - 1);
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      bool bit = (d >> b) & 1u;
      uint64_t bb = ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    uint32_t before = 0;
    if (valid) before = s_cnt[wave][d];
    __builtin_amdgcn_wave_barrier();
    pos[i] = before + popc_below(peers);
    if (valid && (peers >> lane) == 1ull)
      s_cnt[wave][d] = (uint16_t)(before + (uint32_t)__popcll(peers));
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  for (int d = tid; d < R; d += kBlock) {
    uint32_t acc = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      uint32_t c = s_cnt[w][d];
      s_cnt[w][d] = (uint16_t)acc;
      acc += c;
    }
    s_dstart[d] = (uint16_t)acc;
  }
  __syncthreads();
  if constexpr (!STAGE) {
    // Direct scatter from registers: destination = tile's digit start +
    // the wave prefix within the tile + the rank within the wave.
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      uint64_t idx = wbase + (uint64_t)i * kWave + lane;
      if (idx < n) {
        uint32_t d = (uint32_t)(k[i] >> shift) & (R - 1);
        uint64_t o = s_gbase[d] + s_cnt[wave][d] + pos[i];
        kout[o] = k[i];
        vout[o] = v[i];
      }
    }
    return;
  }
  {
    constexpr int PER = (R + kBlock - 1) / kBlock;
    uint32_t loc[PER];
    uint32_t sum = 0;
    for (int j = 0; j < PER; ++j) {
      int d = tid * PER + j;
      loc[j] = d < R ? s_dstart[d] : 0;
      sum += loc[j];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan<uint32_t>(sum, s_scan, &tot);
    for (int j = 0; j < PER; ++j) {
      int d = tid * PER + j;
      if (d < R) s_dstart[d] = (uint16_t)ex;
      ex += loc[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    uint32_t d = (uint32_t)(k[i] >> shift) & (R - 1);
    pos[i] += (uint32_t)s_dstart[d] + (uint32_t)s_cnt[wave][d];
    if (idx < n) s_buf[pos[i]] = k[i];
  }
  __syncthreads();
  const uint32_t cnt = (uint32_t)((n - tile_begin) < (uint64_t)TILE ? (n - tile_begin) : TILE);
  uint64_t dst[IPT];
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const uint32_t j = tid + (uint32_t)i * kBlock;
    if (j < cnt) {
      const uint64_t key = s_buf[j];
      const uint32_t d = (uint32_t)(key >> shift) & (R - 1);
      dst[i] = s_gbase[d] + (j - s_dstart[d]);
      kout[dst[i]] = key;
    }
  }
  __syncthreads();
  float* s_vbuf = reinterpret_cast<float*>(s_buf);
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    if (idx < n) s_vbuf[pos[i]] = v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const uint32_t j = tid + (uint32_t)i * kBlock;
    if (j < cnt) vout[dst[i]] = s_vbuf[j];
  }
}

// Exclusive scan of the uint32 count matrix into uint32 offsets (n < 2^32).
struct CountScanOp {
  const uint32_t* in;
  uint32_t* out;
  __device__ U2 value(uint64_t i) const { return mk2(in[i], 0); }
  __device__ void emit(uint64_t i, U2 excl, U2) const { out[i] = (uint32_t)excl.a; }
};

template <int RB>
constexpr int ipt_for() { return 12; }

// Reduce-then-scan variant for 8-bit digits (SPTC_RTS): tile items per
// thread and LDS staging ("12s" default, "16s", "8s", "12d" = direct scatter).
struct RtsVariant {
  int ipt = 12;
  bool stage = true;
};
static bool use_onesweep() {
  static const bool v = [] {
    const char* e = getenv("SPTC_SORT");
    return e && std::string(e) == "onesweep";
  }();
  return v;
}

static RtsVariant rts_variant() {
  static const RtsVariant v = [] {
    RtsVariant r;
    const char* e = getenv("SPTC_RTS");
    if (!e || use_onesweep()) return r;
    std::string t(e);
    if (t == "16s") r.ipt = 16;
    if (t == "8s") r.ipt = 8;
    if (t == "12d") r.stage = false;
    return r;
  }();
  return v;
}

}  // namespace

void fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s) {
  if (!n) return;
  uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(n, kBlock), 4096);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid), dim3(kBlock), 0, s, p, v, n);
}
void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
  if (!n) return;
  uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(n, kBlock), 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(kBlock), 0, s, p, v, n);
}

int radix_passes(int nbits, int rb) { return nbits <= 0 ? 0 : (nbits + rb - 1) / rb; }

uint64_t sort_tile_items(int rb) {
  if (rb == 8) return (uint64_t)kBlock * rts_variant().ipt;
  switch (rb) {
    case 9: return (uint64_t)kBlock * ipt_for<9>();
    case 10: return (uint64_t)kBlock * ipt_for<10>();
    case 11: return (uint64_t)kBlock * ipt_for<11>();
    default: return (uint64_t)kBlock * ipt_for<8>();
  }
}

bool radix_bits_supported(int rb) { return rb >= 8 && rb <= 11; }

// Fewest passes with digits of at most 10 bits; the narrowest digit among
// those that reach that pass count.
int choose_radix_bits(int nbits) {
  int best = 8, best_p = radix_passes(nbits, 8);
  for (int rb = 9; rb <= 10; ++rb) {
    int p = radix_passes(nbits, rb);
    if (p < best_p) { best = rb; best_p = p; }
  }
  return best;
}

uint64_t SortWorkspace::bytes_per_key(int rb) {
  // alt keys + values, look-back words, and the two count matrices
  return sizeof(uint64_t) + sizeof(float) + (16ull << rb) / sort_tile_items(rb) + 1;
}



void SortWorkspace::ensure(uint64_t n, int rb_, hipStream_t s) {
  const uint64_t R = 1ull << rb_;
  if (counters.size() == 0) {
    counters.reset(64);
    fill_u32(counters.get(), 0, 64, s);
  }
  if (hist.size() < kMaxPasses * R) {
    hist.reset(kMaxPasses * R);
    base.reset(kMaxPasses * R);
  }
  if (alt_keys.size() < n || alt_vals.size() < n) {
    alt_keys.reset(n ? n : 1);
    alt_vals.reset(n ? n : 1);
  }
  const uint64_t tiles = ceil_div(n ? n : 1, sort_tile_items(rb_));
  if (status.size() < tiles * R) {
    status.reset(tiles * R);
    fill_u64(status.get(), 0, tiles * R, s);
    epoch = 0;
  }
  if (rts_cnt.size() < tiles * R) {
    rts_cnt.reset(tiles * R);
    rts_off.reset(tiles * R);
    rts_tmp.reset(scan_tmp_count(tiles * R) + 1);
  }
  rb = rb_;
  cap = alt_keys.size();
  max_tiles = status.size() / R;
}

void SortWorkspace::prepare(int passes, hipStream_t s) {
  uint64_t R = 1ull << rb;
  fill_u64(hist.get(), 0, (uint64_t)passes * R, s);
  fill_u32(counters.get(), 0, kMaxPasses, s);
}

void histogram(const SortWorkspace& ws, const uint64_t* keys, uint64_t n,
               int passes, hipStream_t s) {
  if (!n || !passes) return;
  uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(n, kBlock * 8), 2048);
  switch (ws.rb) {
    case 9: hipLaunchKernelGGL(k_histogram<9>, dim3(grid), dim3(kBlock), 0, s, keys, n, passes,
                               ws.hist.get()); break;
    case 10: hipLaunchKernelGGL(k_histogram<10>, dim3(grid), dim3(kBlock), 0, s, keys, n, passes,
                                ws.hist.get()); break;
    case 11: hipLaunchKernelGGL(k_histogram<11>, dim3(grid), dim3(kBlock), 0, s, keys, n, passes,
                                ws.hist.get()); break;
    default: hipLaunchKernelGGL(k_histogram<8>, dim3(grid), dim3(kBlock), 0, s, keys, n, passes,
                                ws.hist.get()); break;
  }
}

// Sorts keys/vals (n <= ws.cap) on bits [0, nbits) assuming ws.hist already
// holds the digit histograms of all passes.  Result ends in (keys, vals) when
// the pass count is even, else in (ws.alt_keys, ws.alt_vals); the returned
// flag says which (true = alt).
bool sort_after_histogram(SortWorkspace& ws, uint64_t* keys, float* vals,
                          uint64_t n, int nbits, hipStream_t s, KTimer* kt, int cls) {
  const int rb = ws.rb;
  const int passes = radix_passes(nbits, rb);
  if (!n || !passes) return false;
  const uint64_t R = 1ull << rb;
  switch (rb) {
    case 9: hipLaunchKernelGGL(k_digit_scan<9>, dim3(passes), dim3(kBlock), 0, s, ws.hist.get(),
                               ws.base.get()); break;
    case 10: hipLaunchKernelGGL(k_digit_scan<10>, dim3(passes), dim3(kBlock), 0, s, ws.hist.get(),
                                ws.base.get()); break;
    case 11: hipLaunchKernelGGL(k_digit_scan<11>, dim3(passes), dim3(kBlock), 0, s, ws.hist.get(),
                                ws.base.get()); break;
    default: hipLaunchKernelGGL(k_digit_scan<8>, dim3(passes), dim3(kBlock), 0, s, ws.hist.get(),
                                ws.base.get()); break;
  }
  const uint64_t tiles = ceil_div(n, sort_tile_items(rb));
  uint64_t* ka = keys;
  float* va = vals;
  uint64_t* kb = ws.alt_keys.get();
  float* vb = ws.alt_vals.get();
  if (!use_onesweep()) {
    const uint64_t m = tiles * R;
    for (int p = 0; p < passes; ++p) {
      const int shift = p * rb;
      int tok = kt ? kt->begin(s) : -1;
#define SPTC_RTS_UP(RBV, IPTV)                                                          \
  hipLaunchKernelGGL((k_rts_upsweep<RBV, IPTV>), dim3((uint32_t)tiles), dim3(kBlock), 0, s, ka, \
                     n, shift, ws.rts_cnt.get(), tiles)
      const RtsVariant var = rts_variant();
      switch (rb) {
        case 9: SPTC_RTS_UP(9, 12); break;
        case 10: SPTC_RTS_UP(10, 12); break;
        case 11: SPTC_RTS_UP(11, 12); break;
        default:
          if (var.ipt == 16) SPTC_RTS_UP(8, 16);
          else if (var.ipt == 8) SPTC_RTS_UP(8, 8);
          else SPTC_RTS_UP(8, 12);
          break;
      }
#undef SPTC_RTS_UP
      run_scan(CountScanOp{ws.rts_cnt.get(), ws.rts_off.get()}, m, ws.rts_tmp.get(), s);
      if (kt) kt->end(tok, SPTC_K_SORT_COUNT, 8.0 * (double)n + 12.0 * (double)m, s);
      tok = kt ? kt->begin(s) : -1;
#define SPTC_RTS_DOWN(RBV, IPTV, ST)                                                     \
  hipLaunchKernelGGL((k_rts_downsweep<RBV, IPTV, ST>), dim3((uint32_t)tiles), dim3(kBlock), 0, \
                     s, ka, va, kb, vb, n, shift, ws.rts_off.get(), tiles)
      switch (rb) {
        case 9: SPTC_RTS_DOWN(9, 12, true); break;
        case 10: SPTC_RTS_DOWN(10, 12, true); break;
        case 11: SPTC_RTS_DOWN(11, 12, true); break;
        default:
          if (var.ipt == 16) SPTC_RTS_DOWN(8, 16, true);
          else if (var.ipt == 8) SPTC_RTS_DOWN(8, 8, true);
          else if (!var.stage) SPTC_RTS_DOWN(8, 12, false);
          else SPTC_RTS_DOWN(8, 12, true);
          break;
      }
#undef SPTC_RTS_DOWN
      if (kt) kt->end(tok, cls, 24.0 * (double)n, s);
      std::swap(ka, kb);
      std::swap(va, vb);
    }
    return (passes & 1) != 0;
  }
  for (int p = 0; p < passes; ++p) {
    ws.epoch = (ws.epoch + 1) & 0x3FFF;
    if (ws.epoch == 0) {
      fill_u64(ws.status.get(), 0, ws.max_tiles * R, s);
      ws.epoch = 1;
    }
    const int tok = kt ? kt->begin(s) : -1;
#define SPTC_ONESWEEP(RBV)                                                                \
  hipLaunchKernelGGL((k_onesweep<RBV, ipt_for<RBV>()>), dim3((uint32_t)tiles), dim3(kBlock), \
                     0, s, ka, va, kb, vb, n, p * rb, ws.base.get() + p * R,               \
                     ws.status.get(), ws.counters.get() + p, ws.epoch, ws.err())
    switch (rb) {
      case 9: SPTC_ONESWEEP(9); break;
      case 10: SPTC_ONESWEEP(10); break;
      case 11: SPTC_ONESWEEP(11); break;
      default: SPTC_ONESWEEP(8); break;
    }
#undef SPTC_ONESWEEP
    if (kt) kt->end(tok, cls, 24.0 * (double)n, s);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  return (passes & 1) != 0;
}

}  // namespace sptc
