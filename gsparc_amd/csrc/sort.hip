// Stable LSD radix sort of (uint64 key, float value) pairs for gfx950.
//
// Replaces cub::DeviceRadixSort::SortPairs in the reference (sort.cuh:90-108,
// contraction.cuh:506-530, tensor_manager.tpp:779-786).  Design:
//   * Only the significant key bits are sorted: ceil(nbits / RB) passes.
//   * A tile of NT*IPT keys is ranked with wave64 ballot matching (one ballot
//     per digit bit; rounds in order, lanes in order => stable), staged in
//     LDS in digit order and written out in coalesced digit runs -- keys
//     first, then values through the same LDS buffer.  Longer digit runs per
//     tile (bigger tiles) are what make the scatter efficient.
//   * Two ways to find each tile's global digit offsets:
//       reduce-then-scan (default): an upsweep writes per-tile digit counts
//       into a digit-major matrix whose exclusive scan, in memory order, is
//       every tile's start for every digit; no inter-tile waits.
//       onesweep: a single kernel per pass with decoupled look-back over
//       predecessor tiles, using up-front digit histograms that producers
//       (linearize, ESC expand) compute fused.  Look-back status words carry
//       a pass epoch, so the status array is never cleared between passes:
//       [63:62] flag (1 aggregate, 2 inclusive prefix), [61:48] epoch,
//       [47:0] count.
//   * Destinations are 32-bit: every sort here has fewer than 2^32 keys.
#include "common.hpp"
#include "sort.hpp"
#include "scan.hpp"
#include "sort_core.hpp"

#include <cstdlib>
#include <stdexcept>
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

// TAG only separates kernel symbols (0: operand sorts, 1: product sorts) so
// profiles attribute time to the product sort on its own.
template <int RB, int IPT, int NT, int TAG>
__global__ __launch_bounds__(NT) void k_onesweep(
    const uint64_t* __restrict__ kin, const float* __restrict__ vin,
    uint64_t* __restrict__ kout, float* __restrict__ vout, uint64_t n, int shift,
    const uint64_t* __restrict__ digit_base, uint64_t* status,
    uint32_t* tile_counter, uint32_t epoch, uint32_t* err) {
  constexpr int R = 1 << RB;
  constexpr int TILE = NT * IPT;
  static_assert(TILE < 65536, "16-bit tile offsets");
  __shared__ TileSmem<RB, IPT, NT> sm;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) sm.tile = atomicAdd(tile_counter, 1u);
  for (int i = tid; i < (NT / kWave) * R; i += NT) (&sm.cnt[0][0])[i] = 0;
  __syncthreads();
  const uint64_t tile = sm.tile;
  const uint64_t tile_begin = tile * TILE;
  const uint32_t cnt = (uint32_t)(n - tile_begin < (uint64_t)TILE ? n - tile_begin : TILE);
  uint64_t k[IPT];
  float v[IPT];
  uint32_t pos[IPT];
  tile_load_rank<false, RB, IPT, NT>(sm, kin + tile_begin, vin + tile_begin, cnt, shift, k, v,
                                     pos);
  __syncthreads();
  for (int d = tid; d < R; d += NT) {
    const uint64_t total = wave_prefixes<RB, IPT, NT>(sm, d);
    uint64_t excl = 0;
    uint64_t* my = status + tile * R + d;
    if (tile == 0) {
      st_agent(my, pack_status(kFlagPre, epoch, total));
    } else {
      st_agent(my, pack_status(kFlagAgg, epoch, total));
      int64_t j = (int64_t)tile - 1;
      uint32_t spins = 0;
      while (j >= 0) {
        const uint64_t st = ld_agent(status + (uint64_t)j * R + d);
        const uint32_t se = (uint32_t)((st >> 48) & 0x3FFF);
        const uint64_t fl = st & (3ull << 62);
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
    sm.gbase[d] = (uint32_t)(digit_base[d] + excl);
    sm.dstart[d] = (uint16_t)total;
  }
  __syncthreads();
  tile_stage_scatter<false, RB, IPT, NT>(sm, cnt, shift, k, v, pos, kout, vout);
}

template <int RB, int IPT, int NT>
__global__ __launch_bounds__(NT) void k_rts_upsweep(const uint64_t* __restrict__ kin,
                                                    uint64_t n, int shift,
                                                    uint32_t* __restrict__ cnt,
                                                    uint64_t ntiles) {
  constexpr int R = 1 << RB;
  constexpr int TILE = NT * IPT;
  __shared__ uint32_t sh[R];
  for (int i = threadIdx.x; i < R; i += NT) sh[i] = 0;
  __syncthreads();
  const uint64_t tile = blockIdx.x;
  const uint64_t base = tile * TILE;
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const uint64_t idx = base + (uint64_t)i * NT + threadIdx.x;
    const bool valid = idx < n;
    const uint64_t k = valid ? kin[idx] : 0;
    hist_add_wave<RB>(sh, k >> shift, valid, 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < R; d += NT) cnt[(uint64_t)d * ntiles + tile] = sh[d];
}

template <int RB, int IPT, int NT, int TAG>
__global__ __launch_bounds__(NT) void k_rts_downsweep(
    const uint64_t* __restrict__ kin, const float* __restrict__ vin,
    uint64_t* __restrict__ kout, float* __restrict__ vout, uint64_t n, int shift,
    const uint32_t* __restrict__ offs, uint64_t ntiles, uint32_t remap) {
  constexpr int R = 1 << RB;
  constexpr int TILE = NT * IPT;
  static_assert(TILE < 65536, "16-bit tile offsets");
  __shared__ TileSmem<RB, IPT, NT> sm;
  const uint32_t tid = threadIdx.x;
  const uint64_t tile = remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
  for (int i = tid; i < (NT / kWave) * R; i += NT) (&sm.cnt[0][0])[i] = 0;
  for (int d = tid; d < R; d += NT) sm.gbase[d] = offs[(uint64_t)d * ntiles + tile];
  __syncthreads();
  const uint64_t tile_begin = tile * TILE;
  const uint32_t cnt = (uint32_t)(n - tile_begin < (uint64_t)TILE ? n - tile_begin : TILE);
  uint64_t k[IPT];
  float v[IPT];
  uint32_t pos[IPT];
  tile_load_rank<false, RB, IPT, NT>(sm, kin + tile_begin, vin + tile_begin, cnt, shift, k, v,
                                     pos);
  __syncthreads();
  for (int d = tid; d < R; d += NT) sm.dstart[d] = (uint16_t)wave_prefixes<RB, IPT, NT>(sm, d);
  __syncthreads();
  tile_stage_scatter<false, RB, IPT, NT>(sm, cnt, shift, k, v, pos, kout, vout);
}

// Exclusive scan of the uint32 count matrix into uint32 offsets.
struct CountScanOp {
  const uint32_t* in;
  uint32_t* out;
  __device__ U2 value(uint64_t i) const { return mk2(in[i], 0); }
  __device__ void emit(uint64_t i, U2 excl, U2) const { out[i] = (uint32_t)excl.a; }
};

// ---- tile configuration ---------------------------------------------------
// SPTC_SORT selects the pass algorithm ("rts" default, "onesweep");
// SPTC_SORT_TILE the 8-bit-digit tile as threads x items ("512x12" default,
// "512x16", "256x16", "256x12").  9-bit digits use 512x12, wider ones 256x12.
struct SortCfg {
  bool onesweep = false;
  int nt = 512;
  int ipt = 12;
};
const SortCfg& sort_cfg() {
  static const SortCfg c = [] {
    SortCfg r;
    if (const char* e = getenv("SPTC_SORT")) r.onesweep = std::string(e) == "onesweep";
    if (const char* e = getenv("SPTC_SORT_TILE")) {
      const std::string t(e);
      if (t == "256x16") { r.nt = 256; r.ipt = 16; }
      if (t == "256x12") { r.nt = 256; r.ipt = 12; }
      if (t == "512x16") { r.nt = 512; r.ipt = 16; }
    }
    return r;
  }();
  return c;
}

// Calls f.template run<RB, IPT, NT>() for the configured tile of digit width rb.
template <class F>
void with_tile(int rb, F&& f) {
  const SortCfg& c = sort_cfg();
  switch (rb) {
    case 9: f.template run<9, 12, 512>(); return;
    case 10: f.template run<10, 12, 256>(); return;
    case 11: f.template run<11, 12, 256>(); return;
    default:
      if (c.nt == 512 && c.ipt == 16) f.template run<8, 16, 512>();
      else if (c.nt == 512) f.template run<8, 12, 512>();
      else if (c.ipt == 16) f.template run<8, 16, 256>();
      else f.template run<8, 12, 256>();
  }
}

struct PassArgs {
  int tag;
  SortWorkspace* ws;
  uint64_t* ka;
  float* va;
  uint64_t* kb;
  float* vb;
  uint64_t n;
  uint64_t tiles;
  int pass;
  hipStream_t s;
  KTimer* kt;
  int cls;
};

struct OnesweepPass {
  PassArgs a;
  template <int RB, int IPT, int NT>
  void run() {
    const uint64_t R = 1ull << RB;
    const int tok = a.kt ? a.kt->begin(a.s) : -1;
    if (a.tag)
      hipLaunchKernelGGL((k_onesweep<RB, IPT, NT, 1>), dim3((uint32_t)a.tiles), dim3(NT), 0, a.s,
                         a.ka, a.va, a.kb, a.vb, a.n, a.pass * RB, a.ws->base.get() + a.pass * R,
                         a.ws->status.get(), a.ws->counters.get() + a.pass, a.ws->epoch,
                         a.ws->err());
    else
      hipLaunchKernelGGL((k_onesweep<RB, IPT, NT, 0>), dim3((uint32_t)a.tiles), dim3(NT), 0, a.s,
                         a.ka, a.va, a.kb, a.vb, a.n, a.pass * RB, a.ws->base.get() + a.pass * R,
                         a.ws->status.get(), a.ws->counters.get() + a.pass, a.ws->epoch,
                         a.ws->err());
    if (a.kt) a.kt->end(tok, a.cls, 24.0 * (double)a.n, a.s);
  }
};

struct RtsPass {
  PassArgs a;
  template <int RB, int IPT, int NT>
  void run() {
    const uint64_t m = a.tiles << RB;
    const int shift = a.pass * RB;
    int tok = a.kt ? a.kt->begin(a.s) : -1;
    hipLaunchKernelGGL((k_rts_upsweep<RB, IPT, NT>), dim3((uint32_t)a.tiles), dim3(NT), 0, a.s,
                       a.ka, a.n, shift, a.ws->rts_cnt.get(), a.tiles);
    run_scan(CountScanOp{a.ws->rts_cnt.get(), a.ws->rts_off.get()}, m, a.ws->rts_tmp.get(), a.s);
    if (a.kt) a.kt->end(tok, SPTC_K_SORT_COUNT, 8.0 * (double)a.n + 12.0 * (double)m, a.s);
    tok = a.kt ? a.kt->begin(a.s) : -1;
    if (a.tag)
      hipLaunchKernelGGL((k_rts_downsweep<RB, IPT, NT, 1>), dim3((uint32_t)a.tiles), dim3(NT), 0,
                         a.s, a.ka, a.va, a.kb, a.vb, a.n, shift, a.ws->rts_off.get(), a.tiles,
                         xcd_remap_enabled());
    else
      hipLaunchKernelGGL((k_rts_downsweep<RB, IPT, NT, 0>), dim3((uint32_t)a.tiles), dim3(NT), 0,
                         a.s, a.ka, a.va, a.kb, a.vb, a.n, shift, a.ws->rts_off.get(), a.tiles,
                         xcd_remap_enabled());
    if (a.kt) a.kt->end(tok, a.cls, 24.0 * (double)a.n, a.s);
  }
};

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
  if (rb == 8) return (uint64_t)sort_cfg().nt * sort_cfg().ipt;
  if (rb == 9) return 512ull * 12;
  return (uint64_t)kBlock * 12;
}

bool radix_bits_supported(int rb) { return rb >= 8 && rb <= 11; }

// 8-bit digits, or 9-bit ones when that saves a pass (a 9-bit pass on the
// 512x12 tile measured ~10% slower than an 8-bit one, a pass saved is 20-25%
// of the sort).  SPTC_SORT_RB9=0 keeps 8-bit digits throughout.
int choose_radix_bits(int nbits) {
  static const int wide = [] {
    const char* e = getenv("SPTC_SORT_RB9");
    return e && e[0] == '0' ? 0 : 1;
  }();
  if (wide && radix_passes(nbits, 9) < radix_passes(nbits, 8)) return 9;
  return 8;
}

bool sort_needs_histogram() { return sort_cfg().onesweep; }

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
  if (sort_cfg().onesweep && status.size() < tiles * R) {
    status.reset(tiles * R);
    fill_u64(status.get(), 0, tiles * R, s);
    epoch = 0;
  }
  if (!sort_cfg().onesweep && rts_cnt.size() < tiles * R) {
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

// Sorts keys/vals (n <= ws.cap, n < 2^32) on bits [0, nbits).  With the
// onesweep algorithm ws.hist must hold the digit histograms of all passes.
// Result ends in (keys, vals) when the pass count is even, else in
// (ws.alt_keys, ws.alt_vals); the returned flag says which (true = alt).
bool sort_after_histogram(SortWorkspace& ws, uint64_t* keys, float* vals,
                          uint64_t n, int nbits, hipStream_t s, KTimer* kt, int cls) {
  const int rb = ws.rb;
  const int passes = radix_passes(nbits, rb);
  if (!n || !passes) return false;
  if (n >= (1ull << 32)) throw std::length_error("radix sort of 2^32 or more keys");
  const bool onesweep = sort_cfg().onesweep;
  if (onesweep) {
    switch (rb) {
      case 9: hipLaunchKernelGGL(k_digit_scan<9>, dim3(passes), dim3(kBlock), 0, s,
                                 ws.hist.get(), ws.base.get()); break;
      case 10: hipLaunchKernelGGL(k_digit_scan<10>, dim3(passes), dim3(kBlock), 0, s,
                                  ws.hist.get(), ws.base.get()); break;
      case 11: hipLaunchKernelGGL(k_digit_scan<11>, dim3(passes), dim3(kBlock), 0, s,
                                  ws.hist.get(), ws.base.get()); break;
      default: hipLaunchKernelGGL(k_digit_scan<8>, dim3(passes), dim3(kBlock), 0, s,
                                  ws.hist.get(), ws.base.get()); break;
    }
  }
  const uint64_t R = 1ull << rb;
  const uint64_t tiles = ceil_div(n, sort_tile_items(rb));
  const int tag = (cls == SPTC_K_SORT_PROD || cls == SPTC_K_SORT_PROD_WIDE) ? 1 : 0;
  const int kcls = (cls == SPTC_K_SORT_PROD && rb > 8) ? SPTC_K_SORT_PROD_WIDE : cls;
  PassArgs a{tag, &ws, keys, vals, ws.alt_keys.get(), ws.alt_vals.get(), n, tiles, 0, s, kt,
             kcls};
  for (int p = 0; p < passes; ++p) {
    a.pass = p;
    if (onesweep) {
      ws.epoch = (ws.epoch + 1) & 0x3FFF;
      if (ws.epoch == 0) {
        fill_u64(ws.status.get(), 0, ws.max_tiles * R, s);
        ws.epoch = 1;
      }
      with_tile(rb, OnesweepPass{a});
    } else {
      with_tile(rb, RtsPass{a});
    }
    std::swap(a.ka, a.kb);
    std::swap(a.va, a.vb);
  }
  return (passes & 1) != 0;
}

}  // namespace sptc
