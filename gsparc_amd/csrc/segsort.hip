// Segmented LSD radix sort kernels (gfx950); see segsort.hpp.
#include "segsort.hpp"
#include "sort_core.hpp"

#include <algorithm>
#include <stdexcept>

namespace sptc {

namespace {

// ---- segmented reduce-then-scan radix pass (8-bit digits, 512 x 12 tiles) --

// ---- fused product generation --------------------------------------------
struct GenTile {
  uint64_t P;   // global product index of the tile's first product
  uint64_t ja;  // its owner
  uint32_t m;   // owners staged
  uint32_t piece;
};

// Stages the owners of the products at chunk offsets [q0, q0 + cnt) (one
// piece) in LDS: pref[k] = act_prefix[ja + k], k < m.
__device__ __forceinline__ GenTile gen_stage(const ProductGen& g, uint64_t q0, uint32_t cnt,
                                             uint64_t* pref, GenTile* s_t) {
  if (threadIdx.x == 0) {
    uint32_t lo = 0, hi = g.npieces;  // last piece with out0 <= q0
    while (lo + 1 < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (g.pieces[mid].out0 <= q0) lo = mid; else hi = mid;
    }
    const GenPiece pc = g.pieces[lo];
    const uint64_t P = pc.p0 + (q0 - pc.out0);
    uint64_t a = pc.A0, b = pc.A1;  // last j in [A0, A1) with act_prefix[j] <= P
    while (a < b) {
      const uint64_t mid = (a + b) >> 1;
      if (g.act_prefix[mid] <= P) a = mid + 1; else b = mid;
    }
    GenTile t;
    t.P = P;
    t.ja = a - 1;
    t.m = (uint32_t)(pc.A1 - t.ja < cnt ? pc.A1 - t.ja : cnt);
    t.piece = lo;
    *s_t = t;
  }
  __syncthreads();
  const GenTile t = *s_t;
  for (uint32_t k = threadIdx.x; k < t.m; k += blockDim.x) pref[k] = g.act_prefix[t.ja + k];
  __syncthreads();
  return t;
}

// Product at tile offset idx: key and value.  kk: owner search start (the
// thread's previous owner; offsets increase per thread).
__device__ __forceinline__ void gen_product(const ProductGen& g, const GenTile& t,
                                            const GenPiece& pc, const uint64_t* pref,
                                            uint64_t idx, uint32_t& kk, uint64_t& key,
                                            float& val) {
  const uint64_t P = t.P + idx;
  uint32_t lo = kk, hi = t.m;  // last k with pref[k] <= P
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pref[mid] <= P) lo = mid + 1; else hi = mid;
  }
  const uint32_t k = lo - 1;
  kk = k;
  const uint64_t j = t.ja + k;
  const uint64_t y = g.act_ypos[j] + (P - pref[k]);
  const uint32_t row = g.act_row[j];
  const uint64_t cmask =
      g.inst_kyf ? low_mask(g.inst_kyf[find_seg(g.inst_R, g.B, row)]) : g.fmask;
  key = shl64(pc.lrow0 + (row - pc.ra), g.K) | (g.ykeys[y] & cmask);
  val = g.act_xval[j] * g.yvals[y];
}

__device__ __forceinline__ void gen_add_vsum(double v, double* s_red, double* out) {
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  if (lane_id() == 0) s_red[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (uint32_t w = 0; w < blockDim.x / kWave; ++w) t += s_red[w];
    if (t != 0.0) atomicAdd(out, t);
  }
}

template <bool GEN>
__global__ __launch_bounds__(kSegNT) void k_seg_upsweep(const uint64_t* __restrict__ kin,
                                                        int shift, const SegTile* __restrict__ tiles,
                                                        uint32_t* __restrict__ cnt, ProductGen g) {
  constexpr int R = 1 << kSegRB;
  __shared__ uint32_t sh[R];
  __shared__ uint64_t pref[GEN ? kSegTile : 1];
  __shared__ GenTile s_t;
  if (GEN && g.bad && *g.bad) return;
  for (int i = threadIdx.x; i < R; i += kSegNT) sh[i] = 0;
  const SegTile t = tiles[blockIdx.x];
  GenTile gt;
  GenPiece pc;
  if (GEN) {
    gt = gen_stage(g, t.start, t.count, pref, &s_t);  // synchronizes
    pc = g.pieces[gt.piece];
  } else {
    __syncthreads();
  }
  const uint64_t end = t.start + t.count;
  uint32_t kk = 0;
#pragma unroll
  for (int i = 0; i < kSegIPT; ++i) {
    const uint64_t off = (uint64_t)i * kSegNT + threadIdx.x;
    const uint64_t idx = t.start + off;
    const bool valid = idx < end;
    uint64_t k = 0;
    if (valid) {
      if (GEN) {
        float v;
        gen_product(g, gt, pc, pref, off, kk, k, v);
      } else {
        k = kin[idx];
      }
    }
    hist_add_wave<kSegRB>(sh, k >> shift, valid, 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < R; d += kSegNT) cnt[(uint64_t)blockIdx.x * R + d] = sh[d];
}

template <bool GEN>
__global__ __launch_bounds__(kSegNT) void k_seg_downsweep(
    const uint64_t* __restrict__ kin, const float* __restrict__ vin, uint64_t* __restrict__ kout,
    float* __restrict__ vout, int shift, const SegTile* __restrict__ tiles,
    const uint32_t* __restrict__ offs, ProductGen g) {
  constexpr int R = 1 << kSegRB;
  __shared__ TileSmem<kSegRB, kSegIPT, kSegNT> sm;
  __shared__ GenTile s_t;
  __shared__ double s_red[kSegNT / kWave];
  if (GEN && g.bad && *g.bad) return;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = lane_id();
  const SegTile t = tiles[blockIdx.x];
  for (int i = tid; i < (kSegNT / kWave) * R; i += kSegNT) (&sm.cnt[0][0])[i] = 0;
  for (int d = tid; d < R; d += kSegNT) sm.gbase[d] = offs[t.obase + (uint64_t)d * t.ostride];
  __syncthreads();
  const uint64_t end = t.start + t.count;
  const uint64_t wbase = t.start + (uint64_t)(tid / kWave) * kWave * kSegIPT;
  uint64_t k[kSegIPT];
  float v[kSegIPT];
  uint32_t pos[kSegIPT];
  if (GEN) {
    // owners staged in the (not yet used) key staging buffer
    const GenTile gt = gen_stage(g, t.start, t.count, sm.buf, &s_t);
    const GenPiece pc = g.pieces[gt.piece];
    uint32_t kk = 0;
    double vs = 0.0;
#pragma unroll
    for (int i = 0; i < kSegIPT; ++i) {
      const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
      k[i] = 0;
      v[i] = 0.f;
      if (idx < end) {
        gen_product(g, gt, pc, sm.buf, idx - t.start, kk, k[i], v[i]);
        vs += (double)v[i];
      }
    }
    gen_add_vsum(vs, s_red, g.vsum);  // synchronizes: sm.buf is free again
    tile_rank<kSegRB, kSegIPT, kSegNT>(sm, end, shift, wbase, k, pos);
  } else {
    tile_load_rank<kSegRB, kSegIPT, kSegNT>(sm, kin, vin, end, shift, wbase, k, v, pos);
  }
  __syncthreads();
  for (int d = tid; d < R; d += kSegNT)
    sm.dstart[d] = (uint16_t)wave_prefixes<kSegRB, kSegIPT, kSegNT>(sm, d);
  __syncthreads();
  tile_stage_scatter<kSegRB, kSegIPT, kSegNT>(sm, end, shift, t.start, wbase, k, v, pos, kout,
                                              vout);
}

// Per segment (one workgroup, thread d = digit d): absolute destination of
// each tile's run of each digit.
__global__ __launch_bounds__(1 << kSegRB) void k_seg_scan(const SegDesc* __restrict__ segs,
                                                         const uint32_t* __restrict__ cnt,
                                                         uint32_t* __restrict__ off) {
  constexpr int R = 1 << kSegRB;
  __shared__ uint32_t scratch[R / kWave];
  const SegDesc g = segs[blockIdx.x];
  const uint32_t d = threadIdx.x;
  const uint64_t t0 = g.first_tile, t1 = t0 + g.ntiles;
  uint32_t tot = 0;
  for (uint64_t t = t0; t < t1; ++t) tot += cnt[t * R + d];
  uint32_t all;
  uint32_t base = (uint32_t)g.start + block_excl_scan<uint32_t>(tot, scratch, &all);
  uint32_t* o = off + R * t0 + (uint64_t)d * g.ntiles;  // digit-major in the segment
  for (uint64_t t = t0; t < t1; ++t) {
    o[t - t0] = base;
    base += cnt[t * R + d];
  }
}

// ---- row segments: small-row groups (LDS) and big rows (multi-tile) --------
constexpr uint64_t kHalfTile = kSegTile / 2;

__device__ __forceinline__ uint64_t rowseg_locate(const RowRange* ranges, uint32_t nranges,
                                                  uint64_t l, const uint64_t* row_prod,
                                                  uint64_t* first) {
  uint32_t lo = 0, hi = nranges;  // last range with lrow0 <= l
  while (lo + 1 < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ranges[mid].lrow0 <= l) lo = mid; else hi = mid;
  }
  const RowRange rr = ranges[lo];
  const uint64_t g = rr.ra + (l - rr.lrow0);
  *first = rr.out0 + (row_prod[g] - row_prod[rr.ra]);
  return row_prod[g + 1] - row_prod[g];
}

// Per row: small-row products (0 for big rows) and big-row tile counts.
__global__ __launch_bounds__(kBlock) void k_rowseg_sizes(const RowRange* ranges, uint32_t nr,
                                                         uint64_t lrows, const uint64_t* row_prod,
                                                         uint64_t* small_n, uint64_t* big_t,
                                                         uint64_t* big_f) {
  uint64_t l = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; l < lrows; l += stride) {
    uint64_t first;
    const uint64_t n = rowseg_locate(ranges, nr, l, row_prod, &first);
    const bool big = n > kHalfTile;
    small_n[l] = big ? 0 : n;
    big_t[l] = big ? ceil_div(n, kSegTile) : 0;
    big_f[l] = big ? 1 : 0;
  }
}

// Per row: 1 when a small row starts a group (after a big row, or where the
// running small-row total enters a new half-tile window).
__global__ __launch_bounds__(kBlock) void k_rowseg_heads(const RowRange* ranges, uint32_t nr,
                                                         uint64_t lrows, const uint64_t* small_n,
                                                         const uint64_t* big_t,
                                                         const uint64_t* ps, uint64_t* head) {
  uint64_t l = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; l < lrows; l += stride) {
    uint64_t h = 0;
    if (!big_t[l]) {
      uint32_t lo = 0, hi = nr;  // last range with lrow0 <= l
      while (lo + 1 < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ranges[mid].lrow0 <= l) lo = mid; else hi = mid;
      }
      h = l == 0 || l == ranges[lo].lrow0 || big_t[l - 1] != 0 ||
          ps[l] / kHalfTile != ps[l - 1] / kHalfTile;
    }
    head[l] = h;
  }
}

__global__ __launch_bounds__(kBlock) void k_rowseg_emit(
    const RowRange* ranges, uint32_t nr, uint64_t lrows, const uint64_t* row_prod,
    const uint64_t* big_t, const uint64_t* bt, const uint64_t* bs, const uint64_t* head,
    const uint64_t* hs, LocalSeg* local, SegTile* big, SegDesc* bigseg) {
  uint64_t l = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; l < lrows; l += stride) {
    uint64_t first;
    const uint64_t n = rowseg_locate(ranges, nr, l, row_prod, &first);
    const uint64_t nt = big_t[l];
    if (nt) {
      const uint64_t o = bt[l];
      SegDesc g;
      g.start = first;
      g.first_tile = (uint32_t)o;
      g.ntiles = (uint32_t)nt;
      bigseg[bs[l]] = g;
      for (uint64_t t = 0; t < nt; ++t) {
        SegTile d;
        d.start = first + t * kSegTile;
        d.count = (uint32_t)(n - t * kSegTile < kSegTile ? n - t * kSegTile : kSegTile);
        d.obase = (1ull << kSegRB) * o + t;
        d.ostride = (uint32_t)nt;
        big[o + t] = d;
      }
    } else {
      // the group holding row l: its head is the last head at or before l
      const uint64_t gidx = hs[l] + head[l] - 1;
      if (head[l]) local[gidx].start = first;
      if (n) atomicAdd(reinterpret_cast<unsigned int*>(&local[gidx].count), (unsigned int)n);
    }
  }
}

__global__ void k_zero_local(LocalSeg* local, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    local[i].count = 0;
    local[i].pad = 0;
  }
}

// One workgroup per small group: load, LSD passes over the bits that vary
// inside the group entirely in LDS (stable), store.
template <bool GEN>
__global__ __launch_bounds__(kSegNT) void k_local_sort(const uint64_t* __restrict__ kin,
                                                       const float* __restrict__ vin,
                                                       uint64_t* __restrict__ kout,
                                                       float* __restrict__ vout,
                                                       const LocalSeg* __restrict__ segs,
                                                       ProductGen g) {
  constexpr int R = 1 << kSegRB;
  __shared__ TileSmem<kSegRB, kSegIPT, kSegNT> sm;
  __shared__ float vbuf[kSegTile];
  __shared__ uint64_t s_or[kSegNT / kWave];
  __shared__ uint64_t s_k0;
  __shared__ GenTile s_t;
  __shared__ double s_red[kSegNT / kWave];
  if (GEN && g.bad && *g.bad) return;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t wave = tid / kWave;
  const LocalSeg sg = segs[blockIdx.x];
  const uint64_t n = sg.count;
  const uint64_t wbase = (uint64_t)wave * kWave * kSegIPT;  // tile-local
  uint64_t k[kSegIPT];
  float v[kSegIPT];
  uint32_t pos[kSegIPT];
  uint64_t k0;
  if (GEN) {
    if (n) {  // uniform per block
      const GenTile gt = gen_stage(g, sg.start, (uint32_t)n, sm.buf, &s_t);
      const GenPiece pc = g.pieces[gt.piece];
      uint32_t kk = 0;
      double vs = 0.0;
#pragma unroll
      for (int i = 0; i < kSegIPT; ++i) {
        const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
        k[i] = 0;
        v[i] = 0.f;
        if (idx < n) {
          gen_product(g, gt, pc, sm.buf, idx, kk, k[i], v[i]);
          vs += (double)v[i];
        }
      }
      if (tid == 0) s_k0 = k[0];
      gen_add_vsum(vs, s_red, g.vsum);  // synchronizes
      k0 = s_k0;
    } else {
      for (int i = 0; i < kSegIPT; ++i) {
        k[i] = 0;
        v[i] = 0.f;
      }
      k0 = 0;
    }
  } else {
    k0 = n ? kin[sg.start] : 0;
#pragma unroll
    for (int i = 0; i < kSegIPT; ++i) {
      const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
      const bool valid = idx < n;
      k[i] = valid ? kin[sg.start + idx] : 0;
      v[i] = valid ? vin[sg.start + idx] : 0.f;
    }
  }
  uint64_t vary = 0;
#pragma unroll
  for (int i = 0; i < kSegIPT; ++i) {
    const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    if (idx < n) vary |= k[i] ^ k0;
  }
  for (int off = kWave / 2; off > 0; off >>= 1) vary |= __shfl_xor(vary, off, kWave);
  if (lane == 0) s_or[wave] = vary;
  __syncthreads();
  vary = 0;
#pragma unroll
  for (int w = 0; w < kSegNT / kWave; ++w) vary |= s_or[w];
  const int passes = vary ? (64 - __builtin_clzll(vary) + kSegRB - 1) / kSegRB : 0;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * kSegRB;
    for (int i = tid; i < (kSegNT / kWave) * R; i += kSegNT) (&sm.cnt[0][0])[i] = 0;
    __syncthreads();
    tile_rank<kSegRB, kSegIPT, kSegNT>(sm, n, shift, wbase, k, pos);
    __syncthreads();
    for (int d = tid; d < R; d += kSegNT)
      sm.dstart[d] = (uint16_t)wave_prefixes<kSegRB, kSegIPT, kSegNT>(sm, d);
    __syncthreads();
    tile_digit_starts<kSegRB, kSegIPT, kSegNT>(sm);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSegIPT; ++i) {
      const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
      if (idx < n) {
        const uint32_t d = (uint32_t)(k[i] >> shift) & (R - 1);
        const uint32_t q = pos[i] + (uint32_t)sm.dstart[d] + (uint32_t)sm.cnt[wave][d];
        sm.buf[q] = k[i];
        vbuf[q] = v[i];
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSegIPT; ++i) {
      const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
      if (idx < n) {
        k[i] = sm.buf[idx];
        v[i] = vbuf[idx];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < kSegIPT; ++i) {
    const uint64_t idx = wbase + (uint64_t)i * kWave + lane;
    if (idx < n) {
      kout[sg.start + idx] = k[i];
      vout[sg.start + idx] = v[i];
    }
  }
}

}  // namespace

void SegSortPlan::build(const std::vector<uint64_t>& seg_off, hipStream_t s) {
  std::vector<SegTile> tv;
  std::vector<SegDesc> sv;
  for (size_t g = 0; g + 1 < seg_off.size(); ++g) {
    const uint64_t len = seg_off[g + 1] - seg_off[g];
    const uint64_t nt = ceil_div(len, kSegTile);
    if (!nt) continue;
    const uint64_t t0 = tv.size();
    sv.push_back(SegDesc{seg_off[g], (uint32_t)t0, (uint32_t)nt});
    for (uint64_t t = 0; t < nt; ++t) {
      SegTile d;
      d.start = seg_off[g] + t * kSegTile;
      d.count = (uint32_t)std::min<uint64_t>(kSegTile, len - t * kSegTile);
      d.obase = (1ull << kSegRB) * t0 + t;
      d.ostride = (uint32_t)nt;
      tv.push_back(d);
    }
  }
  n = seg_off.empty() ? 0 : seg_off.back();
  if (n >= (1ull << 32)) throw std::length_error("segmented sort of 2^32 or more keys");
  ntiles = tv.size();
  nsegs = sv.size();
  tiles.reset(ntiles ? ntiles : 1);
  segs.reset(nsegs ? nsegs : 1);
  if (ntiles) {
    SPTC_HIP(hipMemcpyAsync(tiles.get(), tv.data(), ntiles * sizeof(SegTile),
                            hipMemcpyHostToDevice, s));
    SPTC_HIP(hipMemcpyAsync(segs.get(), sv.data(), nsegs * sizeof(SegDesc),
                            hipMemcpyHostToDevice, s));
  }
  SPTC_HIP(hipStreamSynchronize(s));  // tv, sv are released on return
}

void segmented_sort_reserve(SortWorkspace& ws, const SegSortPlan& plan, hipStream_t s) {
  segmented_sort_reserve(ws, plan.ntiles, s);
}

void segmented_sort_reserve(SortWorkspace& ws, uint64_t ntiles, hipStream_t s) {
  const uint64_t m = ntiles << kSegRB;
  ws.ensure(1, kSegRB, s);
  if (ws.rts_cnt.size() < m) {
    ws.rts_cnt.reset(m);
    ws.rts_off.reset(m);
  }
}

int segmented_sort_passes(int nbits) { return radix_passes(nbits, kSegRB); }

namespace {
// Passes of a segmented sort; with gen the first pass generates the keys.
bool seg_passes(SortWorkspace& ws, const SegTile* tiles, uint64_t ntiles, const SegDesc* segs,
                uint64_t nsegs, uint64_t n, uint64_t* k0, float* v0, uint64_t* k1, float* v1,
                int nbits, hipStream_t s, KTimer* kt, int cls, const ProductGen* gen) {
  const int passes = radix_passes(nbits, kSegRB);
  if (!n || !passes || !ntiles) return false;
  const uint64_t m = ntiles << kSegRB;
  if (ws.rts_cnt.size() < m) throw std::logic_error("segmented sort workspace not reserved");
  if (ntiles >= (1ull << 31) || nsegs >= (1ull << 31))
    throw std::length_error("segmented sort tile count");
  uint64_t* ka = k0;
  float* va = v0;
  uint64_t* kb = k1;
  float* vb = v1;
  const ProductGen none{};
  for (int p = 0; p < passes; ++p) {
    const int shift = p * kSegRB;
    const bool g = gen && p == 0;
    int tok = kt ? kt->begin(s) : -1;
    if (g)
      hipLaunchKernelGGL((k_seg_upsweep<true>), dim3((uint32_t)ntiles), dim3(kSegNT), 0, s, ka,
                         shift, tiles, ws.rts_cnt.get(), *gen);
    else
      hipLaunchKernelGGL((k_seg_upsweep<false>), dim3((uint32_t)ntiles), dim3(kSegNT), 0, s, ka,
                         shift, tiles, ws.rts_cnt.get(), none);
    hipLaunchKernelGGL(k_seg_scan, dim3((uint32_t)nsegs), dim3(1 << kSegRB), 0, s, segs,
                       ws.rts_cnt.get(), ws.rts_off.get());
    if (kt) kt->end(tok, SPTC_K_SORT_COUNT, 8.0 * (double)n + 16.0 * (double)m, s);
    tok = kt ? kt->begin(s) : -1;
    if (g)
      hipLaunchKernelGGL((k_seg_downsweep<true>), dim3((uint32_t)ntiles), dim3(kSegNT), 0, s, ka,
                         va, kb, vb, shift, tiles, ws.rts_off.get(), *gen);
    else
      hipLaunchKernelGGL((k_seg_downsweep<false>), dim3((uint32_t)ntiles), dim3(kSegNT), 0, s, ka,
                         va, kb, vb, shift, tiles, ws.rts_off.get(), none);
    if (kt) kt->end(tok, cls, 24.0 * (double)n, s);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  return (passes & 1) != 0;
}
}  // namespace

bool segmented_sort(SortWorkspace& ws, const SegTile* tiles, uint64_t ntiles,
                    const SegDesc* segs, uint64_t nsegs, uint64_t n, uint64_t* k0, float* v0,
                    uint64_t* k1, float* v1, int nbits, hipStream_t s, KTimer* kt, int cls) {
  return seg_passes(ws, tiles, ntiles, segs, nsegs, n, k0, v0, k1, v1, nbits, s, kt, cls,
                    nullptr);
}

bool segmented_sort(SortWorkspace& ws, const SegSortPlan& plan, uint64_t* k0, float* v0,
                    uint64_t* k1, float* v1, int nbits, hipStream_t s, KTimer* kt, int cls) {
  return segmented_sort(ws, plan.tiles.get(), plan.ntiles, plan.segs.get(), plan.nsegs, plan.n,
                        k0, v0, k1, v1, nbits, s, kt, cls);
}

void build_row_segments(const RowRange* ranges, uint32_t nranges, uint64_t lrows,
                        const uint64_t* row_prod, RowSegPlan* plan, SortWorkspace& ws, U2* tmp,
                        hipStream_t s) {
  plan->nlocal = plan->nbig_tiles = plan->nbig_rows = plan->nsmall = 0;
  if (!lrows) return;
  const uint64_t m = lrows + 1;
  if (plan->scratch.size() < 8 * m) plan->scratch.reset(8 * m);
  uint64_t* small_n = plan->scratch.get();
  uint64_t* big_t = small_n + m;
  uint64_t* ps = big_t + m;
  uint64_t* bt = ps + m;
  uint64_t* head = bt + m;
  uint64_t* hs = head + m;
  uint64_t* big_f = hs + m;
  uint64_t* bs = big_f + m;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(lrows, kBlock), 4096);
  hipLaunchKernelGGL(k_rowseg_sizes, dim3(grid), dim3(kBlock), 0, s, ranges, nranges, lrows,
                     row_prod, small_n, big_t, big_f);
  run_exclusive_scan(small_n, ps, lrows, tmp, s);
  run_exclusive_scan(big_t, bt, lrows, tmp, s);
  run_exclusive_scan(big_f, bs, lrows, tmp, s);
  hipLaunchKernelGGL(k_rowseg_heads, dim3(grid), dim3(kBlock), 0, s, ranges, nranges, lrows,
                     small_n, big_t, ps, head);
  run_exclusive_scan(head, hs, lrows, tmp, s);
  uint64_t tot[4];
  SPTC_HIP(hipMemcpyAsync(&tot[0], hs + lrows, 8, hipMemcpyDeviceToHost, s));
  SPTC_HIP(hipMemcpyAsync(&tot[1], bt + lrows, 8, hipMemcpyDeviceToHost, s));
  SPTC_HIP(hipMemcpyAsync(&tot[2], ps + lrows, 8, hipMemcpyDeviceToHost, s));
  SPTC_HIP(hipMemcpyAsync(&tot[3], bs + lrows, 8, hipMemcpyDeviceToHost, s));
  SPTC_HIP(hipStreamSynchronize(s));
  plan->nlocal = tot[0];
  plan->nbig_tiles = tot[1];
  plan->nsmall = tot[2];
  plan->nbig_rows = tot[3];
  if (plan->nlocal >= (1ull << 31) || plan->nbig_tiles >= (1ull << 31))
    throw std::length_error("row segments");
  if (plan->local.size() < plan->nlocal) plan->local.reset(plan->nlocal);
  if (plan->big.size() < plan->nbig_tiles) plan->big.reset(plan->nbig_tiles);
  if (plan->bigseg.size() < plan->nbig_rows) plan->bigseg.reset(plan->nbig_rows);
  if (plan->nlocal)
    hipLaunchKernelGGL(k_zero_local, dim3((uint32_t)ceil_div(plan->nlocal, 256)), dim3(256), 0,
                       s, plan->local.get(), plan->nlocal);
  hipLaunchKernelGGL(k_rowseg_emit, dim3(grid), dim3(kBlock), 0, s, ranges, nranges, lrows,
                     row_prod, big_t, bt, bs, head, hs, plan->local.get(), plan->big.get(),
                     plan->bigseg.get());
  if (plan->nbig_tiles) segmented_sort_reserve(ws, plan->nbig_tiles, s);
}

bool sort_row_segments(SortWorkspace& ws, const RowSegPlan& plan, uint64_t n, uint64_t* k0,
                       float* v0, uint64_t* k1, float* v1, int col_bits, hipStream_t s,
                       KTimer* kt, int cls, const ProductGen* gen) {
  bool alt = false;
  if (plan.nbig_tiles)
    alt = seg_passes(ws, plan.big.get(), plan.nbig_tiles, plan.bigseg.get(), plan.nbig_rows,
                     n - plan.nsmall, k0, v0, k1, v1, col_bits, s, kt, cls, gen);
  if (plan.nlocal) {
    int tok = kt ? kt->begin(s) : -1;
    if (gen)
      hipLaunchKernelGGL((k_local_sort<true>), dim3((uint32_t)plan.nlocal), dim3(kSegNT), 0, s,
                         k0, v0, alt ? k1 : k0, alt ? v1 : v0, plan.local.get(), *gen);
    else
      hipLaunchKernelGGL((k_local_sort<false>), dim3((uint32_t)plan.nlocal), dim3(kSegNT), 0, s,
                         k0, v0, alt ? k1 : k0, alt ? v1 : v0, plan.local.get(), ProductGen{});
    if (kt) kt->end(tok, SPTC_K_SORT_LOCAL, 24.0 * (double)plan.nsmall, s);
  }
  return alt;
}

}  // namespace sptc
