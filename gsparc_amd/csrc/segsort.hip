// Segmented LSD radix sort kernels (gfx950); see segsort.hpp.
#include "segsort.hpp"
#include "sort_core.hpp"

#include <algorithm>
#include <stdexcept>

namespace sptc {

namespace {

// ---- segmented reduce-then-scan radix pass (8-bit digits, 512 x 12 tiles) --

__global__ __launch_bounds__(kSegNT) void k_seg_upsweep(const uint64_t* __restrict__ kin,
                                                        int shift, const SegTile* __restrict__ tiles,
                                                        uint32_t* __restrict__ cnt) {
  constexpr int R = 1 << kSegRB;
  __shared__ uint32_t sh[R];
  for (int i = threadIdx.x; i < R; i += kSegNT) sh[i] = 0;
  __syncthreads();
  const SegTile t = tiles[blockIdx.x];
  const uint64_t end = t.start + t.count;
#pragma unroll
  for (int i = 0; i < kSegIPT; ++i) {
    const uint64_t idx = t.start + (uint64_t)i * kSegNT + threadIdx.x;
    const bool valid = idx < end;
    const uint64_t k = valid ? kin[idx] : 0;
    hist_add_wave<kSegRB>(sh, k >> shift, valid, 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < R; d += kSegNT) cnt[t.cbase + (uint64_t)d * t.cstride] = sh[d];
}

__global__ __launch_bounds__(kSegNT) void k_seg_downsweep(
    const uint64_t* __restrict__ kin, const float* __restrict__ vin, uint64_t* __restrict__ kout,
    float* __restrict__ vout, int shift, const SegTile* __restrict__ tiles,
    const uint32_t* __restrict__ offs) {
  constexpr int R = 1 << kSegRB;
  __shared__ TileSmem<kSegRB, kSegIPT, kSegNT> sm;
  const uint32_t tid = threadIdx.x;
  const SegTile t = tiles[blockIdx.x];
  for (int i = tid; i < (kSegNT / kWave) * R; i += kSegNT) (&sm.cnt[0][0])[i] = 0;
  for (int d = tid; d < R; d += kSegNT) sm.gbase[d] = offs[t.cbase + (uint64_t)d * t.cstride];
  __syncthreads();
  const uint64_t end = t.start + t.count;
  const uint64_t wbase = t.start + (uint64_t)(tid / kWave) * kWave * kSegIPT;
  uint64_t k[kSegIPT];
  float v[kSegIPT];
  uint32_t pos[kSegIPT];
  tile_load_rank<kSegRB, kSegIPT, kSegNT>(sm, kin, vin, end, shift, wbase, k, v, pos);
  __syncthreads();
  for (int d = tid; d < R; d += kSegNT)
    sm.dstart[d] = (uint16_t)wave_prefixes<kSegRB, kSegIPT, kSegNT>(sm, d);
  __syncthreads();
  tile_stage_scatter<kSegRB, kSegIPT, kSegNT>(sm, end, shift, t.start, wbase, k, v, pos, kout,
                                              vout);
}

struct SegCountScanOp {
  const uint32_t* in;
  uint32_t* out;
  __device__ U2 value(uint64_t i) const { return mk2(in[i], 0); }
  __device__ void emit(uint64_t i, U2 excl, U2) const { out[i] = (uint32_t)excl.a; }
};

__global__ __launch_bounds__(kBlock) void k_row_tile_count(const RowRange* ranges,
                                                           uint32_t nranges, uint64_t lrows,
                                                           const uint64_t* row_prod,
                                                           uint64_t* cnt) {
  uint64_t l = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; l < lrows; l += stride) {
    uint32_t lo = 0, hi = nranges;  // last range with lrow0 <= l
    while (lo + 1 < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (ranges[mid].lrow0 <= l) lo = mid; else hi = mid;
    }
    const uint64_t g = ranges[lo].ra + (l - ranges[lo].lrow0);
    cnt[l] = ceil_div(row_prod[g + 1] - row_prod[g], kSegTile);
  }
}

__global__ __launch_bounds__(kBlock) void k_row_tiles(const RowRange* ranges, uint32_t nranges,
                                                      uint64_t lrows, const uint64_t* row_prod,
                                                      const uint64_t* cnt, const uint64_t* off,
                                                      SegTile* tiles) {
  constexpr uint64_t R = 1ull << kSegRB;
  uint64_t l = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (; l < lrows; l += stride) {
    uint32_t lo = 0, hi = nranges;
    while (lo + 1 < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (ranges[mid].lrow0 <= l) lo = mid; else hi = mid;
    }
    const RowRange rr = ranges[lo];
    const uint64_t g = rr.ra + (l - rr.lrow0);
    const uint64_t first = rr.out0 + (row_prod[g] - row_prod[rr.ra]);
    const uint64_t len = row_prod[g + 1] - row_prod[g];
    const uint64_t nt = cnt[l], o = off[l];
    for (uint64_t t = 0; t < nt; ++t) {
      SegTile d;
      d.start = first + t * kSegTile;
      d.count = (uint32_t)(len - t * kSegTile < kSegTile ? len - t * kSegTile : kSegTile);
      d.cbase = R * o + t;
      d.cstride = nt;
      d.pad = 0;
      tiles[o + t] = d;
    }
  }
}

}  // namespace

void SegSortPlan::build(const std::vector<uint64_t>& seg_off, hipStream_t s) {
  constexpr uint64_t R = 1ull << kSegRB;
  std::vector<SegTile> tv;
  uint64_t tiles_before = 0;
  for (size_t g = 0; g + 1 < seg_off.size(); ++g) {
    const uint64_t len = seg_off[g + 1] - seg_off[g];
    const uint64_t nt = ceil_div(len, kSegTile);
    for (uint64_t t = 0; t < nt; ++t) {
      SegTile d;
      d.start = seg_off[g] + t * kSegTile;
      d.count = (uint32_t)std::min<uint64_t>(kSegTile, len - t * kSegTile);
      d.cbase = R * tiles_before + t;
      d.cstride = nt;
      d.pad = 0;
      tv.push_back(d);
    }
    tiles_before += nt;
  }
  n = seg_off.empty() ? 0 : seg_off.back();
  if (n >= (1ull << 32)) throw std::length_error("segmented sort of 2^32 or more keys");
  ntiles = tv.size();
  tiles.reset(ntiles ? ntiles : 1);
  if (ntiles)
    SPTC_HIP(hipMemcpyAsync(tiles.get(), tv.data(), ntiles * sizeof(SegTile),
                            hipMemcpyHostToDevice, s));
  SPTC_HIP(hipStreamSynchronize(s));  // tv is released on return
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
  if (ws.rts_tmp.size() < scan_tmp_count(m) + 1) ws.rts_tmp.reset(scan_tmp_count(m) + 1);
}

int segmented_sort_passes(int nbits) { return radix_passes(nbits, kSegRB); }

bool segmented_sort(SortWorkspace& ws, const SegTile* tiles, uint64_t ntiles, uint64_t n,
                    uint64_t* k0, float* v0, uint64_t* k1, float* v1, int nbits, hipStream_t s,
                    KTimer* kt, int cls) {
  const int passes = radix_passes(nbits, kSegRB);
  if (!n || !passes || !ntiles) return false;
  const uint64_t m = ntiles << kSegRB;
  if (ws.rts_cnt.size() < m || ws.rts_tmp.size() < scan_tmp_count(m) + 1)
    throw std::logic_error("segmented sort workspace not reserved");
  if (ntiles >= (1ull << 31)) throw std::length_error("segmented sort tile count");
  uint64_t* ka = k0;
  float* va = v0;
  uint64_t* kb = k1;
  float* vb = v1;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * kSegRB;
    int tok = kt ? kt->begin(s) : -1;
    hipLaunchKernelGGL(k_seg_upsweep, dim3((uint32_t)ntiles), dim3(kSegNT), 0, s, ka, shift,
                       tiles, ws.rts_cnt.get());
    run_scan(SegCountScanOp{ws.rts_cnt.get(), ws.rts_off.get()}, m, ws.rts_tmp.get(), s);
    if (kt) kt->end(tok, SPTC_K_SORT_COUNT, 8.0 * (double)n + 12.0 * (double)m, s);
    tok = kt ? kt->begin(s) : -1;
    hipLaunchKernelGGL(k_seg_downsweep, dim3((uint32_t)ntiles), dim3(kSegNT), 0, s, ka, va, kb,
                       vb, shift, tiles, ws.rts_off.get());
    if (kt) kt->end(tok, cls, 24.0 * (double)n, s);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  return (passes & 1) != 0;
}

bool segmented_sort(SortWorkspace& ws, const SegSortPlan& plan, uint64_t* k0, float* v0,
                    uint64_t* k1, float* v1, int nbits, hipStream_t s, KTimer* kt, int cls) {
  return segmented_sort(ws, plan.tiles.get(), plan.ntiles, plan.n, k0, v0, k1, v1, nbits, s, kt,
                        cls);
}

void launch_row_tile_count(const RowRange* ranges, uint32_t nranges, uint64_t lrows,
                           const uint64_t* row_prod, uint64_t* cnt, hipStream_t s) {
  if (!lrows) return;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(lrows, kBlock), 4096);
  hipLaunchKernelGGL(k_row_tile_count, dim3(grid), dim3(kBlock), 0, s, ranges, nranges, lrows,
                     row_prod, cnt);
}

void launch_row_tiles(const RowRange* ranges, uint32_t nranges, uint64_t lrows,
                      const uint64_t* row_prod, const uint64_t* cnt, const uint64_t* off,
                      SegTile* tiles, hipStream_t s) {
  if (!lrows) return;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(lrows, kBlock), 4096);
  hipLaunchKernelGGL(k_row_tiles, dim3(grid), dim3(kBlock), 0, s, ranges, nranges, lrows,
                     row_prod, cnt, off, tiles);
}

}  // namespace sptc
