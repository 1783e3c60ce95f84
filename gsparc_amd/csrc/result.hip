// Result-side kernels: an output verification pass (FP64 checksum of the
// emitted values and a strict-order check of the (row, column) entries) and
// the decode of output entries to per-mode indices on the device.
#include "result.hpp"

namespace sptc {

namespace {

constexpr int kVerifyItems = 8;  // entries per thread per round

// One launch per chunk.  Entry i must satisfy r0 <= row[i] < r1 and
// (row[i-1], col[i-1]) < (row[i], col[i]); violations are counted.  Each
// workgroup writes its FP64 value sum and violation count to part[block];
// the host adds the partials in order (deterministic).
__global__ void __launch_bounds__(kBlock) k_verify(const uint32_t* __restrict__ row,
                                                   const uint64_t* __restrict__ col,
                                                   const float* __restrict__ val, uint64_t n,
                                                   uint64_t r0, uint64_t r1,
                                                   VerifyPart* __restrict__ part) {
  __shared__ double ssum[kBlock / kWave];
  __shared__ uint64_t sbad[kBlock / kWave];
  double sum = 0.0;
  uint64_t bad = 0;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const uint64_t r = row[i];
    const uint64_t c = col[i];
    sum += (double)val[i];
    bool ok = r >= r0 && r < r1;
    if (i > 0) {
      const uint64_t pr = row[i - 1];
      const uint64_t pc = col[i - 1];
      ok = ok && (pr < r || (pr == r && pc < c));
    }
    bad += ok ? 0 : 1;
  }
  // wave reduction, then the workgroup's four waves in order
  for (int d = kWave / 2; d > 0; d >>= 1) {
    sum += __shfl_down(sum, d, kWave);
    bad += __shfl_down(bad, d, kWave);
  }
  const uint32_t w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    ssum[w] = sum;
    sbad[w] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    uint64_t b = 0;
    for (int k = 0; k < kBlock / kWave; ++k) {
      s += ssum[k];
      b += sbad[k];
    }
    part[blockIdx.x].sum = s;
    part[blockIdx.x].bad = b;
  }
}

// Entry i -> per-mode indices: X free modes from row_key[row[i]], Y free
// modes from col[i].  Fields are (shift, mask) pairs inside the X free key
// and the Y free key.
__global__ void __launch_bounds__(kBlock) k_decode(const uint32_t* __restrict__ row,
                                                   const uint64_t* __restrict__ col,
                                                   const float* __restrict__ val, uint64_t n,
                                                   const uint64_t* __restrict__ row_key,
                                                   DecodeArgs a, uint64_t out0) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const uint64_t xf = row_key[row[i]];
    const uint64_t yf = col[i];
    const uint64_t o = out0 + i;
#pragma unroll 4
    for (int m = 0; m < a.nx; ++m) a.out[m][o] = (uint32_t)(shr64(xf, a.shift[m]) & a.mask[m]);
#pragma unroll 4
    for (int m = a.nx; m < a.nx + a.ny; ++m)
      a.out[m][o] = (uint32_t)(shr64(yf, a.shift[m]) & a.mask[m]);
    if (a.vals) a.vals[o] = val[i];
  }
}

__global__ void __launch_bounds__(kBlock) k_row_extent(const uint32_t* __restrict__ row,
                                                       uint64_t n, uint64_t base,
                                                       uint64_t* __restrict__ start,
                                                       uint64_t* __restrict__ end) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const uint32_t r = row[i];
    if (i == 0 || row[i - 1] != r) start[r] = base + i;
    if (i + 1 == n || row[i + 1] != r) end[r] = base + i + 1;
  }
}

__global__ void __launch_bounds__(kBlock) k_interleave_counts(ExtentSet e, uint64_t R,
                                                              uint64_t* __restrict__ cnt) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  const uint64_t n = R * (uint64_t)e.Q;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const uint64_t r = i / (uint64_t)e.Q;
    const int q = (int)(i - r * (uint64_t)e.Q);
    cnt[i] = e.end[q][r] - e.start[q][r];
  }
}

__global__ void __launch_bounds__(kBlock) k_merge_scatter(
    const uint32_t* __restrict__ row, const uint64_t* __restrict__ col,
    const float* __restrict__ val, uint64_t n, uint64_t base, const uint64_t* __restrict__ start,
    const uint64_t* __restrict__ dst_base, int Q, int q, uint32_t row_off,
    uint32_t* __restrict__ orow, uint64_t* __restrict__ ocol, float* __restrict__ oval) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const uint32_t r = row[i];
    const uint64_t d = dst_base[(uint64_t)r * Q + q] + (base + i - start[r]);
    orow[d] = r + row_off;
    ocol[d] = col[i];
    oval[d] = val[i];
  }
}

uint32_t grid_for(uint64_t n, uint32_t cap) {
  const uint64_t g = ceil_div(n, (uint64_t)kBlock * kVerifyItems);
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

}  // namespace

void launch_row_extent(const uint32_t* row, uint64_t n, uint64_t base, uint64_t* start,
                       uint64_t* end, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_row_extent, dim3(grid_for(n, 8192)), dim3(kBlock), 0, s, row, n, base,
                     start, end);
  SPTC_HIP(hipGetLastError());
}

void launch_interleave_counts(const ExtentSet& e, uint64_t R, uint64_t* cnt, hipStream_t s) {
  if (!R || e.Q < 1 || e.Q > kMaxMergeParts) return;
  hipLaunchKernelGGL(k_interleave_counts, dim3(grid_for(R * (uint64_t)e.Q, 8192)), dim3(kBlock),
                     0, s, e, R, cnt);
  SPTC_HIP(hipGetLastError());
}

void launch_merge_scatter(const uint32_t* row, const uint64_t* col, const float* val,
                          uint64_t n, uint64_t base, const uint64_t* start,
                          const uint64_t* dst_base, int Q, int q, uint32_t row_off,
                          uint32_t* orow, uint64_t* ocol, float* oval, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_merge_scatter, dim3(grid_for(n, 8192)), dim3(kBlock), 0, s, row, col,
                     val, n, base, start, dst_base, Q, q, row_off, orow, ocol, oval);
  SPTC_HIP(hipGetLastError());
}

uint32_t verify_grid(uint64_t n) { return grid_for(n, kVerifyMaxBlocks); }

void launch_verify(const uint32_t* row, const uint64_t* col, const float* val, uint64_t n,
                   uint64_t r0, uint64_t r1, VerifyPart* part, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_verify, dim3(verify_grid(n)), dim3(kBlock), 0, s, row, col, val, n, r0,
                     r1, part);
  SPTC_HIP(hipGetLastError());
}

void launch_decode(const uint32_t* row, const uint64_t* col, const float* val, uint64_t n,
                   const uint64_t* row_key, const DecodeArgs& a, uint64_t out0, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_decode, dim3(grid_for(n, 8192)), dim3(kBlock), 0, s, row, col, val, n,
                     row_key, a, out0);
  SPTC_HIP(hipGetLastError());
}

}  // namespace sptc
