// Seeded synthetic COO tensor generator (host only, no GPU).
//
// The reference has no generator: its inputs are FROSTT `.tns` files read by
// IOManager::parse_metadata_and_data (io_manager.tpp:74-259), which yields
// 0-based SoA indices, FP32 values and dims = observed max index + 1.  This
// generator produces the same SoA form from a fixed seed so every workload is
// reproducible bit-for-bit on any x86-64 host.
//
// Random stream: SplitMix64 over a counter.  Draw k (0,1,2,...) of mode m uses
// counter k*order + m; the value of the j-th accepted (unique) coordinate uses
// its own counter stream, so values do not depend on how many draws were
// rejected as duplicates.  Coordinates are drawn until exactly `nnz` distinct
// ones exist ("duplicates merged"); entries are stored in first-draw order.
//
// Draws are computed in parallel blocks (each draw depends only on its
// counter) and accepted in counter order by one thread, so the output does
// not depend on the thread count.
//
// Zipf(s) over [0, d): inverse CDF over ranks 1..d with P(r) ~ r^-s, rank 1 ->
// index 0.  Uniform ints use a 128-bit multiply-high (exact).  Values are
// uniform on (0,1] in steps of 2^-24 (exact in FP32) or lognormal(0,1) via
// Box-Muller in FP64, cast to FP32.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <thread>

#include "../../include/sptc_gen.h"

namespace {

inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline double unit_open0(uint64_t x) {  // (0, 1]
  return (double)((x >> 11) + 1) * (1.0 / 9007199254740992.0);
}
inline double unit_closed0(uint64_t x) {  // [0, 1)
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

struct ModeSampler {
  int dist;
  uint64_t d;
  std::vector<double> cdf;
  void init(int dist_, uint64_t d_, double s) {
    dist = dist_;
    d = d_;
    if (dist == SPTC_GEN_ZIPF) {
      cdf.resize(d);
      double acc = 0.0;
      for (uint64_t r = 0; r < d; ++r) {
        acc += std::pow((double)(r + 1), -s);
        cdf[r] = acc;
      }
    }
  }
  uint64_t draw(uint64_t x) const {
    if (dist == SPTC_GEN_UNIFORM) {
      return (uint64_t)(((unsigned __int128)x * d) >> 64);
    }
    double u = unit_closed0(x) * cdf[d - 1];
    uint64_t r = (uint64_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    return r < d ? r : d - 1;
  }
};

}  // namespace

extern "C" uint64_t sptc_gen_splitmix64(uint64_t z) { return splitmix64(z); }

extern "C" int sptc_gen_coo(uint64_t seed, int order, const uint64_t* dims,
                            const int* dist, const double* zipf_s, uint64_t nnz,
                            int value_dist, uint32_t* const* idx_out,
                            float* val_out) {
  if (order < 1 || order > SPTC_GEN_MAX_ORDER || !dims || !dist || !idx_out ||
      !val_out)
    return -1;
  unsigned __int128 space = 1;
  for (int m = 0; m < order; ++m) {
    if (dims[m] == 0 || dims[m] > 0xFFFFFFFFull) return -1;
    space *= dims[m];
    if (space > (unsigned __int128)UINT64_MAX) return -2;  // coord won't fit u64
  }
  if ((unsigned __int128)nnz > space) return -3;
  std::vector<ModeSampler> ms(order);
  for (int m = 0; m < order; ++m)
    ms[m].init(dist[m], dims[m], zipf_s ? zipf_s[m] : 1.0);

  uint64_t cap = 1;
  while (cap < 2 * nnz + 16) cap <<= 1;
  std::vector<uint64_t> table(cap, 0);  // stores coord+1; 0 = empty
  const uint64_t mask = cap - 1;
  const uint64_t base = splitmix64(seed);
  const uint64_t vbase = splitmix64(seed ^ 0xA5A5A5A5DEADBEEFull);
  const uint64_t max_draws = 256 * nnz + (1u << 20);
  unsigned nthr = std::thread::hardware_concurrency();
  nthr = nthr < 1 ? 1 : (nthr > 16 ? 16 : nthr);
  const uint64_t block = nnz < (1u << 16) ? 4096 : (1u << 20);
  // draws [k0, k0 + block): per draw its coordinate and per-mode indices
  std::vector<uint64_t> bcoord(block);
  std::vector<uint32_t> bidx((size_t)block * order);
  auto draw_range = [&](uint64_t k0, uint64_t lo, uint64_t hi) {
    for (uint64_t j = lo; j < hi; ++j) {
      const uint64_t k = k0 + j;
      uint64_t coord = 0;
      for (int m = 0; m < order; ++m) {
        const uint64_t c = ms[m].draw(splitmix64(base + k * (uint64_t)order + m));
        bidx[j * order + m] = (uint32_t)c;
        coord = coord * dims[m] + c;
      }
      bcoord[j] = coord;
    }
  };
  uint64_t got = 0;
  for (uint64_t k0 = 0; got < nnz; k0 += block) {
    if (k0 >= max_draws) return -4;  // coordinate space too saturated
    if (nthr == 1 || block < 8192) {
      draw_range(k0, 0, block);
    } else {
      std::vector<std::thread> th;
      const uint64_t per = (block + nthr - 1) / nthr;
      for (unsigned t = 0; t < nthr; ++t) {
        const uint64_t lo = t * per, hi = std::min<uint64_t>(block, lo + per);
        if (lo < hi) th.emplace_back(draw_range, k0, lo, hi);
      }
      for (auto& x : th) x.join();
    }
    for (uint64_t j = 0; j < block && got < nnz; ++j) {
      if (k0 + j >= max_draws) return -4;
      const uint64_t coord = bcoord[j];
      uint64_t h = splitmix64(coord) & mask;
      bool dup = false;
      while (table[h] != 0) {
        if (table[h] == coord + 1) { dup = true; break; }
        h = (h + 1) & mask;
      }
      if (dup) continue;
      table[h] = coord + 1;
      for (int m = 0; m < order; ++m) idx_out[m][got] = bidx[j * order + m];
      ++got;
    }
  }
  // values depend only on the accepted entry's position
  auto values = [&](uint64_t lo, uint64_t hi) {
    for (uint64_t g = lo; g < hi; ++g) {
      uint64_t x1 = splitmix64(vbase + 2 * g);
      if (value_dist == SPTC_GEN_LOGNORMAL) {
        uint64_t x2 = splitmix64(vbase + 2 * g + 1);
        double u1 = unit_open0(x1), u2 = unit_closed0(x2);
        double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        val_out[g] = (float)std::exp(z);
      } else {
        val_out[g] = (float)((double)((x1 >> 40) + 1) * (1.0 / 16777216.0));
      }
    }
  };
  if (nthr == 1 || nnz < (1u << 16)) {
    values(0, nnz);
  } else {
    std::vector<std::thread> th;
    const uint64_t per = (nnz + nthr - 1) / nthr;
    for (unsigned t = 0; t < nthr; ++t) {
      const uint64_t lo = t * per, hi = std::min<uint64_t>(nnz, lo + per);
      if (lo < hi) th.emplace_back(values, lo, hi);
    }
    for (auto& x : th) x.join();
  }
  return 0;
}
