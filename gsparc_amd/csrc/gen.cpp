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
  uint64_t coord_idx[SPTC_GEN_MAX_ORDER];
  uint64_t got = 0;
  for (uint64_t k = 0; got < nnz; ++k) {
    if (k >= max_draws) return -4;  // coordinate space too saturated
    uint64_t coord = 0;
    for (int m = 0; m < order; ++m) {
      coord_idx[m] = ms[m].draw(splitmix64(base + k * (uint64_t)order + m));
      coord = coord * dims[m] + coord_idx[m];
    }
    uint64_t h = splitmix64(coord) & mask;
    bool dup = false;
    while (table[h] != 0) {
      if (table[h] == coord + 1) { dup = true; break; }
      h = (h + 1) & mask;
    }
    if (dup) continue;
    table[h] = coord + 1;
    for (int m = 0; m < order; ++m) idx_out[m][got] = (uint32_t)coord_idx[m];
    uint64_t x1 = splitmix64(vbase + 2 * got);
    if (value_dist == SPTC_GEN_LOGNORMAL) {
      uint64_t x2 = splitmix64(vbase + 2 * got + 1);
      double u1 = unit_open0(x1), u2 = unit_closed0(x2);
      double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
      val_out[got] = (float)std::exp(z);
    } else {
      val_out[got] = (float)((double)((x1 >> 40) + 1) * (1.0 / 16777216.0));
    }
    ++got;
  }
  return 0;
}
