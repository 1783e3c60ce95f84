/* Seeded synthetic COO generator (host only).  Not part of the reference's
 * interface: the reference reads FROSTT .tns files (io_manager.tpp:74-259);
 * this produces the same in-memory SoA form (0-based indices per mode, FP32
 * values) reproducibly, for tests and benchmarks. */
#ifndef SPTC_GEN_H_
#define SPTC_GEN_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPTC_GEN_MAX_ORDER 8
#define SPTC_GEN_UNIFORM 0
#define SPTC_GEN_ZIPF 1
#define SPTC_GEN_UNIFORM01 0
#define SPTC_GEN_LOGNORMAL 1

/* SplitMix64 step, exported so host code can derive per-instance seeds. */
uint64_t sptc_gen_splitmix64(uint64_t z);

/* Draws `nnz` distinct coordinates of an order-`order` tensor with the given
 * dims; dist[m] selects SPTC_GEN_UNIFORM or SPTC_GEN_ZIPF (exponent
 * zipf_s[m]).  idx_out[m] receives nnz uint32 indices, val_out nnz values
 * (value_dist: SPTC_GEN_UNIFORM01 -> (0,1], SPTC_GEN_LOGNORMAL -> exp(N(0,1))).
 * Returns 0 on success; -1 bad args, -2 coordinate space exceeds 2^64,
 * -3 nnz exceeds the coordinate space, -4 too many duplicate draws. */
int sptc_gen_coo(uint64_t seed, int order, const uint64_t* dims,
                 const int* dist, const double* zipf_s, uint64_t nnz,
                 int value_dist, uint32_t* const* idx_out, float* val_out);

#ifdef __cplusplus
}
#endif
#endif
