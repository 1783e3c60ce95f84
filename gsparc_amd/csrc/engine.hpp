// SpTC engine: operand preparation (linearize + sort), index match,
// memory-aware chunking and ESC accumulation.  Host orchestration only; the
// work runs in the kernels of kernels.hip and sort.hip.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "device.hpp"
#include "kernels.hpp"
#include "ktimer.hpp"
#include "segsort.hpp"

namespace sptc {

// Bit layout of one operand's linearized key (the reference's SLITOM layout,
// tensor_manager.tpp:63-186).
struct Layout {
  int order = 0;
  uint64_t dims[SPTC_MAX_ORDER] = {0};
  int shift[SPTC_MAX_ORDER] = {0};
  int width[SPTC_MAX_ORDER] = {0};
  int nc = 0, cmodes[SPTC_MAX_ORDER] = {0};
  int nf = 0, fmodes[SPTC_MAX_ORDER] = {0};
  int nfbits = 0, ncbits = 0, nbits = 0;
  int role = 1;
  // 1: the packed contract field needs more than 64 bits with the free
  // field, so the contract field holds the rank of the contracted
  // coordinates among those of both operands (ncbits = bits(#distinct)),
  // set when the operands are prepared (rank_layout)
  int crank = 0;
};

int mode_bits(uint64_t d);
sptc_status make_layout(int order, const uint64_t* dims, int nc, const int32_t* cmodes,
                        const int32_t* cwidth, int role, Layout* out);
// X and Y layouts of one contraction with one contract-field layout for
// both: the wider of the two dims per contracted pair (equal to the
// reference's layout whenever the dims agree).  msg receives the reason of a
// failure.
sptc_status make_pair_layouts(int xorder, const uint64_t* xdims, int yorder,
                              const uint64_t* ydims, int nc, const int32_t* cx,
                              const int32_t* cy, Layout* xl, Layout* yl, std::string* msg);

// L with its contract field replaced by a rank of rbits bits (free fields
// keep their place relative to it).
Layout rank_layout(const Layout& L, int rbits);

// A linearized, sorted operand resident in device memory.
struct Operand {
  Layout lay;
  uint64_t nnz = 0;
  DevBuf<uint64_t> keys;
  DevBuf<float> vals;
};

// One accumulation chunk of the output (a contiguous X-row range).  A chunk
// either owns its arrays or views a slice of a batch's output buffer.
struct OutChunk {
  uint64_t nnz = 0;
  uint64_t r0 = 0, r1 = 0;  // X-row range [r0, r1) of the chunk
  DevBuf<uint32_t> row;
  DevBuf<uint64_t> col;
  DevBuf<float> val;
  const uint32_t* vrow = nullptr;
  const uint64_t* vcol = nullptr;
  const float* vval = nullptr;
  const uint32_t* R() const { return vrow ? vrow : row.get(); }
  const uint64_t* C() const { return vcol ? vcol : col.get(); }
  const float* V() const { return vval ? vval : val.get(); }
};

}  // namespace sptc

struct sptc_result_s {
  sptc_result_info info;
  sptc::KStat kstats[SPTC_K_COUNT];
  sptc::Layout xl, yl;
  sptc::DevBuf<uint64_t> row_key;
  const uint64_t* vrow_key = nullptr;  // view (batch instance results)
  std::shared_ptr<void> keep;          // owner of the viewed buffers
  std::vector<sptc::OutChunk> chunks;
  // ESC chunk plan: first X row and first product of each chunk, then the
  // totals (empty for the dense accumulator)
  std::vector<uint64_t> part_rows, part_prods;
  // the output as a device COO operand (sptc_result_as_coo): per-mode
  // indices, values and dims, built on first request
  std::vector<sptc::DevBuf<uint32_t>> coo_idx;
  sptc::DevBuf<float> coo_val;
  std::vector<const void*> coo_ptr;
  std::vector<uint64_t> coo_dims;
  const uint64_t* RK() const { return vrow_key ? vrow_key : row_key.get(); }
};

namespace sptc {

struct PhaseTimes {
  double ms[6] = {0, 0, 0, 0, 0, 0};
};

// A COO operand's arrays in device memory (caller's, or staged copies).
struct StagedCOO {
  uint64_t nnz = 0;
  int order = 0;
  int idx_bytes = 4;
  const void* idx[SPTC_MAX_ORDER] = {nullptr};
  const float* val = nullptr;
  std::vector<DevBuf<uint8_t>> own;
};

// Intermediate products of X x Y for two prepared operands (one reduce scan
// over X against a contract table of Y): the partitioner's output-size
// bound, used for admission.  Synchronizes the stream.
uint64_t count_products(const Operand& X, const Operand& Y, hipStream_t s);

// Product sorts are row-segmented (the expansion leaves products grouped by
// X row, so only the column bits need sorting; see build_row_segments) unless
// SPTC_PROD_SORT=full asks for one sort over all key bits.
bool prefer_row_sort();
// SPTC_BIG_TAIL=buckets: big rows of a row-segmented sort take the
// high-digit form (one pass, then an LDS reduce per bucket) where the column
// bits allow; by default they are sorted on all column bits.
bool prefer_bucket_tail();

class Engine {
 public:
  Engine(hipStream_t s, const sptc_opts& o);
  ~Engine();
  // Copies host-resident COO arrays to the device (no-op for device ones).
  void stage(const sptc_coo& t, StagedCOO* out);
  // SPTC_ERR_BAD_ARGS (detail in msg) when an index of t is not below its
  // mode's dim (the linearization would mask it into another coordinate).
  sptc_status check_indices(const StagedCOO& t, const uint64_t* dims, const char* name);
  // Linearize and sort one operand.  With lay.crank, `crank` holds each
  // nonzero's contract rank (index width of t) in place of the contracted
  // modes.
  void prepare(const StagedCOO& t, const Layout& lay, Operand* out,
               const void* crank = nullptr);
  // Both operands of a contraction: for layouts with crank set, first the
  // ranks of the contracted coordinates over X and Y together (one sort
  // per group of modes that fits 64 bits), then each operand with its
  // contract field replaced by the rank.  Returns SPTC_ERR_KEY_OVERFLOW
  // (detail in msg) when a free field and the rank do not fit 64 bits.
  sptc_status prepare_pair(const StagedCOO& sx, const StagedCOO& sy, const Layout& xl,
                           const Layout& yl, Operand* ox, Operand* oy);
  // Match + partition + accumulate.  X and Y are prepared operands.
  sptc_status contract(const Operand& X, const Operand& Y, sptc_result_s* res);
  uint64_t budget() const { return budget_; }
  const sptc_opts& options() const { return o_; }
  // budget bytes above what was live when the engine was made
  uint64_t budget_avail() const { return budget_ - live0_; }
  hipStream_t stream() const { return s_; }
  void record(int i);
  // ms between two recorded events (waits for the later one)
  double elapsed_ms(int a, int b);
  // Collects per-kernel timings (after the stream is synchronized).
  void collect_kernel_stats(sptc::KStat* out) { kt_.collect(out, SPTC_K_COUNT); }
  KTimer& ktimer() { return kt_; }
  PhaseTimes times;
  std::string msg;  // last error detail
  // contract() stops after the ESC chunk plan (res->part_rows/part_prods)
  bool partition_only = false;

 private:
  U2 read_total(const U2* dev);
  // Ranks of the contracted coordinates of X then Y in *rank (nx + ny);
  // returns the number of distinct tuples (0 and msg set on failure).
  uint64_t contract_ranks(const StagedCOO& sx, const StagedCOO& sy, const Layout& xl,
                          const Layout& yl, DevBuf<uint64_t>* rank);
  hipStream_t s_;
  sptc_opts o_;
  uint64_t budget_ = 0;
  uint64_t live0_ = 0;
  int rb_ = 8;
  SortWorkspace ws_;
  KTimer kt_;
  hipEvent_t ev_[8];
  bool ev_used_[8] = {false};
  U2* host_total_ = nullptr;
  uint64_t* host_u64_ = nullptr;
};

}  // namespace sptc
