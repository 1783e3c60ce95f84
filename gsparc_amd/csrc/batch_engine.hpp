// Batched SpTC: many independent contractions through one launch sequence.
//
// plan() validates the instances, lays their operands out back to back and
// runs the whole pipeline once, eagerly, reading back the sizes it needs
// (per-instance row/product/output counts) and allocating every buffer.
// enqueue() replays the same launch sequence with no allocation and no host
// synchronization, so it can be captured into a hipGraph; replays re-read
// the operands' current values (index arrays must be unchanged -- a device
// digest of the sorted operand keys is checked first, and a mismatch turns
// every later kernel into a no-op and is reported by sync()).
#pragma once
#include <string>
#include <vector>

#include "batch.hpp"
#include "engine.hpp"

namespace sptc {

struct BOut {  // one output buffer: a dense slot-width class or an ESC chunk
  uint64_t nnz = 0;
  DevBuf<uint32_t> row;
  DevBuf<uint64_t> col;
  DevBuf<float> val;
  DevBuf<uint64_t> bounds;     // [B+1] first entry of every instance (plan only)
  std::vector<uint64_t> hb;
};

struct DenseClass {
  int eb = 8;
  uint64_t rows = 0;
  uint32_t G = 0;
  DevBuf<uint64_t> grp_rows, grp_row0;
  DevBuf<uint32_t> grp_inst;
  DevBuf<uint32_t> row_cnt;
  DevBuf<uint64_t> out_off;
  int out = -1;
  uint64_t ir = 0;
};

struct EscChunk {
  uint64_t n = 0;      // products
  uint64_t lrows = 0;  // chunk-local rows
  int K = 0, bits = 0, rb = 8;
  std::vector<BPiece> hp;
  std::vector<BETile> ht;
  std::vector<RowRange> hr;  // the pieces as row ranges (row-segmented sort)
  DevBuf<BPiece> pieces;
  DevBuf<BETile> tiles;
  DevBuf<uint32_t> map_row, map_inst;  // chunk-local row -> global row, instance
  bool rowsort = false;
  RowSegPlan rseg;
  DevBuf<uint64_t> row_big;    // chunk-local row -> big segment (row-segmented sort)
  DevBuf<uint64_t> row_start;  // chunk-local row -> first product; [lrows] = n
  DevBuf<uint32_t> tile_row;   // rows at the reduce tiles' starts (NarrowRows)
  int out = -1;
};

struct InstSlice {
  int out;
  uint64_t b, e;    // entry range in the output buffer
  uint64_t r0, r1;  // instance-local row range covered
};

class BatchEngine {
 public:
  BatchEngine();
  ~BatchEngine();
  // pre (optional, 2B entries: X_0..X_{B-1}, Y_0..Y_{B-1}) supplies operands
  // already linearized and sorted with this instance's layout (make_pair_
  // layouts); the batch then copies them instead of linearizing and sorting.
  sptc_status plan(int B, const sptc_coo* X, const sptc_coo* Y, const int32_t* nc,
                   const int32_t* const* cx, const int32_t* const* cy, const sptc_opts& o,
                   hipStream_t s, const Operand* const* pre = nullptr);
  // Launch sequence without allocation or synchronization.
  void enqueue(hipStream_t s) { run(s, false); }
  sptc_status capture();
  sptc_status launch_graph(hipStream_t s);
  // Waits for stream s and reports a stale plan.
  sptc_status sync(hipStream_t s);
  sptc_status instance_result(int i, sptc_result_s* out, hipStream_t s = nullptr);
  sptc_batch_info info;
  std::string msg;
  KTimer& ktimer() { return kt_; }

 private:
  sptc_status run(hipStream_t s, bool planning);
  sptc_status plan_accumulators(hipStream_t s);
  sptc_status build_chunks(hipStream_t s);
  uint64_t readback_u64(const uint64_t* dev, hipStream_t s);
  void readback(void* host, const void* dev, size_t bytes, hipStream_t s);

  sptc_opts o_{};
  uint32_t B_ = 0;
  uint64_t NX_ = 0, NY_ = 0, N_ = 0;
  int nbits_all_ = 0, ynf_max_ = 0;
  uint64_t budget_ = 0;
  std::vector<Layout> xl_, yl_;
  std::vector<StagedCOO> staged_;
  std::vector<uint64_t> seg_off_h_, yoff_h_;
  DevBuf<BLinArgs> lin_;
  DevBuf<uint64_t> seg_off_, yoff_, fmask_;
  DevBuf<int> ynf_, kyf_;
  SegSortPlan plan_all_, plan_y_;
  SortWorkspace ws_;
  bool lin_alt_ = false;
  bool presorted_ = false;
  DevBuf<uint64_t> k0_, k1_, fk_;
  DevBuf<uint32_t> ycol_;  // column field (rank) of every Y key, 32 bits
  DevBuf<uint32_t> ytab32_;  // ytab_ as 32-bit words when every Y free key fits
  DevBuf<float> v0_, v1_, fv_;
  DevBuf<uint64_t> ytab_, yt_off_, ukey_, ustart_, u_off_;
  DevBuf<BInst> inst_;
  DevBuf<uint64_t> act_prefix_, act_ypos_, row_key_, row_prod_, row_act_, itab_;
  DevBuf<float> act_xval_;
  DevBuf<uint32_t> act_row_, dup_, bad_;
  DevBuf<U2> tmp_;
  DevBuf<uint64_t> digest_, scratch_;
  DevBuf<double> vsum_;
  std::vector<uint64_t> iP_, iA_, iR_, ytoff_h_;
  std::vector<int> kyf_h_, acc_;
  std::vector<DenseClass> dense_;
  std::vector<EscChunk> esc_;
  DevBuf<uint64_t> kA_;
  DevBuf<float> vA_;
  std::vector<BOut> outs_;
  std::vector<std::vector<InstSlice>> slices_;
  KTimer kt_;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  hipStream_t cap_stream_ = nullptr;
  bool planned_ = false;
};

}  // namespace sptc
