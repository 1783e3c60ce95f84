// SpTC serving: a request loop over resident tensors with HBM-footprint
// admission and continuous batching.
//
//   submit -> planner thread: prepare (linearize + sort) the operands once per
//             tensor and layout, keep them resident; count the request's
//             intermediate products (the partitioner's output-size bound) ->
//             footprint estimate -> ready queue
//   ready  -> worker threads (one stream and one device pool each): admit
//             requests in arrival order while their footprints fit the free
//             budget; small requests that are ready together form one batch
//             (BatchEngine over the prepared operands), a request above the
//             large-product threshold runs alone as a count-only contraction.
//             Workers run concurrently, so a batch of small requests overlaps
//             a large one on the device.
//   done   -> the request's result (a view into its batch, kept alive by the
//             result) and its latency; the footprint is released when the
//             last result of a batch is freed.
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "batch_engine.hpp"

namespace sptc {

struct ResidentTensor {
  sptc_coo desc{};                 // device descriptor (on_device = 1)
  uint64_t dims[SPTC_MAX_ORDER] = {0};
  const void* idx[SPTC_MAX_ORDER] = {nullptr};
  StagedCOO staged;                // device copies the server owns
  std::map<std::string, std::shared_ptr<Operand>> prepared;  // by layout
};

struct Request {
  uint64_t id = 0;
  int32_t x = 0, y = 0, nc = 0;
  int32_t cx[SPTC_MAX_ORDER] = {0}, cy[SPTC_MAX_ORDER] = {0};
  Layout xl, yl;
  std::shared_ptr<Operand> px, py;
  uint64_t ir = 0;
  uint64_t footprint = 0;
  bool large = false;
  int state = 0;  // 0 submitted, 1 ready, 2 running, 3 done
  sptc_status st = SPTC_OK;
  std::string msg;
  std::shared_ptr<sptc_result_s> result;
  int32_t batch_size = 0, worker = -1;
  std::chrono::steady_clock::time_point t_submit, t_start, t_done;
};

// Budget accounting shared with the batches' results (which may outlive the
// server's queues).
struct Ledger {
  std::mutex mu;
  std::condition_variable cv;
  uint64_t in_use = 0, peak = 0;
  void charge(uint64_t b);
  void release(uint64_t b);
};

class Server {
 public:
  explicit Server(const sptc_server_opts& o);
  ~Server();
  sptc_status add_tensor(const sptc_coo& t, int32_t* id, std::string* msg);
  sptc_status submit(int32_t x, int32_t y, int32_t nc, const int32_t* cx, const int32_t* cy,
                     uint64_t* id, std::string* msg);
  // Blocks until the request is done; hands out its result (count-only
  // requests: info only) and forgets the request.
  sptc_status wait(uint64_t id, sptc_result_s** out, sptc_request_info* info, std::string* msg);
  void stats(sptc_server_stats* s);

 private:
  void planner();
  void worker(int w);
  sptc_status plan_request(Request& r, std::string* msg);
  std::shared_ptr<Operand> prepared(int32_t t, const Layout& lay);
  void run_large(int w, std::shared_ptr<Request> r);
  void run_batch(int w, std::vector<std::shared_ptr<Request>>& rs);
  void finish(const std::shared_ptr<Request>& r);
  void fail_all(std::vector<std::shared_ptr<Request>>& pick, const std::shared_ptr<Request>& big,
                sptc_status st, const std::string& m);

  sptc_server_opts o_{};
  int device_ = 0;
  std::mutex mu_;                      // queues, tensors, request table
  std::condition_variable cv_plan_, cv_done_;
  std::deque<std::shared_ptr<Request>> submitted_, ready_;
  std::map<uint64_t, std::shared_ptr<Request>> requests_;
  std::vector<std::unique_ptr<ResidentTensor>> tensors_;
  std::map<std::string, uint64_t> ir_cache_;
  std::shared_ptr<Ledger> ledger_;
  uint64_t resident_bytes_ = 0;
  uint64_t next_id_ = 1;
  bool stop_ = false;
  // a large request is running (they run one at a time: each is HBM-bound,
  // so two at once only share the bandwidth and delay the earlier one)
  bool large_running_ = false;
  hipStream_t plan_stream_ = nullptr;
  std::vector<hipStream_t> streams_;
  std::vector<std::shared_ptr<DevicePool>> pools_;
  std::thread planner_;
  std::vector<std::thread> workers_;
  // statistics
  uint64_t n_done_ = 0, n_batches_ = 0, n_large_ = 0, n_batched_ = 0;
};

}  // namespace sptc
