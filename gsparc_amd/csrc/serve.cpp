// SpTC serving loop (host orchestration; see serve.hpp).
#include "serve.hpp"

#include <cstring>

namespace sptc {

void Ledger::charge(uint64_t b) {
  std::lock_guard<std::mutex> g(mu);
  in_use += b;
  peak = std::max(peak, in_use);
}

void Ledger::release(uint64_t b) {
  {
    std::lock_guard<std::mutex> g(mu);
    in_use -= std::min(in_use, b);
  }
  cv.notify_all();
}

namespace {

// Owner of one batch's device buffers; the last result of the batch to be
// freed releases them and the batch's budget charge.
struct BatchHolder {
  std::shared_ptr<DevicePool> pool;
  std::shared_ptr<Ledger> ledger;
  std::unique_ptr<BatchEngine> eng;
  uint64_t charge = 0;
  ~BatchHolder() {
    eng.reset();
    ledger->release(charge);
  }
};

std::string layout_key(const Layout& L) {
  std::string k = std::to_string(L.role) + ":" + std::to_string(L.order);
  for (int m = 0; m < L.order; ++m)
    k += "," + std::to_string(L.shift[m]) + "/" + std::to_string(L.width[m]);
  return k;
}

// Device bytes a small request adds to a batch: concatenated operands and
// their ping-pong copies, Y rank and contract tables, match tables, products
// with their sort workspace, and outputs (at most one per product).
uint64_t small_footprint(uint64_t nx, uint64_t ny, uint64_t ir) {
  return 24 * (nx + ny) + 36 * ny + 48 * nx + 29 * ir + 16 * ir + (4ull << 20);
}

}  // namespace

Server::Server(const sptc_server_opts& o) : o_(o) {
  if (!o_.budget_bytes) o_.budget_bytes = 96ull << 30;
  if (o_.workers <= 0) o_.workers = 3;
  if (o_.max_batch <= 0) o_.max_batch = 256;
  if (o_.max_batch > kBatchMaxInstances) o_.max_batch = kBatchMaxInstances;
  if (!o_.large_products) o_.large_products = 1ull << 28;
  if (!o_.large_budget_bytes) o_.large_budget_bytes = o_.budget_bytes / 3;
  SPTC_HIP(hipGetDevice(&device_));
  // the workers allocate from pools of their own: hand back what the
  // process-wide pool caches from earlier calls
  DevicePool::global().trim();
  ledger_ = std::make_shared<Ledger>();
  SPTC_HIP(hipStreamCreateWithFlags(&plan_stream_, hipStreamNonBlocking));
  for (int w = 0; w < o_.workers; ++w) {
    hipStream_t s;
    SPTC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    streams_.push_back(s);
    pools_.push_back(std::make_shared<DevicePool>());
  }
  planner_ = std::thread([this] { planner(); });
  for (int w = 0; w < o_.workers; ++w) workers_.emplace_back([this, w] { worker(w); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_plan_.notify_all();
  ledger_->cv.notify_all();
  cv_done_.notify_all();
  if (planner_.joinable()) planner_.join();
  for (auto& t : workers_)
    if (t.joinable()) t.join();
  requests_.clear();  // results not yet handed out
  submitted_.clear();
  ready_.clear();
  tensors_.clear();
  for (auto s : streams_) (void)hipStreamDestroy(s);
  if (plan_stream_) (void)hipStreamDestroy(plan_stream_);
}

sptc_status Server::add_tensor(const sptc_coo& t, int32_t* id, std::string* msg) {
  auto rt = std::make_unique<ResidentTensor>();
  // Copies in the global pool, synchronized before returning; the planner
  // (the pool's other user while serving) frees only after synchronizing
  // its own stream, so neither reuses a block the other still reads.
  std::lock_guard<std::mutex> g(mu_);
  const uint64_t n = t.nnz;
  for (int m = 0; m < t.order; ++m) {
    rt->dims[m] = t.dims[m];
    rt->staged.own.emplace_back(n ? n * t.idx_bytes : 1);
    if (n)
      SPTC_HIP(hipMemcpyAsync(rt->staged.own.back().get(), t.idx[m], n * t.idx_bytes,
                              hipMemcpyDefault, plan_stream_));
    rt->idx[m] = rt->staged.own.back().get();
    rt->staged.idx[m] = rt->idx[m];
  }
  rt->staged.own.emplace_back(n ? n * 4 : 4);
  if (n)
    SPTC_HIP(hipMemcpyAsync(rt->staged.own.back().get(), t.val, n * 4, hipMemcpyDefault,
                            plan_stream_));
  SPTC_HIP(hipStreamSynchronize(plan_stream_));
  rt->staged.nnz = n;
  rt->staged.order = t.order;
  rt->staged.idx_bytes = t.idx_bytes;
  rt->staged.val = reinterpret_cast<const float*>(rt->staged.own.back().get());
  rt->desc.order = t.order;
  rt->desc.nnz = n;
  rt->desc.dims = rt->dims;
  rt->desc.idx = rt->idx;
  rt->desc.idx_bytes = t.idx_bytes;
  rt->desc.val = rt->staged.val;
  rt->desc.on_device = 1;
  resident_bytes_ += n * (t.order * t.idx_bytes + 4);
  *id = (int32_t)tensors_.size();
  tensors_.push_back(std::move(rt));
  (void)msg;
  return SPTC_OK;
}

sptc_status Server::submit(int32_t x, int32_t y, int32_t nc, const int32_t* cx,
                           const int32_t* cy, uint64_t* id, std::string* msg) {
  auto r = std::make_shared<Request>();
  r->t_submit = std::chrono::steady_clock::now();
  std::lock_guard<std::mutex> g(mu_);
  if (x < 0 || y < 0 || x >= (int32_t)tensors_.size() || y >= (int32_t)tensors_.size()) {
    *msg = "unknown tensor id";
    return SPTC_ERR_BAD_ARGS;
  }
  const sptc_coo& X = tensors_[x]->desc;
  const sptc_coo& Y = tensors_[y]->desc;
  sptc_status st = make_pair_layouts(X.order, X.dims, Y.order, Y.dims, nc, cx, cy, &r->xl,
                                     &r->yl, msg);
  if (st == SPTC_OK && r->xl.crank) {
    *msg = "keys above 64 bits (contract ranks) run through sptc_contract / sptc_plan";
    st = SPTC_ERR_KEY_OVERFLOW;
  }
  if (st != SPTC_OK) return st;
  r->x = x;
  r->y = y;
  r->nc = nc;
  for (int i = 0; i < nc; ++i) {
    r->cx[i] = cx[i];
    r->cy[i] = cy[i];
  }
  r->id = next_id_++;
  requests_[r->id] = r;
  submitted_.push_back(r);
  *id = r->id;
  cv_plan_.notify_one();
  return SPTC_OK;
}

std::shared_ptr<Operand> Server::prepared(int32_t t, const Layout& lay) {
  ResidentTensor* rt;
  {
    std::lock_guard<std::mutex> g(mu_);
    rt = tensors_[t].get();
  }
  const std::string k = layout_key(lay);
  auto it = rt->prepared.find(k);
  if (it != rt->prepared.end()) return it->second;
  auto op = std::make_shared<Operand>();
  sptc_opts o;
  std::memset(&o, 0, sizeof(o));
  {
    Engine e(plan_stream_, o);
    e.prepare(rt->staged, lay, op.get());
    SPTC_HIP(hipStreamSynchronize(plan_stream_));
  }
  rt->prepared[k] = op;
  std::lock_guard<std::mutex> g(mu_);
  resident_bytes_ += 12 * op->nnz;
  return op;
}

sptc_status Server::plan_request(Request& r, std::string* msg) {
  r.px = prepared(r.x, r.xl);
  r.py = prepared(r.y, r.yl);
  const std::string key = std::to_string(r.x) + layout_key(r.xl) + "|" + std::to_string(r.y) +
                          layout_key(r.yl);
  auto it = ir_cache_.find(key);
  if (it != ir_cache_.end()) {
    r.ir = it->second;
  } else {
    r.ir = count_products(*r.px, *r.py, plan_stream_);
    ir_cache_[key] = r.ir;
  }
  r.large = r.ir > o_.large_products;
  r.footprint = r.large ? o_.large_budget_bytes : small_footprint(r.px->nnz, r.py->nnz, r.ir);
  std::lock_guard<std::mutex> g(mu_);
  if (r.footprint + resident_bytes_ > o_.budget_bytes) {
    *msg = "request does not fit the server's memory budget";
    return SPTC_ERR_OOM_BUDGET;
  }
  return SPTC_OK;
}

void Server::planner() {
  (void)hipSetDevice(device_);
  // Prepared operands and the planner's temporaries live in the global pool;
  // only this thread (and add_tensor, under mu_ with a synchronized stream)
  // allocates from it while serving.
  while (true) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_plan_.wait(lk, [&] { return stop_ || !submitted_.empty(); });
      if (stop_) return;
      r = submitted_.front();
      submitted_.pop_front();
    }
    std::string m;
    sptc_status st;
    try {
      st = plan_request(*r, &m);
    } catch (const HipError& e) {
      st = SPTC_ERR_HIP;
      m = std::string(e.what) + ": " + hipGetErrorString(e.err);
    } catch (const std::exception& e) {
      st = SPTC_ERR_INTERNAL;
      m = e.what();
    }
    std::lock_guard<std::mutex> g(mu_);
    if (st != SPTC_OK) {
      r->st = st;
      r->msg = m;
      r->state = 3;
      r->t_done = std::chrono::steady_clock::now();
      cv_done_.notify_all();
    } else {
      r->state = 1;
      ready_.push_back(r);
      ledger_->cv.notify_all();
    }
  }
}

void Server::worker(int w) {
  (void)hipSetDevice(device_);
  DevicePool::Scope scope(pools_[w].get());
  while (true) {
    std::vector<std::shared_ptr<Request>> pick;
    std::shared_ptr<Request> big;
    uint64_t need = 0;
    {
      std::unique_lock<std::mutex> lk(mu_);
      if (stop_) return;
      uint64_t in_use;
      {
        std::lock_guard<std::mutex> g(ledger_->mu);
        in_use = ledger_->in_use;
      }
      const uint64_t cap = o_.budget_bytes > resident_bytes_ ? o_.budget_bytes - resident_bytes_ : 0;
      const uint64_t avail = cap > in_use ? cap - in_use : 0;
      // Arrival order: a large request at the head runs alone when it fits
      // and no other large request is running; small requests are taken in
      // order while they fit (and may pass a large request that waits).
      for (auto it = ready_.begin(); it != ready_.end();) {
        const auto& r = *it;
        if (r->large) {
          if (pick.empty() && !large_running_ && r->footprint <= avail) {
            big = r;
            large_running_ = true;
            ready_.erase(it);
            break;
          }
          ++it;
          continue;
        }
        if ((int)pick.size() >= o_.max_batch || need + r->footprint > avail) break;
        need += r->footprint;
        pick.push_back(r);
        it = ready_.erase(it);
      }
      if (!big && pick.empty()) {
        lk.unlock();
        std::unique_lock<std::mutex> lg(ledger_->mu);
        ledger_->cv.wait_for(lg, std::chrono::milliseconds(2));
        continue;
      }
      const auto now = std::chrono::steady_clock::now();
      for (auto& r : pick) {
        r->state = 2;
        r->t_start = now;
        r->worker = w;
        r->batch_size = (int32_t)pick.size();
      }
      if (big) {
        big->state = 2;
        big->t_start = now;
        big->worker = w;
        big->batch_size = 1;
        need = big->footprint;
      }
      ledger_->charge(need);  // under mu_: the next worker sees this charge
    }
    try {
      if (big) run_large(w, big);
      else run_batch(w, pick);
    } catch (const HipError& e) {
      fail_all(pick, big, SPTC_ERR_HIP,
               std::string(e.what) + ": " + hipGetErrorString(e.err));
    } catch (const std::exception& e) {
      fail_all(pick, big, SPTC_ERR_INTERNAL, e.what());
    }
    std::lock_guard<std::mutex> g(mu_);
    for (auto& r : pick) finish(r);
    if (big) finish(big);
    if (big) large_running_ = false;
    if (big) ++n_large_;
    else {
      ++n_batches_;
      n_batched_ += pick.size();
    }
    cv_done_.notify_all();
  }
}

void Server::fail_all(std::vector<std::shared_ptr<Request>>& pick,
                      const std::shared_ptr<Request>& big, sptc_status st,
                      const std::string& m) {
  for (auto& r : pick) {
    r->st = st;
    r->msg = m;
    r->result.reset();
  }
  if (big) {
    big->st = st;
    big->msg = m;
    big->result.reset();
  }
}

void Server::finish(const std::shared_ptr<Request>& r) {
  r->state = 3;
  r->t_done = std::chrono::steady_clock::now();
  ++n_done_;
}

void Server::run_large(int w, std::shared_ptr<Request> r) {
  hipStream_t s = streams_[w];
  sptc_opts o;
  std::memset(&o, 0, sizeof(o));
  o.count_only = 1;
  o.mem_budget_bytes = o_.large_budget_bytes;
  struct Release {  // the charge goes back however this returns
    Ledger* l;
    uint64_t b;
    ~Release() { l->release(b); }
  } rel{ledger_.get(), r->footprint};
  auto res = std::make_shared<sptc_result_s>();
  std::memset(&res->info, 0, sizeof(res->info));
  sptc_status st;
  {
    Engine e(s, o);
    st = e.contract(*r->px, *r->py, res.get());
    SPTC_HIP(hipStreamSynchronize(s));
    if (st != SPTC_OK) r->msg = e.msg;
  }
  r->st = st;
  if (st == SPTC_OK) r->result = res;
}

void Server::run_batch(int w, std::vector<std::shared_ptr<Request>>& rs) {
  hipStream_t s = streams_[w];
  const int B = (int)rs.size();
  std::vector<sptc_coo> X(B), Y(B);
  std::vector<int32_t> nc(B);
  std::vector<const int32_t*> cx(B), cy(B);
  std::vector<const Operand*> pre(2 * B);
  {
    std::lock_guard<std::mutex> g(mu_);
    for (int i = 0; i < B; ++i) {
      X[i] = tensors_[rs[i]->x]->desc;
      Y[i] = tensors_[rs[i]->y]->desc;
      nc[i] = rs[i]->nc;
      cx[i] = rs[i]->cx;
      cy[i] = rs[i]->cy;
      pre[i] = rs[i]->px.get();
      pre[B + i] = rs[i]->py.get();
    }
  }
  uint64_t charge = 0;
  for (auto& r : rs) charge += r->footprint;
  auto holder = std::make_shared<BatchHolder>();
  holder->pool = pools_[w];
  holder->ledger = ledger_;
  holder->charge = charge;
  holder->eng = std::make_unique<BatchEngine>();
  sptc_opts o;
  std::memset(&o, 0, sizeof(o));
  sptc_status st = holder->eng->plan(B, X.data(), Y.data(), nc.data(), cx.data(), cy.data(), o,
                                     s, pre.data());
  for (int i = 0; i < B; ++i) {
    rs[i]->st = st;
    if (st != SPTC_OK) {
      rs[i]->msg = holder->eng->msg;
      continue;
    }
    auto res = std::make_shared<sptc_result_s>();
    holder->eng->instance_result(i, res.get(), s);
    res->keep = holder;
    rs[i]->result = res;
  }
}

sptc_status Server::wait(uint64_t id, sptc_result_s** out, sptc_request_info* info,
                         std::string* msg) {
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = requests_.find(id);
    if (it == requests_.end()) {
      *msg = "unknown request id";
      return SPTC_ERR_BAD_ARGS;
    }
    r = it->second;
    cv_done_.wait(lk, [&] { return r->state == 3 || stop_; });
    if (r->state != 3) {
      *msg = "server stopped";
      return SPTC_ERR_INTERNAL;
    }
    requests_.erase(it);
  }
  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->id = r->id;
    info->status = r->st;
    info->count_only = r->large ? 1 : 0;
    info->batch_size = r->batch_size;
    info->worker = r->worker;
    info->ir_nnz = r->ir;
    using ms = std::chrono::duration<double, std::milli>;
    if (r->batch_size) info->queue_ms = ms(r->t_start - r->t_submit).count();
    info->latency_ms = ms(r->t_done - r->t_submit).count();
  }
  if (r->st != SPTC_OK) {
    *msg = r->msg;
    return r->st;
  }
  if (out) {
    // hand over a result object of the caller's own (sptc_result_free);
    // it shares the batch's buffers through `keep`
    auto* res = new sptc_result_s();
    res->info = r->result->info;
    res->xl = r->result->xl;
    res->yl = r->result->yl;
    res->vrow_key = r->result->RK();
    res->info.row_key = res->vrow_key;
    res->keep = r->result->keep ? r->result->keep : r->result;
    for (const OutChunk& c : r->result->chunks) {
      OutChunk v;
      v.nnz = c.nnz;
      v.r0 = c.r0;
      v.r1 = c.r1;
      v.vrow = c.R();
      v.vcol = c.C();
      v.vval = c.V();
      res->chunks.push_back(std::move(v));
    }
    std::memcpy(res->kstats, r->result->kstats, sizeof(res->kstats));
    *out = res;
  }
  return SPTC_OK;
}

void Server::stats(sptc_server_stats* st) {
  std::lock_guard<std::mutex> g(mu_);
  std::memset(st, 0, sizeof(*st));
  st->requests_done = n_done_;
  st->batches = n_batches_;
  st->batched_requests = n_batched_;
  st->large_requests = n_large_;
  st->resident_bytes = resident_bytes_;
  st->budget_bytes = o_.budget_bytes;
  std::lock_guard<std::mutex> g2(ledger_->mu);
  st->peak_in_use_bytes = ledger_->peak;
  st->in_use_bytes = ledger_->in_use;
}

}  // namespace sptc
