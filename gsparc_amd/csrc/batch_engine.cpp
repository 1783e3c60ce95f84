// Batched SpTC host orchestration (built with hipcc; launches live in
// batch.hip).  The stages mirror Engine's, over all instances at once.
#include "batch_engine.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace sptc {

namespace {
// Slot-width class of the dense accumulator for a column field of kyf bits.
int dense_class_bits(int kyf) { return kyf <= 8 ? 8 : (kyf <= 10 ? 10 : 13); }
}  // namespace

BatchEngine::BatchEngine() { std::memset(&info, 0, sizeof(info)); }

BatchEngine::~BatchEngine() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
}

void BatchEngine::readback(void* host, const void* dev, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  SPTC_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s));
  SPTC_HIP(hipStreamSynchronize(s));
}

uint64_t BatchEngine::readback_u64(const uint64_t* dev, hipStream_t s) {
  uint64_t v = 0;
  readback(&v, dev, 8, s);
  return v;
}

sptc_status BatchEngine::plan(int B, const sptc_coo* X, const sptc_coo* Y, const int32_t* nc,
                              const int32_t* const* cx, const int32_t* const* cy,
                              const sptc_opts& o, hipStream_t s, const Operand* const* pre) {
  const auto t0 = std::chrono::steady_clock::now();
  o_ = o;
  if (o.count_only) {
    msg = "count_only is not supported for batches";
    return SPTC_ERR_BAD_ARGS;
  }
  if (o.radix_bits && !radix_bits_supported(o.radix_bits)) {
    msg = "radix_bits must be 0 or 8..11";
    return SPTC_ERR_BAD_ARGS;
  }
  B_ = (uint32_t)B;
  xl_.resize(B);
  yl_.resize(B);
  for (int i = 0; i < B; ++i) {
    std::string m;
    sptc_status st = make_pair_layouts(X[i].order, X[i].dims, Y[i].order, Y[i].dims, nc[i],
                                       cx[i], cy[i], &xl_[i], &yl_[i], &m);
    if (st == SPTC_OK && xl_[i].crank) {
      st = SPTC_ERR_KEY_OVERFLOW;
      m = "keys above 64 bits (contract ranks) run through sptc_contract / sptc_plan";
    }
    if (st != SPTC_OK) {
      msg = "instance " + std::to_string(i) + ": " + m;
      return st;
    }
  }
  const uint64_t live = DevicePool::get().live_bytes();
  if (o.mem_budget_bytes) {
    budget_ = live + o.mem_budget_bytes;
  } else {
    const uint64_t fr = device_free_bytes();
    budget_ = live + (fr > (1ull << 30) ? fr - (1ull << 30) : 0);
  }
  kt_.on = o.kernel_timing != 0;

  // ---- operand layout: X_0..X_{B-1}, then Y_0..Y_{B-1} ----------------------
  std::unique_ptr<Engine> stager;  // host operands only
  staged_.resize(2 * B);
  seg_off_h_.assign(2 * B + 1, 0);
  std::vector<BLinArgs> la(2 * B);
  presorted_ = pre != nullptr;
  for (int g = 0; g < 2 * B; ++g) {
    const sptc_coo& t = g < B ? X[g] : Y[g - B];
    const Layout& L = g < B ? xl_[g] : yl_[g - B];
    BLinArgs& a = la[g];
    std::memset(&a, 0, sizeof(a));
    seg_off_h_[g + 1] = seg_off_h_[g] + t.nnz;
    if (pre && pre[g]) {
      if (pre[g]->nnz != t.nnz || pre[g]->lay.nbits != L.nbits) {
        msg = "prepared operand does not match its instance";
        return SPTC_ERR_BAD_ARGS;
      }
      a.skeys = pre[g]->keys.get();
      a.svals = pre[g]->vals.get();
      continue;
    }
    presorted_ = false;
    if (!t.on_device && t.nnz) {
      if (!stager) stager = std::make_unique<Engine>(s, o);
      stager->stage(t, &staged_[g]);
    } else {
      staged_[g].nnz = t.nnz;
      staged_[g].order = t.order;
      staged_[g].idx_bytes = t.idx_bytes;
      for (int m = 0; m < t.order; ++m) staged_[g].idx[m] = t.idx[m];
      staged_[g].val = t.val;
    }
    a.order = L.order;
    a.idx_bytes = t.idx_bytes;
    a.val = staged_[g].val;
    for (int m = 0; m < L.order; ++m) {
      a.idx[m] = staged_[g].idx[m];
      a.shift[m] = L.shift[m];
      a.width[m] = L.width[m];
    }
    nbits_all_ = std::max(nbits_all_, L.nbits);
  }
  if (pre)  // copied segments are sorted already; a mixed batch sorts all
    for (int g = 0; g < 2 * B; ++g) {
      const Layout& L = g < B ? xl_[g] : yl_[g - B];
      nbits_all_ = std::max(nbits_all_, L.nbits);
    }
  NX_ = seg_off_h_[B];
  N_ = seg_off_h_[2 * B];
  NY_ = N_ - NX_;
  if (N_ >= (1ull << 32)) {
    msg = "batch operands hold 2^32 or more nonzeros";
    return SPTC_ERR_BAD_ARGS;
  }
  yoff_h_.assign(B + 1, 0);
  std::vector<uint64_t> fm(B);
  std::vector<int> ynf(B);
  for (int i = 0; i < B; ++i) {
    yoff_h_[i + 1] = seg_off_h_[B + i + 1] - NX_;
    fm[i] = low_mask(yl_[i].nfbits);
    ynf[i] = yl_[i].nfbits;
    ynf_max_ = std::max(ynf_max_, yl_[i].nfbits);
  }
  lin_.reset(2 * B);
  seg_off_.reset(2 * B + 1);
  yoff_.reset(B + 1);
  fmask_.reset(B);
  ynf_.reset(B);
  kyf_.reset(B);
  SPTC_HIP(hipMemcpyAsync(lin_.get(), la.data(), la.size() * sizeof(BLinArgs),
                          hipMemcpyHostToDevice, s));
  SPTC_HIP(hipMemcpyAsync(seg_off_.get(), seg_off_h_.data(), (2 * B + 1) * 8,
                          hipMemcpyHostToDevice, s));
  SPTC_HIP(hipMemcpyAsync(yoff_.get(), yoff_h_.data(), (B + 1) * 8, hipMemcpyHostToDevice, s));
  SPTC_HIP(hipMemcpyAsync(fmask_.get(), fm.data(), B * 8, hipMemcpyHostToDevice, s));
  SPTC_HIP(hipMemcpyAsync(ynf_.get(), ynf.data(), B * sizeof(int), hipMemcpyHostToDevice, s));
  plan_all_.build(seg_off_h_, s);  // synchronizes: the host tables above are released
  plan_y_.build(yoff_h_, s);
  segmented_sort_reserve(ws_, plan_all_, s);
  segmented_sort_reserve(ws_, plan_y_, s);
  lin_alt_ = !presorted_ && (segmented_sort_passes(nbits_all_) & 1) != 0;
  const uint64_t n1 = N_ ? N_ : 1;
  k0_.reset(n1);
  v0_.reset(n1);
  k1_.reset(n1);
  v1_.reset(n1);
  fk_.reset(NY_ ? NY_ : 1);
  fv_.reset(NY_ ? NY_ : 1);
  ycol_.reset(NY_ ? NY_ : 1);
  if (ynf_max_ <= 32) ytab32_.reset(NY_ ? NY_ : 1);
  ytab_.reset(NY_ ? NY_ : 1);
  yt_off_.reset(B + 1);
  ukey_.reset(NY_ + 1);
  ustart_.reset(NY_ + 1);
  u_off_.reset(B + 1);
  inst_.reset(B);
  const uint64_t nx1 = NX_ + 1;
  act_prefix_.reset(nx1);
  act_ypos_.reset(nx1);
  act_xval_.reset(nx1);
  act_row_.reset(nx1);
  row_key_.reset(nx1);
  row_prod_.reset(nx1);
  row_act_.reset(nx1);
  itab_.reset(3 * (B + 1));
  dup_.reset(B);
  bad_.reset(1);
  digest_.reset(2);
  scratch_.reset(4);
  vsum_.reset(1);
  tmp_.reset(scan_tmp_count(n1) + 1);
  slices_.assign(B, {});
  sptc_status st = run(s, true);
  if (st != SPTC_OK) return st;
  SPTC_HIP(hipStreamSynchronize(s));
  planned_ = true;
  info.instances = B;
  info.nnz_in = N_;
  info.plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                     .count();
  info.device_bytes = DevicePool::get().live_bytes() - live;
  return SPTC_OK;
}

// Accumulator per instance, with the plain engine's rule (Engine::contract):
// dense when the column ranks fit an LDS row (<= 2^13), Y has no duplicate
// coordinates and dense_pays (rows dense enough, Y segments long enough,
// few owners per row), or dense is forced.
sptc_status BatchEngine::plan_accumulators(hipStream_t s) {
  const uint32_t B = B_;
  std::vector<uint32_t> dup(B, 0);
  launch_bany_dup(k0_.get() + NX_, yoff_.get(), B, NY_, dup_.get(), s);
  readback(dup.data(), dup_.get(), B * 4, s);
  acc_.assign(B, 0);
  int nd = 0;
  for (uint32_t i = 0; i < B; ++i) {
    const uint64_t ir = iP_[i + 1] - iP_[i], R = iR_[i + 1] - iR_[i];
    const int kyf = kyf_h_[i];
    if (o_.dense_accumulator >= 0 && ir && kyf <= kDenseMaxBits && !dup[i]) {
      const uint64_t E = 1ull << (kyf < 8 ? 8 : kyf);
      const uint64_t A = iA_[i + 1] - iA_[i];
      acc_[i] = (o_.dense_accumulator == 1 || dense_pays(ir, R, A, E)) ? 1 : 0;
    }
    if (o_.dense_accumulator == 1 && ir && !acc_[i]) {
      msg = "instance " + std::to_string(i) +
            ": dense accumulator does not apply (column space above 2^13 or duplicate Y "
            "coordinates)";
      return SPTC_ERR_BAD_ARGS;
    }
    nd += acc_[i];
  }
  info.dense_instances = nd;
  // dense classes by slot width, instances in order
  for (int eb : {8, 10, 13}) {
    std::vector<uint64_t> rows{0}, row0;
    std::vector<uint32_t> inst;
    uint64_t ir = 0;
    for (uint32_t i = 0; i < B; ++i) {
      if (!acc_[i] || dense_class_bits(kyf_h_[i]) != eb) continue;
      const uint64_t R = iR_[i + 1] - iR_[i];
      if (!R) continue;
      rows.push_back(rows.back() + R);
      row0.push_back(iR_[i]);
      inst.push_back(i);
      ir += iP_[i + 1] - iP_[i];
    }
    if (inst.empty()) continue;
    DenseClass c;
    c.eb = eb;
    c.G = (uint32_t)inst.size();
    c.rows = rows.back();
    c.ir = ir;
    c.grp_rows.reset(rows.size());
    c.grp_row0.reset(row0.size());
    c.grp_inst.reset(inst.size());
    c.row_cnt.reset(c.rows);
    c.out_off.reset(c.rows + 1);
    SPTC_HIP(hipMemcpyAsync(c.grp_rows.get(), rows.data(), rows.size() * 8,
                            hipMemcpyHostToDevice, s));
    SPTC_HIP(hipMemcpyAsync(c.grp_row0.get(), row0.data(), row0.size() * 8,
                            hipMemcpyHostToDevice, s));
    SPTC_HIP(hipMemcpyAsync(c.grp_inst.get(), inst.data(), inst.size() * 4,
                            hipMemcpyHostToDevice, s));
    SPTC_HIP(hipStreamSynchronize(s));
    if (tmp_.size() < scan_tmp_count(c.rows) + 1) tmp_.reset(scan_tmp_count(c.rows) + 1);
    dense_.push_back(std::move(c));
  }
  info.dense_classes = (int32_t)dense_.size();
  return SPTC_OK;
}

// ESC chunks over the runs of consecutive ESC instances, each chunk holding
// at most `cap` products (from the memory budget, as in Engine::contract).
sptc_status BatchEngine::build_chunks(hipStream_t s) {
  const uint32_t B = B_;
  uint64_t ir_esc = 0;
  for (uint32_t i = 0; i < B; ++i)
    if (!acc_[i]) ir_esc += iP_[i + 1] - iP_[i];
  if (!ir_esc) return SPTC_OK;
  const uint64_t live = DevicePool::get().live_bytes();
  const uint64_t free_b = budget_ > live ? budget_ - live : 0;
  const uint64_t per_prod = 12 + SortWorkspace::bytes_per_key(10) + 1;
  const uint64_t out_per = 16;
  const uint64_t slack = 64ull << 20;
  uint64_t cap;
  if (free_b > slack + out_per * ir_esc + per_prod * std::min<uint64_t>(ir_esc, 1ull << 20))
    cap = (free_b - slack - out_per * ir_esc) / per_prod;
  else
    cap = free_b > slack ? (free_b - slack) / (per_prod + out_per) : 0;
  if (cap > (1ull << 26)) cap &= ~((1ull << 26) - 1);
  cap = std::min<uint64_t>(cap, std::min<uint64_t>(ir_esc, 4000000000ull));
  if (!cap) {
    msg = "batch products do not fit the memory budget";
    return SPTC_ERR_OOM_BUDGET;
  }
  EscChunk cur;
  auto close = [&]() {
    if (cur.n) esc_.push_back(std::move(cur));
    cur = EscChunk();
  };
  uint32_t i = 0;
  while (i < B) {
    if (acc_[i]) {
      ++i;
      continue;
    }
    uint32_t j = i;
    while (j < B && !acc_[j]) ++j;  // ESC run [i, j)
    uint64_t r = iR_[i], p = iP_[i], a = iA_[i];
    const uint64_t rend = iR_[j];
    while (r < rend) {
      const uint64_t rem = cap - cur.n;
      uint64_t r1, p1, a1;
      if (iP_[j] - p <= rem) {
        r1 = rend;
        p1 = iP_[j];
        a1 = iA_[j];
      } else {
        launch_chunk_end(row_prod_.get(), row_act_.get(), rend, r, rend - r, rem,
                         scratch_.get(), s);
        uint64_t h[3];
        readback(h, scratch_.get(), 24, s);
        r1 = h[0];
        p1 = h[1];
        a1 = h[2];
      }
      if (r1 == r) {
        if (!cur.n) {
          msg = "one X row's products exceed the memory budget";
          return SPTC_ERR_OOM_BUDGET;
        }
        close();
        continue;
      }
      BPiece pc;
      pc.ra = r;
      pc.lrow0 = cur.lrows;
      pc.A0 = a;
      pc.A1 = a1;
      const uint32_t pid = (uint32_t)cur.hp.size();
      cur.hp.push_back(pc);
      cur.hr.push_back(RowRange{r, r1 - r, cur.lrows, cur.n});
      for (uint64_t q = p; q < p1; q += kBExpandTileItems) {
        BETile t;
        t.p0 = q;
        t.out0 = cur.n + (q - p);
        t.cnt = (uint32_t)std::min<uint64_t>(kBExpandTileItems, p1 - q);
        t.piece = pid;
        cur.ht.push_back(t);
      }
      for (uint32_t k = i; k < j; ++k)  // instances whose rows meet [r, r1)
        if (iR_[k] < r1 && iR_[k + 1] > r) cur.K = std::max(cur.K, kyf_h_[k]);
      cur.n += p1 - p;
      cur.lrows += r1 - r;
      r = r1;
      p = p1;
      a = a1;
      if (cur.n == cap) close();
    }
    i = j;
  }
  close();
  uint64_t chunk_cap = 0, max_lrows = 0;
  for (const EscChunk& c : esc_) max_lrows = std::max(max_lrows, c.lrows);
  if (tmp_.size() < scan_tmp_count(max_lrows) + 1) tmp_.reset(scan_tmp_count(max_lrows) + 1);
  for (EscChunk& c : esc_) {
    c.bits = mode_bits(c.lrows) + c.K;
    c.rb = o_.radix_bits ? o_.radix_bits : choose_radix_bits(c.bits);
    if (c.K > 0 && c.K < c.bits && c.K <= 32 && prefer_row_sort()) {
      DevBuf<RowRange> rr(c.hr.size());
      SPTC_HIP(hipMemcpyAsync(rr.get(), c.hr.data(), c.hr.size() * sizeof(RowRange),
                              hipMemcpyHostToDevice, s));
      build_row_segments(rr.get(), (uint32_t)c.hr.size(), c.lrows, row_prod_.get(), c.K,
                         &c.rseg, ws_, tmp_.get(), s, /*msd=*/false, row_act_.get());
      c.rowsort = true;
      // keep the per-row big-segment flags the expansion reads and the row
      // starts of the narrow form, release the rest
      c.row_big.reset(c.lrows);
      SPTC_HIP(hipMemcpyAsync(c.row_big.get(), c.rseg.row_big(), c.lrows * 8,
                              hipMemcpyDeviceToDevice, s));
      c.row_start.reset(c.lrows + 1);
      SPTC_HIP(hipMemcpyAsync(c.row_start.get(), c.rseg.row_start(), (c.lrows + 1) * 8,
                              hipMemcpyDeviceToDevice, s));
      c.tile_row.reset(2 * (narrow_tile_count(c.n) + 1));
      launch_tile_rows(c.row_start.get(), c.lrows, c.n, c.tile_row.get(), s);
      SPTC_HIP(hipStreamSynchronize(s));  // rr and the scratch are released below
      c.rseg.scratch.release();
    }
    c.pieces.reset(c.hp.size());
    c.tiles.reset(c.ht.size());
    SPTC_HIP(hipMemcpyAsync(c.pieces.get(), c.hp.data(), c.hp.size() * sizeof(BPiece),
                            hipMemcpyHostToDevice, s));
    SPTC_HIP(hipMemcpyAsync(c.tiles.get(), c.ht.data(), c.ht.size() * sizeof(BETile),
                            hipMemcpyHostToDevice, s));
    SPTC_HIP(hipStreamSynchronize(s));
    chunk_cap = std::max(chunk_cap, c.n);
    ws_.ensure(c.n, c.rb, s);
  }
  kA_.reset(chunk_cap);
  vA_.reset(chunk_cap);
  const uint64_t ntmp = std::max(reduce_tmp_count(chunk_cap), narrow_tmp_count(chunk_cap));
  if (tmp_.size() < ntmp) tmp_.reset(ntmp);
  info.esc_chunks = (int32_t)esc_.size();
  return SPTC_OK;
}

sptc_status BatchEngine::run(hipStream_t s, bool P) {
  const uint32_t B = B_;
  KTimer* kt = (P && kt_.on) ? &kt_ : nullptr;
  fill_u32(bad_.get(), 0, 1, s);

  // ---- (a)+(b) linearize and sort every operand (segmented) ---------------
  {
    uint64_t* kl = lin_alt_ ? k1_.get() : k0_.get();
    float* vl = lin_alt_ ? v1_.get() : v0_.get();
    uint64_t* ko = lin_alt_ ? k0_.get() : k1_.get();
    float* vo = lin_alt_ ? v0_.get() : v1_.get();
    int tok = kt ? kt->begin(s) : -1;
    launch_blinearize(lin_.get(), seg_off_.get(), 2 * B, N_, kl, vl, s);
    if (kt) kt->end(tok, SPTC_K_LINEARIZE, (double)N_ * 28, s);
    if (!presorted_)
      segmented_sort(ws_, plan_all_, kl, vl, ko, vo, nbits_all_, s, kt, SPTC_K_SORT_INPUT);
  }
  fill_u64(digest_.get(), 0, 1, s);
  launch_digest(k0_.get(), N_, digest_.get(), s);
  if (P)
    SPTC_HIP(hipMemcpyAsync(digest_.get() + 1, digest_.get(), 8, hipMemcpyDeviceToDevice, s));
  else
    launch_check_digest(digest_.get(), digest_.get() + 1, bad_.get(), s);

  // ---- Y free keys -> per-instance ranks ------------------------------------
  uint64_t* yk = k0_.get() + NX_;
  const float* yv = v0_.get() + NX_;
  launch_bfree_extract(yk, yoff_.get(), B, fmask_.get(), NY_, fk_.get(), s);
  const bool falt =
      segmented_sort(ws_, plan_y_, fk_.get(), fv_.get(), k1_.get(), v1_.get(), ynf_max_, s);
  fill_u64(yt_off_.get(), 0, B + 1, s);
  run_bunique(falt ? k1_.get() : fk_.get(), yoff_.get(), B, NY_, ytab_.get(), yt_off_.get(),
              tmp_.get(), s);
  launch_fix_empty(yt_off_.get(), yoff_.get(), B, s);
  if (ynf_max_ <= 32) launch_ycol(ytab_.get(), NY_, 0xFFFFFFFFull, ytab32_.get(), s);
  if (P) {
    ytoff_h_.assign(B + 1, 0);
    readback(ytoff_h_.data(), yt_off_.get(), (B + 1) * 8, s);
    kyf_h_.assign(B, 0);
    std::vector<BInst> bi(B);
    for (uint32_t i = 0; i < B; ++i) {
      kyf_h_[i] = mode_bits(ytoff_h_[i + 1] - ytoff_h_[i]);
      bi[i].cmask = low_mask(xl_[i].ncbits);
      bi[i].ncbits = xl_[i].ncbits;
      bi[i].kyf = kyf_h_[i];
      bi[i].yt_off = ytoff_h_[i];
    }
    SPTC_HIP(hipMemcpyAsync(kyf_.get(), kyf_h_.data(), B * sizeof(int), hipMemcpyHostToDevice,
                            s));
    SPTC_HIP(hipMemcpyAsync(inst_.get(), bi.data(), B * sizeof(BInst), hipMemcpyHostToDevice,
                            s));
    SPTC_HIP(hipStreamSynchronize(s));
  }
  launch_brank(yk, yoff_.get(), B, ynf_.get(), kyf_.get(), ytab_.get(), yt_off_.get(), NY_, yk,
               ycol_.get(), s);

  // ---- (c) contract tables + index match --------------------------------------
  fill_u64(u_off_.get(), 0, B + 1, s);
  run_bycontract(yk, yoff_.get(), B, kyf_.get(), NY_, ukey_.get(), ustart_.get(), u_off_.get(),
                 tmp_.get(), s);
  launch_fix_empty(u_off_.get(), yoff_.get(), B, s);
  uint64_t* iP = itab_.get();
  uint64_t* iA = iP + (B + 1);
  uint64_t* iR = iA + (B + 1);
  fill_u64(itab_.get(), 0, 3 * (B + 1), s);
  BMatchOut mo;
  mo.act_prefix = act_prefix_.get();
  mo.act_ypos = act_ypos_.get();
  mo.act_xval = act_xval_.get();
  mo.act_row = act_row_.get();
  mo.row_key = row_key_.get();
  mo.row_prod = row_prod_.get();
  mo.row_act = row_act_.get();
  mo.inst_P = iP;
  mo.inst_A = iA;
  mo.inst_R = iR;
  mo.bad = bad_.get();
  {
    int tok = kt ? kt->begin(s) : -1;
    run_bmatch(k0_.get(), v0_.get(), seg_off_.get(), B, inst_.get(), ukey_.get(), ustart_.get(),
               u_off_.get(), NX_, mo, tmp_.get(), s);
    if (kt) kt->end(tok, SPTC_K_MATCH, (double)NX_ * (16 + 24 + 24), s);
  }
  launch_fix_empty(iP, seg_off_.get(), B, s);
  launch_fix_empty(iA, seg_off_.get(), B, s);
  launch_fix_empty(iR, seg_off_.get(), B, s);
  if (P) {
    std::vector<uint64_t> h(3 * (B + 1));
    readback(h.data(), itab_.get(), h.size() * 8, s);
    iP_.assign(h.begin(), h.begin() + (B + 1));
    iA_.assign(h.begin() + (B + 1), h.begin() + 2 * (B + 1));
    iR_.assign(h.begin() + 2 * (B + 1), h.end());
    info.ir_nnz = iP_[B];
    sptc_status st = plan_accumulators(s);
    if (st != SPTC_OK) return st;
  }

  // ---- (d) dense classes -------------------------------------------------------
  for (DenseClass& c : dense_) {
    BDenseArgs da;
    std::memset(&da, 0, sizeof(da));
    da.d.act_prefix = act_prefix_.get();
    da.d.act_ypos = act_ypos_.get();
    da.d.act_xval = act_xval_.get();
    da.d.row_act = row_act_.get();
    da.d.ykeys = yk;
    da.d.yvals = yv;
    da.d.row_cnt = c.row_cnt.get();
    da.d.out_off = c.out_off.get();
    da.d.vsum = vsum_.get();
    da.grp_rows = c.grp_rows.get();
    da.grp_row0 = c.grp_row0.get();
    da.grp_inst = c.grp_inst.get();
    da.inst = inst_.get();
    da.ytab = ytab_.get();
    da.G = c.G;
    da.bad = bad_.get();
    int tok = kt ? kt->begin(s) : -1;
    launch_bdense_count(da, c.eb, c.rows, s);
    if (kt) kt->end(tok, SPTC_K_DENSE, 8.0 * (double)c.ir, s);
    run_row_count_scan(c.row_cnt.get(), c.out_off.get(), c.rows, tmp_.get(), s);
    if (P) {
      U2 t;
      readback(&t, tmp_.get() + scan_tmp_count(c.rows), sizeof(U2), s);
      if (DevicePool::get().live_bytes() + t.a * 16 > budget_) {
        msg = "batch output does not fit the memory budget";
        return SPTC_ERR_OOM_BUDGET;
      }
      BOut ob;
      ob.nnz = t.a;
      ob.row.reset(ob.nnz);
      ob.col.reset(ob.nnz);
      ob.val.reset(ob.nnz);
      c.out = (int)outs_.size();
      outs_.push_back(std::move(ob));
    }
    BOut& ob = outs_[c.out];
    da.d.out_row = ob.row.get();
    da.d.out_col = ob.col.get();
    da.d.out_val = ob.val.get();
    tok = kt ? kt->begin(s) : -1;
    launch_bdense_numeric(da, c.eb, c.rows, s);
    if (kt) kt->end(tok, SPTC_K_DENSE, 12.0 * (double)c.ir + 16.0 * (double)ob.nnz, s);
    if (P) {
      ob.bounds.reset(B + 1);
      launch_dense_bounds(c.grp_rows.get(), c.grp_inst.get(), c.G, c.out_off.get(), ob.nnz, B,
                          ob.bounds.get(), bad_.get(), s);
    }
  }

  // ---- (d)+(e) ESC chunks ------------------------------------------------------
  if (P) {
    sptc_status st = build_chunks(s);
    if (st != SPTC_OK) return st;
  }
  for (EscChunk& c : esc_) {
    BExpandArgs ea;
    std::memset(&ea, 0, sizeof(ea));
    ea.act_prefix = act_prefix_.get();
    ea.act_ypos = act_ypos_.get();
    ea.act_xval = act_xval_.get();
    ea.act_row = act_row_.get();
    ea.ycol = ycol_.get();
    ea.yvals = yv;
    ea.pieces = c.pieces.get();
    ea.tiles = c.tiles.get();
    ea.ntiles = (uint32_t)c.ht.size();
    ea.K = c.K;
    ea.out_keys = kA_.get();
    ea.out_vals = vA_.get();
    if (c.rowsort) {
      const NarrowTarget nt =
          narrow_target(c.rseg, c.n, vA_.get(), ws_.alt_keys.get(), ws_.alt_vals.get());
      ea.row_big = c.row_big.get();
      ea.big_col = nt.col;
      ea.big_val = nt.val;
      ea.out_col32 = reinterpret_cast<uint32_t*>(kA_.get());
    }
    ea.bad = bad_.get();
    int tok = kt ? kt->begin(s) : -1;
    launch_bexpand(ea, s);
    if (kt) kt->end(tok, SPTC_K_EXPAND, (c.rowsort ? 16.0 : 24.0) * (double)c.n, s);
    bool alt = false;
    if (c.rowsort) {
      sort_row_segments(ws_, c.rseg, c.n, reinterpret_cast<uint32_t*>(kA_.get()), vA_.get(),
                        ws_.alt_keys.get(), ws_.alt_vals.get(), c.row_start.get(), s, kt,
                        SPTC_K_SORT_PROD);
    } else {
      ws_.ensure(c.n, c.rb, s);
      const int passes = radix_passes(c.bits, c.rb);
      ws_.prepare(passes, s);
      if (sort_needs_histogram()) histogram(ws_, kA_.get(), c.n, passes, s);
      alt = sort_after_histogram(ws_, kA_.get(), vA_.get(), c.n, c.bits, s, kt, SPTC_K_SORT_PROD);
    }
    const uint64_t* sk = alt ? ws_.alt_keys.get() : kA_.get();
    const float* sv = alt ? ws_.alt_vals.get() : vA_.get();
    const NarrowRows nr{reinterpret_cast<const uint32_t*>(kA_.get()), c.row_start.get(),
                        c.lrows, c.tile_row.get()};
    tok = kt ? kt->begin(s) : -1;
    if (c.rowsort) run_reduce_count_narrow(nr, c.n, tmp_.get(), s);
    else run_reduce_count(sk, c.n, tmp_.get(), s);
    if (kt) kt->end(tok, SPTC_K_REDUCE, (c.rowsort ? 4.0 : 8.0) * (double)c.n, s);
    if (P) {
      U2 t;
      readback(&t, tmp_.get() + (c.rowsort ? narrow_tile_count(c.n) : scan_tmp_count(c.n)),
               sizeof(U2), s);
      if (DevicePool::get().live_bytes() + t.a * 16 > budget_) {
        msg = "batch output does not fit the memory budget";
        return SPTC_ERR_OOM_BUDGET;
      }
      BOut ob;
      ob.nnz = t.a;
      ob.row.reset(ob.nnz);
      ob.col.reset(ob.nnz);
      ob.val.reset(ob.nnz);
      c.out = (int)outs_.size();
      outs_.push_back(std::move(ob));
    }
    BOut& ob = outs_[c.out];
    if (P) {
      // chunk-local row -> (global row, instance), and the per-instance
      // output starts that the emit below fills in
      c.map_row.reset(c.lrows);
      c.map_inst.reset(c.lrows);
      launch_chunk_row_maps(c.pieces.get(), (uint32_t)c.hp.size(), c.lrows, iR, B,
                            c.map_row.get(), c.map_inst.get(), s);
      ob.bounds.reset(B + 1);
      fill_u64(ob.bounds.get(), ob.nnz, B + 1, s);
    }
    BReduceArgs rb;
    rb.map_row = c.map_row.get();
    rb.map_inst = c.map_inst.get();
    rb.bounds = P ? ob.bounds.get() : nullptr;
    rb.inst_R = iR;
    rb.inst = inst_.get();
    rb.ytab = ytab_.get();
    rb.ytab32 = ytab32_.size() ? ytab32_.get() : nullptr;
    rb.B = B;
    rb.bad = bad_.get();
    tok = kt ? kt->begin(s) : -1;
    if (c.rowsort)
      run_breduce_emit_narrow(nr, vA_.get(), c.n, ob.row.get(), ob.col.get(), ob.val.get(), rb,
                              tmp_.get(), s);
    else
      run_breduce_emit(sk, sv, c.n, ob.row.get(), ob.col.get(), ob.val.get(), c.K, rb,
                       tmp_.get(), s);
    if (kt)
      kt->end(tok, SPTC_K_REDUCE, (c.rowsort ? 8.0 : 12.0) * (double)c.n + 16.0 * (double)ob.nnz,
              s);
  }

  // ---- per-instance slices (rows are already instance-local) ---------------------
  if (P) {
    for (BOut& ob : outs_) {
      ob.hb.assign(B + 1, 0);
      readback(ob.hb.data(), ob.bounds.get(), (B + 1) * 8, s);
    }
    uint64_t out_nnz = 0;
    auto add_slice = [&](uint32_t i, int k, uint64_t r0, uint64_t r1) {
      const BOut& ob = outs_[k];
      InstSlice sl{k, ob.hb[i], ob.hb[i + 1], r0, r1};
      if (sl.e > sl.b) slices_[i].push_back(sl);
      out_nnz += sl.e - sl.b;
    };
    for (const DenseClass& c : dense_) {
      for (uint32_t i = 0; i < B; ++i)
        if (acc_[i] && dense_class_bits(kyf_h_[i]) == c.eb) add_slice(i, c.out, 0, iR_[i + 1] - iR_[i]);
    }
    for (const EscChunk& c : esc_) {
      // rows of instance i covered by the chunk's pieces
      std::vector<uint64_t> lo(B, ~0ull), hi(B, 0);
      for (size_t q = 0; q < c.hp.size(); ++q) {
        const uint64_t ra = c.hp[q].ra;
        const uint64_t rb = ra + ((q + 1 < c.hp.size() ? c.hp[q + 1].lrow0 : c.lrows) -
                                  c.hp[q].lrow0);
        for (uint32_t i = 0; i < B; ++i) {
          if (iR_[i] >= rb || iR_[i + 1] <= ra) continue;
          lo[i] = std::min(lo[i], std::max(ra, iR_[i]) - iR_[i]);
          hi[i] = std::max(hi[i], std::min(rb, iR_[i + 1]) - iR_[i]);
        }
      }
      for (uint32_t i = 0; i < B; ++i)
        if (lo[i] != ~0ull) add_slice(i, c.out, lo[i], hi[i]);
    }
    // slices of an instance in row order: dense ones come first but an
    // instance is either dense or ESC, and ESC chunks are in row order
    info.out_nnz = out_nnz;
  }
  return SPTC_OK;
}

sptc_status BatchEngine::capture() {
  if (exec_) return SPTC_OK;
  if (!cap_stream_) SPTC_HIP(hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking));
  SPTC_HIP(hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeThreadLocal));
  run(cap_stream_, false);
  SPTC_HIP(hipStreamEndCapture(cap_stream_, &graph_));
  SPTC_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
  size_t n = 0;
  SPTC_HIP(hipGraphGetNodes(graph_, nullptr, &n));
  info.graph_nodes = n;
  return SPTC_OK;
}

sptc_status BatchEngine::launch_graph(hipStream_t s) {
  sptc_status st = capture();
  if (st != SPTC_OK) return st;
  SPTC_HIP(hipGraphLaunch(exec_, s));
  return SPTC_OK;
}

sptc_status BatchEngine::sync(hipStream_t s) {
  uint32_t bad = 0;
  SPTC_HIP(hipMemcpyAsync(&bad, bad_.get(), 4, hipMemcpyDeviceToHost, s));
  SPTC_HIP(hipStreamSynchronize(s));
  if (bad) {
    msg = "operand index arrays changed since the batch was planned";
    return SPTC_ERR_INTERNAL;
  }
  return SPTC_OK;
}

sptc_status BatchEngine::instance_result(int i, sptc_result_s* r, hipStream_t s) {
  const Layout& xl = xl_[i];
  const Layout& yl = yl_[i];
  sptc_result_info& in = r->info;
  std::memset(&in, 0, sizeof(in));
  in.x_nfbits = xl.nfbits;
  in.y_nfbits = yl.nfbits;
  in.ncbits = xl.ncbits;
  in.out_bits = xl.nfbits + yl.nfbits;
  in.out_order = xl.nf + yl.nf;
  in.ir_nnz = iP_[i + 1] - iP_[i];
  in.nrows = iR_[i + 1] - iR_[i];
  in.accumulator = acc_[i];
  r->xl = xl;
  r->yl = yl;
  r->vrow_key = row_key_.get() + iR_[i];
  in.row_key = r->vrow_key;
  DevBuf<double> sum(1);
  fill_u64(reinterpret_cast<uint64_t*>(sum.get()), 0, 1, s);
  for (const InstSlice& sl : slices_[i]) {
    const BOut& ob = outs_[sl.out];
    OutChunk c;
    c.nnz = sl.e - sl.b;
    c.r0 = sl.r0;
    c.r1 = sl.r1;
    c.vrow = ob.row.get() + sl.b;
    c.vcol = ob.col.get() + sl.b;
    c.vval = ob.val.get() + sl.b;
    launch_sum_f64(c.vval, c.nnz, sum.get(), s);
    in.nnz += c.nnz;
    r->chunks.push_back(std::move(c));
  }
  in.nchunks = (int32_t)r->chunks.size();
  SPTC_HIP(hipMemcpyAsync(&in.value_sum, sum.get(), 8, hipMemcpyDeviceToHost, s));
  SPTC_HIP(hipStreamSynchronize(s));
  if (r->chunks.size() == 1) {
    in.row = r->chunks[0].vrow;
    in.col = r->chunks[0].vcol;
    in.val = r->chunks[0].vval;
  }
  return SPTC_OK;
}

}  // namespace sptc
