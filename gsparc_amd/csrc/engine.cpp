// SpTC engine host orchestration (built with hipcc; launches live in .hip).
#include "engine.hpp"

#include "bucket.hpp"

#include "tdense.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace sptc {

int mode_bits(uint64_t d) {
  // 64 - clz(d-1) (tensor_manager.tpp:109); a dim of 1 needs no bits.
  return d <= 1 ? 0 : 64 - __builtin_clzll(d - 1);
}

// Free modes ascending with the last one at bit 0; contracted modes in the
// caller's order with the last one at bit 0 (tensor_manager.tpp:153-174).
// X (role 1): key = f << ncbits | c; Y (role 2): key = c << nfbits | f.
sptc_status make_layout(int order, const uint64_t* dims, int nc, const int32_t* cmodes,
                        const int32_t* cwidth, int role, Layout* L) {
  if (order < 1 || order > SPTC_MAX_ORDER || nc < 0 || nc > order) return SPTC_ERR_BAD_ARGS;
  *L = Layout();
  L->order = order;
  L->role = role;
  bool is_c[SPTC_MAX_ORDER] = {false};
  for (int i = 0; i < nc; ++i) {
    if (cmodes[i] < 0 || cmodes[i] >= order || is_c[cmodes[i]]) return SPTC_ERR_BAD_ARGS;
    is_c[cmodes[i]] = true;
    L->cmodes[i] = cmodes[i];
  }
  L->nc = nc;
  for (int m = 0; m < order; ++m) {
    if (dims[m] == 0) return SPTC_ERR_BAD_ARGS;
    L->dims[m] = dims[m];
    L->width[m] = mode_bits(dims[m]);
    if (!is_c[m]) L->fmodes[L->nf++] = m;
  }
  for (int i = 0; i < nc; ++i)
    if (cwidth && cwidth[i] >= 0) L->width[cmodes[i]] = cwidth[i];
  int s = 0;
  for (int i = L->nf - 1; i >= 0; --i) {
    L->shift[L->fmodes[i]] = s;
    s += L->width[L->fmodes[i]];
  }
  L->nfbits = s;
  s = 0;
  for (int i = nc - 1; i >= 0; --i) {
    L->shift[cmodes[i]] = s;
    s += L->width[cmodes[i]];
  }
  L->ncbits = s;
  L->nbits = L->nfbits + L->ncbits;
  for (int m = 0; m < order; ++m) {
    if (is_c[m] && role == 2) L->shift[m] += L->nfbits;
    if (!is_c[m] && role == 1) L->shift[m] += L->ncbits;
  }
  return L->nbits > 64 ? SPTC_ERR_KEY_OVERFLOW : SPTC_OK;
}

sptc_status make_pair_layouts(int xorder, const uint64_t* xdims, int yorder,
                              const uint64_t* ydims, int nc, const int32_t* cx,
                              const int32_t* cy, Layout* xl, Layout* yl, std::string* msg) {
  if (nc < 0 || nc > xorder || nc > yorder || (nc && (!cx || !cy))) {
    *msg = "bad contracted-mode count";
    return SPTC_ERR_BAD_ARGS;
  }
  int32_t cw[SPTC_MAX_ORDER];
  for (int i = 0; i < nc; ++i) {
    if (cx[i] < 0 || cx[i] >= xorder || cy[i] < 0 || cy[i] >= yorder) {
      *msg = "contracted mode out of range";
      return SPTC_ERR_BAD_ARGS;
    }
    cw[i] = std::max(mode_bits(xdims[cx[i]]), mode_bits(ydims[cy[i]]));
  }
  sptc_status st = make_layout(xorder, xdims, nc, cx, cw, 1, xl);
  if (st == SPTC_ERR_BAD_ARGS) {
    *msg = "bad X modes";
    return st;
  }
  st = make_layout(yorder, ydims, nc, cy, cw, 2, yl);
  if (st == SPTC_ERR_BAD_ARGS) {
    *msg = "bad Y modes";
    return st;
  }
  if (xl->nbits > 64 || yl->nbits > 64) {
    // keys above 64 bits: the contract field becomes a rank of the
    // contracted coordinates (the free fields must fit 64 bits each)
    if (xl->nfbits > 64 || yl->nfbits > 64) {
      *msg = "a free-mode field needs more than 64 bits";
      return SPTC_ERR_KEY_OVERFLOW;
    }
    for (int i = 0; i < nc; ++i)
      if (cw[i] > 64) {
        *msg = "a contracted mode needs more than 64 bits";
        return SPTC_ERR_KEY_OVERFLOW;
      }
    xl->crank = yl->crank = 1;
  }
  return SPTC_OK;
}

Layout rank_layout(const Layout& L, int rbits) {
  Layout R = L;
  R.crank = 1;
  R.ncbits = rbits;
  R.nbits = L.nfbits + rbits;
  if (L.role == 1)
    for (int i = 0; i < L.nf; ++i) R.shift[L.fmodes[i]] = L.shift[L.fmodes[i]] - L.ncbits + rbits;
  for (int i = 0; i < L.nc; ++i) R.shift[L.cmodes[i]] = 0;  // not linearized directly
  return R;
}

bool prefer_row_sort() {
  static const bool rows = [] {
    const char* e = getenv("SPTC_PROD_SORT");
    return !(e && std::string(e) == "full");
  }();
  return rows;
}

bool prefer_bucket_tail() {
  static const bool msd = [] {
    const char* e = getenv("SPTC_BIG_TAIL");
    return e && std::string(e) == "buckets";
  }();
  return msd;
}

uint64_t count_products(const Operand& X, const Operand& Y, hipStream_t s) {
  const uint64_t nx = X.nnz, ny = Y.nnz;
  if (!nx || !ny) return 0;
  const int ncb = X.lay.ncbits, ynf = Y.lay.nfbits;
  DevBuf<uint64_t> tlo, thi, ukey, ustart;
  YTable tab;
  std::memset(&tab, 0, sizeof(tab));
  U2 host;
  if (ncb <= kDirectBits) {
    const uint64_t K = 1ull << ncb;
    tlo.reset(K);
    thi.reset(K);
    fill_u64(tlo.get(), 0, K, s);
    fill_u64(thi.get(), 0, K, s);
    launch_ytable_direct(Y.keys.get(), ny, ynf, K, tlo.get(), thi.get(), s);
    tab.direct = 1;
    tab.lo = tlo.get();
    tab.hi = thi.get();
  } else {
    ukey.reset(ny + 1);
    ustart.reset(ny + 1);
    DevBuf<U2> tmp(scan_tmp_count(ny) + 1);
    run_yunique(Y.keys.get(), ny, ynf, ukey.get(), ustart.get(), tmp.get(), s);
    SPTC_HIP(hipMemcpyAsync(&host, tmp.get() + scan_tmp_count(ny), sizeof(U2),
                            hipMemcpyDeviceToHost, s));
    SPTC_HIP(hipStreamSynchronize(s));
    tab.direct = 0;
    tab.ukey = ukey.get();
    tab.ustart = ustart.get();
    tab.U = host.a;
  }
  DevBuf<U2> tmp(scan_tmp_count(nx) + 1);
  run_match_count(X.keys.get(), nx, ncb, tab, tmp.get(), s);
  SPTC_HIP(hipMemcpyAsync(&host, tmp.get() + scan_tmp_count(nx), sizeof(U2),
                          hipMemcpyDeviceToHost, s));
  SPTC_HIP(hipStreamSynchronize(s));
  return host.a;
}

Engine::Engine(hipStream_t s, const sptc_opts& o) : s_(s), o_(o) {
  for (auto& e : ev_) SPTC_HIP(hipEventCreate(&e));
  SPTC_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_total_), sizeof(U2) * 2, 0));
  SPTC_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_u64_), sizeof(uint64_t) * 8, 0));
  // The budget caps what this call holds at once, on top of what was
  // already live when it started.
  const uint64_t live = DevicePool::get().live_bytes();
  live0_ = live;
  if (o.mem_budget_bytes) {
    budget_ = live + o.mem_budget_bytes;
  } else {
    const uint64_t fr = device_free_bytes();
    const uint64_t margin = 1ull << 30;
    budget_ = live + (fr > margin ? fr - margin : 0);
  }
  rb_ = radix_bits_supported(o.radix_bits) ? o.radix_bits : 0;  // 0: per sort
  kt_.on = o.kernel_timing != 0;
}

Engine::~Engine() {
  for (auto& e : ev_) (void)hipEventDestroy(e);
  (void)hipHostFree(host_total_);
  (void)hipHostFree(host_u64_);
}

void Engine::record(int i) {
  SPTC_HIP(hipEventRecord(ev_[i], s_));
  ev_used_[i] = true;
}

double Engine::elapsed_ms(int a, int b) {
  if (!ev_used_[a] || !ev_used_[b]) return 0.0;
  SPTC_HIP(hipEventSynchronize(ev_[b]));
  float t = 0.f;
  SPTC_HIP(hipEventElapsedTime(&t, ev_[a], ev_[b]));
  return t;
}

U2 Engine::read_total(const U2* dev) {
  SPTC_HIP(hipMemcpyAsync(host_total_, dev, sizeof(U2), hipMemcpyDeviceToHost, s_));
  SPTC_HIP(hipStreamSynchronize(s_));
  return host_total_[0];
}

void Engine::stage(const sptc_coo& t, StagedCOO* out) {
  const uint64_t n = t.nnz;
  out->nnz = n;
  out->order = t.order;
  out->idx_bytes = t.idx_bytes;
  out->own.clear();
  if (t.on_device || n == 0) {
    for (int m = 0; m < t.order; ++m) out->idx[m] = t.idx[m];
    out->val = t.val;
    return;
  }
  for (int m = 0; m < t.order; ++m) {
    out->own.emplace_back(n * t.idx_bytes);
    SPTC_HIP(hipMemcpyAsync(out->own.back().get(), t.idx[m], n * t.idx_bytes,
                            hipMemcpyHostToDevice, s_));
    out->idx[m] = out->own.back().get();
  }
  out->own.emplace_back(n * sizeof(float));
  SPTC_HIP(hipMemcpyAsync(out->own.back().get(), t.val, n * sizeof(float),
                          hipMemcpyHostToDevice, s_));
  out->val = reinterpret_cast<const float*>(out->own.back().get());
}

void Engine::prepare(const StagedCOO& t, const Layout& lay, Operand* out, const void* crank) {
  const uint64_t n = t.nnz;
  out->lay = lay;
  out->nnz = n;
  out->keys.reset(n ? n : 1);
  out->vals.reset(n ? n : 1);
  if (!n) return;
  LinArgs a;
  std::memset(&a, 0, sizeof(a));
  if (lay.crank) {
    // free modes, then the contract rank as one field
    a.order = lay.nf + 1;
    for (int j = 0; j < lay.nf; ++j) {
      const int m = lay.fmodes[j];
      a.shift[j] = lay.shift[m];
      a.width[j] = lay.width[m];
      a.idx[j] = t.idx[m];
    }
    a.shift[lay.nf] = lay.role == 1 ? 0 : lay.nfbits;
    a.width[lay.nf] = lay.ncbits;
    a.idx[lay.nf] = crank;
  } else {
    a.order = lay.order;
    for (int m = 0; m < lay.order; ++m) {
      a.shift[m] = lay.shift[m];
      a.width[m] = lay.width[m];
      a.idx[m] = t.idx[m];
    }
  }
  SPTC_HIP(hipMemcpyAsync(out->vals.get(), t.val, n * sizeof(float),
                          hipMemcpyDeviceToDevice, s_));
  const int rb = rb_ ? rb_ : choose_radix_bits(lay.nbits);
  const int passes = radix_passes(lay.nbits, rb);
  ws_.ensure(n, rb, s_);
  ws_.prepare(passes, s_);
  int tok = kt_.begin(s_);
  const bool hist = sort_needs_histogram();
  launch_linearize(a, t.idx_bytes, n, out->keys.get(), ws_.hist.get(),
                   hist && passes <= kMaxFusedPasses ? passes : 0, rb, s_);
  kt_.end(tok, SPTC_K_LINEARIZE, (double)n * (lay.order * t.idx_bytes + 8), s_);
  if (hist && passes > kMaxFusedPasses) histogram(ws_, out->keys.get(), n, passes, s_);
  if (sort_after_histogram(ws_, out->keys.get(), out->vals.get(), n, lay.nbits, s_, &kt_,
                           SPTC_K_SORT_INPUT)) {
    std::swap(out->keys, ws_.alt_keys);
    std::swap(out->vals, ws_.alt_vals);
  }
}

sptc_status Engine::check_indices(const StagedCOO& t, const uint64_t* dims, const char* name) {
  if (!t.nnz) return SPTC_OK;
  DevBuf<uint64_t> mx(t.order);
  fill_u64(mx.get(), 0, t.order, s_);
  for (int m = 0; m < t.order; ++m) launch_mode_max(t.idx[m], t.idx_bytes, t.nnz, mx.get() + m, s_);
  std::vector<uint64_t> h(t.order);
  SPTC_HIP(hipMemcpyAsync(h.data(), mx.get(), t.order * 8, hipMemcpyDeviceToHost, s_));
  SPTC_HIP(hipStreamSynchronize(s_));
  for (int m = 0; m < t.order; ++m)
    if (h[m] >= dims[m]) {
      msg = std::string(name) + ": mode " + std::to_string(m) + " has index " +
            std::to_string(h[m]) + ", not below its dim " + std::to_string(dims[m]);
      return SPTC_ERR_BAD_ARGS;
    }
  return SPTC_OK;
}

uint64_t Engine::contract_ranks(const StagedCOO& sx, const StagedCOO& sy, const Layout& xl,
                                const Layout& yl, DevBuf<uint64_t>* rank) {
  const uint64_t nx = sx.nnz, ny = sy.nnz, n = nx + ny;
  rank->reset(n ? n : 1);
  if (!n) return 0;
  fill_u64(rank->get(), 0, n, s_);
  const int nc = xl.nc;
  if (!nc) return 1;
  DevBuf<uint64_t> key(n), sorted(n);
  DevBuf<float> payload(n);  // moved with the keys, never read
  DevBuf<uint64_t> uniq(n);
  DevBuf<U2> tmp(scan_tmp_count(n) + 1);
  uint64_t U = 1;
  int rbits = 0;
  bool prev = false;
  for (int pos = 0; pos < nc;) {
    // the next group of contracted modes that fits 64 bits above the rank
    int W = 0, end = pos;
    while (end < nc && rbits + W + xl.width[xl.cmodes[end]] <= 64) W += xl.width[xl.cmodes[end++]];
    if (end == pos) {
      msg = "contracted coordinates do not fit a 64-bit rank key";
      return 0;
    }
    for (int side = 0; side < 2; ++side) {
      const Layout& L = side ? yl : xl;
      const StagedCOO& t = side ? sy : sx;
      TupleArgs ta;
      std::memset(&ta, 0, sizeof(ta));
      int sh = W;
      for (int j = pos; j < end; ++j) {
        const int m = L.cmodes[j];
        sh -= L.width[m];
        ta.idx[j - pos] = t.idx[m];
        ta.width[j - pos] = L.width[m];
        ta.shift[j - pos] = sh;
      }
      ta.n = end - pos;
      ta.bytes = t.idx_bytes;
      ta.prev = prev ? rank->get() + (side ? nx : 0) : nullptr;
      ta.prev_shift = W;
      launch_tuple_key(ta, side ? ny : nx, key.get() + (side ? nx : 0), s_);
    }
    const int bits = rbits + W;
    SPTC_HIP(hipMemcpyAsync(sorted.get(), key.get(), n * 8, hipMemcpyDeviceToDevice, s_));
    const int rb = rb_ ? rb_ : choose_radix_bits(bits);
    ws_.ensure(n, rb, s_);
    ws_.prepare(radix_passes(bits, rb), s_);
    if (sort_needs_histogram()) histogram(ws_, sorted.get(), n, radix_passes(bits, rb), s_);
    const bool alt = sort_after_histogram(ws_, sorted.get(), payload.get(), n, bits, s_);
    run_unique(alt ? ws_.alt_keys.get() : sorted.get(), n, uniq.get(), tmp.get(), s_);
    U = read_total(tmp.get() + scan_tmp_count(n)).a;
    launch_yrank(key.get(), n, 64, ~0ull, uniq.get(), U, 0, rank->get(), s_);
    rbits = mode_bits(U);
    prev = true;
    pos = end;
  }
  return U;
}

sptc_status Engine::prepare_pair(const StagedCOO& sx, const StagedCOO& sy, const Layout& xl,
                                 const Layout& yl, Operand* ox, Operand* oy) {
  if (!xl.crank) {
    prepare(sx, xl, ox);
    prepare(sy, yl, oy);
    return SPTC_OK;
  }
  DevBuf<uint64_t> rank;
  const uint64_t U = contract_ranks(sx, sy, xl, yl, &rank);
  if (!U && sx.nnz + sy.nnz) return SPTC_ERR_KEY_OVERFLOW;
  const int rb = mode_bits(U);
  if (xl.nfbits + rb > 64 || yl.nfbits + rb > 64) {
    msg = "a free-mode field and the contract rank need more than 64 bits";
    return SPTC_ERR_KEY_OVERFLOW;
  }
  const uint64_t nx = sx.nnz, n = sx.nnz + sy.nnz;
  DevBuf<uint32_t> r32;
  if (n && (sx.idx_bytes == 4 || sy.idx_bytes == 4)) {
    r32.reset(n);
    launch_ycol(rank.get(), n, 0xFFFFFFFFull, r32.get(), s_);
  }
  const void* rx = sx.idx_bytes == 4 ? (const void*)r32.get() : (const void*)rank.get();
  const void* ry =
      sy.idx_bytes == 4 ? (const void*)(r32.get() + nx) : (const void*)(rank.get() + nx);
  prepare(sx, rank_layout(xl, rb), ox, rx);
  prepare(sy, rank_layout(yl, rb), oy, ry);
  SPTC_HIP(hipStreamSynchronize(s_));  // the rank buffers are freed on return
  return SPTC_OK;
}

sptc_status Engine::contract(const Operand& X, const Operand& Y, sptc_result_s* res) {
  const Layout& xl = X.lay;
  const Layout& yl = Y.lay;
  const uint64_t nx = X.nnz, ny = Y.nnz;
  const int ncb = xl.ncbits, ynf = yl.nfbits;
  sptc_result_info& info = res->info;
  info.x_nfbits = xl.nfbits;
  info.y_nfbits = ynf;
  info.ncbits = ncb;
  info.out_bits = xl.nfbits + ynf;
  info.out_order = xl.nf + yl.nf;
  res->xl = xl;
  res->yl = yl;

  // ---- Y free-key rank compression (shorter product keys) ----------------
  // kyf: width of the column field inside Y keys and product keys from here
  // on; ytab maps a rank back to the Y free key when the field is a rank.
  int kyf = ynf;
  const uint64_t* ykeys = Y.keys.get();
  DevBuf<uint64_t> ykeys_rank, ytab;
  DevBuf<uint32_t> ytab32;  // ytab as 32-bit words when Y free keys fit them
  if (ynf > 0 && ny > 1) {
    DevBuf<uint64_t> fk(ny);
    DevBuf<float> fv(ny);
    launch_yfree_extract(Y.keys.get(), ny, low_mask(ynf), fk.get(), s_);
    const int rbf = rb_ ? rb_ : choose_radix_bits(ynf);
    ws_.ensure(ny, rbf, s_);
    ws_.prepare(radix_passes(ynf, rbf), s_);
    if (sort_needs_histogram()) histogram(ws_, fk.get(), ny, radix_passes(ynf, rbf), s_);
    const bool alt = sort_after_histogram(ws_, fk.get(), fv.get(), ny, ynf, s_);
    ytab.reset(ny);
    DevBuf<U2> tmp(scan_tmp_count(ny) + 1);
    run_unique(alt ? ws_.alt_keys.get() : fk.get(), ny, ytab.get(), tmp.get(), s_);
    const uint64_t U = read_total(tmp.get() + scan_tmp_count(ny)).a;
    const int yrb = mode_bits(U);
    if (yrb < ynf) {
      ykeys_rank.reset(ny);
      launch_yrank(Y.keys.get(), ny, ynf, low_mask(ynf), ytab.get(), U, yrb, ykeys_rank.get(),
                   s_);
      ykeys = ykeys_rank.get();
      kyf = yrb;
      if (ynf <= 32) {
        ytab32.reset(U);
        launch_ycol(ytab.get(), U, 0xFFFFFFFFull, ytab32.get(), s_);
      }
    } else {
      ytab.release();
    }
  }

  // ---- (c) index match ---------------------------------------------------
  DevBuf<uint64_t> tlo, thi, ukey, ustart;
  YTable tab;
  std::memset(&tab, 0, sizeof(tab));
  if (ncb <= kDirectBits) {
    const uint64_t K = 1ull << ncb;
    tlo.reset(K);
    thi.reset(K);
    fill_u64(tlo.get(), 0, K, s_);
    fill_u64(thi.get(), 0, K, s_);
    launch_ytable_direct(ykeys, ny, kyf, K, tlo.get(), thi.get(), s_);
    tab.direct = 1;
    tab.lo = tlo.get();
    tab.hi = thi.get();
  } else {
    ukey.reset(ny + 1);
    ustart.reset(ny + 1);
    DevBuf<U2> tmp(scan_tmp_count(ny) + 1);
    run_yunique(ykeys, ny, kyf, ukey.get(), ustart.get(), tmp.get(), s_);
    tab.direct = 0;
    tab.ukey = ukey.get();
    tab.ustart = ustart.get();
    tab.U = read_total(tmp.get() + scan_tmp_count(ny)).a;
  }
  DevBuf<uint64_t> act_prefix(nx + 1), act_ypos(nx + 1), row_key(nx + 1), row_prod(nx + 1),
      row_act(nx + 1);
  DevBuf<float> act_xval(nx + 1);
  DevBuf<uint32_t> act_row(nx + 1);
  MatchOut mo;
  std::memset(&mo, 0, sizeof(mo));
  mo.act_prefix = act_prefix.get();
  mo.act_ypos = act_ypos.get();
  mo.act_xval = act_xval.get();
  mo.act_row = act_row.get();
  mo.row_key = row_key.get();
  mo.row_prod = row_prod.get();
  mo.row_act = row_act.get();
  U2 tot;
  {
    DevBuf<U2> tmp(scan_tmp_count(nx) + 1);
    int tok = kt_.begin(s_);
    run_match(X.keys.get(), X.vals.get(), nx, ncb, tab, mo, 0, tmp.get(), s_);
    // keys read twice, plus the compacted active and row tables written
    kt_.end(tok, SPTC_K_MATCH, (double)nx * (16 + 24 + 24), s_);
    tot = read_total(tmp.get() + scan_tmp_count(nx));
  }
  const uint64_t ir = tot.a;
  const uint64_t A = tot.b >> 32;
  const uint64_t R = tot.b & 0xFFFFFFFFull;
  host_u64_[0] = ir;
  host_u64_[1] = A;
  SPTC_HIP(hipMemcpyAsync(row_prod.get() + R, &host_u64_[0], 8, hipMemcpyHostToDevice, s_));
  SPTC_HIP(hipMemcpyAsync(row_act.get() + R, &host_u64_[1], 8, hipMemcpyHostToDevice, s_));
  SPTC_HIP(hipMemcpyAsync(act_prefix.get() + A, &host_u64_[0], 8, hipMemcpyHostToDevice, s_));
  tlo.release();
  thi.release();
  ukey.release();
  ustart.release();
  record(3);
  info.ir_nnz = ir;
  info.nrows = R;
  res->row_key = std::move(row_key);

  // ---- accumulator choice (the reference's -d flag, sptc.cuh:201-218) ------
  // Dense applies when the column space (ranks of Y's free keys) fits an LDS
  // row and Y has no duplicate coordinates; automatically chosen when rows
  // are dense enough for the LDS row to pay off.
  DevBuf<double> vsum(1);
  fill_u64(reinterpret_cast<uint64_t*>(vsum.get()), 0, 1, s_);
  bool use_dense = false;
  if (o_.dense_accumulator >= 0 && ir && kyf <= kDenseMaxBits) {
    DevBuf<uint32_t> flag(1);
    launch_any_dup(Y.keys.get(), ny, flag.get(), s_);
    SPTC_HIP(hipMemcpyAsync(host_u64_, flag.get(), 4, hipMemcpyDeviceToHost, s_));
    SPTC_HIP(hipStreamSynchronize(s_));
    const bool dups = *reinterpret_cast<uint32_t*>(host_u64_) != 0;
    const uint64_t E = 1ull << (kyf < 8 ? 8 : kyf);
    // count-only runs stay on ESC, whose expansion also yields the value sum
    if (!dups)
      use_dense = o_.dense_accumulator == 1 ||
                  (!o_.count_only && dense_pays(ir, R, A, E));
  }
  // Tiled dense (column tiles in LDS, FP64 slots) for column spaces wider
  // than an LDS row: forced with dense_accumulator 2, or by 1 when the LDS
  // row does not apply.
  bool use_tdense = false;
  if (!use_dense && ir && !o_.count_only && kyf > 0 && kyf <= 32)
    use_tdense = o_.dense_accumulator == 2 || o_.dense_accumulator == 1;
  if (partition_only) use_dense = use_tdense = false;  // the plan asked for is ESC's chunking
  if (o_.dense_accumulator >= 1 && ir && !use_dense && !use_tdense) {
    msg = "dense accumulator does not apply (count-only run, or a column field above 32 "
          "bits)";
    return SPTC_ERR_BAD_ARGS;
  }
  info.accumulator = use_dense ? 1 : (use_tdense ? 2 : 0);
  std::vector<uint64_t> bounds_r{0}, bounds_p{0}, bounds_a{0};
  if (use_dense) {
    record(4);
    const int eb = kyf <= 8 ? 8 : (kyf <= 10 ? 10 : 13);
    DevBuf<uint32_t> row_cnt(R);
    DevBuf<uint64_t> out_off(R + 1);
    DevBuf<U2> tmp(scan_tmp_count(R) + 1);
    DenseArgs da;
    std::memset(&da, 0, sizeof(da));
    da.act_prefix = act_prefix.get();
    da.act_ypos = act_ypos.get();
    da.act_xval = act_xval.get();
    da.row_act = row_act.get();
    da.r0 = 0;
    da.ykeys = ykeys;
    da.yvals = Y.vals.get();
    da.fmask = low_mask(kyf);
    da.ytab = ytab.size() ? ytab.get() : nullptr;
    da.row_cnt = row_cnt.get();
    da.out_off = out_off.get();
    da.vsum = vsum.get();
    int tok = kt_.begin(s_);
    launch_dense_count(da, eb, R, s_);
    kt_.end(tok, SPTC_K_DENSE, 8.0 * (double)ir, s_);
    run_row_count_scan(row_cnt.get(), out_off.get(), R, tmp.get(), s_);
    const uint64_t out_n = read_total(tmp.get() + scan_tmp_count(R)).a;
    info.nnz = out_n;
    info.nchunks = 1;
    if (!o_.count_only) {
      if (DevicePool::get().live_bytes() + out_n * 16 > budget_) {
        msg = "output does not fit the memory budget";
        return SPTC_ERR_OOM_BUDGET;
      }
      OutChunk oc;
      oc.nnz = out_n;
      oc.r0 = 0;
      oc.r1 = R;
      oc.row.reset(out_n);
      oc.col.reset(out_n);
      oc.val.reset(out_n);
      da.out_row = oc.row.get();
      da.out_col = oc.col.get();
      da.out_val = oc.val.get();
      tok = kt_.begin(s_);
      launch_dense_numeric(da, eb, R, s_);
      kt_.end(tok, SPTC_K_DENSE, 12.0 * (double)ir + 16.0 * (double)out_n, s_);
      res->chunks.push_back(std::move(oc));
    }
  } else if (use_tdense) {
    record(4);
    const uint64_t fmask = low_mask(kyf);
    DevBuf<uint32_t> ycol32(ny);
    launch_ycol(ykeys, ny, fmask, ycol32.get(), s_);
    DevBuf<uint64_t> cur(A ? A : 1);
    DevBuf<uint32_t> row_cnt(R);
    DevBuf<uint64_t> out_off(R + 1);
    DevBuf<U2> tmp(scan_tmp_count(R) + 1);
    TDenseArgs ta;
    std::memset(&ta, 0, sizeof(ta));
    ta.act_prefix = act_prefix.get();
    ta.act_ypos = act_ypos.get();
    ta.act_xval = act_xval.get();
    ta.row_act = row_act.get();
    ta.ycol = ycol32.get();
    ta.yvals = Y.vals.get();
    ta.ytab = ytab.size() ? ytab.get() : nullptr;
    ta.cur = cur.get();
    ta.row_cnt = row_cnt.get();
    ta.out_off = out_off.get();
    ta.vsum = vsum.get();
    // one workgroup per row, in launches of at most 2^30 rows
    constexpr uint64_t kRowsPerLaunch = 1ull << 30;
    int tok = kt_.begin(s_);
    for (uint64_t b = 0; b < R; b += kRowsPerLaunch) {
      ta.r0 = b;
      ta.row_cnt = row_cnt.get() + b;
      launch_tdense_count(ta, std::min(kRowsPerLaunch, R - b), s_);
    }
    kt_.end(tok, SPTC_K_DENSE, 4.0 * (double)ir, s_);
    run_row_count_scan(row_cnt.get(), out_off.get(), R, tmp.get(), s_);
    const uint64_t out_n = read_total(tmp.get() + scan_tmp_count(R)).a;
    info.nnz = out_n;
    info.nchunks = 1;
    // the whole output stays resident: one chunk
    if (DevicePool::get().live_bytes() + out_n * 16 > budget_) {
      msg = "output does not fit the memory budget";
      return SPTC_ERR_OOM_BUDGET;
    }
    OutChunk oc;
    oc.nnz = out_n;
    oc.r0 = 0;
    oc.r1 = R;
    oc.row.reset(out_n ? out_n : 1);
    oc.col.reset(out_n ? out_n : 1);
    oc.val.reset(out_n ? out_n : 1);
    ta.out_row = oc.row.get();
    ta.out_col = oc.col.get();
    ta.out_val = oc.val.get();
    ta.out0 = 0;
    tok = kt_.begin(s_);
    for (uint64_t b = 0; b < R; b += kRowsPerLaunch) {
      ta.r0 = b;
      ta.out_off = out_off.get() + b;
      launch_tdense_numeric(ta, std::min(kRowsPerLaunch, R - b), s_);
    }
    // per product a Y column and value gathered; per output 16 bytes written
    kt_.end(tok, SPTC_K_DENSE, 8.0 * (double)ir + 16.0 * (double)out_n, s_);
    res->chunks.push_back(std::move(oc));
  } else if (ir) {
    DevBuf<uint64_t> scratch(4);
    launch_max_row(row_prod.get(), R, scratch.get(), s_);
    SPTC_HIP(hipMemcpyAsync(host_u64_, scratch.get(), 8, hipMemcpyDeviceToHost, s_));
    SPTC_HIP(hipStreamSynchronize(s_));
    const uint64_t max_row = host_u64_[0];
    const uint64_t live = DevicePool::get().live_bytes();
    const uint64_t free_b = budget_ > live ? budget_ - live : 0;
    const uint64_t per_prod = 12 + SortWorkspace::bytes_per_key(10) + 1;
    const uint64_t out_per = o_.count_only ? 0 : 16;
    // room for the accumulation's small temporaries (scan tiles, row tables)
    const uint64_t slack = std::min<uint64_t>(64ull << 20, budget_avail() / 8);
    uint64_t cap;
    if (free_b > slack + out_per * ir + per_prod * std::min<uint64_t>(ir, 1ull << 20))
      cap = (free_b - slack - out_per * ir) / per_prod;
    else
      cap = free_b > slack ? (free_b - slack) / (per_prod + out_per) : 0;
    // coarse granularity keeps the chunking identical across repeated calls
    if (cap > (1ull << 26)) cap &= ~((1ull << 26) - 1);
    cap = std::min<uint64_t>(cap, std::min<uint64_t>(ir, 4000000000ull));
    if (cap < max_row) {
      msg = "one X row's products exceed the memory budget";
      return SPTC_ERR_OOM_BUDGET;
    }
    const uint64_t rmax = kyf >= 32 ? (kyf >= 64 ? 1 : (1ull << (64 - kyf))) : (1ull << 32);
    uint64_t r0 = 0;
    while (r0 < R) {
      launch_chunk_end(row_prod.get(), row_act.get(), R, r0, rmax, cap, scratch.get(), s_);
      SPTC_HIP(hipMemcpyAsync(host_u64_, scratch.get(), 24, hipMemcpyDeviceToHost, s_));
      SPTC_HIP(hipStreamSynchronize(s_));
      uint64_t r1 = host_u64_[0];
      if (r1 <= r0) {
        msg = "partitioning made no progress";
        return SPTC_ERR_INTERNAL;
      }
      bounds_r.push_back(r1);
      bounds_p.push_back(host_u64_[1]);
      bounds_a.push_back(host_u64_[2]);
      r0 = r1;
    }
    record(4);
    res->part_rows = bounds_r;
    res->part_prods = bounds_p;
    if (partition_only) {  // the plan only: no chunk is accumulated
      info.nchunks = (int32_t)(bounds_r.size() - 1);
      bounds_r.resize(1);
      bounds_p.resize(1);
      bounds_a.resize(1);
    }

    // ---- (d) ESC accumulate per chunk --------------------------------------
    uint64_t chunk_cap = 0;
    for (size_t c = 1; c < bounds_p.size(); ++c)
      chunk_cap = std::max(chunk_cap, bounds_p[c] - bounds_p[c - 1]);
    DevBuf<uint64_t> kA(chunk_cap);
    DevBuf<float> vA(chunk_cap);
    uint64_t max_rows = 0;
    for (size_t c = 1; c < bounds_r.size(); ++c)
      max_rows = std::max(max_rows, bounds_r[c] - bounds_r[c - 1]);
    DevBuf<U2> tmp(std::max(reduce_tmp_count(std::max(chunk_cap, max_rows)),
                            narrow_tmp_count(chunk_cap)));
    int rb_max = 8;
    for (size_t c = 1; c < bounds_r.size(); ++c) {
      const int b = mode_bits(bounds_r[c] - bounds_r[c - 1]) + kyf;
      rb_max = std::max(rb_max, rb_ ? rb_ : choose_radix_bits(b));
    }
    ws_.ensure(chunk_cap, rb_max, s_);
    const uint64_t fmask = low_mask(kyf);
    // 32-bit column field of Y for the expansion (4 bytes per gathered key
    // instead of 8)
    DevBuf<uint32_t> ycol;
    if (kyf <= 32 && !partition_only) {
      ycol.reset(ny);
      launch_ycol(ykeys, ny, fmask, ycol.get(), s_);
    }
    // row-segmented product sort
    DevBuf<RowRange> rrange(1);
    RowSegPlan rseg;
    DevBuf<uint32_t> trow(2 * (narrow_tile_count(chunk_cap) + 1));
    // verbose: a progress line on stderr at least every 10 s of chunks
    const auto t_loop = std::chrono::steady_clock::now();
    auto t_log = t_loop;
    if (o_.verbose && bounds_r.size() > 1)
      fprintf(stderr, "[sptc] accumulate: %zu chunks, %llu products\n", bounds_r.size() - 1,
              (unsigned long long)ir);
    for (size_t c = 1; c < bounds_r.size(); ++c) {
      const uint64_t r0c = bounds_r[c - 1], r1c = bounds_r[c];
      const uint64_t P0 = bounds_p[c - 1], P1 = bounds_p[c];
      const uint64_t n = P1 - P0;
      if (!n) continue;
      const int bits = mode_bits(r1c - r0c) + kyf;
      const int rb = rb_ ? rb_ : choose_radix_bits(bits);
      const int passes = radix_passes(bits, rb);
      ws_.ensure(n, rb, s_);
      ws_.prepare(passes, s_);
      // row-segmented sort when the column bits leave row bits to skip
      const bool rowsort = kyf > 0 && kyf < bits && kyf <= 32 && prefer_row_sort();
      if (rowsort) {
        const uint64_t lrows = r1c - r0c;
        RowRange* h = reinterpret_cast<RowRange*>(host_u64_ + 4);
        h->ra = r0c;
        h->rows = lrows;
        h->lrow0 = 0;
        h->out0 = 0;
        SPTC_HIP(hipMemcpyAsync(rrange.get(), h, sizeof(RowRange), hipMemcpyHostToDevice, s_));
        build_row_segments(rrange.get(), 1, lrows, row_prod.get(), kyf, &rseg, ws_, tmp.get(),
                           s_, prefer_bucket_tail(), row_act.get());
      }
      ExpandArgs ea;
      ea.act_prefix = act_prefix.get();
      ea.act_ypos = act_ypos.get();
      ea.act_xval = act_xval.get();
      ea.act_row = act_row.get();
      ea.A0 = bounds_a[c - 1];
      ea.A1 = bounds_a[c];
      ea.P0 = P0;
      ea.P1 = P1;
      ea.r0 = (uint32_t)r0c;
      ea.ykeys = ykeys;
      ea.yvals = Y.vals.get();
      ea.ynf = kyf;
      ea.fmask = fmask;
      ea.out_keys = kA.get();
      ea.out_vals = vA.get();
      ea.big_col = nullptr;
      ea.big_val = nullptr;
      ea.out_col32 = nullptr;
      ea.ycol = ycol.size() ? ycol.get() : nullptr;
      // narrow form: sorted columns land in kA's first 4n bytes
      uint32_t* fcol = reinterpret_cast<uint32_t*>(kA.get());
      // count-only runs sort the columns alone: no value is written or moved
      const bool cols_only = rowsort && o_.count_only;
      float* const pv = cols_only ? nullptr : vA.get();
      float* const pv1 = cols_only ? nullptr : ws_.alt_vals.get();
      if (rowsort) {
        const NarrowTarget nt =
            narrow_target(rseg, n, vA.get(), ws_.alt_keys.get(), ws_.alt_vals.get());
        ea.big_col = nt.col;
        ea.big_val = cols_only ? nullptr : nt.val;
        ea.out_col32 = fcol;
        ea.out_vals = pv;
      }
      ea.hist = ws_.hist.get();
      const bool hist = sort_needs_histogram();
      ea.passes = hist && passes <= kMaxFusedPasses ? passes : 0;
      ea.vsum = vsum.get();
      int tok = kt_.begin(s_);
      if (rowsort) launch_expand_items(ea, rseg, ws_, s_);  // ycol is set (kyf <= 32)
      else launch_expand(ea, rb, s_);
      // matched Y (column, value) read; product written: column + value in
      // the narrow form of a row-segmented sort, 64-bit key + value otherwise
      kt_.end(tok, SPTC_K_EXPAND,
              (rowsort ? (cols_only ? 12.0 : 16.0) : 24.0) * (double)n, s_);
      bool alt = false;
      NarrowRows nr{fcol, rowsort ? rseg.row_start() : nullptr, r1c - r0c, trow.get()};
      // high-digit form: big rows are reduced by buckets, the narrow reduce
      // takes the small rows
      const bool buckets = rowsort && rseg.msd && rseg.nbig_rows > 0;
      BucketArgs ba{};
      if (buckets) {
        nr.big = rseg.row_big();
        ba.segs = rseg.bigseg.get();
        ba.tiles = rseg.big.get();
        ba.nsegs = rseg.nbig_rows;
        ba.starts = ws_.rts_cnt.get();
        ba.boff = ws_.rts_off.get();
        ba.cols = fcol;
        ba.vals = vA.get();
        ba.bigrow = rseg.bigrow.get();
        ba.low = kyf - rseg.rb;
        ba.gshift = bucket_group_shift(n - rseg.nsmall, rseg.nbig_rows);
        ba.tile_sums = tmp.get();
        ba.r0 = r0c;
        ba.ytab = ytab.size() ? ytab.get() : nullptr;
        ba.ytab32 = ytab32.size() ? ytab32.get() : nullptr;
      }
      if (rowsort) {
        launch_tile_rows(rseg.row_start(), r1c - r0c, n, trow.get(), s_);
        sort_row_segments(ws_, rseg, n, fcol, pv, ws_.alt_keys.get(), pv1, rseg.row_start(), s_,
                          &kt_, SPTC_K_SORT_PROD, /*counted=*/true);
      } else {
        ws_.ensure(n, rb, s_);
        if (hist && passes > kMaxFusedPasses) histogram(ws_, kA.get(), n, passes, s_);
        alt = sort_after_histogram(ws_, kA.get(), vA.get(), n, bits, s_, &kt_, SPTC_K_SORT_PROD);
      }
      const uint64_t* sk = alt ? ws_.alt_keys.get() : kA.get();
      const float* sv = alt ? ws_.alt_vals.get() : vA.get();
      tok = kt_.begin(s_);
      if (rowsort) {
        run_reduce_count_narrow(nr, n, tmp.get(), s_, /*scan=*/!buckets);
        if (buckets) {
          run_bucket_count(ba, s_);
          scan_narrow_tiles(n, tmp.get(), s_);
        }
      } else {
        run_reduce_count(sk, n, tmp.get(), s_);
      }
      kt_.end(tok, SPTC_K_REDUCE, (rowsort ? 4.0 : 8.0) * (double)n, s_);
      const uint64_t out_n =
          read_total(tmp.get() + (rowsort ? narrow_tile_count(n) : scan_tmp_count(n))).a;
      info.nnz += out_n;
      info.nchunks++;
      if (o_.verbose) {
        const auto now = std::chrono::steady_clock::now();
        if (now - t_log >= std::chrono::seconds(10) || c + 1 == bounds_r.size()) {
          t_log = now;
          fprintf(stderr, "[sptc] chunk %zu/%zu: %llu of %llu products, %llu outputs, %.1f s\n", c,
                  bounds_r.size() - 1, (unsigned long long)P1, (unsigned long long)ir,
                  (unsigned long long)info.nnz,
                  std::chrono::duration<double>(now - t_loop).count());
          fflush(stderr);
        }
      }
      if (o_.count_only) continue;
      if (DevicePool::get().live_bytes() + out_n * 16 > budget_) {
        msg = "output does not fit the memory budget";
        return SPTC_ERR_OOM_BUDGET;
      }
      OutChunk oc;
      oc.nnz = out_n;
      oc.r0 = r0c;
      oc.r1 = r1c;
      oc.row.reset(out_n);
      oc.col.reset(out_n);
      oc.val.reset(out_n);
      tok = kt_.begin(s_);
      if (rowsort) {
        run_reduce_emit_narrow(nr, vA.get(), n, oc.row.get(), oc.col.get(), oc.val.get(), r0c,
                               ytab.size() ? ytab.get() : nullptr,
                               ytab32.size() ? ytab32.get() : nullptr, tmp.get(), s_);
        if (buckets) {
          ba.out_row = oc.row.get();
          ba.out_col = oc.col.get();
          ba.out_val = oc.val.get();
          run_bucket_emit(ba, s_);
        }
      } else
        run_reduce_emit(sk, sv, n, oc.row.get(), oc.col.get(), oc.val.get(), r0c, kyf, fmask,
                        ytab.size() ? ytab.get() : nullptr, tmp.get(), s_);
      kt_.end(tok, SPTC_K_REDUCE, (rowsort ? 8.0 : 12.0) * (double)n + 16.0 * (double)out_n,
              s_);
      res->chunks.push_back(std::move(oc));
    }
  } else {
    record(4);
  }
  SPTC_HIP(hipMemcpyAsync(&info.value_sum, vsum.get(), sizeof(double), hipMemcpyDeviceToHost,
                          s_));
  SPTC_HIP(hipMemcpyAsync(host_u64_, ws_.err(), 4, hipMemcpyDeviceToHost, s_));
  record(5);
  SPTC_HIP(hipStreamSynchronize(s_));
  kt_.collect(res->kstats, SPTC_K_COUNT);
  if (*reinterpret_cast<uint32_t*>(host_u64_) != 0) {
    msg = "radix sort look-back exceeded its spin limit";
    return SPTC_ERR_INTERNAL;
  }
  float t;
  auto el = [&](int a, int b) -> double {
    if (!ev_used_[a] || !ev_used_[b]) return 0.0;
    SPTC_HIP(hipEventElapsedTime(&t, ev_[a], ev_[b]));
    return t;
  };
  times.ms[0] = el(0, 1);
  times.ms[1] = el(1, 2);
  times.ms[2] = el(2, 3);
  times.ms[3] = el(3, 4);
  times.ms[4] = el(4, 5);
  times.ms[5] = el(0, 5);
  if (res->chunks.size() == 1) {
    info.row = res->chunks[0].row.get();
    info.col = res->chunks[0].col.get();
    info.val = res->chunks[0].val.get();
  }
  info.row_key = res->row_key.get();
  return SPTC_OK;
}

}  // namespace sptc
