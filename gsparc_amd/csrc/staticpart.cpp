// Static operand partitioning (host orchestration; the merge runs in
// result.hip's kernels).
#include "staticpart.hpp"

#include <algorithm>
#include <cstring>
#include <thread>

#include "kernels.hpp"
#include "result.hpp"
#include "scan.hpp"

namespace sptc {

namespace {

constexpr int kMaxParts = kMaxMergeParts;
constexpr int kHistBits = 20;

int host_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(h ? h : 1u, 16u));
}

template <typename F>
void parallel_ranges(uint64_t n, int threads, F&& f) {
  std::vector<std::thread> ts;
  const uint64_t per = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    const uint64_t b = std::min(n, per * t), e = std::min(n, per * (t + 1));
    ts.emplace_back([&f, t, b, e] { f(t, b, e); });
  }
  for (auto& th : ts) th.join();
}

inline uint64_t load_idx(const void* a, int bytes, uint64_t i) {
  return bytes == 4 ? static_cast<const uint32_t*>(a)[i] : static_cast<const uint64_t*>(a)[i];
}

// Free key of nonzero i in layout `lay` (the field without its offset).
struct FreeKey {
  int nf = 0;
  int shift[SPTC_MAX_ORDER];
  uint64_t mask[SPTC_MAX_ORDER];
  const void* idx[SPTC_MAX_ORDER];
  int bytes = 4;
  uint64_t operator()(uint64_t i) const {
    uint64_t k = 0;
    for (int j = 0; j < nf; ++j) k |= (load_idx(idx[j], bytes, i) & mask[j]) << shift[j];
    return k;
  }
};

}  // namespace

sptc_coo HostPart::coo(int order, const uint64_t* dims) {
  ptr.resize(order);
  for (int m = 0; m < order; ++m) ptr[m] = idx[m].data();
  sptc_coo c;
  std::memset(&c, 0, sizeof(c));
  c.order = order;
  c.nnz = nnz;
  c.dims = dims;
  c.idx = ptr.data();
  c.idx_bytes = idx_bytes;
  c.val = val.data();
  c.on_device = 0;
  return c;
}

uint64_t resident_bytes(uint64_t n, int role) {
  // X: key + value, then the match tables (active prefix, Y position, row
  // key / product / active prefixes: 8 B each; active value and row: 4 B).
  // Y: key + value, rank-compressed keys and rank table, 32-bit column and
  // rank copies, the contract-key table.
  return role == 1 ? n * (12 + 48) : n * (12 + 8 + 8 + 8 + 16);
}

uint64_t prepare_bytes(uint64_t n, int order, int idx_bytes) {
  // staged indices and values, keys + values, sort workspace (ping-pong
  // keys and values)
  return n * ((uint64_t)order * idx_bytes + 4 + 12 + 12);
}

bool find_partition_num(uint64_t nx, int xorder, uint64_t ny, int yorder, int idx_bytes,
                        uint64_t budget, int* px, int* py) {
  const uint64_t half = budget / 2;
  auto fits = [&](int a, int b) {
    const uint64_t xn = (nx + a - 1) / a, yn = (ny + b - 1) / b;
    const uint64_t peak =
        std::max(prepare_bytes(xn, xorder, idx_bytes) + resident_bytes(yn, 2),
                 prepare_bytes(yn, yorder, idx_bytes) + resident_bytes(xn, 1));
    return resident_bytes(xn, 1) + resident_bytes(yn, 2) <= half && peak <= half;
  };
  *px = *py = 1;
  // Y whole when it fits beside an X part, else both split; the fewest Y
  // parts, then the fewest X parts
  for (int b = 1; b <= kMaxParts; ++b)
    for (int a = 1; a <= kMaxParts; ++a)
      if (fits(a, b)) {
        *px = a;
        *py = b;
        return true;
      }
  return false;
}

void split_operand(const sptc_coo& t0, const Layout& lay, int parts, bool own,
                   OperandSplit* out) {
  *out = OperandSplit();
  const int nfb = lay.nfbits;
  if (parts < 1) parts = 1;
  out->parts = parts;
  out->key_lo.assign(parts + 1, 0);
  out->nnz_lo.assign(parts + 1, 0);
  out->key_lo[parts] = nfb >= 64 ? ~0ull : (1ull << nfb);
  out->nnz_lo[parts] = t0.nnz;
  if (parts == 1 && !own) return;
  // device inputs: copy to the host first
  sptc_coo t = t0;
  std::vector<std::vector<uint8_t>> hidx;
  std::vector<float> hval;
  std::vector<const void*> hptr;
  if (t0.on_device && t0.nnz) {
    hidx.resize(t0.order);
    hptr.resize(t0.order);
    for (int m = 0; m < t0.order; ++m) {
      hidx[m].resize(t0.nnz * t0.idx_bytes);
      SPTC_HIP(hipMemcpy(hidx[m].data(), t0.idx[m], hidx[m].size(), hipMemcpyDeviceToHost));
      hptr[m] = hidx[m].data();
    }
    hval.resize(t0.nnz);
    SPTC_HIP(hipMemcpy(hval.data(), t0.val, t0.nnz * 4, hipMemcpyDeviceToHost));
    t.idx = hptr.data();
    t.val = hval.data();
    t.on_device = 0;
  }
  FreeKey fk;
  fk.bytes = t.idx_bytes;
  const int off = lay.role == 1 ? lay.ncbits : 0;
  for (int j = 0; j < lay.nf; ++j) {
    const int m = lay.fmodes[j];
    fk.shift[j] = lay.shift[m] - off;
    fk.mask[j] = low_mask(lay.width[m]);
    fk.idx[j] = t.idx[m];
  }
  fk.nf = lay.nf;
  const uint64_t n = t.nnz;
  const int hb = std::min(nfb, kHistBits);
  const int hs = nfb - hb;
  const uint64_t H = 1ull << hb;
  const int T = host_threads();
  // histogram of the top free-key bits
  std::vector<std::vector<uint32_t>> lh(T);
  parallel_ranges(n, T, [&](int th, uint64_t b, uint64_t e) {
    lh[th].assign(H, 0);
    for (uint64_t i = b; i < e; ++i) ++lh[th][fk(i) >> hs];
  });
  std::vector<uint64_t> cum(H + 1, 0);
  for (uint64_t h = 0; h < H; ++h) {
    uint64_t c = 0;
    for (int th = 0; th < T; ++th) c += lh[th][h];
    cum[h + 1] = cum[h] + c;
  }
  // cut k at the bucket edge nearest k * n / parts (ascending cuts)
  std::vector<uint64_t> cut(parts + 1, 0);  // bucket index where part k starts
  cut[parts] = H;
  for (int k = 1; k < parts; ++k) {
    const uint64_t want = n * (uint64_t)k / parts;
    uint64_t b = std::lower_bound(cum.begin(), cum.end(), want) - cum.begin();
    if (b > 0 && want - cum[b - 1] < cum[b] - want) --b;
    cut[k] = std::max(cut[k - 1], std::min<uint64_t>(b, H));
  }
  for (int k = 0; k <= parts; ++k) {
    out->key_lo[k] = k == parts ? out->key_lo[parts] : (cut[k] << hs);
    out->nnz_lo[k] = cum[cut[k]];
  }
  std::vector<uint32_t> part_of_bucket(H);
  for (int k = 0; k < parts; ++k)
    for (uint64_t h = cut[k]; h < cut[k + 1]; ++h) part_of_bucket[h] = (uint32_t)k;
  // stable scatter: thread-local counts per part, then positions
  std::vector<std::vector<uint64_t>> lc(T, std::vector<uint64_t>(parts, 0));
  parallel_ranges(n, T, [&](int th, uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) ++lc[th][part_of_bucket[fk(i) >> hs]];
  });
  out->part.resize(parts);
  std::vector<std::vector<uint64_t>> pos(T, std::vector<uint64_t>(parts, 0));
  for (int k = 0; k < parts; ++k) {
    uint64_t c = 0;
    for (int th = 0; th < T; ++th) {
      pos[th][k] = c;
      c += lc[th][k];
    }
    HostPart& hp = out->part[k];
    hp.nnz = c;
    hp.idx_bytes = t.idx_bytes;
    hp.idx.assign(t.order, std::vector<uint8_t>(c * t.idx_bytes));
    hp.val.resize(c);
  }
  const int ib = t.idx_bytes;
  parallel_ranges(n, T, [&](int th, uint64_t b, uint64_t e) {
    std::vector<uint64_t> p = pos[th];
    for (uint64_t i = b; i < e; ++i) {
      const uint32_t k = part_of_bucket[fk(i) >> hs];
      HostPart& hp = out->part[k];
      const uint64_t d = p[k]++;
      for (int m = 0; m < t.order; ++m)
        std::memcpy(hp.idx[m].data() + d * ib, static_cast<const uint8_t*>(t.idx[m]) + i * ib, ib);
      hp.val[d] = t.val[i];
    }
  });
}

namespace {

// Merges the results of one X part against the Y parts (same row table,
// ascending column ranges) into one chunk with rows offset by row_off.
void merge_pair_results(std::vector<sptc_result_s>& sub, uint64_t R, uint64_t row_off,
                        hipStream_t s, OutChunk* out) {
  const int Q = (int)sub.size();
  uint64_t total = 0;
  for (auto& r : sub) total += r.info.nnz;
  out->nnz = total;
  out->r0 = row_off;
  out->r1 = row_off + R;
  out->row.reset(total ? total : 1);
  out->col.reset(total ? total : 1);
  out->val.reset(total ? total : 1);
  if (!total) return;
  std::vector<DevBuf<uint64_t>> start(Q), end(Q);
  ExtentSet es;
  std::memset(&es, 0, sizeof(es));
  es.Q = Q;
  for (int q = 0; q < Q; ++q) {
    start[q].reset(R);
    end[q].reset(R);
    fill_u64(start[q].get(), 0, R, s);
    fill_u64(end[q].get(), 0, R, s);
    uint64_t base = 0;
    for (const OutChunk& c : sub[q].chunks) {
      launch_row_extent(c.R(), c.nnz, base, start[q].get(), end[q].get(), s);
      base += c.nnz;
    }
    es.start[q] = start[q].get();
    es.end[q] = end[q].get();
  }
  const uint64_t RQ = R * (uint64_t)Q;
  DevBuf<uint64_t> cnt(RQ), dst(RQ + 1);
  DevBuf<U2> tmp(scan_tmp_count(RQ) + 1);
  launch_interleave_counts(es, R, cnt.get(), s);
  run_exclusive_scan(cnt.get(), dst.get(), RQ, tmp.get(), s);
  for (int q = 0; q < Q; ++q) {
    uint64_t base = 0;
    for (const OutChunk& c : sub[q].chunks) {
      launch_merge_scatter(c.R(), c.C(), c.V(), c.nnz, base, start[q].get(), dst.get(), Q, q,
                           (uint32_t)row_off, out->row.get(), out->col.get(), out->val.get(), s);
      base += c.nnz;
    }
  }
  SPTC_HIP(hipStreamSynchronize(s));
}

}  // namespace

sptc_status contract_partitioned(Engine& e, const sptc_coo& X, const sptc_coo& Y,
                                 const Layout& xl, const Layout& yl, OperandSplit& xs,
                                 OperandSplit& ys, sptc_result_s* res) {
  hipStream_t s = e.stream();
  const int P = xs.parts, Q = ys.parts;
  sptc_result_info& info = res->info;
  info.x_nfbits = xl.nfbits;
  info.y_nfbits = yl.nfbits;
  info.ncbits = xl.ncbits;
  info.out_bits = xl.nfbits + yl.nfbits;
  info.out_order = xl.nf + yl.nf;
  res->xl = xl;
  res->yl = yl;
  std::vector<DevBuf<uint64_t>> row_tabs;
  std::vector<uint64_t> row_counts;
  uint64_t row_off = 0, prod_off = 0;
  int dense_pairs = 0, pairs = 0;
  double ms[6] = {0, 0, 0, 0, 0, 0};
  res->part_rows.assign(1, 0);
  res->part_prods.assign(1, 0);
  // Y parts with nonzeros (an all-empty Y still runs once, for the rows)
  std::vector<int> yparts;
  for (int q = 0; q < Q; ++q)
    if (Q == 1 || ys.part[q].nnz) yparts.push_back(q);
  if (yparts.empty()) yparts.push_back(0);
  for (int p = 0; p < P; ++p) {
    if (P > 1 && xs.part[p].nnz == 0) continue;
    const sptc_coo xc = xs.part.empty() ? X : xs.part[p].coo(X.order, X.dims);
    // keys above 64 bits rank the contracted coordinates of each pair
    // together, so the X part is prepared with every Y part then
    Operand ox;
    if (!xl.crank) {
      StagedCOO sx;
      e.record(0);
      e.stage(xc, &sx);
      e.record(1);
      e.prepare(sx, xl, &ox);
      e.record(2);
      ms[0] += e.elapsed_ms(0, 1);
      ms[1] += e.elapsed_ms(1, 2);
    }
    std::vector<sptc_result_s> sub(yparts.size());
    for (size_t j = 0; j < yparts.size(); ++j) {
      const int q = yparts[j];
      const sptc_coo yc = ys.part.empty() ? Y : ys.part[q].coo(Y.order, Y.dims);
      Operand oy;
      {
        StagedCOO sx, sy;
        e.record(0);
        if (xl.crank) e.stage(xc, &sx);
        e.stage(yc, &sy);
        e.record(1);
        if (xl.crank) {
          const sptc_status pst = e.prepare_pair(sx, sy, xl, yl, &ox, &oy);
          if (pst != SPTC_OK) return pst;
        } else {
          e.prepare(sy, yl, &oy);
        }
        e.record(2);
        ms[0] += e.elapsed_ms(0, 1);
        ms[1] += e.elapsed_ms(1, 2);
      }
      sptc_result_s& sq = sub[j];
      std::memset(&sq.info, 0, sizeof(sq.info));
      const sptc_status st = e.contract(ox, oy, &sq);
      if (st != SPTC_OK) return st;
      for (int k = 2; k < 5; ++k) ms[k] += e.times.ms[k];
      for (int k = 0; k < SPTC_K_COUNT; ++k) {
        res->kstats[k].launches += sq.kstats[k].launches;
        res->kstats[k].ms += sq.kstats[k].ms;
        res->kstats[k].bytes += sq.kstats[k].bytes;
      }
      info.ir_nnz += sq.info.ir_nnz;
      info.nnz += sq.info.nnz;
      info.value_sum += sq.info.value_sum;
      dense_pairs += sq.info.accumulator;
      ++pairs;
    }
    // every pair result lists the X part's rows
    const uint64_t R = sub[0].info.nrows;
    if (row_off + R > 0xFFFFFFFFull) {
      e.msg = "more than 2^32 X rows";
      return SPTC_ERR_BAD_ARGS;
    }
    uint64_t part_ir = 0;
    for (auto& r : sub) part_ir += r.info.ir_nnz;
    if (!e.options().count_only) {
      uint64_t total = 0;
      for (auto& r : sub) total += r.info.nnz;
      if (DevicePool::get().live_bytes() + total * 16 > e.budget()) {
        e.msg = "merged output of an X part does not fit the memory budget";
        return SPTC_ERR_OOM_BUDGET;
      }
      OutChunk oc;
      merge_pair_results(sub, R, row_off, s, &oc);
      res->chunks.push_back(std::move(oc));
    }
    row_tabs.push_back(std::move(sub[0].row_key));
    row_counts.push_back(R);
    row_off += R;
    prod_off += part_ir;
    res->part_rows.push_back(row_off);
    res->part_prods.push_back(prod_off);
  }
  info.nrows = row_off;
  res->row_key.reset(row_off ? row_off : 1);
  uint64_t o = 0;
  for (size_t k = 0; k < row_tabs.size(); ++k) {
    if (row_counts[k])
      SPTC_HIP(hipMemcpyAsync(res->row_key.get() + o, row_tabs[k].get(), row_counts[k] * 8,
                              hipMemcpyDeviceToDevice, s));
    o += row_counts[k];
  }
  SPTC_HIP(hipStreamSynchronize(s));
  info.nchunks = (int32_t)res->chunks.size();
  if (e.options().count_only) info.nchunks = P;
  info.accumulator = (pairs && dense_pairs == pairs) ? 1 : 0;
  info.row_key = res->row_key.get();
  if (res->chunks.size() == 1) {
    info.row = res->chunks[0].R();
    info.col = res->chunks[0].C();
    info.val = res->chunks[0].V();
  }
  ms[5] = ms[0] + ms[1] + ms[2] + ms[3] + ms[4];
  std::memcpy(info.phase_ms, ms, sizeof(ms));
  return SPTC_OK;
}

}  // namespace sptc
