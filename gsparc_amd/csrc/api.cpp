// extern "C" boundary of libsptc_hip.so (declared in include/sptc.h).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "batch_engine.hpp"
#include "engine.hpp"
#include "result.hpp"
#include "serve.hpp"
#include "staticpart.hpp"

using namespace sptc;

struct sptc_batch_s {
  BatchEngine e;
  KStat kstats[SPTC_K_COUNT];
};

struct sptc_server_s {
  std::unique_ptr<Server> srv;
};

struct sptc_plan_s {
  sptc_opts opts;
  Operand x, y;
  double prepare_ms = 0;
  // statically partitioned plan: host-side parts of both operands, staged
  // and prepared pair by pair at each execution (x, y stay empty)
  bool split = false;
  Layout xl, yl;
  std::vector<uint64_t> xdims, ydims;
  OperandSplit xs, ys;
};

namespace {

thread_local std::string g_last_error;

sptc_status fail(sptc_status s, const std::string& m) {
  g_last_error = m;
  return s;
}

sptc_status from_hip(const HipError& e) {
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s failed: %s (line %d)", e.what, hipGetErrorString(e.err),
                e.line);
  return fail(SPTC_ERR_HIP, buf);
}

sptc_status check_coo(const sptc_coo* t, const char* name) {
  if (!t) return fail(SPTC_ERR_BAD_ARGS, std::string(name) + " is NULL");
  if (t->order < 1 || t->order > SPTC_MAX_ORDER)
    return fail(SPTC_ERR_BAD_ARGS, std::string(name) + ": order out of range");
  if (!t->dims) return fail(SPTC_ERR_BAD_ARGS, std::string(name) + ": dims is NULL");
  if (t->idx_bytes != 4 && t->idx_bytes != 8)
    return fail(SPTC_ERR_BAD_ARGS, std::string(name) + ": idx_bytes must be 4 or 8");
  if (t->nnz) {
    if (!t->idx || !t->val) return fail(SPTC_ERR_BAD_ARGS, std::string(name) + ": NULL arrays");
    for (int m = 0; m < t->order; ++m)
      if (!t->idx[m]) return fail(SPTC_ERR_BAD_ARGS, std::string(name) + ": NULL index array");
  }
  if (t->nnz >= (1ull << 32))
    return fail(SPTC_ERR_BAD_ARGS, std::string(name) + ": nnz must be below 2^32");
  return SPTC_OK;
}

bool device_ok() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return false;
  }
  return true;
}

}  // namespace

namespace sptc {
// Error reporting for entry points defined in other files.
sptc_status api_fail(sptc_status s, const std::string& m) { return fail(s, m); }
}  // namespace sptc

extern "C" {

const char* sptc_status_string(sptc_status s) {
  switch (s) {
    case SPTC_OK: return "ok";
    case SPTC_ERR_BAD_ARGS: return "bad arguments";
    case SPTC_ERR_OOM_BUDGET: return "out of memory budget";
    case SPTC_ERR_KEY_OVERFLOW: return "key needs more than 64 bits";
    case SPTC_ERR_HIP: return "HIP runtime error";
    case SPTC_ERR_NO_DEVICE: return "no usable GPU";
    case SPTC_ERR_INTERNAL: return "internal consistency check failed";
  }
  return "unknown status";
}

const char* sptc_last_error(void) { return g_last_error.c_str(); }

sptc_status sptc_layout(int32_t order, const uint64_t* dims, int32_t nc, const int32_t* cmodes,
                        const int32_t* cwidth, int32_t role, int32_t* shift, int32_t* width,
                        int32_t* nfbits, int32_t* ncbits) {
  if (!dims || (nc && !cmodes) || (role != 1 && role != 2))
    return fail(SPTC_ERR_BAD_ARGS, "sptc_layout: bad arguments");
  Layout L;
  sptc_status st = make_layout(order, dims, nc, cmodes, cwidth, role, &L);
  if (st == SPTC_ERR_BAD_ARGS) return fail(st, "sptc_layout: bad modes or dims");
  for (int m = 0; m < order; ++m) {
    if (shift) shift[m] = L.shift[m];
    if (width) width[m] = L.width[m];
  }
  if (nfbits) *nfbits = L.nfbits;
  if (ncbits) *ncbits = L.ncbits;
  return st;
}

sptc_status sptc_contract(const sptc_coo* X, const sptc_coo* Y, int32_t nc, const int32_t* cx,
                          const int32_t* cy, const sptc_opts* opts, void* stream,
                          sptc_result** out) {
  if (!out) return fail(SPTC_ERR_BAD_ARGS, "out is NULL");
  *out = nullptr;
  sptc_status st;
  if ((st = check_coo(X, "X")) != SPTC_OK) return st;
  if ((st = check_coo(Y, "Y")) != SPTC_OK) return st;
  if (nc < 0 || nc > X->order || nc > Y->order || (nc && (!cx || !cy)))
    return fail(SPTC_ERR_BAD_ARGS, "bad contracted-mode count");
  if (!device_ok()) return fail(SPTC_ERR_NO_DEVICE, "no GPU");
  sptc_opts o;
  std::memset(&o, 0, sizeof(o));
  if (opts) o = *opts;
  Layout xl, yl;
  std::string lmsg;
  st = make_pair_layouts(X->order, X->dims, Y->order, Y->dims, nc, cx, cy, &xl, &yl, &lmsg);
  if (st != SPTC_OK) return fail(st, lmsg);
  sptc_result_s* res = new (std::nothrow) sptc_result_s();
  if (!res) return fail(SPTC_ERR_BAD_ARGS, "host allocation failed");
  std::memset(&res->info, 0, sizeof(res->info));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  try {
    Engine e(s, o);
    int px = 1, py = 1;
    if (!find_partition_num(X->nnz, X->order, Y->nnz, Y->order,
                            std::max(X->idx_bytes, Y->idx_bytes), e.budget_avail(), &px, &py)) {
      delete res;
      return fail(SPTC_ERR_OOM_BUDGET, "operands need more than 64 parts each");
    }
    if (px > 1 || py > 1) {
      OperandSplit xs, ys;
      split_operand(*X, xl, px, false, &xs);
      split_operand(*Y, yl, py, false, &ys);
      st = contract_partitioned(e, *X, *Y, xl, yl, xs, ys, res);
    } else {
      Operand ox, oy;
      e.record(0);
      {
        StagedCOO sx, sy;
        e.stage(*X, &sx);
        e.stage(*Y, &sy);
        e.record(1);
        if (o.check_indices) {
          st = e.check_indices(sx, X->dims, "X");
          if (st == SPTC_OK) st = e.check_indices(sy, Y->dims, "Y");
        }
        if (st == SPTC_OK) st = e.prepare_pair(sx, sy, xl, yl, &ox, &oy);
        e.record(2);
      }
      if (st == SPTC_OK) st = e.contract(ox, oy, res);
      if (st == SPTC_OK) std::memcpy(res->info.phase_ms, e.times.ms, sizeof(res->info.phase_ms));
    }
    if (st != SPTC_OK) {
      delete res;
      return fail(st, e.msg);
    }
    // operand preparation ran before contract() collected: add its launches
    e.collect_kernel_stats(res->kstats);
    if (o.verbose) {
      std::fprintf(stderr,
                   "sptc: ir_nnz %llu out_nnz %llu chunks %d | linearize+sort %.3f ms, "
                   "match %.3f ms, partition %.3f ms, accumulate %.3f ms, total %.3f ms\n",
                   (unsigned long long)res->info.ir_nnz, (unsigned long long)res->info.nnz,
                   res->info.nchunks, e.times.ms[1], e.times.ms[2], e.times.ms[3],
                   e.times.ms[4], e.times.ms[5]);
    }
  } catch (const HipError& he) {
    delete res;
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    delete res;
    return fail(SPTC_ERR_OOM_BUDGET, "host allocation failed");
  } catch (const std::exception& ex) {
    delete res;
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  *out = res;
  return SPTC_OK;
}

sptc_status sptc_plan(const sptc_coo* X, const sptc_coo* Y, int32_t nc, const int32_t* cx,
                      const int32_t* cy, const sptc_opts* opts, void* stream, sptc_plan_t** out) {
  if (!out) return fail(SPTC_ERR_BAD_ARGS, "out is NULL");
  *out = nullptr;
  sptc_status st;
  if ((st = check_coo(X, "X")) != SPTC_OK) return st;
  if ((st = check_coo(Y, "Y")) != SPTC_OK) return st;
  if (nc < 0 || nc > X->order || nc > Y->order || (nc && (!cx || !cy)))
    return fail(SPTC_ERR_BAD_ARGS, "bad contracted-mode count");
  if (!device_ok()) return fail(SPTC_ERR_NO_DEVICE, "no GPU");
  Layout xl, yl;
  std::string lmsg;
  st = make_pair_layouts(X->order, X->dims, Y->order, Y->dims, nc, cx, cy, &xl, &yl, &lmsg);
  if (st != SPTC_OK) return fail(st, lmsg);
  sptc_plan_s* p = new (std::nothrow) sptc_plan_s();
  if (!p) return fail(SPTC_ERR_BAD_ARGS, "host allocation failed");
  std::memset(&p->opts, 0, sizeof(p->opts));
  if (opts) p->opts = *opts;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  try {
    Engine e(s, p->opts);
    int px = 1, py = 1;
    if (!find_partition_num(X->nnz, X->order, Y->nnz, Y->order,
                            std::max(X->idx_bytes, Y->idx_bytes), e.budget_avail(), &px, &py)) {
      delete p;
      return fail(SPTC_ERR_OOM_BUDGET, "operands need more than 64 parts each");
    }
    if (px > 1 || py > 1) {
      // the plan keeps host-side parts (owned copies); pairs are prepared
      // on the device at each execution
      const auto t0 = std::chrono::steady_clock::now();
      p->split = true;
      p->xl = xl;
      p->yl = yl;
      p->xdims.assign(X->dims, X->dims + X->order);
      p->ydims.assign(Y->dims, Y->dims + Y->order);
      split_operand(*X, xl, px, true, &p->xs);
      split_operand(*Y, yl, py, true, &p->ys);
      p->prepare_ms =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    } else {
      e.record(0);
      {
        StagedCOO sx, sy;
        e.stage(*X, &sx);
        e.stage(*Y, &sy);
        sptc_status st2 = SPTC_OK;
        if (p->opts.check_indices) {
          st2 = e.check_indices(sx, X->dims, "X");
          if (st2 == SPTC_OK) st2 = e.check_indices(sy, Y->dims, "Y");
        }
        if (st2 == SPTC_OK) st2 = e.prepare_pair(sx, sy, xl, yl, &p->x, &p->y);
        if (st2 != SPTC_OK) {
          delete p;
          return fail(st2, e.msg);
        }
      }
      e.record(1);
      SPTC_HIP(hipStreamSynchronize(s));
      p->prepare_ms = e.elapsed_ms(0, 1);
    }
  } catch (const HipError& he) {
    delete p;
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    delete p;
    return fail(SPTC_ERR_OOM_BUDGET, "host allocation failed");
  } catch (const std::exception& ex) {
    delete p;
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  *out = p;
  return SPTC_OK;
}

sptc_status sptc_execute(sptc_plan_t* p, const sptc_opts* opts, void* stream,
                         sptc_result** out) {
  if (!p || !out) return fail(SPTC_ERR_BAD_ARGS, "NULL argument");
  *out = nullptr;
  sptc_result_s* res = new (std::nothrow) sptc_result_s();
  if (!res) return fail(SPTC_ERR_BAD_ARGS, "host allocation failed");
  std::memset(&res->info, 0, sizeof(res->info));
  try {
    Engine e(reinterpret_cast<hipStream_t>(stream), opts ? *opts : p->opts);
    sptc_status st;
    if (p->split) {
      sptc_coo X = p->xs.part[0].coo((int)p->xdims.size(), p->xdims.data());
      sptc_coo Y = p->ys.part[0].coo((int)p->ydims.size(), p->ydims.data());
      st = contract_partitioned(e, X, Y, p->xl, p->yl, p->xs, p->ys, res);
    } else {
      e.record(0);
      e.record(1);
      e.record(2);
      st = e.contract(p->x, p->y, res);
      if (st == SPTC_OK) std::memcpy(res->info.phase_ms, e.times.ms, sizeof(res->info.phase_ms));
    }
    if (st != SPTC_OK) {
      delete res;
      return fail(st, e.msg);
    }
  } catch (const HipError& he) {
    delete res;
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    delete res;
    return fail(SPTC_ERR_OOM_BUDGET, "host allocation failed");
  } catch (const std::exception& ex) {
    delete res;
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  *out = res;
  return SPTC_OK;
}

static sptc_status accumulate_forced(sptc_plan_t* p, const sptc_opts* opts, int32_t acc,
                                     void* stream, sptc_result** out) {
  if (!p || !out) return fail(SPTC_ERR_BAD_ARGS, "NULL argument");
  sptc_opts o = opts ? *opts : p->opts;
  o.dense_accumulator = acc;
  return sptc_execute(p, &o, stream, out);
}

sptc_status sptc_accumulate_esc(sptc_plan_t* p, const sptc_opts* opts, void* stream,
                                sptc_result** out) {
  return accumulate_forced(p, opts, -1, stream, out);
}

sptc_status sptc_accumulate_dense(sptc_plan_t* p, const sptc_opts* opts, void* stream,
                                  sptc_result** out) {
  return accumulate_forced(p, opts, 1, stream, out);
}

sptc_status sptc_plan_get_info(const sptc_plan_t* p, sptc_plan_info* info) {
  if (!p || !info) return fail(SPTC_ERR_BAD_ARGS, "NULL argument");
  std::memset(info, 0, sizeof(*info));
  const Layout& xl = p->split ? p->xl : p->x.lay;
  const Layout& yl = p->split ? p->yl : p->y.lay;
  info->x_nnz = p->split ? p->xs.nnz_lo.back() : p->x.nnz;
  info->y_nnz = p->split ? p->ys.nnz_lo.back() : p->y.nnz;
  info->x_nfbits = xl.nfbits;
  info->y_nfbits = yl.nfbits;
  info->ncbits = xl.ncbits;
  info->out_bits = xl.nfbits + yl.nfbits;
  info->out_order = xl.nf + yl.nf;
  info->device_bytes = p->x.keys.size() * 8 + p->x.vals.size() * 4 + p->y.keys.size() * 8 +
                       p->y.vals.size() * 4;
  info->prepare_ms = p->prepare_ms;
  info->x_parts = p->split ? p->xs.parts : 1;
  info->y_parts = p->split ? p->ys.parts : 1;
  return SPTC_OK;
}

sptc_status sptc_plan_get_partitions(const sptc_plan_t* p, int32_t role, int32_t max_parts,
                                     uint64_t* nnz_bounds, uint64_t* key_bounds,
                                     int32_t* nparts) {
  if (!p || !nparts || (role != 1 && role != 2))
    return fail(SPTC_ERR_BAD_ARGS, "sptc_plan_get_partitions: bad arguments");
  const OperandSplit* sp = p->split ? (role == 1 ? &p->xs : &p->ys) : nullptr;
  const int32_t n = sp ? sp->parts : 1;
  *nparts = n;
  if (max_parts < n) return SPTC_OK;
  const Layout& lay = role == 1 ? (p->split ? p->xl : p->x.lay) : (p->split ? p->yl : p->y.lay);
  const uint64_t nnz = role == 1 ? (p->split ? p->xs.nnz_lo.back() : p->x.nnz)
                                 : (p->split ? p->ys.nnz_lo.back() : p->y.nnz);
  for (int32_t k = 0; k <= n; ++k) {
    if (nnz_bounds) nnz_bounds[k] = sp ? sp->nnz_lo[k] : (k ? nnz : 0);
    if (key_bounds)
      key_bounds[k] = sp ? sp->key_lo[k] : (k ? (lay.nfbits >= 64 ? ~0ull : 1ull << lay.nfbits) : 0);
  }
  return SPTC_OK;
}

sptc_status sptc_find_partition_num(uint64_t nnz_x, int32_t order_x, uint64_t nnz_y,
                                    int32_t order_y, int32_t idx_bytes, uint64_t budget_bytes,
                                    int32_t* x_parts, int32_t* y_parts) {
  if (!x_parts || !y_parts || order_x < 1 || order_y < 1 || (idx_bytes != 4 && idx_bytes != 8))
    return fail(SPTC_ERR_BAD_ARGS, "sptc_find_partition_num: bad arguments");
  int px = 1, py = 1;
  if (!find_partition_num(nnz_x, order_x, nnz_y, order_y, idx_bytes, budget_bytes, &px, &py))
    return fail(SPTC_ERR_OOM_BUDGET, "operands need more than 64 parts each");
  *x_parts = px;
  *y_parts = py;
  return SPTC_OK;
}

sptc_status sptc_partition_plan(sptc_plan_t* p, uint64_t budget_bytes, int32_t max_parts,
                                uint64_t* row_bounds, uint64_t* prod_bounds, int32_t* nparts,
                                uint64_t* ir_nnz, void* stream) {
  if (!p || !nparts) return fail(SPTC_ERR_BAD_ARGS, "NULL argument");
  if (p->split)
    return fail(SPTC_ERR_BAD_ARGS,
                "the plan's operands are statically partitioned (sptc_plan_get_partitions)");
  sptc_opts o = p->opts;
  if (budget_bytes) o.mem_budget_bytes = budget_bytes;
  o.dense_accumulator = -1;
  sptc_result_s res;
  std::memset(&res.info, 0, sizeof(res.info));
  try {
    Engine e(reinterpret_cast<hipStream_t>(stream), o);
    e.partition_only = true;
    const sptc_status st = e.contract(p->x, p->y, &res);
    if (st != SPTC_OK) return fail(st, e.msg);
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "host allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  const int32_t n = res.part_rows.empty() ? 0 : (int32_t)(res.part_rows.size() - 1);
  *nparts = n;
  if (ir_nnz) *ir_nnz = res.info.ir_nnz;
  if (max_parts >= n) {
    for (int32_t i = 0; i <= n; ++i) {
      if (row_bounds) row_bounds[i] = n ? res.part_rows[i] : 0;
      if (prod_bounds) prod_bounds[i] = n ? res.part_prods[i] : 0;
    }
  }
  return SPTC_OK;
}

void sptc_plan_destroy(sptc_plan_t* p) { delete p; }

sptc_status sptc_result_get_info(const sptc_result* r, sptc_result_info* info) {
  if (!r || !info) return fail(SPTC_ERR_BAD_ARGS, "NULL argument");
  *info = r->info;
  return SPTC_OK;
}

sptc_status sptc_result_get_chunk(const sptc_result* r, int32_t i, uint64_t* nnz,
                                  const uint32_t** row, const uint64_t** col,
                                  const float** val) {
  if (!r || i < 0 || (size_t)i >= r->chunks.size())
    return fail(SPTC_ERR_BAD_ARGS, "bad chunk index");
  const OutChunk& c = r->chunks[i];
  if (nnz) *nnz = c.nnz;
  if (row) *row = c.R();
  if (col) *col = c.C();
  if (val) *val = c.V();
  return SPTC_OK;
}

sptc_status sptc_result_copy_to_host(const sptc_result* r, uint64_t* keys, uint64_t* xfree,
                                     uint64_t* yfree, float* vals) {
  if (!r) return fail(SPTC_ERR_BAD_ARGS, "NULL result");
  if (keys && r->info.out_bits > 64)
    return fail(SPTC_ERR_KEY_OVERFLOW, "packed output key needs more than 64 bits");
  try {
    std::vector<uint64_t> rk(r->info.nrows);
    if (r->info.nrows)
      SPTC_HIP(hipMemcpy(rk.data(), r->RK(), r->info.nrows * 8, hipMemcpyDeviceToHost));
    uint64_t off = 0;
    std::vector<uint32_t> row;
    std::vector<uint64_t> col;
    for (const OutChunk& c : r->chunks) {
      if (!c.nnz) continue;
      row.resize(c.nnz);
      col.resize(c.nnz);
      SPTC_HIP(hipMemcpy(row.data(), c.R(), c.nnz * 4, hipMemcpyDeviceToHost));
      SPTC_HIP(hipMemcpy(col.data(), c.C(), c.nnz * 8, hipMemcpyDeviceToHost));
      if (vals)
        SPTC_HIP(hipMemcpy(vals + off, c.V(), c.nnz * 4, hipMemcpyDeviceToHost));
      const int ynf = r->info.y_nfbits;
      for (uint64_t i = 0; i < c.nnz; ++i) {
        const uint64_t xf = rk[row[i]];
        if (xfree) xfree[off + i] = xf;
        if (yfree) yfree[off + i] = col[i];
        if (keys) keys[off + i] = shl64(xf, ynf) | col[i];
      }
      off += c.nnz;
    }
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_result_row_keys(const sptc_result* r, uint64_t* row_keys) {
  if (!r || (!row_keys && r->info.nrows)) return fail(SPTC_ERR_BAD_ARGS, "NULL argument");
  try {
    if (r->info.nrows)
      SPTC_HIP(hipMemcpy(row_keys, r->RK(), r->info.nrows * 8, hipMemcpyDeviceToHost));
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_result_row_range(const sptc_result* r, uint64_t row, uint64_t* begin,
                                  uint64_t* end) {
  if (!r || !begin || !end) return fail(SPTC_ERR_BAD_ARGS, "NULL argument");
  try {
    uint64_t off = 0;
    for (const OutChunk& c : r->chunks) {
      if (row >= c.r0 && row < c.r1 && c.nnz) {
        // lower/upper bound of `row` in the chunk's sorted row array
        auto at = [&](uint64_t i) {
          uint32_t v;
          SPTC_HIP(hipMemcpy(&v, c.R() + i, 4, hipMemcpyDeviceToHost));
          return (uint64_t)v;
        };
        uint64_t lo = 0, hi = c.nnz;
        while (lo < hi) {
          uint64_t mid = (lo + hi) / 2;
          if (at(mid) < row) lo = mid + 1; else hi = mid;
        }
        uint64_t b = lo;
        hi = c.nnz;
        while (lo < hi) {
          uint64_t mid = (lo + hi) / 2;
          if (at(mid) <= row) lo = mid + 1; else hi = mid;
        }
        *begin = off + b;
        *end = off + lo;
        return SPTC_OK;
      }
      off += c.nnz;
    }
    *begin = *end = off;
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_result_copy_range(const sptc_result* r, uint64_t begin, uint64_t end,
                                   uint64_t* xfree, uint64_t* yfree, float* vals) {
  if (!r || begin > end || end > r->info.nnz) return fail(SPTC_ERR_BAD_ARGS, "bad range");
  try {
    uint64_t off = 0, o = 0;
    std::vector<uint32_t> row;
    for (const OutChunk& c : r->chunks) {
      const uint64_t cb = std::max(begin, off), ce = std::min(end, off + c.nnz);
      if (cb < ce) {
        const uint64_t n = ce - cb, lo = cb - off;
        if (xfree) {
          row.resize(n);
          SPTC_HIP(hipMemcpy(row.data(), c.R() + lo, n * 4, hipMemcpyDeviceToHost));
          // rows are ascending within the range: fetch that slice of row_key once
          const uint64_t rlo = row.front(), rhi = row.back();
          std::vector<uint64_t> rk(rhi - rlo + 1);
          SPTC_HIP(hipMemcpy(rk.data(), r->RK() + rlo, rk.size() * 8,
                             hipMemcpyDeviceToHost));
          for (uint64_t i = 0; i < n; ++i) xfree[o + i] = rk[row[i] - rlo];
        }
        if (yfree)
          SPTC_HIP(hipMemcpy(yfree + o, c.C() + lo, n * 8, hipMemcpyDeviceToHost));
        if (vals) SPTC_HIP(hipMemcpy(vals + o, c.V() + lo, n * 4, hipMemcpyDeviceToHost));
        o += n;
      }
      off += c.nnz;
    }
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

// Output field layout of a result for the decode kernel; false when a free
// mode is wider than 32 bits (it would not fit the uint32 outputs).
static bool decode_args(const sptc_result* r, DecodeArgs* a) {
  std::memset(a, 0, sizeof(*a));
  const Layout& xl = r->xl;
  const Layout& yl = r->yl;
  a->nx = xl.nf;
  a->ny = yl.nf;
  int o = 0;
  for (int i = 0; i < xl.nf; ++i, ++o) {
    const int m = xl.fmodes[i];
    if (xl.width[m] > 32) return false;
    a->shift[o] = xl.shift[m] - xl.ncbits;
    a->mask[o] = low_mask(xl.width[m]);
  }
  for (int i = 0; i < yl.nf; ++i, ++o) {
    const int m = yl.fmodes[i];
    if (yl.width[m] > 32) return false;
    a->shift[o] = yl.shift[m];
    a->mask[o] = low_mask(yl.width[m]);
  }
  return true;
}

sptc_status sptc_result_decode_device(const sptc_result* r, uint32_t* const* mode_idx,
                                      float* vals, void* stream) {
  if (!r || (!mode_idx && r->info.out_order)) return fail(SPTC_ERR_BAD_ARGS, "NULL argument");
  DecodeArgs a;
  if (!decode_args(r, &a))
    return fail(SPTC_ERR_KEY_OVERFLOW, "an output mode is wider than 32 bits");
  for (int m = 0; m < a.nx + a.ny; ++m) {
    if (!mode_idx[m] && r->info.nnz) return fail(SPTC_ERR_BAD_ARGS, "NULL mode array");
    a.out[m] = mode_idx[m];
  }
  a.vals = vals;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  try {
    uint64_t off = 0;
    for (const OutChunk& c : r->chunks) {
      launch_decode(c.R(), c.C(), c.V(), c.nnz, r->RK(), a, off, s);
      off += c.nnz;
    }
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_result_decode(const sptc_result* r, uint32_t* const* mode_idx, float* vals) {
  if (!r || (!mode_idx && r->info.out_order)) return fail(SPTC_ERR_BAD_ARGS, "NULL argument");
  DecodeArgs a;
  if (!decode_args(r, &a))
    return fail(SPTC_ERR_KEY_OVERFLOW, "an output mode is wider than 32 bits");
  const int order = a.nx + a.ny;
  try {
    // decode each chunk on the device in pieces, then copy the pieces out
    const uint64_t piece = 1ull << 26;
    uint64_t cap = 0;
    for (const OutChunk& c : r->chunks) cap = std::max(cap, std::min(c.nnz, piece));
    DevBuf<uint32_t> dev(cap * (uint64_t)std::max(order, 1));
    DevBuf<float> dval(vals ? cap : 0);
    for (int m = 0; m < order; ++m) a.out[m] = dev.get() + (uint64_t)m * cap;
    a.vals = vals ? dval.get() : nullptr;
    uint64_t off = 0;
    for (const OutChunk& c : r->chunks) {
      for (uint64_t b = 0; b < c.nnz; b += piece) {
        const uint64_t n = std::min(piece, c.nnz - b);
        launch_decode(c.R() + b, c.C() + b, c.V() + b, n, r->RK(), a, 0, nullptr);
        for (int m = 0; m < order; ++m)
          SPTC_HIP(hipMemcpy(mode_idx[m] + off, a.out[m], n * 4, hipMemcpyDeviceToHost));
        if (vals) SPTC_HIP(hipMemcpy(vals + off, dval.get(), n * 4, hipMemcpyDeviceToHost));
        off += n;
      }
    }
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_result_as_coo(sptc_result* r, void* stream, sptc_coo* out) {
  if (!r || !out) return fail(SPTC_ERR_BAD_ARGS, "NULL argument");
  if (r->info.nnz && r->chunks.empty())
    return fail(SPTC_ERR_BAD_ARGS, "a count-only result holds no entries");
  DecodeArgs a;
  if (!decode_args(r, &a))
    return fail(SPTC_ERR_KEY_OVERFLOW, "an output mode is wider than 32 bits");
  const int order = a.nx + a.ny;
  if (order < 1) return fail(SPTC_ERR_BAD_ARGS, "the output has no modes");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  try {
    if (r->coo_ptr.empty()) {
      const uint64_t n = r->info.nnz;
      r->coo_idx.resize(order);
      r->coo_ptr.resize(order);
      for (int m = 0; m < order; ++m) {
        r->coo_idx[m].reset(n ? n : 1);
        a.out[m] = r->coo_idx[m].get();
        r->coo_ptr[m] = r->coo_idx[m].get();
      }
      r->coo_val.reset(n ? n : 1);
      a.vals = r->coo_val.get();
      uint64_t off = 0;
      for (const OutChunk& c : r->chunks) {
        launch_decode(c.R(), c.C(), c.V(), c.nnz, r->RK(), a, off, s);
        off += c.nnz;
      }
      r->coo_dims.clear();
      for (int i = 0; i < r->xl.nf; ++i) r->coo_dims.push_back(r->xl.dims[r->xl.fmodes[i]]);
      for (int i = 0; i < r->yl.nf; ++i) r->coo_dims.push_back(r->yl.dims[r->yl.fmodes[i]]);
      SPTC_HIP(hipStreamSynchronize(s));
    }
  } catch (const HipError& he) {
    r->coo_ptr.clear();
    return from_hip(he);
  } catch (const std::exception& ex) {
    r->coo_ptr.clear();
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  std::memset(out, 0, sizeof(*out));
  out->order = order;
  out->nnz = r->info.nnz;
  out->dims = r->coo_dims.data();
  out->idx = r->coo_ptr.data();
  out->idx_bytes = 4;
  out->val = r->coo_val.get();
  out->on_device = 1;
  return SPTC_OK;
}

sptc_status sptc_result_verify(const sptc_result* r, double* value_sum, uint64_t* violations,
                               void* stream) {
  if (!r) return fail(SPTC_ERR_BAD_ARGS, "NULL result");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  try {
    double sum = 0.0;
    uint64_t bad = 0;
    DevBuf<VerifyPart> part(kVerifyMaxBlocks);
    std::vector<VerifyPart> host(kVerifyMaxBlocks);
    for (const OutChunk& c : r->chunks) {
      if (!c.nnz) continue;
      launch_verify(c.R(), c.C(), c.V(), c.nnz, c.r0, c.r1, part.get(), s);
      const uint32_t g = verify_grid(c.nnz);
      SPTC_HIP(hipMemcpyAsync(host.data(), part.get(), g * sizeof(VerifyPart),
                              hipMemcpyDeviceToHost, s));
      SPTC_HIP(hipStreamSynchronize(s));
      for (uint32_t i = 0; i < g; ++i) {
        sum += host[i].sum;
        bad += host[i].bad;
      }
    }
    if (value_sum) *value_sum = sum;
    if (violations) *violations = bad;
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

void sptc_result_free(sptc_result* r) { delete r; }

sptc_status sptc_result_kernel_stats(const sptc_result* r, int32_t cls, uint64_t* launches,
                                     double* total_ms, double* alg_bytes) {
  if (!r || cls < 0 || cls >= SPTC_K_COUNT) return fail(SPTC_ERR_BAD_ARGS, "bad kernel class");
  if (launches) *launches = r->kstats[cls].launches;
  if (total_ms) *total_ms = r->kstats[cls].ms;
  if (alg_bytes) *alg_bytes = r->kstats[cls].bytes;
  return SPTC_OK;
}

void sptc_pool_stats(uint64_t* live_bytes, uint64_t* reserved_bytes, uint64_t* hip_mallocs) {
  DevicePool& p = DevicePool::get();
  if (live_bytes) *live_bytes = p.live_bytes();
  if (reserved_bytes) *reserved_bytes = p.reserved_bytes();
  if (hip_mallocs) *hip_mallocs = p.hip_mallocs();
}

// ---- per-phase entry points ---------------------------------------------

sptc_status sptc_linearize(int32_t order, uint64_t nnz, const void* const* idx, int32_t idx_bytes,
                           const int32_t* shift, const int32_t* width, uint64_t* keys,
                           void* stream) {
  if (order < 1 || order > SPTC_MAX_ORDER || !shift || (nnz && (!idx || !keys)) ||
      (idx_bytes != 4 && idx_bytes != 8))
    return fail(SPTC_ERR_BAD_ARGS, "sptc_linearize: bad arguments");
  int total = 0;
  for (int m = 0; m < order; ++m) {
    if (shift[m] < 0 || shift[m] > 63) return fail(SPTC_ERR_BAD_ARGS, "shift out of range");
    if (width && (width[m] < 0 || width[m] > 64))
      return fail(SPTC_ERR_BAD_ARGS, "width out of range");
    if (width) total = std::max(total, shift[m] + width[m]);
  }
  if (total > 64) return fail(SPTC_ERR_KEY_OVERFLOW, "fields exceed 64 bits");
  if (!device_ok()) return fail(SPTC_ERR_NO_DEVICE, "no GPU");
  try {
    LinArgs a;
    std::memset(&a, 0, sizeof(a));
    a.order = order;
    for (int m = 0; m < order; ++m) {
      a.idx[m] = idx[m];
      a.shift[m] = shift[m];
      a.width[m] = width ? width[m] : 64;
    }
    launch_linearize(a, idx_bytes, nnz, keys, nullptr, 0, 8,
                     reinterpret_cast<hipStream_t>(stream));
    SPTC_HIP(hipGetLastError());
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_sort_pairs(uint64_t* keys, float* vals, uint64_t n, int32_t nbits,
                            int32_t radix_bits, void* stream) {
  if (nbits < 0 || nbits > 64 || (n && (!keys || !vals)) || n >= (1ull << 32))
    return fail(SPTC_ERR_BAD_ARGS, "sptc_sort_pairs: bad arguments (n must be below 2^32)");
  if (radix_bits != 0 && !radix_bits_supported(radix_bits))
    return fail(SPTC_ERR_BAD_ARGS, "radix_bits must be 0 or 8..11");
  if (!device_ok()) return fail(SPTC_ERR_NO_DEVICE, "no GPU");
  const int rb = radix_bits ? radix_bits : choose_radix_bits(nbits);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  try {
    SortWorkspace ws;
    ws.ensure(n, rb, s);
    const int passes = radix_passes(nbits, rb);
    ws.prepare(passes, s);
    if (sort_needs_histogram()) histogram(ws, keys, n, passes, s);
    if (sort_after_histogram(ws, keys, vals, n, nbits, s)) {
      SPTC_HIP(hipMemcpyAsync(keys, ws.alt_keys.get(), n * 8, hipMemcpyDeviceToDevice, s));
      SPTC_HIP(hipMemcpyAsync(vals, ws.alt_vals.get(), n * 4, hipMemcpyDeviceToDevice, s));
    }
    uint32_t err = 0;
    SPTC_HIP(hipMemcpyAsync(&err, ws.err(), 4, hipMemcpyDeviceToHost, s));
    SPTC_HIP(hipStreamSynchronize(s));
    if (err) return fail(SPTC_ERR_INTERNAL, "look-back spin limit exceeded");
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_index_match(const uint64_t* xkeys, uint64_t nx, const uint64_t* ykeys,
                             uint64_t ny, int32_t ncbits, int32_t y_nfbits, uint64_t* mpos,
                             uint64_t* mcnt, uint64_t* prefix, uint64_t* ir_nnz, void* stream) {
  if (ncbits < 0 || ncbits > 64 || y_nfbits < 0 || y_nfbits > 64 || !prefix ||
      (nx && (!xkeys || !mpos || !mcnt)) || (ny && !ykeys))
    return fail(SPTC_ERR_BAD_ARGS, "sptc_index_match: bad arguments");
  if (!device_ok()) return fail(SPTC_ERR_NO_DEVICE, "no GPU");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  try {
    DevBuf<uint64_t> tlo, thi, ukey, ustart;
    YTable tab;
    std::memset(&tab, 0, sizeof(tab));
    if (ncbits <= kDirectBits) {
      const uint64_t K = 1ull << ncbits;
      tlo.reset(K);
      thi.reset(K);
      fill_u64(tlo.get(), 0, K, s);
      fill_u64(thi.get(), 0, K, s);
      launch_ytable_direct(ykeys, ny, y_nfbits, K, tlo.get(), thi.get(), s);
      tab.direct = 1;
      tab.lo = tlo.get();
      tab.hi = thi.get();
    } else {
      ukey.reset(ny + 1);
      ustart.reset(ny + 1);
      DevBuf<U2> tmp(scan_tmp_count(ny) + 1);
      run_yunique(ykeys, ny, y_nfbits, ukey.get(), ustart.get(), tmp.get(), s);
      U2 t;
      SPTC_HIP(hipMemcpyAsync(&t, tmp.get() + scan_tmp_count(ny), sizeof(U2),
                              hipMemcpyDeviceToHost, s));
      SPTC_HIP(hipStreamSynchronize(s));
      tab.direct = 0;
      tab.ukey = ukey.get();
      tab.ustart = ustart.get();
      tab.U = t.a;
    }
    MatchOut mo;
    std::memset(&mo, 0, sizeof(mo));
    mo.mpos = mpos;
    mo.mcnt = mcnt;
    mo.prefix = prefix;
    DevBuf<U2> tmp(scan_tmp_count(nx) + 1);
    run_match(xkeys, nullptr, nx, ncbits, tab, mo, 1, tmp.get(), s);
    U2 t;
    SPTC_HIP(hipMemcpyAsync(&t, tmp.get() + scan_tmp_count(nx), sizeof(U2),
                            hipMemcpyDeviceToHost, s));
    if (nx == 0) fill_u64(prefix, 0, 1, s);
    SPTC_HIP(hipStreamSynchronize(s));
    if (ir_nnz) *ir_nnz = t.a;
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_exclusive_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n,
                                    void* stream) {
  if (!out || (n && !in)) return fail(SPTC_ERR_BAD_ARGS, "sptc_exclusive_scan_u64: NULL");
  if (!device_ok()) return fail(SPTC_ERR_NO_DEVICE, "no GPU");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  try {
    DevBuf<U2> tmp(scan_tmp_count(n) + 1);
    if (n == 0) fill_u64(out, 0, 1, s);
    else run_exclusive_scan(in, out, n, tmp.get(), s);
    SPTC_HIP(hipStreamSynchronize(s));
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_device_info(int32_t* num_cus, uint64_t* free_bytes, uint64_t* total_bytes,
                             double* peak_hbm_gbs) {
  if (!device_ok()) return fail(SPTC_ERR_NO_DEVICE, "no GPU");
  try {
    const DeviceInfo& d = device_info();
    if (num_cus) *num_cus = d.num_cus;
    if (total_bytes) *total_bytes = d.total_bytes;
    if (peak_hbm_gbs) *peak_hbm_gbs = d.peak_hbm_gbs;
    if (free_bytes) *free_bytes = device_free_bytes();
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

void sptc_trim_cache(void) { DevicePool::get().trim(); }

// ---- batched contractions -------------------------------------------------

sptc_status sptc_batch_create(int32_t n, const sptc_coo* X, const sptc_coo* Y,
                              const int32_t* nc, const int32_t* const* cx,
                              const int32_t* const* cy, const sptc_opts* opts, void* stream,
                              sptc_batch** out) {
  if (!out) return fail(SPTC_ERR_BAD_ARGS, "out is NULL");
  *out = nullptr;
  if (n < 1 || n > kBatchMaxInstances || !X || !Y || !nc || !cx || !cy)
    return fail(SPTC_ERR_BAD_ARGS, "bad batch arguments (1 <= n <= 4096, non-NULL arrays)");
  for (int i = 0; i < n; ++i) {
    sptc_status st;
    if ((st = check_coo(&X[i], "X")) != SPTC_OK) return st;
    if ((st = check_coo(&Y[i], "Y")) != SPTC_OK) return st;
  }
  if (!device_ok()) return fail(SPTC_ERR_NO_DEVICE, "no GPU");
  sptc_opts o;
  std::memset(&o, 0, sizeof(o));
  if (opts) o = *opts;
  sptc_batch_s* b = new (std::nothrow) sptc_batch_s();
  if (!b) return fail(SPTC_ERR_BAD_ARGS, "host allocation failed");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  try {
    sptc_status st = b->e.plan(n, X, Y, nc, cx, cy, o, s);
    if (st != SPTC_OK) {
      std::string m = b->e.msg;
      delete b;
      return fail(st, m);
    }
    b->e.ktimer().collect(b->kstats, SPTC_K_COUNT);
  } catch (const HipError& he) {
    delete b;
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    delete b;
    return fail(SPTC_ERR_OOM_BUDGET, "host allocation failed");
  } catch (const std::exception& ex) {
    delete b;
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  *out = b;
  return SPTC_OK;
}

sptc_status sptc_batch_run(sptc_batch* b, int32_t use_graph, void* stream) {
  if (!b) return fail(SPTC_ERR_BAD_ARGS, "NULL batch");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  try {
    if (use_graph) return b->e.launch_graph(s);
    b->e.enqueue(s);
    SPTC_HIP(hipGetLastError());
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_batch_sync(sptc_batch* b, void* stream) {
  if (!b) return fail(SPTC_ERR_BAD_ARGS, "NULL batch");
  try {
    sptc_status st = b->e.sync(reinterpret_cast<hipStream_t>(stream));
    if (st != SPTC_OK) return fail(st, b->e.msg);
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_batch_get_info(const sptc_batch* b, sptc_batch_info* info) {
  if (!b || !info) return fail(SPTC_ERR_BAD_ARGS, "NULL argument");
  *info = b->e.info;
  return SPTC_OK;
}

sptc_status sptc_batch_result(sptc_batch* b, int32_t i, sptc_result** out) {
  if (!b || !out || i < 0 || i >= b->e.info.instances)
    return fail(SPTC_ERR_BAD_ARGS, "bad instance index");
  *out = nullptr;
  sptc_result_s* r = new (std::nothrow) sptc_result_s();
  if (!r) return fail(SPTC_ERR_BAD_ARGS, "host allocation failed");
  try {
    sptc_status st = b->e.instance_result(i, r);
    if (st != SPTC_OK) {
      delete r;
      return fail(st, b->e.msg);
    }
  } catch (const HipError& he) {
    delete r;
    return from_hip(he);
  } catch (const std::exception& ex) {
    delete r;
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  *out = r;
  return SPTC_OK;
}

sptc_status sptc_batch_kernel_stats(const sptc_batch* b, int32_t cls, uint64_t* launches,
                                    double* total_ms, double* alg_bytes) {
  if (!b || cls < 0 || cls >= SPTC_K_COUNT) return fail(SPTC_ERR_BAD_ARGS, "bad kernel class");
  if (launches) *launches = b->kstats[cls].launches;
  if (total_ms) *total_ms = b->kstats[cls].ms;
  if (alg_bytes) *alg_bytes = b->kstats[cls].bytes;
  return SPTC_OK;
}

void sptc_batch_free(sptc_batch* b) { delete b; }

// ---- serving ----------------------------------------------------------------

sptc_status sptc_server_create(const sptc_server_opts* opts, sptc_server** out) {
  if (!out) return fail(SPTC_ERR_BAD_ARGS, "out is NULL");
  *out = nullptr;
  if (!device_ok()) return fail(SPTC_ERR_NO_DEVICE, "no GPU");
  sptc_server_opts o;
  std::memset(&o, 0, sizeof(o));
  if (opts) o = *opts;
  if (o.workers > 16 || o.max_batch < 0 || o.workers < 0)
    return fail(SPTC_ERR_BAD_ARGS, "workers must be 0..16, max_batch >= 0");
  try {
    auto* s = new sptc_server_s();
    s->srv = std::make_unique<Server>(o);
    *out = s;
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_server_add_tensor(sptc_server* s, const sptc_coo* t, int32_t* id) {
  if (!s || !id) return fail(SPTC_ERR_BAD_ARGS, "NULL argument");
  sptc_status st = check_coo(t, "tensor");
  if (st != SPTC_OK) return st;
  try {
    std::string m;
    st = s->srv->add_tensor(*t, id, &m);
    if (st != SPTC_OK) return fail(st, m);
  } catch (const HipError& he) {
    return from_hip(he);
  } catch (const std::bad_alloc&) {
    return fail(SPTC_ERR_OOM_BUDGET, "allocation failed");
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_server_submit(sptc_server* s, int32_t x, int32_t y, int32_t nc,
                               const int32_t* cx, const int32_t* cy, uint64_t* request) {
  if (!s || !request || (nc && (!cx || !cy))) return fail(SPTC_ERR_BAD_ARGS, "NULL argument");
  std::string m;
  try {
    sptc_status st = s->srv->submit(x, y, nc, cx, cy, request, &m);
    if (st != SPTC_OK) return fail(st, m);
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_server_wait(sptc_server* s, uint64_t request, sptc_result** out,
                             sptc_request_info* info) {
  if (!s) return fail(SPTC_ERR_BAD_ARGS, "NULL server");
  if (out) *out = nullptr;
  std::string m;
  try {
    sptc_status st = s->srv->wait(request, out, info, &m);
    if (st != SPTC_OK) return fail(st, m);
  } catch (const std::exception& ex) {
    return fail(SPTC_ERR_INTERNAL, ex.what());
  }
  return SPTC_OK;
}

sptc_status sptc_server_get_stats(sptc_server* s, sptc_server_stats* stats) {
  if (!s || !stats) return fail(SPTC_ERR_BAD_ARGS, "NULL argument");
  s->srv->stats(stats);
  return SPTC_OK;
}

void sptc_server_free(sptc_server* s) { delete s; }

}  // extern "C"
