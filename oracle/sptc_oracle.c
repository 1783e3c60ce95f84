/*
 * CPU oracle for the SpTC path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, the smoke check and the benchmark's CPU-baseline leg may load
 * this library, and only as the checker.  The product (libsptc_hip.so) never
 * links or calls it.
 *
 * It restates the reference's algorithm on the host, function by function:
 *   orc_mode_bits        ComputeNBits / compute_nbtis  (sparse_tensor.tpp:39-47,
 *                        tensor_manager.tpp:90-133); bits(1) := 0 (the
 *                        reference's clz(0) is undefined there, :109)
 *   orc_layout           get_fmode + setup_slitom + create_mask
 *                        (tensor_manager.tpp:63-88, 135-210)
 *   orc_pdep             common::pdep (bitops.hpp:9-13), portable bit loop
 *   orc_linearize        ConvertTensor (tensor_manager.tpp:212-298)
 *   orc_sort_pairs       sort_64 / sort_tensor_cpu (sort.cuh:39-159,
 *                        tensor_manager.tpp:408-472) as a STABLE sort
 *   orc_index_match      compute_result_size_cuda + InclusiveSum
 *                        (indexMatch.cuh:20-118, 231-287)
 *   orc_esc              contraction_esc_kernel + SortPairs + ReduceByKey
 *                        (contraction.cuh:100-141, 364-599)
 *   orc_dense_nnz        Dense_Contraction's nonzero count
 *                        (contraction.cuh:58-98, 247-362)
 *   orc_partition        TensorManager::Partition (tensor_manager.tpp:633-749)
 *   orc_dynamic_partition Dynamic_partition (contraction.cuh:143-245)
 *   orc_join_fp64        independent FP64 hash-join SpTC (no keys, no sort of
 *                        products): the value oracle.
 *
 * Pins (tests/test_oracle.py): orc_join_fp64 and orc_esc against a dense
 * numpy.tensordot on tiny tensors; orc_linearize against the closed form
 * sum(idx << shift) and a BMI2 _pdep_u64 twin (tests/golden/pdep_kat.json);
 * orc_sort_pairs against numpy's stable argsort; orc_index_match against a
 * brute-force scan and the identity ir = sum_k cX_k * cY_k; partition
 * functions against their invariants (coverage, no split row, budget fit).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_ORDER 8

/* ------------------------------------------------------------------ */
/* Layout                                                              */

int orc_mode_bits(uint64_t d) {
  if (d <= 1) return 0;
  return 64 - __builtin_clzll(d - 1);
}

/*
 * role 1 = X (key = f << ncbits | c), role 2 = Y (key = c << nfbits | f).
 * cw[i] overrides the width of contracted mode cmodes[i] when >= 0 (so X and
 * Y share one contract-field layout; with equal dims this is the reference's
 * layout exactly).  Outputs: shift[m] = bit offset of mode m inside the key,
 * width[m] = its width, fpos (free modes ascending), *nf, *nc.
 */
int orc_layout(int order, const uint64_t* dims, int nc, const int* cmodes,
               const int* cw, int role, int* shift, int* width, int* fpos,
               int* nfbits, int* ncbits) {
  int is_c[ORC_MAX_ORDER] = {0};
  int f = 0, nfb = 0, ncb = 0;
  if (order > ORC_MAX_ORDER || nc < 0 || nc > order) return -1;
  for (int i = 0; i < nc; ++i) is_c[cmodes[i]] = 1;
  for (int m = 0; m < order; ++m) {
    width[m] = orc_mode_bits(dims[m]);
    if (!is_c[m]) fpos[f++] = m;
  }
  for (int i = 0; i < nc; ++i)
    if (cw && cw[i] >= 0) width[cmodes[i]] = cw[i];
  /* Free modes: last free mode at bit 0 (tensor_manager.tpp:153-161). */
  int s = 0;
  for (int i = f - 1; i >= 0; --i) {
    shift[fpos[i]] = s;
    s += width[fpos[i]];
  }
  nfb = s;
  /* Contract modes in user order, last at bit 0 (:163-174). */
  s = 0;
  for (int i = nc - 1; i >= 0; --i) {
    shift[cmodes[i]] = s;
    s += width[cmodes[i]];
  }
  ncb = s;
  for (int m = 0; m < order; ++m) {
    if (is_c[m]) {
      if (role == 2) shift[m] += nfb;
    } else {
      if (role == 1) shift[m] += ncb;
    }
  }
  *nfbits = nfb;
  *ncbits = ncb;
  return 0;
}

uint64_t orc_pdep(uint64_t x, uint64_t mask) {
  uint64_t r = 0;
  for (uint64_t bb = 1; mask; bb += bb) {
    if (x & bb) r |= mask & (0 - mask);
    mask &= mask - 1;
  }
  return r;
}

/*
 * Key for every nonzero: each mode index is deposited into its field with
 * pdep (as ConvertTensor does), fields OR-ed together.  idx is SoA uint32.
 */
int orc_linearize(int order, uint64_t nnz, const uint32_t* const* idx,
                  const int* shift, const int* width, uint64_t* keys) {
  for (uint64_t i = 0; i < nnz; ++i) {
    uint64_t k = 0;
    for (int m = 0; m < order; ++m) {
      if (width[m] == 0) continue;
      uint64_t mask = (width[m] >= 64 ? ~0ull : ((1ull << width[m]) - 1)) << shift[m];
      k |= orc_pdep(idx[m][i], mask);
    }
    keys[i] = k;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Stable sort of (key, value) pairs (bottom-up merge sort).           */

int orc_sort_pairs(uint64_t* keys, float* vals, uint64_t n) {
  if (n < 2) return 0;
  uint64_t* tk = (uint64_t*)malloc(n * sizeof(uint64_t));
  float* tv = (float*)malloc(n * sizeof(float));
  if (!tk || !tv) { free(tk); free(tv); return -1; }
  uint64_t *sk = keys, *dk = tk;
  float *sv = vals, *dv = tv;
  for (uint64_t w = 1; w < n; w *= 2) {
    for (uint64_t lo = 0; lo < n; lo += 2 * w) {
      uint64_t mid = lo + w < n ? lo + w : n;
      uint64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
      uint64_t a = lo, b = mid, o = lo;
      while (a < mid && b < hi) {
        if (sk[b] < sk[a]) { dk[o] = sk[b]; dv[o++] = sv[b++]; }
        else { dk[o] = sk[a]; dv[o++] = sv[a++]; }
      }
      while (a < mid) { dk[o] = sk[a]; dv[o++] = sv[a++]; }
      while (b < hi) { dk[o] = sk[b]; dv[o++] = sv[b++]; }
    }
    uint64_t* t = sk; sk = dk; dk = t;
    float* u = sv; sv = dv; dv = u;
  }
  if (sk != keys) {
    memcpy(keys, sk, n * sizeof(uint64_t));
    memcpy(vals, sv, n * sizeof(float));
  }
  free(tk);
  free(tv);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Index match: lower_bound of X's contract key over Y's (key >> nfbits),
 * then the run length; exclusive prefix with prefix[0] = 0 and
 * prefix[nx] = ir_nnz.  Returns ir_nnz.                                */

uint64_t orc_index_match(const uint64_t* xkeys, uint64_t nx,
                         const uint64_t* ykeys, uint64_t ny, int ncbits,
                         int y_nfbits, uint64_t* mpos, uint64_t* mcnt,
                         uint64_t* prefix) {
  uint64_t cmask = ncbits >= 64 ? ~0ull : ((1ull << ncbits) - 1);
  uint64_t acc = 0;
  for (uint64_t x = 0; x < nx; ++x) {
    uint64_t c = xkeys[x] & cmask;
    uint64_t lo = 0, hi = ny;
    while (lo < hi) {
      uint64_t mid = lo + (hi - lo) / 2;
      if ((ykeys[mid] >> y_nfbits) < c) lo = mid + 1; else hi = mid;
    }
    uint64_t cnt = 0;
    while (lo + cnt < ny && (ykeys[lo + cnt] >> y_nfbits) == c) ++cnt;
    mpos[x] = cnt ? lo : 0;
    mcnt[x] = cnt;
    prefix[x] = acc;
    acc += cnt;
  }
  prefix[nx] = acc;
  return acc;
}

/* ------------------------------------------------------------------ */
/* ESC accumulation over X rows [x_begin, x_end) (whole rows).
 * Product p of X nonzero x and Y nonzero mpos[x]+o has output key
 * (xf, yf) = (xkey >> ncbits, ykey & (2^y_nfbits - 1)) and value
 * vx*vy rounded to FP32 (contraction.cuh:132-136).  Products are stably
 * sorted by (xf, yf) and equal keys reduced: out_f32 with a sequential FP32
 * sum in sorted order, out_f64 with an FP64 sum of the exact products.
 * Output buffers are malloc'ed; free with orc_free.  Returns out nnz.    */

typedef struct { uint64_t xf, yf; float v32; double v64; } orc_prod;

static int prod_cmp(const void* a, const void* b) {
  const orc_prod* p = (const orc_prod*)a;
  const orc_prod* q = (const orc_prod*)b;
  if (p->xf != q->xf) return p->xf < q->xf ? -1 : 1;
  if (p->yf != q->yf) return p->yf < q->yf ? -1 : 1;
  return 0;
}

static void prod_msort(orc_prod* a, orc_prod* t, uint64_t n) {
  if (n < 2) return;
  uint64_t h = n / 2;
  prod_msort(a, t, h);
  prod_msort(a + h, t, n - h);
  uint64_t i = 0, j = h, o = 0;
  while (i < h && j < n) {
    if (prod_cmp(&a[j], &a[i]) < 0) t[o++] = a[j++]; else t[o++] = a[i++];
  }
  while (i < h) t[o++] = a[i++];
  while (j < n) t[o++] = a[j++];
  memcpy(a, t, n * sizeof(orc_prod));
}

int64_t orc_esc(const uint64_t* xkeys, const float* xvals, uint64_t x_begin,
                uint64_t x_end, const uint64_t* ykeys, const float* yvals,
                const uint64_t* mpos, const uint64_t* mcnt, int ncbits,
                int y_nfbits, uint64_t** out_xf, uint64_t** out_yf,
                float** out_f32, double** out_f64) {
  uint64_t np = 0;
  for (uint64_t x = x_begin; x < x_end; ++x) np += mcnt[x];
  orc_prod* p = (orc_prod*)malloc((np ? np : 1) * sizeof(orc_prod));
  orc_prod* t = (orc_prod*)malloc((np ? np : 1) * sizeof(orc_prod));
  if (!p || !t) { free(p); free(t); return -1; }
  uint64_t fmask = y_nfbits >= 64 ? ~0ull : ((1ull << y_nfbits) - 1);
  uint64_t k = 0;
  for (uint64_t x = x_begin; x < x_end; ++x) {
    for (uint64_t o = 0; o < mcnt[x]; ++o) {
      uint64_t y = mpos[x] + o;
      p[k].xf = xkeys[x] >> ncbits;
      p[k].yf = ykeys[y] & fmask;
      p[k].v32 = xvals[x] * yvals[y];
      p[k].v64 = (double)xvals[x] * (double)yvals[y];
      ++k;
    }
  }
  prod_msort(p, t, np);
  uint64_t nout = 0;
  for (uint64_t i = 0; i < np; ++i)
    if (i == 0 || prod_cmp(&p[i], &p[i - 1]) != 0) ++nout;
  *out_xf = (uint64_t*)malloc((nout ? nout : 1) * sizeof(uint64_t));
  *out_yf = (uint64_t*)malloc((nout ? nout : 1) * sizeof(uint64_t));
  *out_f32 = (float*)malloc((nout ? nout : 1) * sizeof(float));
  *out_f64 = (double*)malloc((nout ? nout : 1) * sizeof(double));
  int64_t o = -1;
  for (uint64_t i = 0; i < np; ++i) {
    if (i == 0 || prod_cmp(&p[i], &p[i - 1]) != 0) {
      ++o;
      (*out_xf)[o] = p[i].xf;
      (*out_yf)[o] = p[i].yf;
      (*out_f32)[o] = p[i].v32;
      (*out_f64)[o] = p[i].v64;
    } else {
      (*out_f32)[o] += p[i].v32;
      (*out_f64)[o] += p[i].v64;
    }
  }
  free(p);
  free(t);
  return (int64_t)nout;
}

/* Dense path: the number of slots whose FP32 sum is nonzero
 * (contraction.cuh:320-327 counts value != 0; exact-zero sums drop out). */
uint64_t orc_dense_nnz(const float* out_f32, uint64_t n) {
  uint64_t c = 0;
  for (uint64_t i = 0; i < n; ++i) c += out_f32[i] != 0.0f;
  return c;
}

/* ------------------------------------------------------------------ */
/* Static partition: nnz/prtn_num-sized ranges, extended so that no value
 * of (key >> shift) straddles a boundary (tensor_manager.tpp:633-749).
 * prtn_idx must hold prtn_num+1 entries; returns the partition count.   */

int orc_partition(const uint64_t* keys, uint64_t nnz, int prtn_num, int shift,
                  uint64_t* prtn_idx) {
  int np = 0;
  prtn_idx[0] = 0;
  if (prtn_num <= 1) { prtn_idx[1] = nnz; return 1; }
  uint64_t size = nnz / (uint64_t)prtn_num, rem = nnz % (uint64_t)prtn_num;
  uint64_t end = 0;
  for (int i = 0; i < prtn_num; ++i) {
    uint64_t start = end;
    uint64_t e = start + size < nnz ? start + size : nnz;
    if ((uint64_t)i < rem) e++;
    if (e > nnz) e = nnz;
    while (e < nnz && e > 0 && (keys[e] >> shift) == (keys[e - 1] >> shift)) e++;
    prtn_idx[++np] = e;
    end = e;
    if (e == nnz) break;
  }
  if (prtn_idx[np] != nnz) prtn_idx[++np] = nnz;
  return np;
}

/* Dynamic partition of X rows so each range's ESC buffers fit `avail`
 * bytes: start with half the remaining nonzeros, extend to a row boundary,
 * accept if 2*(ir_range*(8+4)*2 + 8) <= avail, else halve
 * (contraction.cuh:143-245).  ends[] receives inclusive range ends.
 * Returns the range count, or -1 when a single row cannot fit.          */

int64_t orc_dynamic_partition(const uint64_t* xkeys, uint64_t nx,
                              uint64_t ir, const uint64_t* prefix, int ncbits,
                              uint64_t avail, uint64_t* ends,
                              uint64_t max_ends, uint64_t* max_range_ir) {
  uint64_t start = 0;
  int64_t n = 0;
  *max_range_ir = 0;
  while (start < nx) {
    uint64_t cand_nnz = (nx - start) / 2;
    uint64_t valid_end = start;
    for (;;) {
      if ((xkeys[start] >> ncbits) == (xkeys[nx - 1] >> ncbits)) {
        valid_end = nx - 1;
        uint64_t r = ir - prefix[start];
        if (r > *max_range_ir) *max_range_ir = r;
        break;
      }
      if (cand_nnz < 1) return -1;
      uint64_t c = start + cand_nnz;
      if (c > nx - 1) c = nx - 1;
      while (c < nx - 1 && (xkeys[c] >> ncbits) == (xkeys[c + 1] >> ncbits)) c++;
      uint64_t r = (c == nx - 1) ? ir - prefix[start] : prefix[c + 1] - prefix[start];
      uint64_t est = 2 * (r * 12 * 2 + 8);
      if (est <= avail) {
        valid_end = c;
        if (r > *max_range_ir) *max_range_ir = r;
        break;
      }
      cand_nnz -= cand_nnz / 2 > 1 ? cand_nnz / 2 : 1;
    }
    if ((uint64_t)n >= max_ends) return -2;
    ends[n++] = valid_end;
    start = valid_end + 1;
  }
  return n;
}

/* ------------------------------------------------------------------ */
/* Independent FP64 value oracle: hash join on the contracted coordinates
 * (no key packing), restricted to X nonzeros whose free coordinates are
 * listed in the optional filter (sorted X-free linear ids; NULL = all).
 * Emits (xfree_lin, yfree_lin, value) sorted by (xfree_lin, yfree_lin),
 * where *_lin are mixed-radix linear ids over the free modes in ascending
 * mode order (row-major).                                              */

typedef struct { uint64_t c; uint64_t f; double v; } orc_yent;
typedef struct { uint64_t xf; uint64_t yf; double v; } orc_zent;

static int yent_cmp(const void* a, const void* b) {
  const orc_yent* p = (const orc_yent*)a;
  const orc_yent* q = (const orc_yent*)b;
  if (p->c != q->c) return p->c < q->c ? -1 : 1;
  return 0;
}
static int zent_cmp(const void* a, const void* b) {
  const orc_zent* p = (const orc_zent*)a;
  const orc_zent* q = (const orc_zent*)b;
  if (p->xf != q->xf) return p->xf < q->xf ? -1 : 1;
  if (p->yf != q->yf) return p->yf < q->yf ? -1 : 1;
  return 0;
}

static int bsearch_u64(const uint64_t* a, uint64_t n, uint64_t v) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    uint64_t mid = (lo + hi) / 2;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo < n && a[lo] == v;
}

int64_t orc_join_fp64(int xorder, const uint64_t* xdims, uint64_t xnnz,
                      const uint32_t* const* xidx, const float* xval,
                      int yorder, const uint64_t* ydims, uint64_t ynnz,
                      const uint32_t* const* yidx, const float* yval, int nc,
                      const int* cx, const int* cy, const uint64_t* xf_filter,
                      uint64_t n_filter, uint64_t** out_xf, uint64_t** out_yf,
                      double** out_v) {
  int isx[ORC_MAX_ORDER] = {0}, isy[ORC_MAX_ORDER] = {0};
  for (int i = 0; i < nc; ++i) { isx[cx[i]] = 1; isy[cy[i]] = 1; }
  /* contract linear id: mixed radix over cx order with radix max(dx, dy) */
  uint64_t crad[ORC_MAX_ORDER];
  for (int i = 0; i < nc; ++i)
    crad[i] = xdims[cx[i]] > ydims[cy[i]] ? xdims[cx[i]] : ydims[cy[i]];
  orc_yent* ye = (orc_yent*)malloc((ynnz ? ynnz : 1) * sizeof(orc_yent));
  if (!ye) return -1;
  for (uint64_t j = 0; j < ynnz; ++j) {
    uint64_t c = 0, f = 0;
    for (int i = 0; i < nc; ++i) c = c * crad[i] + yidx[cy[i]][j];
    for (int m = 0; m < yorder; ++m)
      if (!isy[m]) f = f * ydims[m] + yidx[m][j];
    ye[j].c = c; ye[j].f = f; ye[j].v = (double)yval[j];
  }
  qsort(ye, ynnz, sizeof(orc_yent), yent_cmp);
  uint64_t cap = 1 << 16, nz = 0;
  orc_zent* z = (orc_zent*)malloc(cap * sizeof(orc_zent));
  for (uint64_t x = 0; x < xnnz; ++x) {
    uint64_t c = 0, f = 0;
    for (int m = 0; m < xorder; ++m)
      if (!isx[m]) f = f * xdims[m] + xidx[m][x];
    if (xf_filter && !bsearch_u64(xf_filter, n_filter, f)) continue;
    for (int i = 0; i < nc; ++i) c = c * crad[i] + xidx[cx[i]][x];
    uint64_t lo = 0, hi = ynnz;
    while (lo < hi) {
      uint64_t mid = (lo + hi) / 2;
      if (ye[mid].c < c) lo = mid + 1; else hi = mid;
    }
    for (uint64_t j = lo; j < ynnz && ye[j].c == c; ++j) {
      if (nz == cap) {
        cap *= 2;
        orc_zent* zz = (orc_zent*)realloc(z, cap * sizeof(orc_zent));
        if (!zz) { free(z); free(ye); return -1; }
        z = zz;
      }
      z[nz].xf = f; z[nz].yf = ye[j].f;
      z[nz].v = (double)xval[x] * ye[j].v;
      ++nz;
    }
  }
  free(ye);
  qsort(z, nz, sizeof(orc_zent), zent_cmp);
  uint64_t nout = 0;
  for (uint64_t i = 0; i < nz; ++i)
    if (i == 0 || zent_cmp(&z[i], &z[i - 1]) != 0) ++nout;
  *out_xf = (uint64_t*)malloc((nout ? nout : 1) * sizeof(uint64_t));
  *out_yf = (uint64_t*)malloc((nout ? nout : 1) * sizeof(uint64_t));
  *out_v = (double*)malloc((nout ? nout : 1) * sizeof(double));
  int64_t o = -1;
  for (uint64_t i = 0; i < nz; ++i) {
    if (i == 0 || zent_cmp(&z[i], &z[i - 1]) != 0) {
      ++o;
      (*out_xf)[o] = z[i].xf; (*out_yf)[o] = z[i].yf; (*out_v)[o] = z[i].v;
    } else {
      (*out_v)[o] += z[i].v;
    }
  }
  free(z);
  return (int64_t)nout;
}

void orc_free(void* p) { free(p); }
