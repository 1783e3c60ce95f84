/*
 * sptc.h -- C ABI of libsptc_hip.so, the MI355X (gfx950) sparse tensor
 * contraction engine.
 *
 * Drop-in boundary for GSPARC's SpTC path.  The reference entry point is the
 * C++ template
 *     gsparc::SpTC<TensorType,SLITOMType>(SX, SY, SZ, pools, gpu_count,
 *         copy_timers, idxmatch_timers, dprtn_timers, contraction_timers,
 *         dense_accumulator)                          (include/gsparc/sptc.cuh:16-26)
 * preceded by TensorManager::{FindPartitionNum, ConvertTensor, SortSlitomXY,
 * Partition} (main.cu:293-321) on SparseTensor<uint64_t,float> inputs
 * (sparse_tensor.hpp:12-44).  Here the whole chain -- linearize, sort,
 * partition, index match, accumulate -- sits behind sptc_contract(), which
 * takes the same COO tensors and contracted-mode lists and returns the
 * contraction's COO output.  The per-phase entry points below replace the
 * individual reference routines named in their comments.
 *
 * Conventions: plain pointers and sizes; every call returns sptc_status (the
 * reference exit()s instead, cuda_helper.cpp:83-92, cuda_memory.cpp:75-79).
 * `stream` is a hipStream_t passed as void* (NULL = the default stream).
 * Device pointers are HIP device addresses.  No call allocates from, or
 * frees, memory that the caller owns.
 */
#ifndef SPTC_H_
#define SPTC_H_
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SPTC_OK = 0,
  SPTC_ERR_BAD_ARGS = 1,     /* malformed tensor, modes or options        */
  SPTC_ERR_OOM_BUDGET = 2,   /* work does not fit the memory budget       */
  SPTC_ERR_KEY_OVERFLOW = 3, /* a linearized key needs more than 64 bits  */
  SPTC_ERR_HIP = 4,          /* a HIP runtime call failed                 */
  SPTC_ERR_NO_DEVICE = 5,    /* no usable gfx950 device                   */
  SPTC_ERR_INTERNAL = 6      /* a device-side consistency check failed    */
} sptc_status;

#define SPTC_MAX_ORDER 8

/* COO tensor in SoA form (the reference's SparseTensor<uint64_t,float>,
 * sparse_tensor.hpp:12-44): 0-based per-mode index arrays and FP32 values.
 * dims[m] sets the bit width of mode m (64 - clz(dims[m]-1),
 * tensor_manager.tpp:109; the .tns reader sets dims = max index + 1). */
typedef struct {
  int32_t order;
  uint64_t nnz;
  const uint64_t* dims;     /* host array [order]                        */
  const void* const* idx;   /* [order] index arrays (host array of ptrs) */
  int32_t idx_bytes;        /* 4 (uint32) or 8 (uint64)                  */
  const float* val;         /* [nnz]                                     */
  int32_t on_device;        /* 1: idx[m], val are device pointers        */
} sptc_coo;

/* Options.  dense_accumulator follows the reference's -d flag
 * (sptc.cuh:201-218): 1 forces the dense accumulator (the LDS row when the
 * column space fits one, else the tiled form; error if neither applies),
 * -1 forces ESC, 0 lets the planner choose, 2 forces the tiled dense
 * accumulator. */
typedef struct {
  int32_t dense_accumulator;
  int32_t count_only;        /* 1: return nnz and checksums, no entries   */
  uint64_t mem_budget_bytes; /* device bytes this call may hold at once
                                (operands, workspaces, output); 0: free
                                device memory minus 1 GiB               */
  int32_t radix_bits;        /* 0: per-sort choice; else 8..11            */
  int32_t verbose;           /* 1: phase log on stderr                    */
  int32_t kernel_timing;     /* 1: HIP-event time every hot-kernel launch
                                (see sptc_result_kernel_stats)          */
  int32_t check_indices;     /* 1: one device pass per mode checks that
                                every index is below its dim, else
                                SPTC_ERR_BAD_ARGS; 0: an index >= its dim
                                is masked to the mode's bit width, as the
                                reference's pdep linearization does
                                (tensor_manager.tpp:246-294)            */
} sptc_opts;

/* Result of a contraction (device-resident).  Entry i has output free
 * coordinates (X free modes ascending, then Y free modes ascending):
 *   X part = row_key[row[i]]  (X free field of the reference's X key,
 *            key >> ncbits, tensor_manager.tpp:266)
 *   Y part = col[i]           (Y free field, key & (2^Ynf - 1))
 * Entries are sorted by (row, col): for out_bits <= 64 that is ascending
 * packed key (row_key << y_nfbits | col), the order of the reference's ESC
 * output (contraction.cuh:506-537). */
typedef struct {
  uint64_t nnz;          /* output nonzeros                                */
  uint64_t ir_nnz;       /* intermediate products (index-match total)      */
  uint64_t nrows;        /* entries in row_key (X free rows of X)          */
  int32_t x_nfbits, y_nfbits, ncbits, out_bits;
  int32_t out_order;
  int32_t nchunks;       /* accumulation chunks (dynamic partitions)       */
  int32_t accumulator;   /* 0 = ESC, 1 = dense LDS row, 2 = tiled dense    */
  double value_sum;      /* sum of all output values (FP64 on device)      */
  const uint32_t* row;   /* device [nnz]                                   */
  const uint64_t* col;   /* device [nnz]                                   */
  const float* val;      /* device [nnz]                                   */
  const uint64_t* row_key; /* device [nrows]                               */
  /* phase times in ms (HIP events): 0 upload, 1 linearize+sort, 2 match,
   * 3 partition, 4 accumulate (expand+sort+reduce), 5 total */
  double phase_ms[6];
} sptc_result_info;

typedef struct sptc_result_s sptc_result;

/* Full SpTC: Z = X x_{cx,cy} Y.  Replaces TensorManager::ConvertTensor +
 * SortSlitomXY + Partition + gsparc::SpTC (main.cu:293-376).  On success
 * *out holds the result (free with sptc_result_free). */
sptc_status sptc_contract(const sptc_coo* X, const sptc_coo* Y, int32_t nc,
                          const int32_t* cx, const int32_t* cy,
                          const sptc_opts* opts, void* stream,
                          sptc_result** out);

sptc_status sptc_result_get_info(const sptc_result* r, sptc_result_info* info);

/* Chunk i (0 <= i < info.nchunks) of the result: outputs of one contiguous
 * X-row range, in ascending order; chunks are in ascending order too.  The
 * reference keeps one ESC partition's output at a time (contraction.cuh:
 * 440-562); here every chunk stays resident. */
sptc_status sptc_result_get_chunk(const sptc_result* r, int32_t i, uint64_t* nnz,
                                  const uint32_t** row, const uint64_t** col,
                                  const float** val);

/* Copies entries to host arrays of r->nnz elements (any may be NULL):
 * packed keys (row_key << y_nfbits | col; only when out_bits <= 64, else
 * SPTC_ERR_KEY_OVERFLOW), the row/col parts, values. */
sptc_status sptc_result_copy_to_host(const sptc_result* r, uint64_t* keys,
                                     uint64_t* xfree, uint64_t* yfree,
                                     float* vals);

/* Row table (X free keys, ascending) to a host array of info.nrows. */
sptc_status sptc_result_row_keys(const sptc_result* r, uint64_t* row_keys);

/* Global entry range [*begin, *end) of row `row` (an index into the row
 * table); empty when the row has no outputs. */
sptc_status sptc_result_row_range(const sptc_result* r, uint64_t row,
                                  uint64_t* begin, uint64_t* end);

/* Copies entries [begin, end) (global entry order) to host arrays of
 * end-begin elements; any output pointer may be NULL. */
sptc_status sptc_result_copy_range(const sptc_result* r, uint64_t begin,
                                   uint64_t end, uint64_t* xfree,
                                   uint64_t* yfree, float* vals);

/* Decodes entries to 0-based per-mode uint32 indices (mode_idx[out_order]
 * host arrays of nnz each), X free modes ascending then Y free modes; the
 * decode runs on the device, pieces are copied out.  SPTC_ERR_KEY_OVERFLOW
 * when a free mode is wider than 32 bits. */
sptc_status sptc_result_decode(const sptc_result* r, uint32_t* const* mode_idx,
                               float* vals);

/* Decodes entries to 0-based per-mode uint32 indices on the device:
 * mode_idx is a host array of out_order DEVICE pointers (nnz entries each),
 * vals a device array (may be NULL).  Asynchronous on `stream`.  The
 * reference has no output decode (its COO output is discarded,
 * contraction.cuh:353-357,596-597).  SPTC_ERR_KEY_OVERFLOW when a free mode
 * is wider than 32 bits (as for sptc_result_decode). */
sptc_status sptc_result_decode_device(const sptc_result* r, uint32_t* const* mode_idx,
                                      float* vals, void* stream);

/* The output as a device-resident COO operand, for chaining contractions
 * without leaving the GPU (the reference's multi_SpTC keeps Z resident and
 * feeds it to the next SpTC, sptc.cuh:362-698, main.cu:323-371): order =
 * out_order, dims = the X free modes' dims then the Y free modes', 0-based
 * uint32 indices and FP32 values in device buffers owned by the result
 * (built by one decode on the first call, freed with the result).  *out can
 * be passed as X or Y to sptc_contract / sptc_plan.  SPTC_ERR_KEY_OVERFLOW
 * when a free mode is wider than 32 bits; SPTC_ERR_BAD_ARGS for a
 * count-only result. */
sptc_status sptc_result_as_coo(sptc_result* r, void* stream, sptc_coo* out);

/* One pass over the device-resident output: *value_sum = FP64 sum of the
 * emitted values, *violations = entries that are not strictly ascending in
 * (row, col) within their chunk or lie outside the chunk's row range
 * (0 for a well-formed result).  Synchronizes the stream.  No reference
 * counterpart: a check of the result's own invariants. */
sptc_status sptc_result_verify(const sptc_result* r, double* value_sum, uint64_t* violations,
                               void* stream);

/* Kernel classes timed when opts.kernel_timing is set. */
typedef enum {
  SPTC_K_LINEARIZE = 0,   /* linearize + fused digit histograms           */
  SPTC_K_SORT_INPUT = 1,  /* radix pass (scatter) over operand keys       */
  SPTC_K_MATCH = 2,       /* index match (reduce + emit)                  */
  SPTC_K_EXPAND = 3,      /* ESC product expansion                        */
  SPTC_K_SORT_PROD = 4,   /* radix pass (scatter) over product keys:
                             8-bit digits, or the row-segmented passes  */
  SPTC_K_REDUCE = 5,      /* reduce-by-key (count + emit)                 */
  SPTC_K_DENSE = 6,       /* dense per-row accumulation                   */
  SPTC_K_SORT_COUNT = 7,  /* radix pass digit counting + offset scan      */
  SPTC_K_SORT_PROD_WIDE = 8, /* radix pass (scatter) over product keys,
                                digits wider than 8 bits                 */
  SPTC_K_SORT_LOCAL = 9,  /* product groups of small rows sorted in LDS     */
  SPTC_K_COUNT = 10
} sptc_kernel_class;

/* Launch count, summed HIP-event time (ms) and algorithmic HBM bytes of one
 * kernel class over this contraction (zeros unless opts.kernel_timing). */
sptc_status sptc_result_kernel_stats(const sptc_result* r, int32_t cls,
                                     uint64_t* launches, double* total_ms,
                                     double* alg_bytes);

/* Device arena statistics: bytes held by live buffers, bytes reserved from
 * the HIP runtime, and hipMalloc calls so far. */
void sptc_pool_stats(uint64_t* live_bytes, uint64_t* reserved_bytes,
                     uint64_t* hip_mallocs);

void sptc_result_free(sptc_result* r);

/* ---- batched contractions ------------------------------------------------
 * Many independent contractions through one launch sequence: operands of
 * all instances are laid out back to back and every phase (linearize,
 * segmented sort, match, dense accumulation, ESC chunks) runs as one launch
 * over all of them.  Replaces n separate gsparc::SpTC calls (sptc.cuh:16-26,
 * each driven as in main.cu:293-376); the reference has no batched form. */
typedef struct sptc_batch_s sptc_batch;

typedef struct {
  int32_t instances;
  int32_t dense_instances;  /* instances on the dense accumulator          */
  int32_t dense_classes;    /* dense launches (one per LDS slot width)     */
  int32_t esc_chunks;       /* ESC chunks (each spans many instances)      */
  uint64_t nnz_in;          /* operand nonzeros, all instances             */
  uint64_t ir_nnz;          /* intermediate products, all instances        */
  uint64_t out_nnz;         /* output nonzeros, all instances              */
  uint64_t graph_nodes;     /* nodes of the captured graph (0 until used)  */
  uint64_t device_bytes;    /* device bytes the batch holds                */
  double plan_ms;           /* host wall time of sptc_batch_create         */
} sptc_batch_info;

/* Plans and runs the batch once: instance i is X[i] x_{cx[i],cy[i]} Y[i]
 * with nc[i] contracted modes.  Options apply to every instance
 * (count_only is not supported; dense_accumulator is the per-instance
 * rule).  Results are available after the call (sptc_batch_result). */
sptc_status sptc_batch_create(int32_t n, const sptc_coo* X, const sptc_coo* Y,
                              const int32_t* nc, const int32_t* const* cx,
                              const int32_t* const* cy, const sptc_opts* opts,
                              void* stream, sptc_batch** out);

/* Re-runs the batch on the operands' current values (host operands were
 * copied at creation); index arrays must be unchanged.  use_graph = 1
 * launches the captured hipGraph of the whole sequence (captured on first
 * use), 0 enqueues the kernels directly.  Asynchronous: call
 * sptc_batch_sync before reading results.  Results are overwritten in place. */
sptc_status sptc_batch_run(sptc_batch* b, int32_t use_graph, void* stream);

/* Waits for the stream; SPTC_ERR_INTERNAL when the operands' index arrays
 * changed since creation (the run then left the results untouched). */
sptc_status sptc_batch_sync(sptc_batch* b, void* stream);

sptc_status sptc_batch_get_info(const sptc_batch* b, sptc_batch_info* info);

/* Result of instance i as a view into the batch's buffers (valid until
 * sptc_batch_free; free the handle with sptc_result_free).  All sptc_result_*
 * accessors apply; rows index the instance's own row table. */
sptc_status sptc_batch_result(sptc_batch* b, int32_t i, sptc_result** out);

/* Kernel statistics of the planning run (opts.kernel_timing). */
sptc_status sptc_batch_kernel_stats(const sptc_batch* b, int32_t cls,
                                    uint64_t* launches, double* total_ms,
                                    double* alg_bytes);

void sptc_batch_free(sptc_batch* b);

/* ---- serving ----------------------------------------------------------------
 * A request loop over tensors kept resident on the device.  Requests
 * (contractions of two registered tensors) are planned on a planner thread
 * -- operands linearized and sorted once per tensor and layout and kept
 * resident, intermediate products counted -- then admitted in arrival order
 * while their estimated device footprint fits the budget.  Small requests
 * that are ready together run as one batch (sptc_batch_* machinery); a
 * request above the large-product threshold runs alone, count-only (nnz and
 * value checksum: the reference's plain path likewise returns only the nnz,
 * main.cu:375), one large request at a time.  Several workers, each with
 * its own stream, overlap their launches.  No reference counterpart. */
typedef struct sptc_server_s sptc_server;

typedef struct {
  uint64_t budget_bytes;       /* device bytes the server may hold: resident
                                  tensors, prepared operands and admitted
                                  work; 0: 96 GiB                         */
  int32_t workers;             /* concurrent launch streams; 0: 3         */
  int32_t max_batch;           /* requests per batched launch; 0: 256     */
  uint64_t large_products;     /* requests with more intermediate products
                                  run alone, count-only; 0: 2^28          */
  uint64_t large_budget_bytes; /* device bytes one large request may hold;
                                  0: budget / 3                           */
} sptc_server_opts;

typedef struct {
  uint64_t id;
  int32_t status;              /* sptc_status of the request             */
  int32_t count_only;          /* 1: result holds nnz and value_sum only */
  int32_t batch_size;          /* requests in the launch that served it  */
  int32_t worker;
  uint64_t ir_nnz;
  double queue_ms;             /* submit -> its launch started           */
  double latency_ms;           /* submit -> result ready                 */
} sptc_request_info;

typedef struct {
  uint64_t requests_done;
  uint64_t batches;            /* batched launches                       */
  uint64_t batched_requests;   /* requests they served                   */
  uint64_t large_requests;     /* requests served alone, count-only      */
  uint64_t resident_bytes;     /* tensors + prepared operands            */
  uint64_t peak_in_use_bytes;  /* admitted footprints, peak              */
  uint64_t in_use_bytes;
  uint64_t budget_bytes;
} sptc_server_stats;

sptc_status sptc_server_create(const sptc_server_opts* opts, sptc_server** out);

/* Copies a COO tensor (host or device arrays) into the server; it stays
 * resident until sptc_server_free.  *id names it in requests. */
sptc_status sptc_server_add_tensor(sptc_server* s, const sptc_coo* t, int32_t* id);

/* Queues Z = T[x] x_{cx,cy} T[y]; returns at once with the request id. */
sptc_status sptc_server_submit(sptc_server* s, int32_t x, int32_t y, int32_t nc,
                               const int32_t* cx, const int32_t* cy, uint64_t* request);

/* Blocks until the request is done.  *out (may be NULL) receives its result
 * (free with sptc_result_free; a batched request's result keeps its batch's
 * buffers -- and their share of the budget -- until freed).  info (may be
 * NULL) receives status, batch size and latency.  Returns the request's
 * status. */
sptc_status sptc_server_wait(sptc_server* s, uint64_t request, sptc_result** out,
                             sptc_request_info* info);

sptc_status sptc_server_get_stats(sptc_server* s, sptc_server_stats* stats);

/* Stops the workers (queued requests are dropped) and frees the tensors. */
void sptc_server_free(sptc_server* s);

/* ---- planned form: operands converted and sorted once -------------------- */

typedef struct sptc_plan_s sptc_plan_t;

typedef struct {
  uint64_t x_nnz, y_nnz;
  int32_t x_nfbits, y_nfbits, ncbits, out_bits;
  int32_t out_order;
  uint64_t device_bytes;   /* resident operand bytes held by the plan      */
  double prepare_ms;       /* upload + linearize + sort of both operands
                              (a partitioned plan: the host-side split)   */
  int32_t x_parts, y_parts;/* static operand parts (1: resident whole)    */
} sptc_plan_info;

/* TensorManager::ConvertTensor (x2) + SortSlitomXY + Partition
 * (main.cu:293-321, tensor_manager.tpp:212-298,487-522,633-749): lays out,
 * linearizes and sorts both operands into device memory owned by the plan
 * (the reference's SLITOM d_indices/d_values).  Host inputs are copied;
 * device inputs are only read during the call.
 * When the operands do not fit half of the memory budget
 * (sptc_find_partition_num), the plan instead keeps both operands split
 * into free-key ranges on the host (Partition, tensor_manager.tpp:633-749:
 * X by free rows, Y by output columns) and every execution runs the
 * reference's X x Y partition-pair loop (sptc.cuh:42-81): each Y part
 * re-keyed to the Y layout and sorted on the device (convert_and_sort_64,
 * tensor_manager.tpp:751-798), contracted with each X part, and the pair
 * results merged into row order -- the same entries and values as an
 * unpartitioned run.  sptc_contract partitions the same way. */
sptc_status sptc_plan(const sptc_coo* X, const sptc_coo* Y, int32_t nc,
                      const int32_t* cx, const int32_t* cy,
                      const sptc_opts* opts, void* stream, sptc_plan_t** out);

/* gsparc::SpTC (sptc.cuh:16-360) on a plan's resident operands: index
 * match, dynamic partitioning, accumulation.  Each call returns a new
 * result; phase_ms[0] and [1] are 0 (the operands were prepared by
 * sptc_plan).  opts (NULL: the plan's) apply to this call; as in the
 * reference the accumulator is chosen here: dense_accumulator = 1 / -1
 * runs the dense / ESC accumulator (Dense_Contraction / ESC_Contraction,
 * contraction.cuh:247-599) alone. */
sptc_status sptc_execute(sptc_plan_t* p, const sptc_opts* opts, void* stream,
                         sptc_result** out);

/* The accumulators on a plan's sorted operands, each forced regardless of
 * opts->dense_accumulator (which is otherwise honoured as in sptc_execute):
 *   sptc_accumulate_esc   -- ESC_Contraction (contraction.cuh:364-599):
 *                            expand, row-segmented sort, reduce;
 *   sptc_accumulate_dense -- Dense_Contraction (contraction.cuh:247-362):
 *                            the LDS row when the column space fits it and Y
 *                            has no duplicate coordinates, else the tiled
 *                            dense accumulator (column tiles through LDS).
 * The reference's accumulators receive the index match's mPos / mCnt /
 * prefix arrays (sptc.cuh:118-318); here each call runs the match on the
 * resident operands first (its time is phase_ms[2]). */
sptc_status sptc_accumulate_esc(sptc_plan_t* p, const sptc_opts* opts, void* stream,
                                sptc_result** out);
sptc_status sptc_accumulate_dense(sptc_plan_t* p, const sptc_opts* opts, void* stream,
                                  sptc_result** out);

sptc_status sptc_plan_get_info(const sptc_plan_t* p, sptc_plan_info* info);

/* Dynamic_partition (contraction.cuh:143-245) for a plan under a device
 * budget (0: the plan's): runs the index match and cuts X's rows into
 * chunks whose products (and, unless count_only, outputs) fit, never
 * splitting a row.  row_bounds/prod_bounds receive *nparts + 1 entries
 * (first rows and first products of each chunk, then the totals) when
 * max_parts is large enough; *nparts is always set.  *ir_nnz receives the
 * product count.  SPTC_ERR_OOM_BUDGET when one row's products do not fit. */
sptc_status sptc_partition_plan(sptc_plan_t* p, uint64_t budget_bytes,
                                int32_t max_parts, uint64_t* row_bounds,
                                uint64_t* prod_bounds, int32_t* nparts,
                                uint64_t* ir_nnz, void* stream);

/* FindPartitionNum (tensor_manager.tpp:32-61) for this engine: the numbers
 * of X and Y parts that sptc_plan / sptc_contract use for operands of these
 * sizes under a device budget (1 and 1 when both stay resident whole).
 * SPTC_ERR_OOM_BUDGET when more than 64 parts per operand would be needed. */
sptc_status sptc_find_partition_num(uint64_t nnz_x, int32_t order_x, uint64_t nnz_y,
                                    int32_t order_y, int32_t idx_bytes,
                                    uint64_t budget_bytes, int32_t* x_parts,
                                    int32_t* y_parts);

/* Partition's output (the reference's SLITOM prtn_idx / prtn_coord,
 * tensor_manager.tpp:633-749) for operand role 1 (X) or 2 (Y) of a plan:
 * *nparts parts; when max_parts >= *nparts, nnz_bounds[k] = nonzeros in
 * parts before k and key_bounds[k] = the first free key of part k
 * (*nparts + 1 entries each, the last the totals / 2^nfbits). */
sptc_status sptc_plan_get_partitions(const sptc_plan_t* p, int32_t role, int32_t max_parts,
                                     uint64_t* nnz_bounds, uint64_t* key_bounds,
                                     int32_t* nparts);

void sptc_plan_destroy(sptc_plan_t* p);

/* ---- per-phase entry points (device pointers, caller-owned buffers) ---- */

/* Key layout of one operand (get_fmode + compute_nbtis + setup_slitom,
 * tensor_manager.tpp:63-186).  role 1 = X: key = f << ncbits | c; role 2 =
 * Y: key = c << nfbits | f.  cwidth[i] >= 0 overrides the width of
 * contracted mode cmodes[i] (NULL: widths from dims).  shift/width are
 * [order] outputs. */
sptc_status sptc_layout(int32_t order, const uint64_t* dims, int32_t nc,
                        const int32_t* cmodes, const int32_t* cwidth,
                        int32_t role, int32_t* shift, int32_t* width,
                        int32_t* nfbits, int32_t* ncbits);

/* Linearize (ConvertTensor, tensor_manager.tpp:212-298, on the GPU):
 * keys[i] = OR_m idx[m][i] << shift[m].  idx are device pointers. */
sptc_status sptc_linearize(int32_t order, uint64_t nnz, const void* const* idx,
                           int32_t idx_bytes, const int32_t* shift,
                           const int32_t* width, uint64_t* keys, void* stream);

/* Stable LSD radix sort of (key, value) pairs on key bits [0, nbits)
 * (sort_64's cub::DeviceRadixSort::SortPairs, sort.cuh:90-108).  Sorts
 * in place; radix_bits 0 = autotune. */
sptc_status sptc_sort_pairs(uint64_t* keys, float* vals, uint64_t n,
                            int32_t nbits, int32_t radix_bits, void* stream);

/* Index match (IndexMatch, indexMatch.cuh:231-287): for each X nonzero,
 * mpos = first Y index whose key >> y_nfbits equals the X contract field,
 * mcnt = run length, prefix = exclusive scan of mcnt ([nx] entries + 1,
 * prefix[nx] = ir).  *ir_nnz receives the total. */
sptc_status sptc_index_match(const uint64_t* xkeys, uint64_t nx,
                             const uint64_t* ykeys, uint64_t ny,
                             int32_t ncbits, int32_t y_nfbits, uint64_t* mpos,
                             uint64_t* mcnt, uint64_t* prefix,
                             uint64_t* ir_nnz, void* stream);

/* sptc_index_match with no allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 * arguments and results as sptc_index_match; its working memory (the Y
 * contract table and the scan's tile sums) is a fixed 33 MiB device array
 * that each GPU holds until the process ends.  Limits: a contract field of
 * at most 21 bits (direct table) or fewer than 2^21 Y nonzeros (distinct-key
 * table), and at most 65535 * 2048 nonzeros per operand; larger inputs
 * return SPTC_ERR_BAD_ARGS before any device work.  Calls on one GPU share
 * the array: calls that overlap on one GPU (two streams or threads) race. */
sptc_status sptc_index_match_noalloc(const uint64_t* xkeys, uint64_t nx,
                                     const uint64_t* ykeys, uint64_t ny,
                                     int32_t ncbits, int32_t y_nfbits, uint64_t* mpos,
                                     uint64_t* mcnt, uint64_t* prefix,
                                     uint64_t* ir_nnz, void* stream);

/* Exclusive scan of n uint64 (cub::DeviceScan::InclusiveSum shifted,
 * indexMatch.cuh:275-279); out has n+1 entries, out[n] = total. */
sptc_status sptc_exclusive_scan_u64(const uint64_t* in, uint64_t* out,
                                    uint64_t n, void* stream);

/* Device properties used by the planner and the benchmark. */
sptc_status sptc_device_info(int32_t* num_cus, uint64_t* free_bytes,
                             uint64_t* total_bytes, double* peak_hbm_gbs);

const char* sptc_status_string(sptc_status s);

/* Detail of the last failure on the calling thread (empty when none). */
const char* sptc_last_error(void);

/* Returns cached device blocks to the HIP runtime. */
void sptc_trim_cache(void);

#ifdef __cplusplus
}
#endif
#endif
