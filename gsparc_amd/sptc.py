"""Python mirror of the reference SpTC entry point over libsptc_hip.so.

``SpTC(X, Y, cmodes_x, cmodes_y, dense_accumulator=0)`` mirrors
``gsparc::SpTC`` (include/gsparc/sptc.cuh:16-26) together with the
TensorManager preprocessing that precedes it in main.cu:293-321: it takes COO
tensors and contracted-mode lists, and returns the contraction's COO output
(which the reference computes and then discards; only its nnz is printed,
main.cu:428).  Errors raise :class:`SpTCError` carrying the library status
instead of exiting the process (the reference's exit(1) paths,
cuda_memory.cpp:75-79, sptc.cuh:206-211).

Everything runs on the GPU through the C ABI of include/sptc.h; there is no
CPU fallback.
"""
import ctypes
from ctypes import (POINTER, Structure, byref, c_char_p, c_double, c_int32, c_uint32,
                    c_uint64, c_void_p)

import numpy as np

from ._lib import hip_lib

OK, ERR_BAD_ARGS, ERR_OOM_BUDGET, ERR_KEY_OVERFLOW, ERR_HIP, ERR_NO_DEVICE, ERR_INTERNAL = range(7)
MAX_ORDER = 8


class SpTCError(RuntimeError):
    def __init__(self, status, detail=""):
        self.status = status
        super().__init__(f"sptc status {status}: {detail}")


class _COO(Structure):
    _fields_ = [("order", c_int32), ("nnz", c_uint64), ("dims", POINTER(c_uint64)),
                ("idx", POINTER(c_void_p)), ("idx_bytes", c_int32),
                ("val", c_void_p), ("on_device", c_int32)]


class _Opts(Structure):
    _fields_ = [("dense_accumulator", c_int32), ("count_only", c_int32),
                ("mem_budget_bytes", c_uint64), ("radix_bits", c_int32),
                ("verbose", c_int32), ("kernel_timing", c_int32), ("check_indices", c_int32)]


KERNEL_CLASSES = ["linearize", "sort_input", "match", "expand", "sort_products", "reduce",
                  "dense", "sort_count", "sort_products_wide", "sort_local"]


class _Info(Structure):
    _fields_ = [("nnz", c_uint64), ("ir_nnz", c_uint64), ("nrows", c_uint64),
                ("x_nfbits", c_int32), ("y_nfbits", c_int32), ("ncbits", c_int32),
                ("out_bits", c_int32), ("out_order", c_int32), ("nchunks", c_int32),
                ("accumulator", c_int32), ("value_sum", c_double),
                ("row", c_void_p), ("col", c_void_p), ("val", c_void_p),
                ("row_key", c_void_p), ("phase_ms", c_double * 6)]


class _BatchInfo(Structure):
    _fields_ = [("instances", c_int32), ("dense_instances", c_int32),
                ("dense_classes", c_int32), ("esc_chunks", c_int32), ("nnz_in", c_uint64),
                ("ir_nnz", c_uint64), ("out_nnz", c_uint64), ("graph_nodes", c_uint64),
                ("device_bytes", c_uint64), ("plan_ms", c_double)]


class _PlanInfo(Structure):
    _fields_ = [("x_nnz", c_uint64), ("y_nnz", c_uint64), ("x_nfbits", c_int32),
                ("y_nfbits", c_int32), ("ncbits", c_int32), ("out_bits", c_int32),
                ("out_order", c_int32), ("device_bytes", c_uint64), ("prepare_ms", c_double),
                ("x_parts", c_int32), ("y_parts", c_int32)]


class _ServerOpts(Structure):
    _fields_ = [("budget_bytes", c_uint64), ("workers", c_int32), ("max_batch", c_int32),
                ("large_products", c_uint64), ("large_budget_bytes", c_uint64)]


class _ReqInfo(Structure):
    _fields_ = [("id", c_uint64), ("status", c_int32), ("count_only", c_int32),
                ("batch_size", c_int32), ("worker", c_int32), ("ir_nnz", c_uint64),
                ("queue_ms", c_double), ("latency_ms", c_double)]


class _ServerStats(Structure):
    _fields_ = [(n, c_uint64) for n in ("requests_done", "batches", "batched_requests",
                                        "large_requests", "resident_bytes",
                                        "peak_in_use_bytes", "in_use_bytes", "budget_bytes")]


_bound = False


def lib():
    global _bound
    L = hip_lib()
    if not _bound:
        L.sptc_contract.argtypes = [POINTER(_COO), POINTER(_COO), c_int32, POINTER(c_int32),
                                    POINTER(c_int32), POINTER(_Opts), c_void_p,
                                    POINTER(c_void_p)]
        L.sptc_result_get_info.argtypes = [c_void_p, POINTER(_Info)]
        L.sptc_result_get_chunk.argtypes = [c_void_p, c_int32, POINTER(c_uint64),
                                            POINTER(c_void_p), POINTER(c_void_p),
                                            POINTER(c_void_p)]
        L.sptc_result_copy_to_host.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
        L.sptc_result_decode.argtypes = [c_void_p, POINTER(c_void_p), c_void_p]
        L.sptc_result_row_keys.argtypes = [c_void_p, c_void_p]
        L.sptc_result_row_range.argtypes = [c_void_p, c_uint64, POINTER(c_uint64),
                                            POINTER(c_uint64)]
        L.sptc_result_copy_range.argtypes = [c_void_p, c_uint64, c_uint64, c_void_p, c_void_p,
                                             c_void_p]
        L.sptc_result_decode_device.argtypes = [c_void_p, POINTER(c_void_p), c_void_p,
                                                c_void_p]
        L.sptc_result_verify.argtypes = [c_void_p, POINTER(c_double), POINTER(c_uint64),
                                         c_void_p]
        L.sptc_result_free.argtypes = [c_void_p]
        L.sptc_result_free.restype = None
        L.sptc_layout.argtypes = [c_int32, POINTER(c_uint64), c_int32, POINTER(c_int32),
                                  POINTER(c_int32), c_int32, POINTER(c_int32),
                                  POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)]
        L.sptc_linearize.argtypes = [c_int32, c_uint64, POINTER(c_void_p), c_int32,
                                     POINTER(c_int32), POINTER(c_int32), c_void_p, c_void_p]
        L.sptc_sort_pairs.argtypes = [c_void_p, c_void_p, c_uint64, c_int32, c_int32, c_void_p]
        L.sptc_index_match.argtypes = [c_void_p, c_uint64, c_void_p, c_uint64, c_int32,
                                       c_int32, c_void_p, c_void_p, c_void_p,
                                       POINTER(c_uint64), c_void_p]
        L.sptc_exclusive_scan_u64.argtypes = [c_void_p, c_void_p, c_uint64, c_void_p]
        L.sptc_index_match_noalloc.argtypes = L.sptc_index_match.argtypes
        L.sptc_device_info.argtypes = [POINTER(c_int32), POINTER(c_uint64), POINTER(c_uint64),
                                       POINTER(c_double)]
        L.sptc_status_string.argtypes = [c_int32]
        L.sptc_status_string.restype = c_char_p
        L.sptc_last_error.argtypes = []
        L.sptc_last_error.restype = c_char_p
        L.sptc_plan_get_partitions.argtypes = [c_void_p, c_int32, c_int32, c_void_p,
                                               c_void_p, POINTER(c_int32)]
        L.sptc_find_partition_num.argtypes = [c_uint64, c_int32, c_uint64, c_int32, c_int32,
                                              c_uint64, POINTER(c_int32), POINTER(c_int32)]
        L.sptc_result_as_coo.argtypes = [c_void_p, c_void_p, POINTER(_COO)]
        L.sptc_accumulate_esc.argtypes = [c_void_p, POINTER(_Opts), c_void_p, POINTER(c_void_p)]
        L.sptc_accumulate_dense.argtypes = [c_void_p, POINTER(_Opts), c_void_p,
                                            POINTER(c_void_p)]
        L.sptc_trim_cache.argtypes = []
        L.sptc_trim_cache.restype = None
        L.sptc_result_kernel_stats.argtypes = [c_void_p, c_int32, POINTER(c_uint64),
                                               POINTER(c_double), POINTER(c_double)]
        L.sptc_result_kernel_stats.restype = c_int32
        L.sptc_pool_stats.argtypes = [POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64)]
        L.sptc_pool_stats.restype = None
        L.sptc_batch_create.argtypes = [c_int32, POINTER(_COO), POINTER(_COO),
                                        POINTER(c_int32), POINTER(POINTER(c_int32)),
                                        POINTER(POINTER(c_int32)), POINTER(_Opts), c_void_p,
                                        POINTER(c_void_p)]
        L.sptc_batch_run.argtypes = [c_void_p, c_int32, c_void_p]
        L.sptc_batch_sync.argtypes = [c_void_p, c_void_p]
        L.sptc_batch_get_info.argtypes = [c_void_p, POINTER(_BatchInfo)]
        L.sptc_batch_result.argtypes = [c_void_p, c_int32, POINTER(c_void_p)]
        L.sptc_batch_kernel_stats.argtypes = [c_void_p, c_int32, POINTER(c_uint64),
                                              POINTER(c_double), POINTER(c_double)]
        L.sptc_batch_free.argtypes = [c_void_p]
        L.sptc_batch_free.restype = None
        L.sptc_plan.argtypes = [POINTER(_COO), POINTER(_COO), c_int32, POINTER(c_int32),
                                POINTER(c_int32), POINTER(_Opts), c_void_p, POINTER(c_void_p)]
        L.sptc_execute.argtypes = [c_void_p, POINTER(_Opts), c_void_p, POINTER(c_void_p)]
        L.sptc_plan_get_info.argtypes = [c_void_p, POINTER(_PlanInfo)]
        L.sptc_partition_plan.argtypes = [c_void_p, c_uint64, c_int32, c_void_p, c_void_p,
                                          POINTER(c_int32), POINTER(c_uint64), c_void_p]
        L.sptc_plan_destroy.argtypes = [c_void_p]
        L.sptc_plan_destroy.restype = None
        for name in ("sptc_plan", "sptc_execute", "sptc_plan_get_info", "sptc_partition_plan"):
            getattr(L, name).restype = c_int32
        L.sptc_server_create.argtypes = [POINTER(_ServerOpts), POINTER(c_void_p)]
        L.sptc_server_add_tensor.argtypes = [c_void_p, POINTER(_COO), POINTER(c_int32)]
        L.sptc_server_submit.argtypes = [c_void_p, c_int32, c_int32, c_int32, POINTER(c_int32),
                                         POINTER(c_int32), POINTER(c_uint64)]
        L.sptc_server_wait.argtypes = [c_void_p, c_uint64, POINTER(c_void_p), POINTER(_ReqInfo)]
        L.sptc_server_get_stats.argtypes = [c_void_p, POINTER(_ServerStats)]
        L.sptc_server_free.argtypes = [c_void_p]
        L.sptc_server_free.restype = None
        for name in ("sptc_server_create", "sptc_server_add_tensor", "sptc_server_submit",
                     "sptc_server_wait", "sptc_server_get_stats"):
            getattr(L, name).restype = c_int32
        for name in ("sptc_batch_create", "sptc_batch_run", "sptc_batch_sync",
                     "sptc_batch_get_info", "sptc_batch_result", "sptc_batch_kernel_stats"):
            getattr(L, name).restype = c_int32
        for name in ("sptc_contract", "sptc_result_get_info", "sptc_result_get_chunk",
                     "sptc_result_copy_to_host", "sptc_result_decode", "sptc_layout",
                     "sptc_result_row_keys", "sptc_result_row_range", "sptc_result_copy_range",
                     "sptc_linearize", "sptc_sort_pairs", "sptc_index_match",
                     "sptc_exclusive_scan_u64", "sptc_device_info",
                     "sptc_result_decode_device", "sptc_result_verify"):
            getattr(L, name).restype = c_int32
        _bound = True
    return L


def check(status):
    if status != OK:
        L = lib()
        raise SpTCError(status, (L.sptc_last_error() or b"").decode())


def layout(dims, cmodes, role, cwidth=None):
    """Key layout of one operand: (shift[], width[], nfbits, ncbits)."""
    L = lib()
    order = len(dims)
    d = (c_uint64 * order)(*dims)
    cm = (c_int32 * max(1, len(cmodes)))(*cmodes)
    cw = (c_int32 * max(1, len(cmodes)))(*(cwidth or [-1] * len(cmodes)))
    sh = (c_int32 * order)()
    wd = (c_int32 * order)()
    nf, nc = c_int32(), c_int32()
    st = L.sptc_layout(order, d, len(cmodes), cm, cw if cwidth else None, role, sh, wd,
                       byref(nf), byref(nc))
    check(st)
    return list(sh), list(wd), nf.value, nc.value


class _ResultCOO:
    """A result used as an operand: its device COO view (sptc_result_as_coo),
    kept alive with the result."""

    def __init__(self, r):
        self.result = r
        self.s = _COO()
        check(lib().sptc_result_as_coo(r._h, None, byref(self.s)))


def _operand(t, on_device):
    return _ResultCOO(t) if isinstance(t, Result) else _HostCOO(t, on_device)


class _HostCOO:
    """Keeps the ctypes views of a COOTensor alive for one call."""

    def __init__(self, t, on_device=False):
        self.order = t.order
        self.dims = (c_uint64 * t.order)(*t.dims)
        if on_device:
            self.arrays = t.idx
            ptrs = [a.data_ptr() for a in t.idx]
            val_ptr = t.val.data_ptr()
            idx_bytes = t.idx[0].element_size() if t.order else 4
        else:
            self.arrays = [np.ascontiguousarray(a) for a in t.idx]
            ptrs = [a.ctypes.data for a in self.arrays]
            self.val = np.ascontiguousarray(t.val, dtype=np.float32)
            val_ptr = self.val.ctypes.data
            idx_bytes = self.arrays[0].dtype.itemsize if t.order else 4
        self.ptrs = (c_void_p * t.order)(*ptrs)
        self.s = _COO(t.order, t.nnz if not on_device else int(t.val.numel()), self.dims,
                      self.ptrs, idx_bytes, val_ptr, 1 if on_device else 0)


class Result:
    """Device-resident contraction output (freed with the object)."""

    def __init__(self, handle, owner=None):
        self._h = c_void_p(handle)
        self._owner = owner  # a Batch whose buffers this result views
        self.info = _Info()
        check(lib().sptc_result_get_info(self._h, byref(self.info)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().sptc_result_free(h)
            self._h = c_void_p(0)

    @property
    def nnz(self):
        return int(self.info.nnz)

    @property
    def ir_nnz(self):
        return int(self.info.ir_nnz)

    @property
    def phase_ms(self):
        return list(self.info.phase_ms)

    def chunks(self):
        out = []
        for i in range(self.info.nchunks):
            n, r, c, v = c_uint64(), c_void_p(), c_void_p(), c_void_p()
            check(lib().sptc_result_get_chunk(self._h, i, byref(n), byref(r), byref(c), byref(v)))
            out.append((n.value, r.value, c.value, v.value))
        return out

    def to_host(self):
        """(xfree u64, yfree u64, vals f32) sorted by (xfree, yfree)."""
        n = self.nnz
        xf = np.empty(n, np.uint64)
        yf = np.empty(n, np.uint64)
        v = np.empty(n, np.float32)
        check(lib().sptc_result_copy_to_host(self._h, None, xf.ctypes.data, yf.ctypes.data,
                                             v.ctypes.data))
        return xf, yf, v

    def packed_keys(self):
        """Packed 64-bit keys xfree << y_nfbits | yfree (the reference's
        output key, contraction.cuh:135) and values."""
        n = self.nnz
        k = np.empty(n, np.uint64)
        v = np.empty(n, np.float32)
        check(lib().sptc_result_copy_to_host(self._h, k.ctypes.data, None, None, v.ctypes.data))
        return k, v

    def kernel_stats(self):
        """{class: (launches, total_ms, algorithmic_bytes)} (needs
        kernel_timing=True)."""
        out = {}
        for i, name in enumerate(KERNEL_CLASSES):
            n, ms, b = c_uint64(), c_double(), c_double()
            check(lib().sptc_result_kernel_stats(self._h, i, byref(n), byref(ms), byref(b)))
            out[name] = (n.value, ms.value, b.value)
        return out

    def row_keys(self):
        rk = np.empty(int(self.info.nrows), np.uint64)
        check(lib().sptc_result_row_keys(self._h, rk.ctypes.data))
        return rk

    def row_range(self, row):
        b, e = c_uint64(), c_uint64()
        check(lib().sptc_result_row_range(self._h, int(row), byref(b), byref(e)))
        return b.value, e.value

    def copy_range(self, begin, end):
        n = end - begin
        xf = np.empty(n, np.uint64)
        yf = np.empty(n, np.uint64)
        v = np.empty(n, np.float32)
        check(lib().sptc_result_copy_range(self._h, begin, end, xf.ctypes.data, yf.ctypes.data,
                                           v.ctypes.data))
        return xf, yf, v

    def verify(self, stream=None):
        """(FP64 sum of the emitted values, order violations) from one pass
        over the device-resident output."""
        vs, bad = c_double(), c_uint64()
        check(lib().sptc_result_verify(self._h, byref(vs), byref(bad), c_void_p(stream or 0)))
        return vs.value, bad.value

    def decode_device(self, mode_ptrs, vals_ptr=None, stream=None):
        """Decodes into caller-owned device arrays (one uint32 array of nnz
        per output mode, given as device addresses)."""
        order = self.info.out_order
        ptrs = (c_void_p * max(1, order))(*mode_ptrs)
        check(lib().sptc_result_decode_device(self._h, ptrs, c_void_p(vals_ptr or 0),
                                              c_void_p(stream or 0)))

    def decode(self):
        """Per-mode 0-based indices (X free modes ascending, then Y free
        modes ascending) and values."""
        n = self.nnz
        order = self.info.out_order
        idx = [np.empty(n, np.uint32) for _ in range(order)]
        v = np.empty(n, np.float32)
        ptrs = (c_void_p * max(1, order))(*[a.ctypes.data for a in idx])
        check(lib().sptc_result_decode(self._h, ptrs, v.ctypes.data))
        return idx, v


def SpTC(X, Y, cmodes_x, cmodes_y, dense_accumulator=0, count_only=False,
         mem_budget_bytes=0, radix_bits=0, verbose=False, stream=None, on_device=False,
         kernel_timing=False, check_indices=False):
    """Z = X x_{cmodes_x, cmodes_y} Y on the GPU; returns a :class:`Result`.

    X and Y are COOTensor (host numpy arrays), with on_device=True objects
    with .dims, .idx (torch device tensors) and .val, or the Result of an
    earlier contraction, used from device memory (chained contractions:
    its modes are the X free modes, then the Y free modes)."""
    L = lib()
    hx = _operand(X, on_device)
    hy = X is Y and hx or _operand(Y, on_device)
    nc = len(cmodes_x)
    if len(cmodes_y) != nc:
        raise SpTCError(ERR_BAD_ARGS, "contracted mode lists differ in length")
    cx = (c_int32 * max(1, nc))(*cmodes_x)
    cy = (c_int32 * max(1, nc))(*cmodes_y)
    o = _Opts(int(dense_accumulator), int(bool(count_only)), int(mem_budget_bytes),
              int(radix_bits), int(bool(verbose)), int(bool(kernel_timing)),
              int(bool(check_indices)))
    out = c_void_p()
    st = L.sptc_contract(byref(hx.s), byref(hy.s), nc, cx, cy, byref(o),
                         c_void_p(stream or 0), byref(out))
    check(st)
    return Result(out.value)


class Plan:
    """Operands laid out, linearized and sorted once (sptc_plan: the
    reference's ConvertTensor + SortSlitomXY + Partition); :meth:`execute` runs
    the contraction on them (gsparc::SpTC) as often as wanted and
    :meth:`partitions` returns the dynamic chunk plan under a budget."""

    def __init__(self, X, Y, cmodes_x, cmodes_y, dense_accumulator=0, count_only=False,
                 mem_budget_bytes=0, radix_bits=0, kernel_timing=False, stream=None,
                 on_device=False, check_indices=False):
        L = lib()
        hx = _operand(X, on_device)
        hy = X is Y and hx or _operand(Y, on_device)
        nc = len(cmodes_x)
        if len(cmodes_y) != nc:
            raise SpTCError(ERR_BAD_ARGS, "contracted mode lists differ in length")
        cx = (c_int32 * max(1, nc))(*cmodes_x)
        cy = (c_int32 * max(1, nc))(*cmodes_y)
        o = _Opts(int(dense_accumulator), int(bool(count_only)), int(mem_budget_bytes),
                  int(radix_bits), 0, int(bool(kernel_timing)), int(bool(check_indices)))
        h = c_void_p()
        check(L.sptc_plan(byref(hx.s), byref(hy.s), nc, cx, cy, byref(o),
                          c_void_p(stream or 0), byref(h)))
        self._h = h
        self._opts = o

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().sptc_plan_destroy(h)
            self._h = c_void_p(0)

    @property
    def info(self):
        pi = _PlanInfo()
        check(lib().sptc_plan_get_info(self._h, byref(pi)))
        return {k: getattr(pi, k) for k, _ in _PlanInfo._fields_}

    def execute(self, dense_accumulator=None, count_only=None, stream=None):
        """Runs the contraction; the arguments override the plan's options
        for this call (the reference picks the accumulator per SpTC call)."""
        o = _Opts.from_buffer_copy(self._opts)
        if dense_accumulator is not None:
            o.dense_accumulator = int(dense_accumulator)
        if count_only is not None:
            o.count_only = int(bool(count_only))
        out = c_void_p()
        check(lib().sptc_execute(self._h, byref(o), c_void_p(stream or 0), byref(out)))
        return Result(out.value)

    def accumulate_esc(self, stream=None):
        """ESC_Contraction on the plan's operands (sptc_accumulate_esc)."""
        out = c_void_p()
        check(lib().sptc_accumulate_esc(self._h, byref(self._opts), c_void_p(stream or 0),
                                        byref(out)))
        return Result(out.value)

    def accumulate_dense(self, stream=None):
        """Dense_Contraction on the plan's operands (sptc_accumulate_dense):
        the LDS row or, for wider column spaces, the tiled form."""
        out = c_void_p()
        check(lib().sptc_accumulate_dense(self._h, byref(self._opts), c_void_p(stream or 0),
                                          byref(out)))
        return Result(out.value)

    def static_partitions(self, role):
        """(nnz_bounds, key_bounds) of the plan's static parts of X (role 1)
        or Y (role 2): nonzeros before each part and its first free key."""
        L = lib()
        n = c_int32()
        check(L.sptc_plan_get_partitions(self._h, int(role), 0, None, None, byref(n)))
        nb = np.zeros(n.value + 1, np.uint64)
        kb = np.zeros(n.value + 1, np.uint64)
        check(L.sptc_plan_get_partitions(self._h, int(role), n.value, nb.ctypes.data,
                                         kb.ctypes.data, byref(n)))
        return nb, kb

    def partitions(self, budget_bytes=0, stream=None):
        """(row_bounds, prod_bounds, ir_nnz): first X row and first product
        of each ESC chunk, then the totals."""
        L = lib()
        n, ir = c_int32(), c_uint64()
        check(L.sptc_partition_plan(self._h, int(budget_bytes), 0, None, None, byref(n),
                                    byref(ir), c_void_p(stream or 0)))
        rows = np.zeros(n.value + 1, np.uint64)
        prods = np.zeros(n.value + 1, np.uint64)
        check(L.sptc_partition_plan(self._h, int(budget_bytes), n.value,
                                    rows.ctypes.data, prods.ctypes.data, byref(n), byref(ir),
                                    c_void_p(stream or 0)))
        return rows[:n.value + 1], prods[:n.value + 1], ir.value


class Batch:
    """Many independent contractions planned and run as one launch sequence
    (sptc_batch_*).  ``items`` is a list of (X, Y, cmodes_x, cmodes_y).
    Creating the batch runs it once; :meth:`run` re-runs it (on the
    operands' current values) directly or as a captured hipGraph."""

    def __init__(self, items, dense_accumulator=0, mem_budget_bytes=0, radix_bits=0,
                 kernel_timing=False, stream=None, on_device=False):
        L = lib()
        n = len(items)
        self._keep = []
        xs = (_COO * n)()
        ys = (_COO * n)()
        ncs = (c_int32 * n)()
        cxs = (POINTER(c_int32) * n)()
        cys = (POINTER(c_int32) * n)()
        for i, (X, Y, cmx, cmy) in enumerate(items):
            if len(cmx) != len(cmy):
                raise SpTCError(ERR_BAD_ARGS, "contracted mode lists differ in length")
            hx = _HostCOO(X, on_device)
            hy = hx if X is Y else _HostCOO(Y, on_device)
            cx = (c_int32 * max(1, len(cmx)))(*cmx)
            cy = (c_int32 * max(1, len(cmy)))(*cmy)
            self._keep += [hx, hy, cx, cy]
            xs[i] = hx.s
            ys[i] = hy.s
            ncs[i] = len(cmx)
            cxs[i] = ctypes.cast(cx, POINTER(c_int32))
            cys[i] = ctypes.cast(cy, POINTER(c_int32))
        o = _Opts(int(dense_accumulator), 0, int(mem_budget_bytes), int(radix_bits), 0,
                  int(bool(kernel_timing)))
        h = c_void_p()
        check(L.sptc_batch_create(n, xs, ys, ncs, cxs, cys, byref(o), c_void_p(stream or 0),
                                  byref(h)))
        self._h = h
        if not on_device:
            self._keep = []  # host operands were copied to the device

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().sptc_batch_free(h)
            self._h = c_void_p(0)

    def run(self, use_graph=True, stream=None):
        check(lib().sptc_batch_run(self._h, int(bool(use_graph)), c_void_p(stream or 0)))

    def sync(self, stream=None):
        check(lib().sptc_batch_sync(self._h, c_void_p(stream or 0)))

    @property
    def info(self):
        bi = _BatchInfo()
        check(lib().sptc_batch_get_info(self._h, byref(bi)))
        return {k: getattr(bi, k) for k, _ in _BatchInfo._fields_}

    def __len__(self):
        return self.info["instances"]

    def result(self, i):
        out = c_void_p()
        check(lib().sptc_batch_result(self._h, int(i), byref(out)))
        return Result(out.value, owner=self)

    def kernel_stats(self):
        out = {}
        for i, name in enumerate(KERNEL_CLASSES):
            n, ms, b = c_uint64(), c_double(), c_double()
            check(lib().sptc_batch_kernel_stats(self._h, i, byref(n), byref(ms), byref(b)))
            out[name] = (n.value, ms.value, b.value)
        return out


class Server:
    """Request loop over resident tensors (sptc_server_*): HBM-footprint
    admission, continuous batching of small requests, large requests served
    alone and count-only."""

    def __init__(self, budget_bytes=0, workers=0, max_batch=0, large_products=0,
                 large_budget_bytes=0):
        o = _ServerOpts(int(budget_bytes), int(workers), int(max_batch), int(large_products),
                        int(large_budget_bytes))
        h = c_void_p()
        check(lib().sptc_server_create(byref(o), byref(h)))
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().sptc_server_free(h)
            self._h = c_void_p(0)

    __del__ = close

    def add_tensor(self, t, on_device=False):
        ht = _HostCOO(t, on_device)
        tid = c_int32()
        check(lib().sptc_server_add_tensor(self._h, byref(ht.s), byref(tid)))
        return tid.value

    def submit(self, x, y, cmodes_x, cmodes_y):
        nc = len(cmodes_x)
        cx = (c_int32 * max(1, nc))(*cmodes_x)
        cy = (c_int32 * max(1, nc))(*cmodes_y)
        rid = c_uint64()
        check(lib().sptc_server_submit(self._h, int(x), int(y), nc, cx, cy, byref(rid)))
        return rid.value

    def wait(self, request, want_result=True):
        """(Result or None, info dict); raises SpTCError when the request
        failed."""
        out = c_void_p()
        info = _ReqInfo()
        st = lib().sptc_server_wait(self._h, int(request), byref(out) if want_result else None,
                                    byref(info))
        check(st)
        d = {k: getattr(info, k) for k, _ in _ReqInfo._fields_}
        return (Result(out.value) if out.value else None), d

    def stats(self):
        st = _ServerStats()
        check(lib().sptc_server_get_stats(self._h, byref(st)))
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_}


def find_partition_num(nnz_x, order_x, nnz_y, order_y, idx_bytes, budget_bytes):
    """(x_parts, y_parts) the engine splits operands of these sizes into
    under a device budget (FindPartitionNum); (1, 1) when both fit whole."""
    px, py = c_int32(), c_int32()
    check(lib().sptc_find_partition_num(int(nnz_x), int(order_x), int(nnz_y), int(order_y),
                                        int(idx_bytes), int(budget_bytes), byref(px),
                                        byref(py)))
    return px.value, py.value


def pool_stats():
    live, res, n = c_uint64(), c_uint64(), c_uint64()
    lib().sptc_pool_stats(byref(live), byref(res), byref(n))
    return {"live_bytes": live.value, "reserved_bytes": res.value, "hip_mallocs": n.value}


def device_info():
    cus, fr, tot, peak = c_int32(), c_uint64(), c_uint64(), c_double()
    check(lib().sptc_device_info(byref(cus), byref(fr), byref(tot), byref(peak)))
    return {"num_cus": cus.value, "free_bytes": fr.value, "total_bytes": tot.value,
            "peak_hbm_gbs": peak.value}
