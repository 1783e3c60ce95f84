"""CPU oracle for the SpTC path (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this package, and only as the checker.  It wraps liboracle.so (built
from sptc_oracle.c by this folder's Makefile), a host restatement of the
reference algorithm; see the function list in sptc_oracle.c.
"""
import ctypes
import os
from ctypes import POINTER, byref, c_double, c_int, c_int64, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_TTT = os.path.join(HERE, "_ref", "ttt")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C {HERE}`")
        L = ctypes.CDLL(LIB)
        L.orc_mode_bits.argtypes = [c_uint64]
        L.orc_mode_bits.restype = c_int
        L.orc_layout.argtypes = [c_int, POINTER(c_uint64), c_int, POINTER(c_int), POINTER(c_int),
                                 c_int, POINTER(c_int), POINTER(c_int), POINTER(c_int),
                                 POINTER(c_int), POINTER(c_int)]
        L.orc_pdep.argtypes = [c_uint64, c_uint64]
        L.orc_pdep.restype = c_uint64
        L.orc_linearize.argtypes = [c_int, c_uint64, POINTER(c_void_p), POINTER(c_int),
                                    POINTER(c_int), c_void_p]
        L.orc_sort_pairs.argtypes = [c_void_p, c_void_p, c_uint64]
        L.orc_index_match.argtypes = [c_void_p, c_uint64, c_void_p, c_uint64, c_int, c_int,
                                      c_void_p, c_void_p, c_void_p]
        L.orc_index_match.restype = c_uint64
        L.orc_esc.argtypes = [c_void_p, c_void_p, c_uint64, c_uint64, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_int, c_int, POINTER(c_void_p),
                              POINTER(c_void_p), POINTER(c_void_p), POINTER(c_void_p)]
        L.orc_esc.restype = c_int64
        L.orc_dense_nnz.argtypes = [c_void_p, c_uint64]
        L.orc_dense_nnz.restype = c_uint64
        L.orc_partition.argtypes = [c_void_p, c_uint64, c_int, c_int, c_void_p]
        L.orc_partition.restype = c_int
        L.orc_dynamic_partition.argtypes = [c_void_p, c_uint64, c_uint64, c_void_p, c_int,
                                            c_uint64, c_void_p, c_uint64, POINTER(c_uint64)]
        L.orc_dynamic_partition.restype = c_int64
        L.orc_join_fp64.argtypes = [c_int, POINTER(c_uint64), c_uint64, POINTER(c_void_p),
                                    c_void_p, c_int, POINTER(c_uint64), c_uint64,
                                    POINTER(c_void_p), c_void_p, c_int, POINTER(c_int),
                                    POINTER(c_int), c_void_p, c_uint64, POINTER(c_void_p),
                                    POINTER(c_void_p), POINTER(c_void_p)]
        L.orc_join_fp64.restype = c_int64
        L.orc_free.argtypes = [c_void_p]
        L.orc_free.restype = None
        _lib = L
    return _lib


def _take(ptr, n, dtype):
    """Copy n items from a malloc'ed buffer into numpy and free it."""
    if n:
        buf = (ctypes.c_char * (n * np.dtype(dtype).itemsize)).from_address(ptr.value)
        a = np.frombuffer(buf, dtype=dtype).copy()
    else:
        a = np.empty(0, dtype=dtype)
    lib().orc_free(ptr)
    return a


def mode_bits(d):
    return lib().orc_mode_bits(d)


def layout(dims, cmodes, role, cwidth=None):
    """-> (shift[], width[], fpos[], nfbits, ncbits)"""
    order = len(dims)
    d = (c_uint64 * order)(*dims)
    cm = (c_int * max(1, len(cmodes)))(*cmodes)
    cw = (c_int * max(1, len(cmodes)))(*(cwidth if cwidth else [-1] * len(cmodes)))
    sh, wd, fp = (c_int * order)(), (c_int * order)(), (c_int * order)()
    nf, nc = c_int(), c_int()
    rc = lib().orc_layout(order, d, len(cmodes), cm, cw, role, sh, wd, fp, byref(nf), byref(nc))
    assert rc == 0
    nfree = order - len(cmodes)
    return list(sh), list(wd), list(fp)[:nfree], nf.value, nc.value


def common_cwidth(xdims, ydims, cx, cy):
    return [max(mode_bits(xdims[a]), mode_bits(ydims[b])) for a, b in zip(cx, cy)]


def pdep(x, mask):
    return lib().orc_pdep(x, mask)


def linearize(idx, shift, width):
    order = len(idx)
    n = len(idx[0]) if order else 0
    arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in idx]
    ptrs = (c_void_p * order)(*[a.ctypes.data for a in arrs])
    keys = np.empty(n, np.uint64)
    lib().orc_linearize(order, n, ptrs, (c_int * order)(*shift), (c_int * order)(*width),
                        keys.ctypes.data)
    return keys


def sort_pairs(keys, vals):
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    v = np.ascontiguousarray(vals, dtype=np.float32).copy()
    assert lib().orc_sort_pairs(k.ctypes.data, v.ctypes.data, len(k)) == 0
    return k, v


def index_match(xkeys, ykeys, ncbits, y_nfbits):
    nx = len(xkeys)
    mpos = np.empty(nx, np.uint64)
    mcnt = np.empty(nx, np.uint64)
    prefix = np.empty(nx + 1, np.uint64)
    xk = np.ascontiguousarray(xkeys, dtype=np.uint64)
    yk = np.ascontiguousarray(ykeys, dtype=np.uint64)
    ir = lib().orc_index_match(xk.ctypes.data, nx, yk.ctypes.data, len(yk), ncbits, y_nfbits,
                               mpos.ctypes.data, mcnt.ctypes.data, prefix.ctypes.data)
    return mpos, mcnt, prefix, int(ir)


def esc(xkeys, xvals, x_begin, x_end, ykeys, yvals, mpos, mcnt, ncbits, y_nfbits):
    """-> (xf, yf, value_f32_sequential, value_f64) sorted by (xf, yf)."""
    ptrs = [c_void_p() for _ in range(4)]
    args = [np.ascontiguousarray(a) for a in (xkeys, xvals, ykeys, yvals, mpos, mcnt)]
    n = lib().orc_esc(args[0].ctypes.data, args[1].ctypes.data, x_begin, x_end,
                      args[2].ctypes.data, args[3].ctypes.data, args[4].ctypes.data,
                      args[5].ctypes.data, ncbits, y_nfbits, *[byref(p) for p in ptrs])
    assert n >= 0
    return (_take(ptrs[0], n, np.uint64), _take(ptrs[1], n, np.uint64),
            _take(ptrs[2], n, np.float32), _take(ptrs[3], n, np.float64))


def dense_nnz(vals_f32):
    v = np.ascontiguousarray(vals_f32, dtype=np.float32)
    return int(lib().orc_dense_nnz(v.ctypes.data, len(v)))


def partition(keys, prtn_num, shift):
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.zeros(prtn_num + 2, np.uint64)
    n = lib().orc_partition(k.ctypes.data, len(k), prtn_num, shift, out.ctypes.data)
    return out[: n + 1]


def dynamic_partition(xkeys, ir, prefix, ncbits, avail):
    k = np.ascontiguousarray(xkeys, dtype=np.uint64)
    p = np.ascontiguousarray(prefix, dtype=np.uint64)
    ends = np.zeros(len(k) + 1, np.uint64)
    mx = c_uint64()
    n = lib().orc_dynamic_partition(k.ctypes.data, len(k), ir, p.ctypes.data, ncbits, avail,
                                    ends.ctypes.data, len(ends), byref(mx))
    if n < 0:
        return None, 0
    return ends[:n], mx.value


def join_fp64(X, Y, cx, cy, xf_filter=None):
    """Independent FP64 SpTC; -> (xfree_lin, yfree_lin, value_f64) where
    *_lin are row-major linear ids over the free modes (ascending)."""
    def coo(t):
        arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in t.idx]
        return (arrs, (c_void_p * t.order)(*[a.ctypes.data for a in arrs]),
                (c_uint64 * t.order)(*t.dims), np.ascontiguousarray(t.val, np.float32))
    xa, xp, xd, xv = coo(X)
    ya, yp, yd, yv = coo(Y)
    nc = len(cx)
    filt = None if xf_filter is None else np.ascontiguousarray(np.sort(xf_filter), np.uint64)
    ptrs = [c_void_p() for _ in range(3)]
    n = lib().orc_join_fp64(X.order, xd, X.nnz, xp, xv.ctypes.data, Y.order, yd, Y.nnz, yp,
                            yv.ctypes.data, nc, (c_int * max(1, nc))(*cx),
                            (c_int * max(1, nc))(*cy),
                            filt.ctypes.data if filt is not None else None,
                            0 if filt is None else len(filt), *[byref(p) for p in ptrs])
    assert n >= 0
    return (_take(ptrs[0], n, np.uint64), _take(ptrs[1], n, np.uint64),
            _take(ptrs[2], n, np.float64))


def pipeline(X, Y, cx, cy, x_rows=None):
    """The reference pipeline restated on the CPU: layout (shared contract
    widths), linearize, stable sort, index match, ESC.  x_rows optionally
    restricts to X free rows given as X-layout free keys.
    -> dict(xf, yf, v32, v64, ir, xl, yl)"""
    cw = common_cwidth(X.dims, Y.dims, cx, cy)
    xl = layout(X.dims, cx, 1, cw)
    yl = layout(Y.dims, cy, 2, cw)
    xk = linearize(X.idx, xl[0], xl[1])
    yk = linearize(Y.idx, yl[0], yl[1])
    xk, xv = sort_pairs(xk, X.val)
    yk, yv = sort_pairs(yk, Y.val)
    ncb, ynf = xl[4], yl[3]
    if x_rows is not None:
        sel = np.isin(xk >> np.uint64(ncb) if ncb < 64 else np.zeros_like(xk),
                      np.asarray(x_rows, np.uint64))
        xk, xv = xk[sel], xv[sel]
    mpos, mcnt, prefix, ir = index_match(xk, yk, ncb, ynf)
    xf, yf, v32, v64 = esc(xk, xv, 0, len(xk), yk, yv, mpos, mcnt, ncb, ynf)
    return dict(xf=xf, yf=yf, v32=v32, v64=v64, ir=ir, xl=xl, yl=yl, xkeys=xk, ykeys=yk,
                prefix=prefix)
