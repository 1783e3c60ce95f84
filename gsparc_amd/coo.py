"""In-memory COO tensors and the seeded generator.

A :class:`COOTensor` mirrors the reference's ``SparseTensor<uint64_t,float>``
(sparse_tensor.hpp:12-44): SoA 0-based indices per mode, FP32 values and
dims.  Indices are held as uint32 (every dim handled here is < 2^32).
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from ._lib import gen_lib

UNIFORM, ZIPF = 0, 1
VAL_UNIFORM01, VAL_LOGNORMAL = 0, 1


@dataclass
class COOTensor:
    dims: tuple
    idx: list  # list of np.uint32 arrays, one per mode
    val: np.ndarray  # float32
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def order(self):
        return len(self.dims)

    @property
    def nnz(self):
        return int(self.val.shape[0])

    def index_matrix(self):
        return np.stack(self.idx, axis=1) if self.idx else np.zeros((0, 0))


def splitmix64(z):
    lib = gen_lib()
    lib.sptc_gen_splitmix64.restype = ctypes.c_uint64
    lib.sptc_gen_splitmix64.argtypes = [ctypes.c_uint64]
    return int(lib.sptc_gen_splitmix64(ctypes.c_uint64(z & (2**64 - 1))))


def generate(seed, dims, dist, zipf_s, nnz, value_dist=VAL_UNIFORM01, name=""):
    """Draw `nnz` distinct coordinates (see gsparc_amd/csrc/gen.cpp)."""
    order = len(dims)
    lib = gen_lib()
    lib.sptc_gen_coo.restype = ctypes.c_int
    u64a = (ctypes.c_uint64 * order)(*[int(d) for d in dims])
    dista = (ctypes.c_int * order)(*[int(x) for x in dist])
    sa = (ctypes.c_double * order)(*[float(x) for x in zipf_s])
    idx = [np.empty(int(nnz), dtype=np.uint32) for _ in range(order)]
    val = np.empty(int(nnz), dtype=np.float32)
    ptrs = (ctypes.c_void_p * order)(*[a.ctypes.data for a in idx])
    rc = lib.sptc_gen_coo(ctypes.c_uint64(seed & (2**64 - 1)), ctypes.c_int(order),
                          u64a, dista, sa, ctypes.c_uint64(int(nnz)),
                          ctypes.c_int(value_dist), ptrs,
                          val.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise ValueError(f"sptc_gen_coo failed with {rc} for dims={dims} nnz={nnz}")
    return COOTensor(tuple(int(d) for d in dims), idx, val, name=name)


def observed_dims(t):
    """dims = max index + 1 per mode, as the .tns reader defines them
    (io_manager.tpp:213, 250-256)."""
    return tuple(int(a.max()) + 1 if a.size else 1 for a in t.idx)


def write_tns(path, t, based=1):
    """Write a FROSTT .tns text file (1-based indices, value last)."""
    cols = [a.astype(np.int64) + based for a in t.idx]
    with open(path, "w") as f:
        for i in range(t.nnz):
            f.write(" ".join(str(int(c[i])) for c in cols))
            f.write(" %.9g\n" % float(t.val[i]))


def write_tns_fast(path, t, based=1):
    """Vectorised .tns writer for large tensors."""
    fmt = " ".join(["%d"] * t.order) + " %.9g"
    arr = np.column_stack([a.astype(np.int64) + based for a in t.idx] +
                          [t.val.astype(np.float64)])
    np.savetxt(path, arr, fmt=fmt)


def read_tns(path, based=1):
    """Read a .tns file: whitespace separated, `order` indices then a value;
    1-based by default; dims = observed max + 1 (io_manager.tpp:74-259)."""
    rows = np.loadtxt(path, dtype=np.float64, ndmin=2)
    order = rows.shape[1] - 1
    idx = [(rows[:, m].astype(np.int64) - based).astype(np.uint32) for m in range(order)]
    val = rows[:, order].astype(np.float32)
    t = COOTensor(tuple([0] * order), idx, val)
    t.dims = observed_dims(t)
    return t
