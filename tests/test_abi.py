"""The C ABI library: loads, exports every function include/*.h declares, and
its host-only layout routine agrees with the oracle.  CPU only (no GPU
compute is invoked)."""
import os
import re

import numpy as np
import pytest

import oracle as orc
from gsparc_amd import _lib, sptc

INCLUDE = os.path.join(_lib.REPO_DIR, "include")


def _declared(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sptc_[a-z0-9_]+)\s*\(", text)))


def _exported(path):
    import subprocess
    out = subprocess.check_output(["nm", "-D", "--defined-only", path]).decode()
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


@pytest.mark.parametrize("header,lib", [("sptc.h", _lib.HIP_LIB), ("sptc_gen.h", _lib.GEN_LIB)])
def test_every_declared_symbol_is_exported(header, lib):
    names = _declared(header)
    assert len(names) >= 2
    exported = _exported(lib)
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_library_loads():
    L = sptc.lib()
    assert L.sptc_status_string(0) == b"ok"
    assert L.sptc_status_string(3) == b"key needs more than 64 bits"


@pytest.mark.parametrize("seed", range(15))
def test_layout_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    order = int(rng.integers(1, 8))
    dims = [int(x) for x in rng.integers(1, 100000, size=order)]
    nc = int(rng.integers(0, order + 1))
    cm = [int(x) for x in rng.permutation(order)[:nc]]
    for role in (1, 2):
        try:
            got = sptc.layout(dims, cm, role)
        except sptc.SpTCError as e:
            assert e.status == sptc.ERR_KEY_OVERFLOW
            assert sum(orc.mode_bits(d) for d in dims) > 64
            continue
        sh, wd, fp, nf, ncb = orc.layout(dims, cm, role)
        assert got == (sh, wd, nf, ncb)


def test_layout_rejects_bad_modes():
    with pytest.raises(sptc.SpTCError) as e:
        sptc.layout([10, 10], [2], 1)
    assert e.value.status == sptc.ERR_BAD_ARGS
    with pytest.raises(sptc.SpTCError):
        sptc.layout([10, 10], [0, 0], 1)


def test_find_partition_num_host_only():
    """FindPartitionNum's part counts: whole operands when they fit half the
    budget, more parts as the budget shrinks (never fewer), an error when 64
    parts per operand are not enough; a larger Y is split before X grows
    past it."""
    MB = 1 << 20
    assert sptc.find_partition_num(10_000, 3, 10_000, 3, 4, 1 << 30) == (1, 1)
    prev = (1, 1)
    for budget in (2048, 1024, 512, 256, 128, 64):
        px, py = sptc.find_partition_num(3_000_000, 3, 3_000_000, 3, 4, budget * MB)
        assert px >= 1 and py >= 1
        assert px * py >= prev[0] * prev[1]
        prev = (px, py)
    assert prev[0] * prev[1] > 1
    with pytest.raises(sptc.SpTCError) as e:
        sptc.find_partition_num(3_000_000, 3, 3_000_000, 3, 4, 1 * MB)
    assert e.value.status == sptc.ERR_OOM_BUDGET
    # Y alone too large: Y is split, X stays whole
    assert sptc.find_partition_num(1000, 3, 20_000_000, 3, 4, 1800 * MB)[0] == 1
