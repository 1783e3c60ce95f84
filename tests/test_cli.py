"""The reference-shaped host side: the command line (gsparc_amd/bin/gsparc,
flags and stderr lines of cmdline_opts.cpp:9-171 and main.cu:415-432), the
readers of include/gsparc/io_manager.hpp (.tns and quantum directory,
io_manager.tpp:74-484), chained contractions (multi_SpTC, -m 1) and the
reference's own main.cu SpTC block compiled unchanged against
include/gsparc/ (bin/main_block)."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as orc
from gsparc_amd import build
from gsparc_amd.coo import COOTensor, write_tns
from helpers import random_coo

BIN = os.path.join(os.path.dirname(build.__file__), "bin", "gsparc")
MAIN_BLOCK = os.path.join(os.path.dirname(build.__file__), "bin", "main_block")
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LINES = ["--------Break Down--------", "SLITOM time:", "Partition time:", "Idxmatch time:",
         "Contraction time:", "Dynamic Partition time:", "--------Elapsed Time--------",
         "Result NNZ:", "SpTC(IdxMatch + Contraction + Data Copy) time:", "Total time:"]


@pytest.fixture(scope="module")
def cli():
    if not os.path.exists(BIN):
        build.build_cli()
    return BIN


@pytest.fixture(scope="module")
def reader(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("rt") / "read_tensor")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-pthread", "-I",
                    os.path.join(REPO, "include"), "-o", out,
                    os.path.join(REPO, "tests", "dropin", "read_tensor.cpp")], check=True)
    return out


def run(cli, *args):
    return subprocess.run([cli, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                          text=True, timeout=120)


def write_quantum(d, t):
    """The quantum directory format: indices.txt holds one line per mode with
    the 1-based indices, values.txt the values."""
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "indices.txt"), "w") as f:
        for a in t.idx:
            f.write(" ".join(str(int(v) + 1) for v in a) + "\n")
    with open(os.path.join(d, "values.txt"), "w") as f:
        f.write("\n".join(repr(float(v)) for v in t.val) + "\n")


def observed(t):
    """The reader's dims: observed maximum index + 1."""
    return COOTensor(tuple(int(a.max()) + 1 for a in t.idx), t.idx, t.val)


def parse_reader(out):
    lines = out.split("\n")
    order, nnz = map(int, re.match(r"order (\d+) nnz (\d+)", lines[0]).groups())
    dims, sums = [], []
    for m in range(order):
        d, s = map(int, re.match(r"mode \d+ dim (\d+) sum (\d+)", lines[1 + m]).groups())
        dims.append(d)
        sums.append(s)
    return order, nnz, dims, sums, float(lines[1 + order].split()[1])


def test_usage_and_bad_input(cli):
    assert run(cli, "-h").returncode == 1
    p = run(cli, "-X", "/nonexistent.tns", "-Y", "/nonexistent.tns", "-c", "1", "-x", "0",
            "-y", "0")
    assert p.returncode == 1 and "Cannot open file" in p.stderr
    # the reference's echo lines come out while parsing
    assert "tensor X: /nonexistent.tns" in p.stderr
    assert "# contract modes of X: 1, contract modes: 0" in p.stderr
    p = run(cli, "-X", "a", "-Y", "b", "-c", "1", "-x", "0", "-y", "0", "-g", "2")
    assert p.returncode == 1 and "one GPU" in p.stderr


@pytest.mark.parametrize("quantum", [False, True])
def test_readers(reader, tmp_path, quantum):
    """Both readers return the written tensor: 0-based indices, dims =
    observed maximum + 1, values (the FP64 sum of the FP32 values)."""
    rng = np.random.default_rng(5)
    dims = (30, 4000, 7, 100)
    idx = [rng.integers(0, d, 200_000).astype(np.uint32) for d in dims]
    t = COOTensor(dims, idx, rng.random(200_000).astype(np.float32))
    if quantum:
        path = str(tmp_path / "qt")
        write_quantum(path, t)
        p = subprocess.run([reader, path, "-q"], capture_output=True, text=True)
    else:
        path = str(tmp_path / "t.tns")
        with open(path, "w") as f:  # a comment and a blank line are skipped
            f.write("# comment\n\n")
        with open(path, "a") as f:
            for i in range(t.nnz):
                f.write(" ".join(str(int(a[i]) + 1) for a in idx) + f" {float(t.val[i])!r}\n")
        p = subprocess.run([reader, path], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    order, nnz, rdims, sums, vsum = parse_reader(p.stdout)
    assert (order, nnz) == (4, t.nnz)
    assert rdims == [int(a.max()) + 1 for a in idx]
    assert sums == [int(a.astype(np.int64).sum()) for a in idx]
    assert vsum == pytest.approx(float(t.val.astype(np.float64).sum()), rel=1e-12)


def test_reader_rejects_malformed(reader, tmp_path):
    bad = tmp_path / "bad.tns"
    bad.write_text("1 2 3 0.5\n1 2 3\n")  # second line: indices and no value
    p = subprocess.run([reader, str(bad)], capture_output=True, text=True)
    assert p.returncode == 1 and "missing value" in p.stderr
    short = tmp_path / "short.tns"
    short.write_text("1 2 3 0.5\n1 2\n")  # second line: one index short
    p = subprocess.run([reader, str(short)], capture_output=True, text=True)
    assert p.returncode == 1 and "bad index" in p.stderr
    zero = tmp_path / "zero.tns"
    zero.write_text("0 2 3 0.5\n")  # 1-based: 0 is not an index
    p = subprocess.run([reader, str(zero)], capture_output=True, text=True)
    assert p.returncode == 1 and "bad index" in p.stderr


@pytest.mark.skipif(not os.path.exists(build.REF_MAIN), reason="reference checkout absent")
def test_reference_main_block_compiles_unchanged(tmp_path):
    """main.cu's SpTC block (the 64-bit path with its multi-SpTC branch and
    the >64-bit extra:: path), extracted verbatim from the reference
    checkout, compiles and links against include/gsparc/ and the library."""
    out = build.build_main_block(out=str(tmp_path / "main_block"))
    assert os.path.exists(out)
    first, last = build.extract_main_block(build.REF_MAIN, str(tmp_path / "block.inc"))
    text = (tmp_path / "block.inc").read_text()
    for call in ("FindPartitionNum", "ConvertTensor_extra", "SortSlitomXY", "Partition_128",
                 "convert_and_sort_64", "gsparc::multi_SpTC", "gsparc::extra::SpTC",
                 "gsparc::SpTC<tensor_t, slitom_t>"):
        assert call in text
    assert last - first > 250


def _tns_pair(tmp_path):
    rng = np.random.default_rng(31)
    X = random_coo(rng, (40, 30, 20, 8), 6000)
    Y = random_coo(rng, (8, 20, 25), 3000)
    px, py = tmp_path / "x.tns", tmp_path / "y.tns"
    write_tns(str(px), X)
    write_tns(str(py), Y)
    return X, Y, str(px), str(py)


@pytest.mark.gpu
def test_result_nnz_matches_the_library(gpu, cli, tmp_path):
    from gsparc_amd import sptc
    X, Y, px, py = _tns_pair(tmp_path)
    p = run(cli, "-X", px, "-Y", py, "-c", "2", "-x", "3", "2", "-y", "0", "1", "-n", "2",
            "-d", "-1")
    assert p.returncode == 0, p.stderr
    for line in LINES:
        assert line in p.stderr
    nnz = int(re.search(r"Result NNZ: (\d+)", p.stderr).group(1))
    assert nnz == sptc.SpTC(observed(X), observed(Y), [3, 2], [0, 1]).nnz > 0
    # the quantum directory form of the same tensors gives the same result
    write_quantum(str(tmp_path / "qx"), X)
    write_quantum(str(tmp_path / "qy"), Y)
    q = run(cli, "-X", str(tmp_path / "qx"), "-Y", str(tmp_path / "qy"), "-c", "2", "-x", "3",
            "2", "-y", "0", "1", "-q")
    assert q.returncode == 0, q.stderr
    assert int(re.search(r"Result NNZ: (\d+)", q.stderr).group(1)) == nnz


@pytest.mark.gpu
def test_chained_contraction(gpu, cli, tmp_path):
    """-m 1 (Table 4's multi-SpTC): Z1 = X x_{1,2} X, then Z1's trailing two
    modes (X's free modes 0 and 3, from the second factor) contracted with
    X's modes 0 and 3.  The command line, the library chaining on device
    results, and the oracle on the decoded intermediate agree."""
    from gsparc_amd import sptc
    rng = np.random.default_rng(8)
    X = observed(random_coo(rng, (30, 25, 20, 6), 4000))
    path = tmp_path / "x.tns"
    write_tns(str(path), X)
    p = run(cli, "-X", str(path), "-Y", str(path), "-c", "2", "-a", "2", "-x", "1", "2", "-y",
            "1", "2", "-z", "0", "3", "-m", "1")
    assert p.returncode == 0, p.stderr
    assert "First SpTC time:" in p.stderr
    nnz = int(re.search(r"Result NNZ: (\d+)", p.stderr).group(1))
    z1 = sptc.SpTC(X, X, [1, 2], [1, 2])
    z2 = sptc.SpTC(z1, X, [2, 3], [0, 3])  # z1 used from device memory
    assert nnz == z2.nnz > 0
    idx, v = z1.decode()
    Z1 = COOTensor(tuple(int(d) for d in (X.dims[0], X.dims[3], X.dims[0], X.dims[3])),
                   [a.astype(np.uint32) for a in idx], v)
    ref = orc.pipeline(Z1, X, [2, 3], [0, 3])
    xf, yf, vals = z2.to_host()
    np.testing.assert_array_equal(xf, ref["xf"])
    np.testing.assert_array_equal(yf, ref["yf"])
    rel = np.abs(vals.astype(np.float64) - ref["v64"]) / np.abs(ref["v64"])
    assert rel.max() <= 1e-5


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(MAIN_BLOCK), reason="bin/main_block not built")
def test_reference_main_block_runs(gpu, cli, tmp_path):
    """The reference's own main.cu SpTC block, compiled unchanged, reports
    the same Result NNZ as the command line, plain and chained."""
    X, Y, px, py = _tns_pair(tmp_path)
    args = ["-X", px, "-Y", py, "-c", "2", "-x", "3", "2", "-y", "0", "1"]
    a, b = run(MAIN_BLOCK, *args), run(cli, *args)
    assert a.returncode == 0 and b.returncode == 0, a.stderr + b.stderr
    pick = lambda s: int(re.search(r"Result NNZ: (\d+)", s).group(1))  # noqa: E731
    assert pick(a.stderr) == pick(b.stderr) > 0
    chain = ["-X", px, "-Y", px, "-c", "2", "-a", "2", "-x", "1", "2", "-y", "1", "2", "-z",
             "0", "3", "-m", "1"]
    a, b = run(MAIN_BLOCK, *chain), run(cli, *chain)
    assert a.returncode == 0 and b.returncode == 0, a.stderr + b.stderr
    assert pick(a.stderr) == pick(b.stderr) > 0
