"""The reference-compatible command line (gsparc_amd/bin/gsparc): flags and
stderr lines of the reference's driver (cmdline_opts.cpp:9-171,
main.cu:415-432) over the MI355X engine."""
import os
import re
import subprocess

import numpy as np
import pytest

from gsparc_amd import build
from gsparc_amd.coo import COOTensor, write_tns
from helpers import random_coo

BIN = os.path.join(os.path.dirname(build.__file__), "bin", "gsparc")
LINES = ["--------Break Down--------", "SLITOM time:", "Partition time:", "Idxmatch time:",
         "Contraction time:", "Dynamic Partition time:", "--------Elapsed Time--------",
         "Result NNZ:", "SpTC(IdxMatch + Contraction + Data Copy) time:", "Total time:"]


@pytest.fixture(scope="module")
def cli():
    if not os.path.exists(BIN):
        build.build_cli()
    return BIN


def run(cli, *args):
    return subprocess.run([cli, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                          text=True, timeout=120)


def test_usage_and_rejected_flags(cli):
    assert run(cli, "-h").returncode == 1
    p = run(cli, "-X", "a.tns", "-Y", "b.tns", "-c", "1", "-x", "0", "-y", "0", "-q")
    assert p.returncode == 1 and "not supported" in p.stderr
    p = run(cli, "-X", "/nonexistent.tns", "-Y", "/nonexistent.tns", "-c", "1", "-x", "0",
            "-y", "0")
    assert p.returncode == 1 and "cannot read" in p.stderr
    # the reference's echo lines come out while parsing
    assert "tensor X: /nonexistent.tns" in p.stderr
    assert "# contract modes of X: 1, contract modes: 0" in p.stderr


@pytest.mark.gpu
def test_result_nnz_matches_the_library(gpu, cli, tmp_path):
    from gsparc_amd import sptc
    rng = np.random.default_rng(31)
    X = random_coo(rng, (40, 30, 20, 8), 6000)
    Y = random_coo(rng, (8, 20, 25), 3000)
    px, py = tmp_path / "x.tns", tmp_path / "y.tns"
    write_tns(str(px), X)
    write_tns(str(py), Y)
    p = run(cli, "-X", str(px), "-Y", str(py), "-c", "2", "-x", "3", "2", "-y", "0", "1",
            "-n", "2", "-d", "-1")
    assert p.returncode == 0, p.stderr
    for line in LINES:
        assert line in p.stderr
    nnz = int(re.search(r"Result NNZ: (\d+)", p.stderr).group(1))
    # the .tns reader takes dims from the observed maximum index
    Xo = COOTensor(tuple(int(a.max()) + 1 for a in X.idx), X.idx, X.val)
    Yo = COOTensor(tuple(int(a.max()) + 1 for a in Y.idx), Y.idx, Y.val)
    assert nnz == sptc.SpTC(Xo, Yo, [3, 2], [0, 1]).nnz > 0
