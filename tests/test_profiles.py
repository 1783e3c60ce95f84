"""The committed rocprofv3 summaries agree with themselves and with what
bench.py reads from them: the `bench-traffic` line of a workload's counter
summary (bytes per launch of the dominant kernel class) must equal its
FETCH_SIZE and WRITE_SIZE table rows for that kernel, scaled by the
calibration factors the summary states.  CPU only."""
import importlib.util
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROFILES = os.path.join(REPO, "profiles")
# kernel class -> kernel whose counters the class's launches are
CLASS_KERNEL = {"sort_products": "k_seg_downsweep<unsigned int, unsigned int, 10, 16"}


def _section(text, counter):
    start = text.index(f"## {counter}")
    nxt = text.find("\n## ", start + 3)
    return text[start:nxt if nxt >= 0 else len(text)]


def _factor(section):
    return float(re.search(r"\(factor ([0-9.]+) = bytes / \(KiB reading\)\)", section).group(1))


def _avg_kib(section, kernel):
    for line in section.splitlines():
        if line.startswith("| `" + kernel):
            return float(line.split("|")[3])
    raise AssertionError(f"{kernel} not in the table")


@pytest.mark.parametrize("name", sorted(f for f in os.listdir(PROFILES) if f.endswith("_pmc.md")))
def test_bench_traffic_matches_counter_tables(name):
    text = open(os.path.join(PROFILES, name)).read()
    lines = [l for l in text.splitlines() if l.startswith("bench-traffic:")]
    assert len(lines) == 1
    _, cls, value = lines[0].split()
    kernel = CLASS_KERNEL[cls]
    total = 0.0
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        sec = _section(text, counter)
        total += _avg_kib(sec, kernel) * _factor(sec) * 1024.0
    assert abs(float(value) - total) <= 1e-3 * total, (value, total)


def test_bench_reads_the_traffic_line():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(REPO, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    text = open(os.path.join(PROFILES, "nips_pmc.md")).read()
    value = float(re.search(r"^bench-traffic: sort_products ([0-9.e+]+)", text, re.M).group(1))
    assert bench.measured_traffic("nips", "sort_products") == value
    assert bench.measured_traffic("nips", "no_such_class") is None
