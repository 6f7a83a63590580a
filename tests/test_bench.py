"""bench.py's contract on the CPU: the one-GPU guard, and the committed counter summary that
`roofline.traffic` is read from (its walk traffic must follow from the summary's own counter row
and copy calibration)."""
import importlib.util
import json
import os
import re
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(REPO, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_walk_traffic_follows_from_the_committed_counters():
    bench = load_bench()
    traffic = bench.walk_traffic(512, "zeldovich", {"tuned": True})
    assert traffic is not None and traffic > 3.2e9  # at least the sorted positions, read once
    text = open(os.path.join(REPO, bench.TRAFFIC_PROFILE)).read()
    fetch_copy = float(re.search(r"FETCH_SIZE = ([\d,.]+) KiB", text).group(1).replace(",", ""))
    write_copy = float(re.search(r"WRITE_SIZE = ([\d,.]+) KiB", text).group(1).replace(",", ""))
    row = re.search(r"\| void tpm::k_group_walk<false> \| ([\d.]+) \| ([\d.]+) \|", text)
    fetch, write = float(row.group(1)), float(row.group(2))
    expect = (fetch * 1048576.0 / fetch_copy + write * 1048576.0 / write_copy) * 1024.0
    assert abs(traffic - expect) <= 1e-6 * expect
    # Other workloads have no committed summary: no traffic figure rather than a borrowed one.
    assert bench.walk_traffic(256, "zeldovich", {"tuned": True}) is None
    assert bench.walk_traffic(512, "nfw", {"tuned": True}) is None
