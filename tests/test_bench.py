"""bench.py's contract on the host: one GPU only, the reference FLOP counts,
and the HBM traffic constant it reports against the profile it cites."""
import os
import re
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


import pytest


@pytest.mark.parametrize("extra", [[], ["--workload", "batch"], ["--workload", "clustered"], ["--copy-probe"]])
def test_more_than_one_gpu_is_refused(extra):
    """Every workload refuses --gpus other than 1 before touching a GPU."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", *extra],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_reference_flop_counts():
    # src/forceAccel.cpp:764-802 (pair 58, epilogue 21), src/diffEqSolver.cpp:131-174 (step ~50)
    assert (bench.FLOP_PAIR, bench.FLOP_EPI, bench.FLOP_VERLET) == (58, 21, 50)
    res = {"round_evals": [10], "round_steps": [7]}
    assert bench.kernel_flops(2000, res) == 10 * (2000 * 58 + 21) + 7 * 50


def test_traffic_constant_matches_its_profile():
    with open(os.path.join(ROOT, "profiles", "traj_grp_mixed_ion2000.md")) as f:
        text = f.read()
    m = re.search(r"`traffic` = ([0-9,]+) B", text)
    assert m, "the profile states the traffic bench.py reports"
    assert int(m.group(1).replace(",", "")) == bench.TRAFFIC_BYTES
    fetch = re.search(r"FETCH_SIZE ([0-9,]+) KiB \(x 2\.0", text)
    write = re.search(r"WRITE_SIZE ([0-9,]+) KiB \(", text)
    assert fetch and write
    kib = 2 * int(fetch.group(1).replace(",", "")) + int(write.group(1).replace(",", ""))
    assert kib * 1024 == bench.TRAFFIC_BYTES


def test_cpu_ratio_constant_matches_its_profile():
    with open(os.path.join(ROOT, "profiles", "cpu_reference_vs_oracle_ion128.md")) as f:
        text = f.read()
    rows = dict(re.findall(r"^\| (\d+) \|.*\| ([0-9.]+) \|$", text, re.M))
    assert float(rows["1"]) == bench.ORACLE_OVER_REFERENCE["threads_1"]
    assert float(rows["8"]) == bench.ORACLE_OVER_REFERENCE["threads_8"]
