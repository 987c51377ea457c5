"""The command-line program on the GPU against the reference's example logs.

With the reference's initial-condition stream and the FP64
lane-per-trajectory kernel (-p fp64 -r reference -l 1), the program's
omega_<name>.txt must reproduce the example log (examples/*/omega_*.txt)
line for line: settings, atom types, boundary fit, clusters, velocity
window, every progress line, every CCS integral and the final statistics
(dates, wall times and paths aside).  The default mode (mixed precision,
Philox stream, tuned launch) draws other trajectories, so its final CCS is
checked against the reference's within the two runs' combined standard error.
"""
import math
import os
import re

import pytest

from conftest import LOGS, run_cli_example

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("entry", LOGS, ids=[e["log"].split("/")[-1] for e in LOGS])
def test_cli_reproduces_example_log(entry, tmp_path):
    from golden.logfilter import log_lines
    print(f"cli fp64/reference/lane: {entry['log']}", flush=True)
    rc, out, log = run_cli_example(entry, str(tmp_path), ["-p", "fp64", "-r", "reference", "-l", "1"])
    assert rc == 0, out[-2000:]
    got = log_lines(open(log).read())
    want = entry["lines"]
    for i, (g, w) in enumerate(zip(got, want)):
        assert g == w, f"line {i}: {g!r} != {w!r}"
    assert len(got) == len(want)


def test_cli_default_mode_agrees_statistically(tmp_path):
    entry = next(e for e in LOGS if e["log"].endswith("omega_acau.txt"))
    print("cli default mode: acau charged", flush=True)
    rc, out, log = run_cli_example(entry, str(tmp_path))
    assert rc == 0, out[-2000:]
    text = open(log).read()
    assert " Engine: mixed precision, Philox initial conditions" in text
    wm = float(re.search(r"Weighted CCS Mean:\s+([\d.]+)", text).group(1))
    we = float(re.search(r"Weighted Error:\s+([\d.]+)", text).group(1))
    ref_wm, ref_we = float(entry["final"]["wmean"]), float(entry["final"]["werr"])
    assert abs(wm - ref_wm) <= 4 * math.hypot(we, ref_we), (wm, we, ref_wm, ref_we)
    assert len(re.findall(r"^ Omega:", text, re.M)) == entry["iters"]
    assert len(re.findall(r"^ progress:", text, re.M)) == 2 * entry["iters"]
