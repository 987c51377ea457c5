"""Times the compiled reference against the FP64 oracle on the CPU.

TEST INFRASTRUCTURE ONLY: it puts the oracle's CPU timing (the baseline
bench.py reports from the GPU box, where the reference never goes) in
context.  It runs where oracle/_ref/cosims exists (built from the reference
sources by ``make -C oracle ref``), never on the GPU box.

Workload: the 128-atom synthetic ion of tests/golden/ion128.mfj, exact
electrostatics + full LJ (force type 4), He at 298 K, 3 CCS integrals x
20,000 trajectories.  The reference runs its own command line and reports
the trajectory phase's wall time in its log ("Total time", summed over
integrals: src/monteTraj.cpp:309, 382-383, 397); the oracle integrates the
hits among the same number of Philox-sampled trajectories (its setup and
sampling are outside the timing, as the reference's are).  Both at 1 thread
and at every thread of the machine.  Prints one JSON line.

    python oracle/time_ref.py [--threads N]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

REF_BIN = os.path.join(HERE, "_ref", "cosims")
MOL = os.path.join(ROOT, "tests", "golden", "ion128.mfj")
ITERS, TRAJ = 3, 20000
INPUT = ("temp 298\ntraj 20\niter 3\nmultipole false\ndispersion_cutoff false\n"
         "multipole_radius 20\ndispersion_radius 23.43015325287593\n")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_reference(threads):
    work = tempfile.mkdtemp(prefix="cosims_tref_")
    shutil.copy(MOL, work)
    with open(os.path.join(work, "in.txt"), "w") as f:
        f.write(INPUT)
    t0 = time.perf_counter()
    r = subprocess.run([REF_BIN, "-m", "ion128.mfj", "-i", "in.txt", "-n", str(threads), "-s", "123456789"],
                       cwd=work, capture_output=True, text=True, timeout=3600)
    wall = time.perf_counter() - t0
    if r.returncode != 0:
        raise SystemExit(f"reference run failed: {r.stderr[-500:]}")
    logs = [f for f in os.listdir(work) if f.startswith("omega_")]
    text = open(os.path.join(work, logs[0])).read()
    phase = [float(x) for x in re.findall(r"Total time:\s+([\d.]+);", text)]
    omega = [float(x) for x in re.findall(r"Omega:\s+([\d.]+) \+/-", text)]
    shutil.rmtree(work, ignore_errors=True)
    return dict(traj_phase_s=phase[-1], wall_s=wall, omega=omega)


def time_oracle(threads):
    import oracle
    from cosims_amd import engine as E
    p = E.default_params(temp=298.0, traj=TRAJ, iter=ITERS, multipole=0, dispersion_cutoff=0, seed=123456789,
                         setup_gpu=0)
    m = E.Molecule.from_file(MOL, params=p)  # host-only setup (no GPU call)
    inf = m.info()
    A = m.atoms()
    mo = oracle.Mol(A["xyz"], A["q"], A["sig6"], A["eps24m"], A["eps4"], inf["charged"], inf["mu"],
                    inf["force_type"])
    cfg = dict(seed=123456789, job=0, iteration=0, gmin=inf["gmin"], gmax=inf["gmax"], max_val=inf["max_val"],
               vd_a=inf["vd_a"], vd_norm=inf["vd_norm"], bmax=inf["bmax"], ea=inf["a"], eb=inf["b"], ec=inf["c"],
               projection=1)
    total, hits = 0.0, 0
    for k in range(ITERS):
        cfg["iteration"] = k
        s = oracle.sample_ic(cfg, np.arange(TRAJ))
        h = s["hit"] == 1
        hits += int(h.sum())
        t0 = time.perf_counter()
        mo.verlet(p.dt, s["p0"][h], s["v0"][h], (inf["a"], inf["b"], inf["c"]), nthreads=threads)
        total += time.perf_counter() - t0
    return dict(traj_phase_s=total, hits=hits)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    if not os.path.exists(REF_BIN):
        raise SystemExit("oracle/_ref/cosims missing: run `make -C oracle ref` where the reference sources are")
    n = ITERS * TRAJ
    out = {"workload": "128-atom ion, type 4, He 298 K, 3 x 20,000 trajectories", "cpu": cpu_model(),
           "trajectories": n}
    for t in sorted({1, a.threads}):
        ref = time_reference(t)
        orc = time_oracle(t)
        out[f"threads_{t}"] = {
            "reference_traj_per_s": n / ref["traj_phase_s"], "oracle_traj_per_s": n / orc["traj_phase_s"],
            "oracle_over_reference": ref["traj_phase_s"] / orc["traj_phase_s"],
            "reference_s": ref["traj_phase_s"], "oracle_s": orc["traj_phase_s"], "reference_omega": ref["omega"]}
        print(f"{t} thread(s): reference {ref['traj_phase_s']:.2f} s, oracle {orc['traj_phase_s']:.2f} s",
              file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
