#!/usr/bin/env python
"""He trajectories/s at 2,000 atoms, mixed precision vs FP64, on one MI355X.

Workload: the synthetic 2,000-atom globular ion (C/H/N/O/P 30:50:10:8:2,
radius 18 A, q ~ N(0, 0.3 e) shifted to +8 e), He at 298 K, 100 CCS
integrals x 10,000 trajectories = 1,000,000 trajectories per step, exact
electrostatics + full LJ (force type 4), dt 0.01 ps.  One step = one
complete CCS run (initial-condition sampling, every Verlet attempt round,
energy checks, reduction) with the molecule already prepared and resident in
HBM.  The FP64 run on the same (paired) initial conditions gives the CCS
deviation of the mixed-precision result.

Prints one JSON line (see keys below).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "He trajectories/s at 2k atoms (mixed vs FP64, ≤0.1% CCS); p50/p99 CCS latency"
FLOP_PAIR, FLOP_EPI, FLOP_VERLET = 58, 21, 50   # algorithmic FLOPs, type-4 pair / epilogue / step
PEAK_FP32 = 157.3          # TFLOP/s, MI355X FP32 vector (packed) peak
PEAK_FP32_UNPACKED = 78.6  # TFLOP/s, FP32 vector without packed math (= FP64 vector)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_molecule(E, synth, traj, iters, seed):
    ion = synth.config_ion("ion2000")
    d = tempfile.mkdtemp(prefix="cosims_bench_")
    path = os.path.join(d, "ion2000.mfj")
    synth.write_mfj(path, "ion2000", ion)
    p = E.default_params(temp=298.0, traj=traj, iter=iters, multipole=0, dispersion_cutoff=0, seed=seed)
    p.print_rate = traj // 5
    return E.Molecule.from_file(path, params=p)


def kernel_flops(n_atoms, res, r=0):
    return res["round_evals"][r] * (n_atoms * FLOP_PAIR + FLOP_EPI) + res["round_steps"][r] * FLOP_VERLET


def cpu_baseline(m, budget_s, threads):
    """The FP64 oracle on the same workload's initial conditions, bounded."""
    import oracle
    inf = m.info()
    A = m.atoms()
    mo = oracle.Mol(A["xyz"], A["q"], A["sig6"], A["eps24m"], A["eps4"], inf["charged"], inf["mu"],
                    inf["force_type"])
    s = m.sample(m.params.seed, 0, 0, 0, 64 * threads)
    h = s["hit"] == 1
    done, t_used, batches = 0, 0.0, 0
    while t_used < budget_s and batches < 64:
        t0 = time.perf_counter()
        mo.verlet(m.params.dt, s["p0"][h], s["v0"][h], (inf["a"], inf["b"], inf["c"]), nthreads=threads)
        t_used += time.perf_counter() - t0
        done += int(s["hit"].size)  # sampled trajectories (misses cost nothing)
        batches += 1
    return dict(value=done / t_used, unit="trajectories/s", cores=threads, kind="port",
                sample=f"{done} trajectories of the 2,000-atom ion (FP64 oracle, OpenMP, {t_used:.1f} s)")


def copy_probe(nbytes=256 << 20, reps=3):
    """A plain copy kernel of known size for calibrating the memory counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copied {nbytes} bytes x{reps}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--traj", type=int, default=10000, help="trajectories per CCS integral")
    ap.add_argument("--iters", type=int, default=100, help="CCS integrals per step")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run 3 device-to-device copies of 256 MiB (scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    a = ap.parse_args()
    if a.copy_probe:
        copy_probe()
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); trajectories shard "
              "trivially across GPUs but no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)

    from cosims_amd import engine as E
    from cosims_amd import synth

    name, cus = E.device_info()
    log(f"device {name} ({cus} CUs)")
    t0 = time.perf_counter()
    m = make_molecule(E, synth, a.traj, a.iters, seed=20261203)
    m.upload()
    inf = m.info()
    log(f"setup {time.perf_counter() - t0:.2f} s: n={inf['n']} type={inf['force_type']} "
        f"axes=({inf['a']:.2f},{inf['b']:.2f},{inf['c']:.2f})")
    n_traj = a.traj * a.iters

    for w in range(a.warmup):
        r = m.run(E.PREC_MIXED, job=0)
        log(f"warmup {w}: {r['run_ms']:.1f} ms")
    E.device_sync()
    times, runs = [], []
    t_all = time.perf_counter()
    for k in range(a.steps):
        t1 = time.perf_counter()
        r = m.run(E.PREC_MIXED, job=0)
        times.append(time.perf_counter() - t1)
        runs.append(r)
        log(f"step {k}: {times[-1] * 1e3:.1f} ms, traj kernel {r['traj_kernel_ms']} ms, "
            f"CCS {r['wmean']:.3f} +- {r['werr']:.3f}, lanes {r['lanes']} wg {r['wg_size']}")
    E.device_sync()
    elapsed = time.perf_counter() - t_all
    value = n_traj * a.steps / elapsed
    mixed = runs[-1]

    fp64 = None
    if not a.no_fp64:
        m.run(E.PREC_FP64, job=1, iters=1, traj=2000)   # tunes the FP64 launch shape outside the timing
        E.device_sync()
        t1 = time.perf_counter()
        fp64 = m.run(E.PREC_FP64, job=0)
        fp64_s = time.perf_counter() - t1
        log(f"fp64: {fp64_s * 1e3:.1f} ms, CCS {fp64['wmean']:.3f} +- {fp64['werr']:.3f}, "
            f"lanes {fp64['lanes']} wg {fp64['wg_size']}")

    # roofline of the dominant kernel (first trajectory launch of each step)
    kms = np.array([r["traj_kernel_ms"][0] for r in runs])
    flops = np.array([kernel_flops(inf["n"], r) for r in runs])
    achieved = float(np.mean(flops / (kms * 1e-3)) / 1e12)

    out = {
        "metric": METRIC, "value": value, "unit": "trajectories/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": elapsed / a.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 force sum / fp64 integrator",
        "data": "synthetic",
        "config": {"workload": "2,000-atom synthetic ion, +8 e, He 298 K, 100 x 10,000 trajectories",
                   "atoms": inf["n"], "trajectories_per_step": n_traj, "force_type": inf["force_type"],
                   "precision": "mixed"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32, "frac_unpacked_fp32": achieved / PEAK_FP32_UNPACKED,
                     "traffic": None, "kernel": "k_traj<mixed,type4>",
                     "kernel_ms": float(kms.mean())},
        "ccs": {"mixed_wmean": mixed["wmean"], "mixed_werr": mixed["werr"]},
        "launch": {"lanes_per_trajectory": mixed["lanes"], "wg_size": mixed["wg_size"],
                   "n_retry_or_restart": mixed["n_retry"], "n_fail": mixed["n_fail"],
                   "rounds": mixed["n_rounds"], "sample_ms": mixed["sample_kernel_ms"]},
        "steps_per_s": mixed["total_steps"] / (times[-1]),
        "pair_evals_per_s": mixed["pair_evals"] / times[-1],
    }
    if fp64 is not None:
        d = mixed["omega"] - fp64["omega"]
        out["ccs"].update(fp64_wmean=fp64["wmean"], fp64_werr=fp64["werr"],
                          rel_dev=abs(mixed["wmean"] - fp64["wmean"]) / fp64["wmean"],
                          paired_rel_dev=float(d.mean() / fp64["omega"].mean()),
                          paired_rel_se=float(d.std(ddof=1) / np.sqrt(d.size) / fp64["omega"].mean()))
        out["fp64_value"] = n_traj / (fp64_s)
        out["mixed_over_fp64"] = value / out["fp64_value"]
        out["fp64_kernel_ms"] = fp64["traj_kernel_ms"][0]
        out["fp64_launch"] = {"lanes_per_trajectory": fp64["lanes"], "wg_size": fp64["wg_size"]}
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")), 16)
    if a.cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline(m, a.cpu_seconds, threads)
    out["served_latency_ms"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
