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
# TFLOP/s, MI355X FP32 vector peak.  Reachable without packed FP32: gfx950 issues an FP32 VALU
# instruction whose operands are all VGPRs in 2 cycles a wavefront (one with an SGPR operand in 4)
PEAK_FP32 = 157.3
PEAK_FP64 = 78.6  # TFLOP/s, MI355X FP64 vector peak
# HBM bytes per launch of the timed mixed kernel (rocprofv3 --pmc FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0,
# scale factors from a 256 MiB device copy; 16 lanes, workgroup 256), from
# profiles/traj_grp_mixed_ion2000.md
TRAFFIC_BYTES = (147044 * 2 + 185893) * 1024


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


# the FP64 oracle against the compiled reference on the same CPU and workload (128-atom ion,
# 3 x 20,000 trajectories): oracle speed / reference speed at 1 and 8 threads, from
# profiles/cpu_reference_vs_oracle_ion128.md (oracle/time_ref.py)
ORACLE_OVER_REFERENCE = {"threads_1": 1.05, "threads_8": 1.21}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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
                sample=f"{done} trajectories of the 2,000-atom ion (FP64 oracle, OpenMP, {t_used:.1f} s)",
                cpu=cpu_model(), nproc=os.cpu_count(),
                port_speed_over_reference=ORACLE_OVER_REFERENCE,
                port_speed_source="profiles/cpu_reference_vs_oracle_ion128.md")


def pct(a, q):
    return float(np.percentile(np.asarray(a, dtype=float), q)) if len(a) else None


def serving(E, synth, n_requests, rate, target, seed=5000, policy=0, fp64_max_atoms=0, fp64_max_requests=40,
            clustered=False, **server_kw):
    """Poisson request stream through the CCS server; latency = completion -
    submission on the server's clock (setup, queueing and trajectories).
    The arrival times are replayed in real time.  clustered: the requests
    keep the reference's default cut-offs (multipole and dispersion on: PCM
    clustering, clustered force type) instead of exact electrostatics."""
    import tempfile
    reqs = synth.serving_stream(count=n_requests, rate=rate, seed=seed)
    d = tempfile.mkdtemp(prefix="cosims_serve_")
    mols = []
    for k, r in enumerate(reqs):
        path = os.path.join(d, f"req{k}.mfj")
        synth.write_mfj(path, f"req{k}", r["ion"])
        mols.append(E.read_molfile(path))

    def params(k):
        cut = 1 if clustered else 0
        return E.default_params(temp=reqs[k]["temp"], seed=seed + k, multipole=cut, dispersion_cutoff=cut, setup_gpu=1)
    # launch shape: lanes per trajectory by molecule size (the size rule), fixed for every request
    W = 512
    # first issue sized for a per-trajectory relative spread of 2.0 (the spread of the CCS terms runs
    # 2.3-2.9 on this stream), then the whole estimate of what is still needed each slice
    server_kw.setdefault("prior_rel_sd", 2.0)
    server_kw.setdefault("issue_fraction", 1.0)
    # workgroups of 1024 for the two largest size classes (768 atoms and up: fewer, larger LDS
    # tiles and one LDS staging shared by more groups); 950-960 k traj/s against 800-890 k at 512
    # on the same box
    server_kw.setdefault("class_wg", [0, 0, 0, 1024, 1024])
    # twice the size rule's lanes below 1,536 atoms: a served slice is a few hundred thousand
    # trajectories, so its drain weighs more than in one large run (p50 2.37-2.39 s against
    # 2.57-2.58 s, p99 and throughput no worse, same box; half the rule's lanes: 840 k against
    # 971 k traj/s)
    server_kw.setdefault("class_lanes", [8, 16, 32, 64, 64])
    log(f"serving: {n_requests} requests at {rate}/s, target {target}, lanes by size, wg {W}, {server_kw}")
    srv = E.Server(prec=E.PREC_MIXED, lanes=0, wg_size=W, block_traj=10000, min_blocks=3, max_blocks=1000,
                   policy=policy, **server_kw)
    t0 = time.perf_counter()
    tickets = []
    for k, r in enumerate(reqs):
        delay = t0 + r["t"] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        tickets.append(srv.submit(mols[k], params(k), job=k, target=target))
    res = []
    for k, t in enumerate(tickets):
        r = srv.wait(t)
        res.append(r)
        log(f"  request {k}: N={reqs[k]['n']} T={reqs[k]['temp']:.0f} K  {r['blocks']} blocks  "
            f"CCS {r['wmean']:.1f} +- {100 * r['rel_err']:.2f}%  latency {r['t_done_ms'] - r['t_submit_ms']:.0f} ms")
    st = srv.stats()
    srv.close()
    vs64 = None
    if fp64_max_atoms > 0:
        # CCS error against FP64: the same blocks of the request's stream (cosims_run over
        # iter = blocks, the request's lanes) in FP64, for the requests of up to fp64_max_atoms atoms
        dev = []
        for k, r in enumerate(res):
            if reqs[k]["n"] > fp64_max_atoms or len(dev) >= fp64_max_requests:
                continue
            m = E.Molecule.from_file(os.path.join(d, f"req{k}.mfj"), params=params(k))
            f = m.run(E.PREC_FP64, job=k, iters=r["blocks"], traj=10000, lanes=r["lanes"], wg_size=W)
            dev.append((r["wmean"] - f["wmean"]) / f["wmean"])
            m.close()
        dev = np.array(dev)
        vs64 = {"requests": int(dev.size), "max_atoms": fp64_max_atoms, "mean_rel_dev": float(dev.mean()),
                "mean_rel_dev_se": float(dev.std(ddof=1) / np.sqrt(dev.size)) if dev.size > 1 else None,
                "max_abs_rel_dev": float(np.abs(dev).max())}
        log(f"serving vs FP64: {dev.size} requests, mean rel dev {dev.mean():.2e}, max |dev| {np.abs(dev).max():.2e}")
    lat = [r["t_done_ms"] - r["t_submit_ms"] for r in res]
    makespan = (max(r["t_done_ms"] for r in res) - min(r["t_submit_ms"] for r in res)) / 1e3
    traj = sum(r["n_traj"] for r in res)
    # latency against arrival order: p50 and p99 of each fifth of the stream (a backlog shows as a
    # rise from fifth to fifth)
    fifths = np.array_split(np.arange(len(lat)), 5)
    by_arrival = [{"requests": f"{int(f[0])}-{int(f[-1])}", "p50": pct([lat[i] for i in f], 50),
                   "p99": pct([lat[i] for i in f], 99)} for f in fifths if len(f)]
    return {
        "workload": f"Poisson {rate} req/s, {n_requests} requests, N log-uniform 100-20,000, "
                    f"q round(N/800)+1 e, T 200-400 K, to {100 * target:.1f}% relative SE"
                    + (", PCM-clustered force field (multipole and dispersion cut-offs on)" if clustered
                       else ", exact electrostatics and LJ"),
        "requests": n_requests, "offered_rate": rate, "served_rate": n_requests / makespan,
        "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "mean": float(np.mean(lat)), "max": max(lat)},
        "latency_by_arrival_fifth_ms": by_arrival,
        "setup_ms_p50": pct([r["setup_ms"] for r in res], 50),
        # setup phases of a request: boundary marches (GPU), PCM clustering (host threads), multipoles,
        # device upload
        "setup_phase_ms_p50": {k: pct([r[f"setup_{k}_ms"] for r in res], 50)
                               for k in ("boundary", "cluster", "moments", "upload")},
        "setup_phase_ms_max": {k: max(r[f"setup_{k}_ms"] for r in res)
                               for k in ("boundary", "cluster", "moments", "upload")},
        "queue_ms_p50": pct([r["t_admit_ms"] - r["t_ready_ms"] for r in res], 50),
        "trajectories_per_s": traj / makespan, "trajectories": traj, "makespan_s": makespan,
        "trajectories_issued": st["traj"],
        "rel_err_max": max(r["rel_err"] for r in res), "blocks_mean": float(np.mean([r["blocks"] for r in res])),
        "slices": st["slices"], "kernel_s": st["kernel_ms"] / 1e3, "busy_s": st["busy_ms"] / 1e3,
        # reference FLOPs of every trajectory issued (blocks past a stopping point included) over the
        # time some slice was in flight
        "achieved_tflops_busy": st["ref_flops"] / (st["busy_ms"] * 1e-3) / 1e12,
        "lanes_per_trajectory": "by size class", "class_lanes": server_kw.get("class_lanes"), "wg_size": W,
        "class_wg": server_kw["class_wg"],
        "prior_rel_sd": server_kw["prior_rel_sd"], "issue_fraction": server_kw["issue_fraction"],
        "policy": "shortest remaining work first" if policy == E.SERVE_SHORTEST_FIRST else "equal share",
        "slice_launches": "a launch of its own per request" if server_kw.get("own_launches")
                          else "shared, molecules staged in LDS",
        "ccs_vs_fp64": vs64,
    }


def batch_workload(E, synth, count, traj, steps, fp64_every=16):
    """256 synthetic conformers (N log-uniform 50-5,000), 50,000 trajectories
    each in one call: the default mode (one persistent launch per launch
    shape, each workgroup staging the molecule of the conformer it works on
    in LDS) and own launches (a launch per conformer); per-size-class launch
    shapes from the batch tuner for each mode, tuned before the timing."""
    import tempfile
    d = tempfile.mkdtemp(prefix="cosims_batch_")
    mols = []
    t0 = time.perf_counter()
    for k, (n, ion) in enumerate(synth.uneven_conformers(count=count)):
        path = os.path.join(d, f"c{k}.mfj")
        synth.write_mfj(path, f"c{k}", ion)
        p = E.default_params(temp=298.0, seed=7000 + k, multipole=0, dispersion_cutoff=0, setup_gpu=1,
                             traj=traj, iter=1)
        mols.append(E.Molecule.from_file(path, params=p))
    log(f"batch: {count} conformers set up in {time.perf_counter() - t0:.1f} s")
    jobs = [(m, dict(job=k, iters=1, traj=traj)) for k, m in enumerate(mols)]
    CL, CW = E.tune_batch(jobs, E.PREC_MIXED)
    log(f"batch: tuned lanes {CL} wg {CW}")
    n_atoms = [m.info()["n"] for m in mols]
    out = {}
    # one persistent launch per launch shape for all conformers (molecules through L2), its class
    # shapes tuned for that mode in an untimed call
    _, st_own = E.run_batch(jobs, E.PREC_MIXED, lanes=-1, own_launches=True)
    OL, OW = st_own["class_lanes"], st_own["class_wg"]
    log(f"batch: own-launch tuned lanes {OL} wg {OW}")
    for label, kw in (("tuned", dict(lanes=-1, class_lanes=CL, class_wg=CW)), ("size_rule", dict(lanes=0)),
                      ("own_launches_tuned", dict(lanes=-1, class_lanes=OL, class_wg=OW, own_launches=True))):
        times = []
        for k in range(steps):
            E.device_sync()
            t1 = time.perf_counter()
            res, st = E.run_batch(jobs, prec=E.PREC_MIXED, **kw)
            if label == "size_rule":
                res_rule = res
            E.device_sync()
            times.append(time.perf_counter() - t1)
            log(f"batch {label} {k}: {times[-1]:.2f} s, kernel {st['kernel_ms']:.0f} ms, {st['launches']} launches")
        flops = st["pair_evals"] * FLOP_PAIR + st["force_evals"] * FLOP_EPI + st["total_steps"] * FLOP_VERLET
        out[label] = {"trajectories_per_s": count * traj / float(np.median(times)), "s_per_step": float(np.median(times)),
                      "kernel_s": st["kernel_ms"] / 1e3, "launches": st["launches"],
                      "achieved_tflops": flops / (st["kernel_ms"] * 1e-3) / 1e12,
                      "class_lanes": st["class_lanes"], "class_wg": st["class_wg"]}
    out["workload"] = (f"{count} synthetic conformers, N log-uniform 50-5,000 (total {sum(n_atoms)} atoms), "
                       f"q round(N/250)+1 e, He 298 K, {traj} trajectories each, mixed precision")
    out["modes"] = ("tuned / size_rule (default mode): all conformers of a launch shape in one persistent launch, "
                    "each workgroup staging the molecule of the conformer it works on in LDS; "
                    "own_launches_tuned: a launch of its own per conformer")
    out["trajectories"] = count * traj
    # CCS error against FP64 on paired initial conditions (same jobs and streams), on every
    # fp64_every-th conformer, against the size-rule mixed results
    sub = list(range(0, count, fp64_every))
    f64, _ = E.run_batch([jobs[k] for k in sub], prec=E.PREC_FP64, lanes=0)
    dev = np.array([(res_rule[k]["wmean"] - f["wmean"]) / f["wmean"] for k, f in zip(sub, f64)])
    rse = np.array([f["werr"] / f["wmean"] for f in f64])
    out["ccs_vs_fp64"] = {"conformers": len(sub), "mean_rel_dev": float(dev.mean()),
                          "mean_rel_dev_se": float(dev.std(ddof=1) / np.sqrt(dev.size)) if dev.size > 1 else None,
                          "max_abs_rel_dev": float(np.abs(dev).max()), "median_rel_se_per_job": float(np.median(rse)),
                          "sample": f"every {fp64_every}th conformer, size-rule lanes, FP64 group kernels"}
    log(f"batch vs FP64: {len(sub)} conformers, mean rel dev {dev.mean():.2e}, max |dev| {np.abs(dev).max():.2e} "
        f"(per-job rel SE {np.median(rse):.2e})")
    for m in mols:
        m.close()
    return out


def clustered_workload(E, synth, traj, oracle_sample):
    """The 20,000-atom ion (+25 e): clustered (force type 1, PCM clusters +
    multipoles) against exact (type 4), mixed precision, and the exact FP64
    GPU run on the same initial conditions; an FP64 oracle subsample pins the
    FP64 run per trajectory."""
    import tempfile
    ion = synth.config_ion("ion20000")
    d = tempfile.mkdtemp(prefix="cosims_c4_")
    path = os.path.join(d, "ion20000.mfj")
    synth.write_mfj(path, "ion20000", ion)
    out = {"workload": f"20,000-atom synthetic ion, +25 e, He 298 K, {traj} trajectories, mixed precision"}
    runs = {}
    for label, mp in (("clustered", 1), ("exact", 0)):
        p = E.default_params(temp=298.0, seed=4242, multipole=mp, dispersion_cutoff=mp, setup_gpu=1,
                             traj=traj // 10, iter=10)
        t0 = time.perf_counter()
        m = E.Molecule.from_file(path, params=p)
        setup = time.perf_counter() - t0
        inf = m.info()
        log(f"c4 {label}: setup {setup:.2f} s; tuning the launch shape")
        m.run(E.PREC_MIXED, job=99)  # tunes for this run size, outside the timing
        E.device_sync()
        t1 = time.perf_counter()
        r = m.run(E.PREC_MIXED, job=0, keep=True)
        dt = time.perf_counter() - t1
        log(f"c4 {label}: type {inf['force_type']} ncl {inf['ncl']} setup {setup:.2f} s, run {dt:.2f} s, "
            f"CCS {r['wmean']:.2f} +- {r['werr']:.2f}, lanes {r['lanes']}")
        runs[label] = (m, r)
        out[label] = {"force_type": inf["force_type"], "clusters": inf["ncl"], "setup_s": setup,
                      "trajectories_per_s": traj / dt, "ccs": r["wmean"], "ccs_err": r["werr"],
                      "lanes": r["lanes"], "wg_size": r["wg_size"], "kernel_s": r["traj_kernel_ms"][0] / 1e3,
                      **launch_roofline(r, PEAK_FP32)}
    m4, r4 = runs["exact"]
    t1 = time.perf_counter()
    f = m4.run(E.PREC_FP64, job=0, keep=True, lanes=64, wg_size=512)
    dt = time.perf_counter() - t1
    log(f"c4 exact FP64: {dt:.2f} s, CCS {f['wmean']:.2f} +- {f['werr']:.2f}")
    out["exact_fp64"] = {"trajectories_per_s": traj / dt, "ccs": f["wmean"], "ccs_err": f["werr"],
                         "lanes": f["lanes"], **launch_roofline(f, PEAK_FP64)}
    for label in ("clustered", "exact"):
        r = runs[label][1]
        dd = r["omega"] - f["omega"]
        out[label]["rel_dev_vs_fp64"] = (r["wmean"] - f["wmean"]) / f["wmean"]
        out[label]["paired_rel_dev_vs_fp64"] = float(dd.mean() / f["omega"].mean())
        out[label]["paired_rel_se"] = float(dd.std(ddof=1) / np.sqrt(dd.size) / f["omega"].mean())
    # the clustered field in FP64 on the same initial conditions: the error of mixed precision on
    # its own force model (the deviation from the exact sum above is the model's)
    m1, r1 = runs["clustered"]
    t1 = time.perf_counter()
    f1 = m1.run(E.PREC_FP64, job=0, keep=True, lanes=64, wg_size=512)
    dt = time.perf_counter() - t1
    d1 = r1["omega"] - f1["omega"]
    log(f"c4 clustered FP64: {dt:.2f} s, CCS {f1['wmean']:.2f} +- {f1['werr']:.2f}")
    out["clustered_fp64"] = {"trajectories_per_s": traj / dt, "ccs": f1["wmean"], "ccs_err": f1["werr"],
                             "lanes": f1["lanes"], **launch_roofline(f1, PEAK_FP64)}
    out["clustered"]["paired_rel_dev_vs_clustered_fp64"] = float(d1.mean() / f1["omega"].mean())
    out["clustered"]["paired_rel_se_vs_clustered_fp64"] = float(d1.std(ddof=1) / np.sqrt(d1.size)
                                                                / f1["omega"].mean())
    if oracle_sample > 0:
        import oracle
        A = m4.atoms()
        inf = m4.info()
        mo = oracle.Mol(A["xyz"], A["q"], A["sig6"], A["eps24m"], A["eps4"], inf["charged"], inf["mu"],
                        inf["force_type"])
        s = m4.sample(m4.params.seed, 0, 0, 0, oracle_sample)
        h = s["hit"] == 1
        t1 = time.perf_counter()
        o = mo.verlet(m4.params.dt, s["p0"][h], s["v0"][h], (inf["a"], inf["b"], inf["c"]), nthreads=16)
        dt = time.perf_counter() - t1
        g = f["chi"][:oracle_sample][h]
        ok = (o["ok"] == 1) & (f["status"][:oracle_sample][h] == 1)
        dchi = np.abs(g[ok] - o["chi"][ok])
        out["exact_fp64"]["oracle_sample"] = {"trajectories": int(h.sum()), "median_abs_dchi": float(np.median(dchi)),
                                              "frac_dchi_le_1e-8": float(np.mean(dchi <= 1e-8)), "oracle_s": dt}
        log(f"c4 oracle sample: {int(h.sum())} trajectories, median |dchi| {np.median(dchi):.2e}")
    for m, _ in runs.values():
        m.close()
    return out


def launch_roofline(r, peak):
    """Reference FLOPs of a run's trajectory launch (the library's per-branch
    count, cosims_ccs.ref_flops) over its HIP-event time, against a peak."""
    if r["n_rounds"] != 1 or not r["ref_flops"]:
        return {}
    tf = r["ref_flops"] / (r["traj_kernel_ms"][0] * 1e-3) / 1e12
    cl = dict(zip(("clusters_examined", "far", "far_dipole", "atoms_charge_lj", "atoms_charge", "atoms_lj"),
                  r["cl_counts"])) if any(r["cl_counts"]) else None
    return {"achieved_tflops": tf, "peak_tflops": peak, "frac": tf / peak, "ref_flops": r["ref_flops"],
            "branch_counts": cl}


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


def serving_block(E, synth, a):
    """The serving streams of the bench line (warm-up, the configuration's
    stream, the sustained and the clustered streams): the keys they add."""
    out = {"served_latency_ms": None}
    cwg = [int(x) for x in a.serve_class_wg.split(",")]
    cl = [int(x) for x in a.serve_class_lanes.split(",")]
    own = bool(a.serve_own_launches)
    if a.serve_warmup > 0:
        wu = serving(E, synth, a.serve_warmup, 20.0, a.serve_target, seed=7000, class_wg=cwg, class_lanes=cl,
                     own_launches=own)
        log(f"serving warm-up: {a.serve_warmup} requests, {wu['trajectories_per_s']:.0f} traj/s (not reported)")
    sv = serving(E, synth, a.serve_requests, a.serve_rate, a.serve_target, fp64_max_atoms=a.serve_fp64_atoms,
                 class_wg=cwg, class_lanes=cl, own_launches=own)
    out["serving"] = sv
    # the config's stream (20 req/s, several times what one GPU serves: its percentiles measure
    # the backlog) and a stream at a rate the GPU sustains
    out["served_latency_ms"] = {"p50": sv["latency_ms"]["p50"], "p99": sv["latency_ms"]["p99"],
                                "offered_rate": a.serve_rate, "requests": a.serve_requests}
    if a.serve_sustained_rate > 0 and a.serve_sustained_requests > 0:
        su = serving(E, synth, a.serve_sustained_requests, a.serve_sustained_rate, a.serve_target, seed=6000,
                     class_wg=cwg, class_lanes=cl, own_launches=own)
        out["serving_sustained"] = su
        out["served_latency_ms"]["sustained"] = {"p50": su["latency_ms"]["p50"], "p99": su["latency_ms"]["p99"],
                                                 "offered_rate": a.serve_sustained_rate,
                                                 "requests": a.serve_sustained_requests}
    if a.serve_compare:
        eq = serving(E, synth, a.serve_requests, a.serve_rate, a.serve_target, policy=E.SERVE_EQUAL_SHARE,
                     class_wg=cwg, class_lanes=cl, own_launches=own)
        out["serving_equal_share"] = {k: eq[k] for k in ("latency_ms", "served_rate", "trajectories_per_s",
                                                         "makespan_s", "slices", "trajectories_issued")}
    if a.serve_clustered_requests > 0:
        # requests with the reference's default cut-offs: PCM clustering in each request's setup
        # and the clustered force type in its trajectories
        cu = serving(E, synth, a.serve_clustered_requests, a.serve_clustered_rate, a.serve_target, seed=7000,
                     fp64_max_atoms=a.serve_fp64_atoms, fp64_max_requests=8, clustered=True, class_wg=cwg,
                     class_lanes=cl, own_launches=own)
        out["serving_clustered"] = cu
        out["served_latency_ms"]["clustered"] = {"p50": cu["latency_ms"]["p50"], "p99": cu["latency_ms"]["p99"],
                                                 "offered_rate": a.serve_clustered_rate,
                                                 "requests": a.serve_clustered_requests}
    return out


def serving_child(a):
    """The serving block in a process of its own, as a server is deployed (a
    process that only serves: the streams ran 10 % slower after the headline
    legs in the same process); its JSON line is merged into this one."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__)] + sys.argv[1:] + ["--serve-child"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True)
    if r.returncode != 0:
        raise SystemExit(f"serving child failed with status {r.returncode}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--traj", type=int, default=10000, help="trajectories per CCS integral")
    ap.add_argument("--iters", type=int, default=100, help="CCS integrals per step")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--lanes", type=int, default=0,
                    help="lanes per trajectory of the mixed run (0: the autotuner's; fixes the shape for profiling)")
    ap.add_argument("--wg", type=int, default=0, help="workgroup size of the mixed run (0: the autotuner's)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--serve-requests", type=int, default=200, help="requests of the serving stream (0: none)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-sustained-rate", type=float, default=1.2,
                    help="offered rate of a second stream the GPU can sustain (0: none); one MI355X serves "
                         "1.6-2.1 req/s of this mix")
    ap.add_argument("--serve-sustained-requests", type=int, default=60)
    ap.add_argument("--serve-own-launches", type=int, default=0,
                    help="serving slices with a launch of their own per request (1) or shared LDS launches (0)")
    ap.add_argument("--serve-clustered-requests", type=int, default=16,
                    help="requests of a stream with the reference's default (clustered) force field (0: none)")
    ap.add_argument("--serve-clustered-rate", type=float, default=1.0)
    ap.add_argument("--serve-target", type=float, default=0.005)
    ap.add_argument("--serve-warmup", type=int, default=40,
                    help="requests of an unmeasured stream (20 req/s, its own seed) before the measured ones: a "
                         "process's first stream runs ~20 %% slower while the device cache fills (0: none)")
    ap.add_argument("--serve-fp64-atoms", type=int, default=1000,
                    help="rerun served requests of up to this many atoms in FP64 over the same blocks (0: none)")
    ap.add_argument("--serve-class-wg", type=str, default="0,0,0,1024,1024",
                    help="workgroup size per size class of the serving stream (0: the server's default)")
    ap.add_argument("--serve-class-lanes", type=str, default="8,16,32,64,64",
                    help="lanes per trajectory per size class of the serving stream (0: the size rule, "
                         "4 x 2^class)")
    ap.add_argument("--serve-compare", type=int, default=0,
                    help="also run the stream with equal-share slices (1) or not (0)")
    ap.add_argument("--workload", choices=("headline", "batch", "clustered"), default="headline",
                    help="headline: the 2,000-atom ion + serving stream (the JSON contract line); batch: 256 uneven "
                         "conformers in one call; clustered: the 20,000-atom ion, clustered vs exact")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--batch-traj", type=int, default=50000)
    ap.add_argument("--c4-traj", type=int, default=500000)
    ap.add_argument("--oracle-sample", type=int, default=300)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run 3 device-to-device copies of 256 MiB (scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--ceiling-probe", action="store_true",
                    help="only one mixed run of the headline workload at --lanes/--wg and the force-loop ceiling "
                         "probe at the same shape (for rocprofv3 counter passes over both kernels)")
    ap.add_argument("--serve-in-process", action="store_true",
                    help="run the serving streams in this process after the headline legs (default: in a child "
                         "process that only serves)")
    ap.add_argument("--serve-child", action="store_true", help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); trajectories shard "
              "trivially across GPUs but no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)
    if a.copy_probe:
        copy_probe()
        return
    if a.serve_child:  # the serving block alone, its keys as one JSON line (see serving_child)
        from cosims_amd import engine as E
        from cosims_amd import synth
        E.lib()
        print(json.dumps(serving_block(E, synth, a)), flush=True)
        return
    if a.ceiling_probe:
        from cosims_amd import engine as E
        from cosims_amd import synth
        m = make_molecule(E, synth, a.traj, a.iters, seed=20261203)
        r = m.run(E.PREC_MIXED, job=0, lanes=a.lanes, wg_size=a.wg)
        c = m.force_rate(E.PREC_MIXED, r["lanes"], r["wg_size"])
        log(f"trajectory launch {r['traj_kernel_ms'][0]:.1f} ms ({kernel_flops(m.info()['n'], r) / r['traj_kernel_ms'][0] / 1e9:.1f} "
            f"TFLOP/s), lanes {r['lanes']} wg {r['wg_size']}; force-loop ceiling {c['tflops']:.1f} TFLOP/s "
            f"({c['kernel_ms']:.1f} ms a launch)")
        return
    if a.workload != "headline":
        from cosims_amd import engine as E
        from cosims_amd import synth
        if a.workload == "batch":
            res = batch_workload(E, synth, a.batch_count, a.batch_traj, max(1, a.steps))
        else:
            res = clustered_workload(E, synth, a.c4_traj, a.oracle_sample)
        print(json.dumps({"workload": a.workload, **res}), flush=True)
        return

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

    shape = dict(lanes=a.lanes, wg_size=a.wg)
    for w in range(a.warmup):
        r = m.run(E.PREC_MIXED, job=0, **shape)
        log(f"warmup {w}: {r['run_ms']:.1f} ms")
    E.device_sync()
    times, runs = [], []
    t_all = time.perf_counter()
    for k in range(a.steps):
        t1 = time.perf_counter()
        r = m.run(E.PREC_MIXED, job=0, **shape)
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
        m.run(E.PREC_FP64, job=1, iters=1, traj=cus * 1024)  # tunes the FP64 launch shape outside the timing
        E.device_sync()
        t1 = time.perf_counter()
        fp64 = m.run(E.PREC_FP64, job=0)
        fp64_s = time.perf_counter() - t1
        log(f"fp64: {fp64_s * 1e3:.1f} ms, CCS {fp64['wmean']:.3f} +- {fp64['werr']:.3f}, "
            f"lanes {fp64['lanes']} wg {fp64['wg_size']}")

    fma = E.device_fma_rate()  # FP32 FMA rate of this build's code (VGPR operands), measured
    # roofline of the dominant kernel (first trajectory launch of each step)
    kms = np.array([r["traj_kernel_ms"][0] for r in runs])
    flops = np.array([kernel_flops(inf["n"], r) for r in runs])
    achieved = float(np.mean(flops / (kms * 1e-3)) / 1e12)
    # the measured ceiling of the kernel's instruction mix: its force call alone (molecule in LDS,
    # full occupancy, no Verlet bookkeeping, no drain) at the run's lanes, the best of the three
    # workgroup sizes (at the run's own workgroup size the probe varies by ~10 % from box to box)
    ceil_w = {w: m.force_rate(E.PREC_MIXED, mixed["lanes"], w)["tflops"] for w in (256, 512, 1024)
              if w >= mixed["lanes"]}
    ceil_wg = max(ceil_w, key=ceil_w.get)
    log(f"force-loop ceiling: {ceil_w[ceil_wg]:.1f} TFLOP/s (wg {ceil_wg}), by wg {ceil_w}; "
        f"FMA probe {fma:.1f} TFLOP/s")

    out = {
        "metric": METRIC, "value": value, "unit": "trajectories/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": elapsed / a.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 force sum / fp64 integrator",
        "data": "synthetic",
        "config": {"workload": "2,000-atom synthetic ion, +8 e, He 298 K, 100 x 10,000 trajectories",
                   "atoms": inf["n"], "trajectories_per_step": n_traj, "force_type": inf["force_type"],
                   "precision": "mixed"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32,
                     "ceiling_measured": ceil_w[ceil_wg], "frac_of_ceiling": achieved / ceil_w[ceil_wg],
                     "ceiling_by_wg": {str(w): t for w, t in ceil_w.items()},
                     "ceiling_source": "cosims_force_rate at the run's lanes, best workgroup size: the kernel's "
                                       "force call alone, molecule in LDS, full occupancy, no Verlet bookkeeping, "
                                       "no drain",
                     "fma_probe_tflops": fma,
                     "traffic": TRAFFIC_BYTES, "traffic_source": "profiles/traj_grp_mixed_ion2000.md",
                     "kernel": f"k_traj_grp<mixed,type4,{mixed['lanes']}>",
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
        out["fp64_launch"] = {"lanes_per_trajectory": fp64["lanes"], "wg_size": fp64["wg_size"],
                              "kernel": f"k_traj_grp<fp64,type4,{fp64['lanes']}>",
                              **launch_roofline(fp64, PEAK_FP64)}
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")), 16)
    if a.cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline(m, a.cpu_seconds, threads)
    if a.serve_requests > 0:
        out.update(serving_block(E, synth, a) if a.serve_in_process else serving_child(a))
    else:
        out["served_latency_ms"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
