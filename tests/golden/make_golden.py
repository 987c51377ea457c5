"""Write the golden vectors in tests/golden/*.json from the compiled reference.

Runs only where the reference sources exist (it needs oracle/_ref, built by
``make -C oracle ref`` from /root/reference/src).  Inputs are the example
molecule files copied next to this script and seeded synthetic ions; outputs
are what the reference's own classes return for them:

  * setup: prepared atoms (rotated coordinates, sigma^6, 24 eps/mu, 4 eps),
    reduced mass, ellipsoid axes, velocity window, CCS constant, PCM clusters;
  * forceAccel::getReducedVec / getPotential at seeded probe points for each
    force type the case exercises;
  * diffEqSolver::verletVel on seeded initial conditions (chi, success,
    steps, energy deviation);
  * the example runs' logs (examples/*/omega_*.txt): settings, the CCS
    integrals with their standard errors, the final statistics, and the
    lines logfilter.log_lines keeps.

The 20,000-atom ion (ion20000) is stored compactly: its molecule file is
regenerated by cosims_amd.synth (its SHA-256 is recorded so a test can
check the input first), and the prepared atoms and PCM clusters are
recorded as SHA-256 digests of their little-endian float64 / int32 bytes
plus a few sums, instead of every value.

Floats are written with repr() (round-trip exact).
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
import logfilter  # noqa: E402  (tests/golden/logfilter.py)
from cosims_amd import synth  # noqa: E402

REFLIB = os.path.join(ROOT, "oracle", "_ref", "libcosims_ref.so")

CASES = [
    # name, molecule file, input-file lines, atoms.dat, extra force flags, n_verlet
    dict(name="acau_charged", mol="acau.mfj", inp=None, atoms=None,
         flags=[(1, 1, 1, 1), (1, 1, 0, 1), (1, 0, 1, 1), (1, 0, 0, 1), (1, 1, 1, 2)],
         nverlet=400),
    dict(name="acau_neutral", mol="acau.pdb", inp="name\tacau_neutral\n", atoms=None,
         flags=[(0, 1, 1, 1), (0, 0, 0, 1)], nverlet=400),
    dict(name="acau_phos_sigma_chg", mol="acau_phos_sigma_chg.mfj", inp="name\tacau_phos_sigma_chg\n",
         atoms="atoms.dat", flags=[(1, 1, 1, 1), (1, 0, 0, 1)], nverlet=200),
    dict(name="acau_phos_sigma", mol="acau_phos_sigma.pdb", inp="name\tacau_phos_sigma\n",
         atoms="atoms.dat", flags=[(0, 1, 1, 1)], nverlet=200),
    dict(name="ion128", mol="ion128.mfj",
         inp="temp 298\ntraj 20\niter 3\nmultipole false\ndispersion_cutoff false\n"
             "multipole_radius 20\ndispersion_radius 23.43015325287593\n", atoms=None,
         flags=[(1, 0, 0, 1), (1, 1, 1, 1), (0, 0, 0, 1)], nverlet=1000),
    dict(name="ion2000", mol="ion2000.mfj",
         inp="temp 298\ntraj 10\niter 100\nmultipole false\ndispersion_cutoff false\n"
             "multipole_radius 20\ndispersion_radius 23.43015325287593\n", atoms=None,
         flags=[(1, 0, 0, 1), (1, 1, 1, 1)], nverlet=300),
    # the large ion: exact (type 4) and clustered (type 1, the reference's
    # defaults for a charged molecule) arms, 64 sampled lines each
    dict(name="ion20000", mol="ion20000.mfj", synthetic=True, compact=True,
         inp="temp 298\ntraj 500\niter 1\nmultipole true\ndispersion_cutoff true\n"
             "multipole_radius 20\ndispersion_radius 23.43015325287593\n", atoms=None,
         flags=[(1, 0, 0, 1), (1, 1, 1, 1)], nverlet=[64, 64]),
]

SEED = 20261203


def load_ref():
    L = C.CDLL(REFLIB)
    d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    L.ref_setup.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_char_p]
    L.ref_setup.restype = C.c_int
    L.ref_scalars.argtypes = [d]
    L.ref_atoms.argtypes = [d] * 6
    L.ref_clusters.argtypes = [i, i, d, d]
    L.ref_clusters.restype = C.c_int
    L.ref_set_flags.argtypes = [C.c_int] * 4
    L.ref_accel.argtypes = [C.c_int, d, d]
    L.ref_potential.argtypes = [C.c_int, d, d]
    L.ref_verlet.argtypes = [C.c_int, d, d, C.c_double, C.c_double, C.c_double, d, i, i, d]
    L.ref_sample.argtypes = [C.c_longlong, C.c_int, C.c_int, d, d, d, i, i]
    return L


SCALAR_KEYS = ["mu", "dipole_scalar", "a", "b", "c", "bmax", "gmin", "gmax", "mean", "mode", "sigma",
               "norm", "max_val", "cdf_frac", "ccs_scale", "min_kin", "dispersion_radius",
               "multipole_radius", "total_mass", "charged", "multipole", "dispersion_cutoff",
               "multipole_order", "max_cluster_size", "projection", "dt", "temp", "max_x", "max_y",
               "max_z", "total_charge", "ncl"]


def probe_points(xyz, bmax, n, rng):
    pts = []
    while len(pts) < n:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        p = d * rng.uniform(1.0, 1.3 * bmax)
        if np.min(np.linalg.norm(xyz - p, axis=1)) > 1.5:
            pts.append(p)
    return np.array(pts)


def ff(a):
    return [float(x) for x in np.asarray(a, dtype=np.float64).reshape(-1)]


def digest(a, dtype):
    """SHA-256 of an array's little-endian bytes (the compact record)."""
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(np.asarray(a, dtype=dtype).reshape(-1)).astype(
        np.dtype(dtype).newbyteorder("<")).tobytes()).hexdigest()


def file_digest(path):
    import hashlib
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def run_case(L, case, work):
    molpath = os.path.join(work, case["mol"])
    argv = ["cosims", "-m", molpath, "-s", "123456789", "-n", "8"]
    if case["inp"] is not None:
        ip = os.path.join(work, case["name"] + ".in")
        open(ip, "w").write(case["inp"])
        argv += ["-i", ip]
    if case["atoms"]:
        argv += ["-a", os.path.join(work, case["atoms"])]
    arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
    n = L.ref_setup(len(argv), arr, work.encode())
    assert n > 0, f"reference setup failed for {case['name']}"
    s = np.empty(32)
    L.ref_scalars(s)
    sc = dict(zip(SCALAR_KEYS, [float(x) for x in s]))
    xyz = np.empty(3 * n); q = np.empty(n); sig6 = np.empty(n); e24 = np.empty(n); e4 = np.empty(n)
    mass = np.empty(n)
    L.ref_atoms(xyz, q, sig6, e24, e4, mass)
    start = np.empty(n + 2, np.int32); atoms = np.empty(n, np.int32)
    cen = np.empty(3 * n + 3); md = np.empty(n + 1)
    ncl = L.ref_clusters(start, atoms, cen, md)
    out = dict(case=case["name"], mol=case["mol"], input=case["inp"], atoms_dat=case["atoms"],
               scalars=sc)
    if case.get("compact"):
        out["mol_sha256"] = file_digest(molpath)
        out["atoms_sha256"] = dict(xyz=digest(xyz, np.float64), q=digest(q, np.float64),
                                   sig6=digest(sig6, np.float64), eps24m=digest(e24, np.float64),
                                   eps4=digest(e4, np.float64))
        out["atoms_fsum"] = dict(xyz=[math.fsum(xyz[k::3]) for k in range(3)], q=math.fsum(q),
                                 sig6=math.fsum(sig6), eps24m=math.fsum(e24), eps4=math.fsum(e4))
        out["pcm_sha256"] = dict(start=digest(start[: ncl + 1], np.int32), atoms=digest(atoms, np.int32),
                                 center=digest(cen[: 3 * ncl], np.float64), maxdist=digest(md[:ncl], np.float64))
    else:
        out["atoms"] = dict(xyz=ff(xyz), q=ff(q), sig6=ff(sig6), eps24m=ff(e24), eps4=ff(e4), mass=ff(mass))
        out["pcm"] = dict(start=[int(x) for x in start[: ncl + 1]], atoms=[int(x) for x in atoms],
                          center=ff(cen[: 3 * ncl]), maxdist=ff(md[:ncl]))
    rng = np.random.default_rng(SEED + n)
    X = xyz.reshape(-1, 3)
    pts = probe_points(X, sc["bmax"], 48, rng)
    vd = oracle.veldist(sc["mu"], sc["temp"])
    cfg = dict(seed=SEED, job=7, iteration=0, gmin=sc["gmin"], gmax=sc["gmax"], max_val=sc["max_val"],
               vd_a=vd["a"], vd_norm=sc["norm"], bmax=sc["bmax"], ea=sc["a"], eb=sc["b"], ec=sc["c"],
               projection=1)
    forces = []
    for fl in case["flags"]:
        L.ref_set_flags(*fl)
        acc = np.empty(3 * len(pts)); pot = np.empty(len(pts))
        P = np.ascontiguousarray(pts.reshape(-1))
        L.ref_accel(len(pts), P, acc)
        L.ref_potential(len(pts), P, pot)
        ft = oracle.lib().or_force_type(fl[0], fl[1], fl[2])
        entry = dict(flags=list(fl), force_type=int(ft), points=ff(pts), acc=ff(acc), pot=ff(pot))
        forces.append(entry)
    verlets = []
    for k, fl in enumerate(case["flags"][:2]):
        if isinstance(case["nverlet"], list):
            nv = case["nverlet"][k]
        else:
            nv = case["nverlet"] if k == 0 else max(100, case["nverlet"] // 4)
        L.ref_set_flags(*fl)
        ic = oracle.sample_ic(cfg, np.arange(nv) + 1000 * k)
        hit = ic["hit"].astype(bool)
        p0 = np.ascontiguousarray(ic["p0"][hit].reshape(-1)); v0 = np.ascontiguousarray(ic["v0"][hit].reshape(-1))
        m = int(hit.sum())
        chi = np.empty(m); ok = np.empty(m, np.int32); st = np.empty(m, np.int32); eg = np.empty(m)
        L.ref_verlet(m, p0, v0, sc["a"], sc["b"], sc["c"], chi, ok, st, eg)
        verlets.append(dict(flags=list(fl), force_type=int(oracle.lib().or_force_type(*fl[:3])),
                            p0=ff(p0), v0=ff(v0), chi=ff(chi), ok=[int(x) for x in ok],
                            steps=[int(x) for x in st], ergy=ff(eg)))
    out["forces"] = forces
    out["verlet"] = verlets
    return out


def stream_case(L, work):
    """The reference's own initial-condition stream (setUpTraj driven as a
    single-thread runTraj) for the neutral example, 3 integrals x 1500."""
    argv = ["cosims", "-m", os.path.join(work, "acau.pdb"), "-s", "123456789", "-n", "1",
            "-i", os.path.join(work, "neutral.in")]
    open(os.path.join(work, "neutral.in"), "w").write("name\tacau_neutral\n")
    arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
    n = L.ref_setup(len(argv), arr, work.encode())
    assert n > 0
    iters, traj = 3, 1500
    m = iters * traj
    p0 = np.empty(3 * m); v0 = np.empty(3 * m); g = np.empty(m)
    hit = np.empty(m, np.int32); tries = np.empty(m, np.int32)
    L.ref_sample(123456789, iters, traj, p0, v0, g, hit, tries)
    P = p0.reshape(-1, 3); V = v0.reshape(-1, 3)
    h = hit == 1
    first = [k * traj + i for k in range(iters) for i in range(200)]
    return dict(case="acau_neutral_stream", mol="acau.pdb", input="name\tacau_neutral\n", seed=123456789,
                iters=iters, traj=traj, hit="".join(str(int(x)) for x in hit), tries=[int(x) for x in tries],
                g=ff(g), first=first, p0_first=ff(P[first]), v0_first=ff(V[first]),
                hit_p0_fsum=[math.fsum(P[h][:, k]) for k in range(3)],
                hit_v0_fsum=[math.fsum(V[h][:, k]) for k in range(3)])


LOGS = [
    # example directory, log, molecule file, input-file text, atoms.dat
    ("acau_neutral", "omega_acau_neutral_0.txt", "acau.pdb", "name\tacau_neutral\n", None),
    ("acau_charged", "omega_acau.txt", "acau.mfj", None, None),
    ("change_atom_type", "omega_acau_phos_sigma_0.txt", "acau_phos_sigma.pdb", "name\tacau_phos_sigma\n",
     "atoms.dat"),
    ("change_atom_type", "omega_acau_phos_sigma_chg_0.txt", "acau_phos_sigma_chg.mfj",
     "name\tacau_phos_sigma_chg\n", "atoms.dat"),
]


def example_logs():
    """Recorded results of the reference's example runs (text logs)."""
    import re
    out = []
    for d, log, mol, inp, atoms in LOGS:
        text = open(f"/root/reference/examples/{d}/{log}").read()
        seed = int(re.search(r"Mersenne-Twister:\s+(\d+)", text).group(1))
        iters = int(re.search(r"Number of CCS integrals:\s+(\d+)", text).group(1))
        traj = int(re.search(r"trajectories per CCS integral:\s+(\d+)", text).group(1))
        om = re.findall(r"Omega:\s+([\d.]+) \+/- ([\d.]+)", text)
        fin = {k: re.search(pat + r"\s+([\d.]+)", text).group(1) for k, pat in
               (("mean", "Mean Cross-section:"), ("sd", "Standard Dev:"), ("wmean", "Weighted CCS Mean:"),
                ("werr", "Weighted Error:"), ("max_ergy", "Max Energy Deviation:"))}
        out.append(dict(log=f"examples/{d}/{log}", mol=mol, input=inp, atoms_dat=atoms, seed=seed, iters=iters,
                        traj=traj, omega=[o[0] for o in om], omega_err=[o[1] for o in om], final=fin,
                        lines=logfilter.log_lines(text)))
    return out


def main(names=None):
    if not os.path.exists(REFLIB):
        raise SystemExit("oracle/_ref/libcosims_ref.so missing: run `make -C oracle ref` first")
    for key in ("ion128", "ion2000"):
        path = os.path.join(HERE, key + ".mfj")
        if not os.path.exists(path):
            synth.write_mfj(path, key, synth.config_ion(key))
    if not names or "logs" in names:
        with open(os.path.join(HERE, "example_logs.json"), "w") as f:
            json.dump(example_logs(), f, indent=1)
        print("example_logs written", flush=True)
    jobs = [c for c in CASES if not names or c["name"] in names]
    if not names or "stream" in names:
        jobs.append(dict(name="acau_neutral_stream", stream=True))
    for case in jobs:
        # one process per case: the reference keeps its state in globals
        pid = os.fork()
        if pid == 0:
            work = tempfile.mkdtemp(prefix="cosims_golden_")
            for f in os.listdir(HERE):
                if f.endswith((".mfj", ".pdb", ".dat")):
                    os.symlink(os.path.join(HERE, f), os.path.join(work, f))
            if case.get("synthetic"):  # written here, not kept next to this script
                key = case["mol"][:-4]
                synth.write_mfj(os.path.join(work, case["mol"]), key, synth.config_ion(key))
            L = load_ref()
            devnull = os.open(os.devnull, os.O_WRONLY)
            saved = os.dup(1)
            os.dup2(devnull, 1)
            try:
                res = stream_case(L, work) if case.get("stream") else run_case(L, case, work)
            finally:
                os.dup2(saved, 1)
            with open(os.path.join(HERE, case["name"] + ".json"), "w") as f:
                json.dump(res, f, separators=(",", ":"))
            print(case["name"], "written", flush=True)
            os._exit(0)
        _, st = os.waitpid(pid, 0)
        if st != 0:
            raise SystemExit(f"case {case['name']} failed ({st})")


if __name__ == "__main__":
    main(sys.argv[1:] or None)
