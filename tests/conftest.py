import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)

CASES = ["acau_charged", "acau_neutral", "acau_phos_sigma_chg", "acau_phos_sigma", "ion128", "ion2000"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def golden(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        return json.load(f)


def golden_clusters(g):
    """Reference PCM groups + the oracle's multipoles of them."""
    import oracle
    A = g["atoms"]
    sc = g["scalars"]
    xyz = np.array(A["xyz"])
    cl = oracle.pcm_clusters(xyz, A["q"], int(sc["max_cluster_size"]))
    return cl


def golden_oracle_mol(g, flags):
    """Oracle molecule built from the reference's prepared atoms."""
    import oracle
    A = g["atoms"]
    sc = g["scalars"]
    ch, mp, co, order = flags
    ft = oracle.lib().or_force_type(ch, mp, co)
    cl = golden_clusters(g)
    return oracle.Mol(np.array(A["xyz"]).reshape(-1, 3), A["q"], A["sig6"], A["eps24m"], A["eps4"], ch,
                      sc["mu"], ft, cl, sc["multipole_radius"], sc["dispersion_radius"], order)


def golden_engine_mol(g, flags):
    """Engine molecule injected with the reference's prepared atoms/clusters."""
    from cosims_amd import engine as E
    A = g["atoms"]
    sc = g["scalars"]
    ch, mp, co, order = flags
    p = E.default_params(temp=sc["temp"], dt=sc["dt"], charge=ch, multipole=mp, dispersion_cutoff=co,
                         multipole_order=order, max_cluster_size=int(sc["max_cluster_size"]))
    pcm = g["pcm"]
    clusters = dict(start=pcm["start"], atoms=pcm["atoms"], center=pcm["center"], maxdist=pcm["maxdist"])
    return E.Molecule.from_prepared(p, np.array(A["xyz"]), A["q"], A["sig6"], A["eps24m"], A["eps4"],
                                    sc["mu"], (sc["a"], sc["b"], sc["c"]), clusters,
                                    sc["dispersion_radius"], sc["multipole_radius"])


def paired_ccs_dev(chi_a, st_a, chi_b, st_b):
    """Relative CCS difference of run a against run b on the same initial
    conditions, and its standard error.  Every sampled trajectory (status
    != 0: accepted, or failed at chi = pi/2) adds the same constant times
    1 - cos(chi) to the CCS sum (src/monteTraj.cpp:471-477), so the paired
    difference is the sum of the per-trajectory differences of 1 - cos(chi)
    over the sum of run b's, and its standard error follows from their
    spread.  Mixed precision and FP64 agree to ~1e-9 rad on almost every
    trajectory; the difference is carried by the few chaotic ones (about one
    in a thousand) whose end states part, so it shrinks as 1/sqrt(n)."""
    wa = np.where(np.asarray(st_a) != 0, 1.0 - np.cos(chi_a), 0.0)
    wb = np.where(np.asarray(st_b) != 0, 1.0 - np.cos(chi_b), 0.0)
    d = wa - wb
    return d.sum() / wb.sum(), d.std(ddof=1) * np.sqrt(d.size) / wb.sum()


def golden_engine_setup(g, threads=0, setup_gpu=0):
    """Engine molecule set up from the case's raw input files, as the
    reference's main.cpp reads them."""
    from cosims_amd import engine as E
    f = E.read_molfile(os.path.join(GOLDEN, g["mol"]),
                       os.path.join(GOLDEN, g["atoms_dat"]) if g["atoms_dat"] else None)
    p = E.default_params()
    if g["input"]:
        import tempfile
        with tempfile.NamedTemporaryFile("w", suffix=".in", delete=False) as t:
            t.write(g["input"])
        E.read_params(t.name, p)
        os.unlink(t.name)
    p.charge = int(f["charged"])
    p.threads = threads
    p.setup_gpu = setup_gpu
    return E.Molecule.create(p, f["xyz"], f["mass"], f["sigma"], f["epsilon"],
                             f["charge"] if f["charged"] else None, f["sigma_max"])


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    return oracle.lib()


CLI = os.path.join(ROOT, "cosims_amd", "cosims")
LOGS = json.load(open(os.path.join(GOLDEN, "example_logs.json")))


def run_cli_example(entry, work, extra=(), timeout=600):
    """Runs the command-line program on one of the reference's example runs
    (its molecule, input file, atoms.dat and seed) in directory `work`.
    Returns (exit code, console text, path of the omega log it named)."""
    import shutil
    import subprocess
    shutil.copy(os.path.join(GOLDEN, entry["mol"]), work)
    args = [CLI, "-m", entry["mol"]]
    if entry["input"]:
        with open(os.path.join(work, "input.txt"), "w") as f:
            f.write(entry["input"])
        args += ["-i", "input.txt"]
    if entry["atoms_dat"]:
        shutil.copy(os.path.join(GOLDEN, entry["atoms_dat"]), work)
        args += ["-a", entry["atoms_dat"]]
    args += ["-s", str(entry["seed"]), *extra]
    r = subprocess.run(args, cwd=work, capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout + r.stderr, os.path.join(work, os.path.basename(entry["log"]))


def digest(a, dtype):
    """SHA-256 of an array's little-endian bytes: how the large ion's golden
    record (tests/golden/make_golden.py) stores prepared atoms and clusters."""
    import hashlib
    b = np.ascontiguousarray(np.asarray(a, dtype=dtype).reshape(-1)).astype(np.dtype(dtype).newbyteorder("<"))
    return hashlib.sha256(b.tobytes()).hexdigest()


_LARGE = {}


def large_ion_file():
    """The 20,000-atom synthetic ion (radius 40 A, +25 e, q ~ N(0, 0.4 e))
    written by cosims_amd.synth, checked against the SHA-256 its golden
    record holds before anything compares against it."""
    import hashlib
    import tempfile
    if "path" not in _LARGE:
        from cosims_amd import synth
        path = os.path.join(tempfile.mkdtemp(prefix="cosims_c4_"), "ion20000.mfj")
        synth.write_mfj(path, "ion20000", synth.config_ion("ion20000"))
        want = golden("ion20000")["mol_sha256"]
        assert hashlib.sha256(open(path, "rb").read()).hexdigest() == want, "synthetic input changed"
        _LARGE["path"] = path
    return _LARGE["path"]


def large_ion_params(g, multipole, cutoff, **kw):
    """Engine run parameters of the large ion's golden input, with the
    force-type switches set."""
    from cosims_amd import engine as E
    sc = g["scalars"]
    return E.default_params(temp=sc["temp"], dt=sc["dt"], multipole=multipole, dispersion_cutoff=cutoff,
                            multipole_radius=sc["multipole_radius"], dispersion_radius=sc["dispersion_radius"],
                            max_cluster_size=int(sc["max_cluster_size"]), multipole_order=int(sc["multipole_order"]),
                            seed=20261203, **kw)


def check_large_ion_setup(g, m, clusters=True):
    """Engine setup of the large ion against the reference's record."""
    sc = g["scalars"]
    inf = m.info()
    assert inf["n"] == 20000
    assert (inf["a"], inf["b"], inf["c"], inf["bmax"], inf["mu"]) == (sc["a"], sc["b"], sc["c"], sc["bmax"], sc["mu"])
    A = m.atoms()
    for k in ("xyz", "q", "sig6", "eps24m", "eps4"):
        assert digest(A[k], np.float64) == g["atoms_sha256"][k], k
    if clusters:
        assert inf["ncl"] == int(sc["ncl"])
        cl = m.clusters()
        assert digest(cl["start"], np.int32) == g["pcm_sha256"]["start"]
        assert digest(cl["atoms"], np.int32) == g["pcm_sha256"]["atoms"]
        assert digest(cl["center"], np.float64) == g["pcm_sha256"]["center"]
        assert digest(cl["maxdist"], np.float64) == g["pcm_sha256"]["maxdist"]


def no_crossing_molecule(n, seed=7):
    """A neutral molecule whose Lennard-Jones well (epsilon 1e-9 kcal/mol) is
    far shallower than 1 % of the slowest gas particle's kinetic energy: the
    potential never falls below the boundary threshold and the z-axis walk
    of impactParam::setBoundry (src/impactParam.cpp:640-655) never crosses."""
    xyz = np.random.default_rng(seed).normal(scale=6.0, size=(n, 3))
    return xyz, np.full(n, 12.0), np.full(n, 3.4), np.full(n, 1e-9)
