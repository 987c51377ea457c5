"""Engine host logic on the CPU: the library loads and exports the C ABI,
its readers and input-file parser, and its molecule setup (no GPU calls)."""
import os
import re

import numpy as np
import pytest

import oracle
from oracle import molio
from cosims_amd import engine as E

from conftest import CASES, GOLDEN, ROOT, golden, golden_engine_setup, no_crossing_molecule


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "cosims_engine.h")).read()
    declared = set(re.findall(r"\b(cosims_\w+)\s*\(", hdr))
    assert declared, "no declarations found"
    L = E.lib()
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert declared == set(E.EXPORTS)
    limits = dict(re.findall(r"#define COSIMS_NOALLOC_(MAX_\w+) (\d+)", hdr))
    assert (int(limits["MAX_TRAJ"]), int(limits["MAX_ITER"])) == (E.NOALLOC_MAX_TRAJ, E.NOALLOC_MAX_ITER)


def test_params_defaults_and_input_file(tmp_path):
    p = E.default_params()
    assert (p.temp, p.dt, p.traj, p.iter, p.multipole, p.dispersion_cutoff) == (298, 0.01, 50000, 10, 1, 1)
    assert (p.multipole_order, p.max_cluster_size, p.projection, p.print_rate) == (1, 4, 1, 20000)
    f = tmp_path / "in.txt"
    f.write_text("# comment line\nname myion\ntemp 350\ntraj 12.5\niter 4\nmultipole false\n"
                 "dispersion_cutoff 0\nmax_cluster_size 4.5\nseed 42\ncharge 3\nbogus 1\ntemp 300\n"
                 "incomplete\n")
    E.read_params(str(f), p)
    assert p.name == b"myion"
    assert p.temp == 300.0            # last value wins
    assert p.traj == 12500            # thousands
    assert p.iter == 4 and p.seed == 42
    assert p.multipole == 0 and p.dispersion_cutoff == 0
    assert p.max_cluster_size == 4    # atoi
    assert p.charge == 1 and p.total_charge == 3.0
    assert p.print_rate == 2500       # traj/5 when not given


@pytest.mark.parametrize("mol,atoms", [("acau.mfj", None), ("acau.pdb", None),
                                       ("acau_phos_sigma_chg.mfj", "atoms.dat"),
                                       ("acau_phos_sigma.pdb", "atoms.dat"), ("ion128.mfj", None)])
def test_molfile_reader_matches_oracle_reader(mol, atoms):
    f = E.read_molfile(os.path.join(GOLDEN, mol), os.path.join(GOLDEN, atoms) if atoms else None)
    tab = molio.default_table()
    if atoms:
        molio.apply_atoms_file(tab, os.path.join(GOLDEN, atoms))
    path = os.path.join(GOLDEN, mol)
    m = molio.read_pdb(path) if mol.endswith(".pdb") else molio.read_mfj(path, tab)
    par = molio.resolve(tab, m["symbols"])
    assert f["n"] == len(par)
    assert np.array_equal(f["xyz"], np.array(m["xyz"]))
    assert np.array_equal(f["mass"], [p[0] for p in par])
    assert np.array_equal(f["sigma"], [p[1] for p in par])
    assert np.array_equal(f["epsilon"], [p[2] for p in par])
    assert f["charged"] == m["charged"]
    if m["charged"]:
        assert np.array_equal(f["charge"], m["charge"])
        assert f["total_charge"] == m["total_charge"]
    assert f["sigma_max"] == max(r[3] for r in tab)


def _engine_setup(g, threads=0):
    return golden_engine_setup(g, threads=threads)


@pytest.mark.parametrize("case", CASES)
def test_engine_setup_matches_reference(case):
    g = golden(case)
    sc = g["scalars"]
    m = _engine_setup(g, threads=4)
    inf = m.info()
    A = m.atoms()
    assert np.array_equal(A["xyz"].reshape(-1), g["atoms"]["xyz"])
    for k in ("q", "sig6", "eps24m", "eps4"):
        assert np.array_equal(A[k], g["atoms"][k]), k
    for k in ("mu", "a", "b", "c", "bmax", "gmin", "gmax", "max_val", "dispersion_radius", "multipole_radius"):
        assert inf[k] == sc[k], k
    assert inf["vd_norm"] == sc["norm"]
    assert inf["ccs_const"] == sc["ccs_scale"] * (sc["bmax"] * sc["bmax"])
    assert inf["dipole_scalar"] == sc["dipole_scalar"]
    cl = m.clusters()
    if sc["multipole"] or sc["dispersion_cutoff"]:
        assert list(cl["start"]) == g["pcm"]["start"]
        assert list(cl["atoms"]) == g["pcm"]["atoms"]
        assert np.array_equal(cl["center"], g["pcm"]["center"])
        assert np.array_equal(cl["maxdist"], g["pcm"]["maxdist"])
        # multipoles and centre-of-charge replacement against the oracle
        o = oracle.pcm_clusters(A["xyz"], A["q"], int(sc["max_cluster_size"]))
        for k, ko in (("q", "q"), ("dip", "dip"), ("quad", "quad"), ("pos", "pos"), ("replace", "replace")):
            assert np.array_equal(cl[k], o[ko]), k
    else:
        assert cl is None


def test_setup_thread_count_does_not_change_result():
    g = golden("acau_charged")
    a = _engine_setup(g, threads=1).info()
    b = _engine_setup(g, threads=8).info()
    assert (a["a"], a["b"], a["c"], a["n_boundary"]) == (b["a"], b["b"], b["c"], b["n_boundary"])


def test_errors_are_reported():
    p = E.default_params()
    with pytest.raises(E.CosimsError, match="invalid argument"):
        E.Molecule.create(p, np.zeros((0, 3)), [], [], [])
    with pytest.raises(E.CosimsError):
        E.read_molfile(os.path.join(GOLDEN, "does_not_exist.mfj"))
    # cut-off force types need clusters when injected
    p = E.default_params(charge=1, multipole=1, dispersion_cutoff=1)
    with pytest.raises(E.CosimsError, match="clusters"):
        E.Molecule.from_prepared(p, np.zeros((1, 3)), [1.0], [1.0], [1.0], [1.0], 3.9, (10, 10, 10))


def test_boundary_walk_without_crossing_fails_within_its_bound():
    """The z-axis walk stops 10^4 A past the far side of the molecule; a walk
    that finds no crossing is a setup error after ~2 x 10^4 points, not after
    the 2^22 points of the rays' limit (2,000 atoms: ~20 s of potential
    evaluations on one core before the bound)."""
    import time
    p = E.default_params(temp=298.0, multipole=0, dispersion_cutoff=0, setup_gpu=0, threads=1)
    t = time.perf_counter()
    with pytest.raises(E.CosimsError, match="no potential crossing on the z axis"):
        E.Molecule.create(p, *no_crossing_molecule(2000))
    assert time.perf_counter() - t < 5.0


def test_unknown_atom_type_is_rejected(tmp_path):
    f = tmp_path / "x.mfj"
    f.write_text("x\n1\n2\nang\ncalc\n1.0\n0 0 0 12 0.5\n1.5 0 0 99 0.5\n")
    with pytest.raises(E.CosimsError, match="type not in the force field"):
        E.read_molfile(str(f))


def test_reference_stream_matches_setUpTraj():
    """The engine's reference-stream sampler against the reference's own
    monteTraj::setUpTraj driven as a single-thread runTraj (golden vectors):
    identical acceptance (draw counts), speeds and start points."""
    import math
    s = golden("acau_neutral_stream")
    g = golden("acau_neutral")
    m = _engine_setup(g)
    r = m.sample_reference(s["seed"], s["iters"], s["traj"])
    assert "".join(str(int(x)) for x in r["hit"]) == s["hit"]
    assert np.array_equal(r["g"], s["g"])
    first = np.array(s["first"])
    h = r["hit"] == 1
    # misses: the reference leaves the previous start point in place; only hits are compared
    hf = h[first]
    assert np.array_equal(r["p0"][first][hf], np.array(s["p0_first"]).reshape(-1, 3)[hf])
    assert np.array_equal(r["v0"][first][hf], np.array(s["v0_first"]).reshape(-1, 3)[hf])
    assert [math.fsum(r["p0"][h][:, k]) for k in range(3)] == s["hit_p0_fsum"]
    assert [math.fsum(r["v0"][h][:, k]) for k in range(3)] == s["hit_v0_fsum"]


def test_pcm_threaded_splits_match_serial_oracle(tmp_path):
    """PCM clustering of a 2,400-atom synthetic ion (radius 19 A), large
    enough that the engine splits the clusters of each pass on 8 host
    threads: membership, order, centres and extents equal the oracle's
    serial restatement of PCM_Cluster::clusterAtoms (src/PCM_Cluster.cpp:
    38-249) on the same prepared coordinates, and the phase is timed."""
    from cosims_amd import synth
    path = str(tmp_path / "ion2400.mfj")
    synth.write_mfj(path, "ion2400", synth.synthetic_ion(2400, 19.0, 0.3, 8.0, seed=2401))
    p = E.default_params(multipole=1, dispersion_cutoff=1, setup_gpu=0)
    p.threads = 8
    m = E.Molecule.from_file(path, params=p)
    inf = m.info()
    assert inf["ncl"] > 100 and inf["setup_cluster_ms"] > 0
    assert inf["setup_boundary_ms"] + inf["setup_cluster_ms"] <= inf["setup_ms"]
    cl, A = m.clusters(), m.atoms()
    o = oracle.pcm_clusters(A["xyz"], A["q"], p.max_cluster_size)
    assert np.array_equal(cl["start"], o["start"])
    assert np.array_equal(cl["atoms"], o["atoms"])
    assert np.array_equal(cl["center"], o["center"])
    assert np.array_equal(cl["maxdist"], o["rad"])
