"""The oracle against the compiled reference's golden vectors (CPU).

Every comparison is exact: the oracle restates the reference's FP64
arithmetic in the same operand order, so setup, clusters, forces of all six
force types, potentials and whole trajectories reproduce bit for bit.
"""
import numpy as np
import pytest

import oracle
from oracle import molio

from conftest import CASES, GOLDEN, golden, golden_oracle_mol


def _input_flags(g):
    sc = g["scalars"]
    return dict(charged=bool(sc["charged"]), multipole=bool(sc["multipole"]),
                dispersion_cutoff=bool(sc["dispersion_cutoff"]), max_cluster_size=int(sc["max_cluster_size"]),
                dispersion_radius=sc["dispersion_radius"] if not sc["dispersion_cutoff"] else 0.0,
                multipole_radius=sc["multipole_radius"] if not sc["multipole"] else 0.0)


def _oracle_setup(g):
    tab = molio.default_table()
    if g["atoms_dat"]:
        molio.apply_atoms_file(tab, f"{GOLDEN}/{g['atoms_dat']}")
    path = f"{GOLDEN}/{g['mol']}"
    m = molio.read_pdb(path) if path.endswith(".pdb") else molio.read_mfj(path, tab)
    par = molio.resolve(tab, m["symbols"])
    fl = _input_flags(g)
    return oracle.prepare(m["xyz"], [p[0] for p in par], [p[1] for p in par], [p[2] for p in par],
                          m["charge"], fl["charged"], temp=g["scalars"]["temp"], multipole=fl["multipole"],
                          dispersion_cutoff=fl["dispersion_cutoff"], max_cluster_size=fl["max_cluster_size"],
                          dispersion_radius=fl["dispersion_radius"], multipole_radius=fl["multipole_radius"],
                          sigma_max=max(r[3] for r in tab))


@pytest.mark.parametrize("case", CASES)
def test_setup_matches_reference(case):
    g = golden(case)
    sc = g["scalars"]
    prep = _oracle_setup(g)
    M = prep["mol"]
    A = g["atoms"]
    xyz = np.stack([M.x, M.y, M.z], 1).reshape(-1)
    assert np.array_equal(xyz, A["xyz"])
    for k in ("sig6", "eps24m", "eps4", "q"):
        assert np.array_equal(getattr(M, k), A[k]), k
    assert prep["mu"] == sc["mu"]
    assert (prep["a"], prep["b"], prep["c"]) == (sc["a"], sc["b"], sc["c"])
    assert prep["bmax"] == sc["bmax"]
    vd = prep["vd"]
    for k in ("gmin", "gmax", "norm", "max_val", "min_kin", "mean", "mode", "sigma", "cdf_frac"):
        assert vd[k] == sc[k], k
    assert vd["ccs_scale"] == sc["ccs_scale"]
    assert prep["ccs_const"] == sc["ccs_scale"] * (sc["bmax"] * sc["bmax"])


@pytest.mark.parametrize("case", CASES)
def test_clusters_match_reference(case):
    g = golden(case)
    cl = oracle.pcm_clusters(np.array(g["atoms"]["xyz"]), g["atoms"]["q"], int(g["scalars"]["max_cluster_size"]))
    assert list(cl["start"]) == g["pcm"]["start"]
    assert list(cl["atoms"]) == g["pcm"]["atoms"]
    assert np.array_equal(cl["center"], g["pcm"]["center"])
    assert np.array_equal(cl["rad"], g["pcm"]["maxdist"])


@pytest.mark.parametrize("case", CASES)
def test_forces_match_reference(case):
    g = golden(case)
    for f in g["forces"]:
        M = golden_oracle_mol(g, f["flags"])
        pts = np.array(f["points"]).reshape(-1, 3)
        assert np.array_equal(M.accel(pts), np.array(f["acc"]).reshape(-1, 3)), f["flags"]
        assert np.array_equal(M.potential(pts), f["pot"]), f["flags"]


@pytest.mark.parametrize("case", CASES)
def test_trajectories_match_reference(case):
    g = golden(case)
    sc = g["scalars"]
    for v in g["verlet"]:
        M = golden_oracle_mol(g, v["flags"])
        r = M.verlet(sc["dt"], v["p0"], v["v0"], (sc["a"], sc["b"], sc["c"]))
        assert np.array_equal(r["chi"], v["chi"])
        assert np.array_equal(r["ok"], v["ok"])
        assert np.array_equal(r["steps"], v["steps"])
        assert np.array_equal(r["ergy"], v["ergy"])
