"""The 20,000-atom ion against the compiled reference (CPU).

The golden record (tests/golden/ion20000.json) holds the reference's setup
of the synthetic ion as SHA-256 digests (prepared atoms, PCM clusters), its
forces and potentials at 48 probe points for the exact (type 4) and
clustered (type 1, multipole + dispersion cut-off) force fields, and 64
sampled trajectories of each.  The engine's host setup must reproduce the
digests (bit-identical atoms and clusters); the oracle, given those atoms,
must reproduce every force, potential and trajectory bit for bit.
"""
import numpy as np
import pytest

import oracle
from cosims_amd import engine as E

from conftest import check_large_ion_setup, golden, large_ion_file, large_ion_params


@pytest.fixture(scope="module")
def large():
    g = golden("ion20000")
    p = large_ion_params(g, 1, 1, setup_gpu=0, threads=0)
    m = E.Molecule.from_file(large_ion_file(), params=p)
    return g, m


def test_large_ion_engine_setup_matches_reference(large):
    g, m = large
    check_large_ion_setup(g, m, clusters=True)


def _oracle(g, m, flags):
    sc = g["scalars"]
    A = m.atoms()
    cl = oracle.pcm_clusters(A["xyz"], A["q"], int(sc["max_cluster_size"]))
    ft = oracle.lib().or_force_type(*flags[:3])
    return oracle.Mol(A["xyz"], A["q"], A["sig6"], A["eps24m"], A["eps4"], flags[0], sc["mu"], ft, cl,
                      sc["multipole_radius"], sc["dispersion_radius"], flags[3])


def test_large_ion_oracle_forces_match_reference(large):
    g, m = large
    assert [f["force_type"] for f in g["forces"]] == [4, 1]
    for f in g["forces"]:
        M = _oracle(g, m, f["flags"])
        pts = np.array(f["points"]).reshape(-1, 3)
        assert np.array_equal(M.accel(pts), np.array(f["acc"]).reshape(-1, 3)), f["flags"]
        assert np.array_equal(M.potential(pts), f["pot"]), f["flags"]


def test_large_ion_oracle_trajectories_match_reference(large):
    g, m = large
    sc = g["scalars"]
    for v in g["verlet"]:
        M = _oracle(g, m, v["flags"])
        r = M.verlet(sc["dt"], v["p0"], v["v0"], (sc["a"], sc["b"], sc["c"]), nthreads=8)
        assert np.array_equal(r["chi"], v["chi"]), v["flags"]
        assert np.array_equal(r["ok"], v["ok"])
        assert np.array_equal(r["steps"], v["steps"])
        assert np.array_equal(r["ergy"], v["ergy"])
