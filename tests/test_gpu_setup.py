"""Boundary search on the GPU (MI355X): the potential marches of
impactParam::setBoundry (src/impactParam.cpp:602-736) run one wavefront per
ray; the points, and so the fitted and enlarged ellipsoid, must be
bit-identical to the host loop and to the reference's golden values."""
import os
import tempfile
import time

import numpy as np
import pytest

from cosims_amd import engine as E
from cosims_amd import synth

from conftest import CASES, golden, golden_engine_setup, no_crossing_molecule

pytestmark = pytest.mark.gpu

KEYS = ("a", "b", "c", "bmax", "fit_rmsd", "n_boundary", "enlarge", "ccs_const")


@pytest.mark.parametrize("case", CASES)
def test_gpu_boundary_matches_reference(case):
    g = golden(case)
    sc = g["scalars"]
    gpu = golden_engine_setup(g, setup_gpu=1).info()
    for k in ("a", "b", "c", "bmax"):
        assert gpu[k] == sc[k], k
    host = golden_engine_setup(g, threads=8).info()
    for k in KEYS:
        assert gpu[k] == host[k], k
    assert list(gpu["fit_abc"]) == list(host["fit_abc"])


def test_gpu_boundary_large_ion():
    """20,000 atoms, +25 e: the ion whose host setup dominates a request."""
    ion = synth.config_ion("ion20000")
    path = os.path.join(tempfile.mkdtemp(prefix="cosims_t_"), "ion20000.mfj")
    synth.write_mfj(path, "ion20000", ion)
    out = {}
    for gpu in (1, 0):
        p = E.default_params(temp=298.0, multipole=1, dispersion_cutoff=1, setup_gpu=gpu)
        t = time.perf_counter()
        m = E.Molecule.from_file(path, params=p)
        out[gpu] = (m.info(), time.perf_counter() - t)
        print(f"setup_gpu={gpu}: {out[gpu][1]:.2f} s, n_boundary {out[gpu][0]['n_boundary']}", flush=True)
    for k in KEYS:
        assert out[1][0][k] == out[0][0][k], k
    assert out[1][0]["ncl"] == out[0][0]["ncl"] > 0


def test_gpu_boundary_walk_without_crossing_fails_within_its_bound():
    """The device z-axis walk has the host loop's bound: a molecule whose
    potential never crosses the threshold is a setup error, and the one
    wavefront that walks it stops after ~2 x 10^4 points."""
    p = E.default_params(temp=298.0, multipole=0, dispersion_cutoff=0, setup_gpu=1)
    t = time.perf_counter()
    with pytest.raises(E.CosimsError, match="no potential crossing on the z axis"):
        E.Molecule.create(p, *no_crossing_molecule(2000))
    assert time.perf_counter() - t < 5.0


def test_concurrent_gpu_setups_are_deterministic():
    """Molecule setups on 8 host threads at once (as the request server's
    setup workers run them), four rounds over 256 conformers of 50-5,000
    atoms in different orders, every molecule destroyed before the next
    round so the next setups draw recycled device blocks: every boundary
    search succeeds and gives the same fit, bit for bit, each round (the
    marches are deterministic; a read of another molecule's data would show
    as a failed search or a different fit)."""
    import concurrent.futures as cf
    import random
    confs = synth.uneven_conformers(count=256, seed=4100)
    d = tempfile.mkdtemp(prefix="cosims_cc_")
    paths = []
    for k, (n, ion) in enumerate(confs):
        paths.append(os.path.join(d, f"c{k}.mfj"))
        synth.write_mfj(paths[-1], f"c{k}", ion)

    def setup(k):
        p = E.default_params(temp=298.0, seed=11 + k, multipole=0, dispersion_cutoff=0, setup_gpu=1)
        m = E.Molecule.from_file(paths[k], params=p)
        inf = m.info()
        out = tuple(inf[key] for key in ("a", "b", "c", "bmax", "n_boundary"))
        m.close()
        return out

    ref = None
    for rnd in range(4):
        order = list(range(len(paths)))
        random.Random(rnd).shuffle(order)
        with cf.ThreadPoolExecutor(8) as ex:
            res = dict(zip(order, ex.map(setup, order)))
        got = [res[k] for k in range(len(paths))]
        if ref is None:
            ref = got
        else:
            bad = [k for k in range(len(paths)) if got[k] != ref[k]]
            assert not bad, (rnd, bad[:5])
