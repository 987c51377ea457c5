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

from conftest import CASES, golden, golden_engine_setup

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
