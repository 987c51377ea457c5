"""HIP engine against the reference golden vectors and the oracle (MI355X).

Tolerances:
  FP64 path  forces / potentials: bit-exact to the reference (same IEEE
             operations in the same order); trajectories: median |dchi| <=
             1e-8 rad over non-failed trajectories (the only differences left
             are libm vs device acos/sin/cos rounding, amplified by chaos).
  mixed path forces: |da| <= 2e-4 * max|a| per probe set; CCS checked in
             test_gpu_run.py.
"""
import numpy as np
import pytest

import oracle
from cosims_amd import engine as E

from conftest import CASES, golden, golden_engine_mol

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", CASES)
def test_accel_fp64_bitexact(case):
    g = golden(case)
    for f in g["forces"]:
        m = golden_engine_mol(g, f["flags"])
        pts = np.array(f["points"]).reshape(-1, 3)
        acc = m.accel(pts, E.PREC_FP64)
        ref = np.array(f["acc"]).reshape(-1, 3)
        assert np.array_equal(acc, ref), (f["flags"], np.abs(acc - ref).max())
        assert np.array_equal(m.potential(pts), f["pot"]), f["flags"]


@pytest.mark.parametrize("case", CASES)
def test_accel_mixed_close(case):
    g = golden(case)
    for f in g["forces"]:
        m = golden_engine_mol(g, f["flags"])
        pts = np.array(f["points"]).reshape(-1, 3)
        acc = m.accel(pts, E.PREC_MIXED)
        ref = np.array(f["acc"]).reshape(-1, 3)
        scale = np.abs(ref).max(axis=1, keepdims=True)
        assert np.all(np.abs(acc - ref) <= 2e-4 * scale + 1e-12), (f["flags"], np.abs(acc - ref).max())


@pytest.mark.parametrize("case", CASES)
def test_verlet_fp64_matches_reference(case):
    g = golden(case)
    for v in g["verlet"]:
        m = golden_engine_mol(g, v["flags"])
        r = m.verlet(v["p0"], v["v0"], E.PREC_FP64)
        ok_ref = np.array(v["ok"])
        # identical success flags and step counts except where chaos has moved
        # a trajectory across a retry threshold
        assert np.mean(r["ok"] == ok_ref) >= 0.99
        assert np.mean(r["steps"] == np.array(v["steps"])) >= 0.98
        both = (r["ok"] == 1) & (ok_ref == 1)
        d = np.abs(r["chi"][both] - np.array(v["chi"])[both])
        assert np.median(d) <= 1e-8, np.median(d)
        assert np.mean(d <= 1e-8) >= 0.95


def test_sampler_matches_oracle():
    g = golden("ion128")
    m = golden_engine_mol(g, (1, 0, 0, 1))
    inf = m.info()
    n = 4096
    s = m.sample(20261203, 3, 1, 0, n)
    cfg = dict(seed=20261203, job=3, iteration=1, gmin=inf["gmin"], gmax=inf["gmax"], max_val=inf["max_val"],
               vd_a=inf["vd_a"], vd_norm=inf["vd_norm"], bmax=inf["bmax"], ea=inf["a"], eb=inf["b"],
               ec=inf["c"], projection=1)
    o = oracle.sample_ic(cfg, np.arange(n))
    assert np.array_equal(s["hit"], o["hit"])
    np.testing.assert_allclose(s["g"], o["g"], rtol=1e-14, atol=0)
    h = o["hit"] == 1
    np.testing.assert_allclose(s["p0"][h], o["p0"][h], rtol=0, atol=1e-11)
    np.testing.assert_allclose(s["v0"][h], o["v0"][h], rtol=0, atol=1e-12)


def test_device_sqrt_div_correctly_rounded():
    """Potential of one atom at many points: exercises IEEE div and sqrt in
    FP64 against the host (any mis-rounding shows as a bit difference)."""
    from cosims_amd import engine as E
    rng = np.random.default_rng(5)
    p = E.default_params(charge=1, multipole=0, dispersion_cutoff=0)
    m = E.Molecule.from_prepared(p, np.zeros((1, 3)), [0.7], [3.043 ** 6], [0.3], [0.5], 3.99, (30, 30, 30))
    pts = rng.uniform(-20, 20, (20000, 3))
    mo = oracle.Mol(np.zeros((1, 3)), [0.7], [3.043 ** 6], [0.3], [0.5], 1, 3.99, 4)
    assert np.array_equal(m.potential(pts), mo.potential(pts))
    assert np.array_equal(m.accel(pts, E.PREC_FP64), mo.accel(pts))


def test_device_fma_rate_is_plausible():
    """The measured non-packed FP32 FMA rate with VGPR operands: gfx950
    issues such an instruction in 2 cycles a wavefront, so the rate lies
    above the 4-cycle figure (78.6 TFLOP/s at 2.4 GHz, less a throttled
    clock) and at most the published 157.3 TFLOP/s."""
    t = E.device_fma_rate()
    assert 100.0 < t <= 157.3, t


def test_force_rate_probe():
    """The force-sum ceiling probe on the 2,000-atom ion: mixed precision
    between the 4-cycle FP32 figure's half and the published peak, FP64
    below the FP64 vector peak; molecules it cannot stage are refused."""
    from conftest import golden_engine_mol
    m = golden_engine_mol(golden("ion2000"), (1, 0, 0, 1))
    mix = m.force_rate(E.PREC_MIXED, 16, 512)
    assert 40.0 < mix["tflops"] <= 157.3, mix
    assert mix["pairs_per_s"] * 58 <= mix["tflops"] * 1e12 * 1.001
    f64 = m.force_rate(E.PREC_FP64, 32, 512)
    assert 2.0 < f64["tflops"] <= 78.6, f64
    with pytest.raises(E.CosimsError, match="fits LDS only"):
        golden_engine_mol(golden("ion2000"), (1, 1, 1, 1)).force_rate(E.PREC_MIXED, 16, 512)
