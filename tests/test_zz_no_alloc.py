"""cosims_run_noalloc against cosims_run on the GPU (MI355X).

cosims_run_noalloc runs cosims_run's group-kernel path with every device
buffer carved from a static device array instead of allocated.  Through the
C ABI: on two calls in a row, the second with a smaller run, its CCS per
integral, statistics, counters and per-trajectory outcomes equal cosims_run's
bit for bit (integer and index outputs exactly, floating point bit for bit:
both runs execute the same kernels on the same Philox initial conditions),
for the exact force type (molecule staged in LDS, mixed and FP64) and a
clustered type (table kernel); a run of exactly the scratch's capacity
matches too; runs beyond it, molecules not yet on the device and an untuned
shape are refused; repeated calls leave the device's free memory unchanged.
The workload is the 128-atom synthetic ion of tests/golden/ion128.mfj.
"""
import ctypes as C
import os

import numpy as np
import pytest

from cosims_amd import engine as E

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FIELDS = ("mean", "sd", "wmean", "werr", "max_ergy", "n_traj", "n_hit", "n_fail", "n_retry", "total_steps",
          "pair_evals", "pot_evals", "n_rounds", "round_steps", "round_evals", "round_items", "lanes", "wg_size",
          "cl_counts", "ref_flops")


def _mol(clustered):
    p = E.default_params(temp=298.0, multipole=int(clustered), dispersion_cutoff=int(clustered), seed=20261203)
    return E.Molecule.from_file(os.path.join(GOLD, "ion128.mfj"), params=p)


def _same(a, b, what):
    for k in FIELDS:
        assert a[k] == b[k], (what, k, a[k], b[k])
    assert np.array_equal(a["omega"], b["omega"]), what
    assert np.array_equal(a["omega_err"], b["omega_err"]), what
    assert np.array_equal(a["status"], b["status"]), what
    run = a["status"] != 0
    assert np.array_equal(a["chi"][run], b["chi"][run]), what
    assert np.array_equal(a["g"], b["g"]), what


def _free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("clustered,prec", [(False, E.PREC_MIXED), (False, E.PREC_FP64), (True, E.PREC_MIXED)],
                         ids=["exact-mixed", "exact-fp64", "clustered-mixed"])
def test_noalloc_equals_run_twice(clustered, prec):
    m = _mol(clustered)
    ft = m.info()["force_type"]
    assert (ft in (1, 2, 3, 5)) if clustered else ft == 4
    with pytest.raises(E.CosimsError, match="cosims_mol_upload"):
        m.run(prec, job=5, iters=1, traj=100, lanes=16, wg_size=256, noalloc=True)
    m.upload()
    runs = [dict(job=5, iters=4, traj=20000, lanes=16, wg_size=256),
            dict(job=6, iters=2, traj=3000, lanes=8, wg_size=512)]
    ref = [m.run(prec, keep=True, **kw) for kw in runs]
    got = [m.run(prec, keep=True, noalloc=True, **kw) for kw in runs]  # two calls in a row, the second smaller
    for k, (a, b) in enumerate(zip(got, ref)):
        print(f"force type {ft} prec {prec} run {k}: CCS {a['wmean']:.4f} +- {a['werr']:.4f}, "
              f"{a['n_hit']} hits, {a['total_steps']} steps")
        _same(a, b, (ft, prec, k))
        assert a["n_hit"] > 0 and a["traj_kernel_ms"] == [0.0]


def test_noalloc_full_scratch_and_limits():
    m = _mol(False)
    m.upload()
    kw = dict(job=9, iters=1, traj=E.NOALLOC_MAX_TRAJ, lanes=16, wg_size=256)
    _same(m.run(E.PREC_MIXED, keep=True, noalloc=True, **kw), m.run(E.PREC_MIXED, keep=True, **kw), "full")
    for it, T in [(1, E.NOALLOC_MAX_TRAJ + 1), (2, E.NOALLOC_MAX_TRAJ // 2 + 1), (E.NOALLOC_MAX_ITER + 1, 100)]:
        with pytest.raises(E.CosimsError, match="static scratch"):
            m.run(E.PREC_MIXED, job=9, iters=it, traj=T, lanes=16, wg_size=256, noalloc=True)
    with pytest.raises(E.CosimsError, match="lanes"):
        m.run(E.PREC_MIXED, job=9, iters=1, traj=100, lanes=1, wg_size=256, noalloc=True)
    with pytest.raises(E.CosimsError, match="Philox"):
        m.run(E.PREC_MIXED, job=9, iters=1, traj=100, lanes=16, wg_size=256, noalloc=True, rng=E.RNG_REFERENCE)


def test_noalloc_tuned_shape_and_free_memory():
    m = _mol(False)
    m.upload()
    kw = dict(job=11, iters=3, traj=20000)
    with pytest.raises(E.CosimsError, match="tuned"):
        m.run(E.PREC_MIXED, noalloc=True, **kw)
    ref = m.run(E.PREC_MIXED, keep=True, **kw)  # tunes the shape for runs of this size
    first = m.run(E.PREC_MIXED, keep=True, noalloc=True, **kw)
    _same(first, ref, "tuned")
    before = _free_bytes()
    for k in range(3):  # same run size (the tuned shape is kept per size bucket), other streams
        m.run(E.PREC_MIXED, noalloc=True, **dict(kw, job=20 + k))
    assert _free_bytes() == before
