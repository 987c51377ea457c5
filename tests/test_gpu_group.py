"""Group kernels beyond the LDS-staged exact types (clustered types through
the table kernel, molecules beyond LDS through the tiled kernel), and
batched runs (MI355X).

The lane-per-trajectory FP64 kernel (lanes=1) sums in the reference's atom /
cluster order and is pinned bit-for-bit to the reference in
test_gpu_parity.py; here the group kernels are checked against it on the
same Philox initial conditions.

Tolerances: FP64 group vs FP64 lane kernel, per trajectory, median |dchi|
<= 1e-8 rad on trajectories both accept, success flags equal on >= 99 %,
per-integral CCS within 5e-4; mixed precision vs FP64 on paired initial
conditions: median |dchi| <= 1e-6 rad, >= 99 % of trajectories within 1e-3
rad, and the paired CCS difference within 1e-3 (its standard error is
printed beside it).  Batched runs: every job's per-integral CCS is
bit-identical to the same job run alone (same launch shape), and to
cosims_run of the same molecule.
"""
import os

import numpy as np
import pytest

from cosims_amd import engine as E
from cosims_amd import synth

from conftest import golden, golden_engine_mol, paired_ccs_dev

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")

# (case, flags) -> force type: [1,1,1] 1, [1,1,0] 2, [1,0,1] 3, [0,1,1] 5
CLUSTERED = [("acau_charged", [1, 1, 1, 1]), ("acau_charged", [1, 1, 0, 1]), ("acau_charged", [1, 0, 1, 1]),
             ("acau_charged", [1, 1, 1, 2]), ("acau_neutral", [0, 1, 1, 1]), ("ion2000", [1, 1, 1, 1])]


def _paired(a, b):
    both = (a["status"] == 1) & (b["status"] == 1)
    return both, np.abs(a["chi"][both] - b["chi"][both])


@pytest.mark.parametrize("case,flags", CLUSTERED, ids=[f"{c}-{''.join(map(str, f))}" for c, f in CLUSTERED])
def test_clustered_group_fp64_matches_lane_kernel(case, flags):
    m = golden_engine_mol(golden(case), flags)
    ft = m.info()["force_type"]
    assert ft in (1, 2, 3, 5)
    kw = dict(job=3, iters=2, traj=2000, keep=True)
    lane = m.run(E.PREC_FP64, lanes=1, **kw)
    for L in (4, 16, 64):  # 64: one group per wavefront, clusters sorted into LDS lists
        grp = m.run(E.PREC_FP64, lanes=L, wg_size=256, **kw)
        assert grp["lanes"] == L
        assert np.array_equal(grp["status"] != 0, lane["status"] != 0)
        assert np.mean(grp["status"] == lane["status"]) >= 0.99
        both, d = _paired(grp, lane)
        assert np.median(d) <= 1e-8, (ft, L, np.median(d))
        np.testing.assert_allclose(grp["omega"], lane["omega"], rtol=5e-4)


@pytest.mark.parametrize("case,flags", CLUSTERED, ids=[f"{c}-{''.join(map(str, f))}" for c, f in CLUSTERED])
def test_clustered_group_mixed_close_to_fp64(case, flags):
    m = golden_engine_mol(golden(case), flags)
    kw = dict(job=4, iters=2, traj=4000, keep=True)
    f64 = m.run(E.PREC_FP64, lanes=1, **kw)
    for lanes in (0, 64):
        mix = m.run(E.PREC_MIXED, lanes=lanes, **kw)
        assert mix["lanes"] >= 4
        assert np.array_equal(mix["status"] != 0, f64["status"] != 0)
        both, d = _paired(mix, f64)
        assert np.median(d) <= 1e-6, (lanes, np.median(d))
        assert np.mean(d <= 1e-3) >= 0.99, lanes
        dev, se = paired_ccs_dev(mix["chi"], mix["status"], f64["chi"], f64["status"])
        print(f"{case} {flags} lanes {lanes}: paired CCS dev {dev:.2e} +- {se:.2e}")
        assert abs(dev) <= 1e-3, (lanes, dev, se)


@pytest.fixture(scope="module")
def ion128():
    p = E.default_params(temp=298.0, traj=3000, iter=2, multipole=0, dispersion_cutoff=0, seed=77)
    return E.Molecule.from_file(os.path.join(GOLD, "ion128.mfj"), params=p)


@pytest.fixture(scope="module")
def ion2000():
    p = E.default_params(temp=298.0, traj=2000, iter=2, multipole=0, dispersion_cutoff=0, seed=78)
    return E.Molecule.from_file(os.path.join(GOLD, "ion2000.mfj"), params=p)


@pytest.mark.parametrize("prec", [E.PREC_MIXED, E.PREC_FP64])
def test_batch_of_one_equals_run(ion128, prec):
    """A batch of one job (shared LDS launch) and cosims_run (LDS kernel)
    give identical results."""
    shape = dict(lanes=8, wg_size=256)
    single = ion128.run(prec, job=11, **shape)
    (b,), st = E.run_batch([(ion128, dict(job=11))], prec=prec, **shape)
    assert st["lanes"] == 8 and st["launches"] == 1
    assert np.array_equal(b["omega"], single["omega"])
    assert b["n_hit"] == single["n_hit"] and b["total_steps"] == single["total_steps"]


@pytest.mark.parametrize("case,flags", CLUSTERED, ids=[f"{c}-{''.join(map(str, f))}" for c, f in CLUSTERED])
def test_clustered_lds_headers_equal_table_kernel(case, flags):
    """Mixed precision, 64 lanes: cosims_run of a clustered molecule stages
    its cluster headers in LDS (k_traj_cl_lds, multipole order <= 1); a batch
    of that one job reads them through L2 (table kernel, k_traj_multi).  Same
    terms in the same order: the same bits, trajectory by trajectory.  (Order
    2 runs the table kernel both ways.)"""
    m = golden_engine_mol(golden(case), flags)
    kw = dict(job=7, iters=2, traj=1500)
    run = m.run(E.PREC_MIXED, lanes=64, wg_size=1024, keep=True, **kw)
    (bat,), _ = E.run_batch([(m, kw)], prec=E.PREC_MIXED, lanes=64, wg_size=256)
    assert run["lanes"] == bat["lanes"] == 64
    assert np.array_equal(run["omega"], bat["omega"])
    assert run["total_steps"] == bat["total_steps"]
    assert run["wmean"] == bat["wmean"]


@pytest.mark.parametrize("own", [False, True], ids=["shared", "own"])
def test_batch_alone_equals_mix(ion128, ion2000, own):
    """Uneven molecules and two force types in one call: each job's result
    is the same alone and in the mix.  The exact-type jobs fit LDS: one
    shared LDS launch for all four (default) or one LDS-kernel launch each
    (own launches); the clustered job runs the table kernel."""
    g = golden("acau_charged")
    ac1 = golden_engine_mol(g, [1, 1, 1, 1])   # clustered, type 1
    ac4 = golden_engine_mol(g, [1, 0, 0, 1])   # exact, type 4
    jobs = [(ion128, dict(job=1)), (ion2000, dict(job=2, iters=1, traj=3000)),
            (ac1, dict(job=3, iters=2, traj=1500)), (ac4, dict(job=4, iters=1, traj=2500)),
            (ion128, dict(job=5, iters=1, traj=1000))]
    shape = dict(lanes=16, wg_size=512)
    mix, st = E.run_batch(jobs, prec=E.PREC_MIXED, own_launches=own, **shape)
    assert st["launches"] == (5 if own else 2)
    assert st["n_traj"] == sum(j[1].get("iters", j[0].params.iter) * j[1].get("traj", j[0].params.traj)
                               for j in jobs)
    for k, job in enumerate(jobs):
        (alone,), _ = E.run_batch([job], prec=E.PREC_MIXED, own_launches=own, **shape)
        assert np.array_equal(mix[k]["omega"], alone["omega"]), k
        assert mix[k]["wmean"] == alone["wmean"]


def test_batch_size_rule_alone_equals_mix(ion128, ion2000):
    """Lanes by molecule size: one launch per (force type, lanes), run
    concurrently; each job's result is the same alone and in the mix."""
    g = golden("acau_charged")
    ac6 = golden_engine_mol(g, [0, 0, 0, 1])   # neutral exact, type 6, 252 atoms
    jobs = [(ion128, dict(job=31, iters=1, traj=3000)), (ion2000, dict(job=32, iters=1, traj=3000)),
            (ac6, dict(job=33, iters=1, traj=3000))]
    mix, st = E.run_batch(jobs, prec=E.PREC_MIXED, lanes=0)
    assert st["launches"] == 3
    assert [r["lanes"] for r in mix] == [4, 64, 8]  # size classes 0, 4, 1
    for k, job in enumerate(jobs):
        (alone,), _ = E.run_batch([job], prec=E.PREC_MIXED, lanes=0)
        assert np.array_equal(mix[k]["omega"], alone["omega"]), k
        single = job[0].run(E.PREC_MIXED, job=job[1]["job"], iters=1, traj=3000, lanes=mix[k]["lanes"],
                            wg_size=512)
        assert np.array_equal(mix[k]["omega"], single["omega"]), k


def test_batch_tuner_and_stats(ion128, ion2000):
    jobs = [(ion128, dict(job=21, iters=1, traj=2000)), (ion2000, dict(job=22, iters=1, traj=2000))]
    CL, CW = E.tune_batch(jobs, E.PREC_MIXED)
    present = {E.size_class(128), E.size_class(2000)}
    for c in range(E.SIZE_CLASSES):
        if c in present:
            assert CL[c] in (4, 8, 16, 32, 64) and CW[c] in (256, 512, 1024), (c, CL, CW)
        else:
            assert CL[c] == 0 and CW[c] == 0
    res, st = E.run_batch(jobs, prec=E.PREC_MIXED, lanes=-1, class_lanes=CL, class_wg=CW)
    assert st["class_lanes"] == CL
    assert [r["lanes"] for r in res] == [CL[E.size_class(128)], CL[E.size_class(2000)]]
    assert st["kernel_ms"] > 0 and st["force_evals"] > st["total_steps"] > 0
    # every force evaluation of these exact-type molecules is a full atom sweep
    assert st["pair_evals"] >= 128 * st["force_evals"]
    for r in res:
        assert r["n_traj"] == 2000 and 0 < r["n_hit"] <= 2000 and np.isfinite(r["wmean"])
    _, st2 = E.run_batch(jobs, prec=E.PREC_MIXED, lanes=-1)  # tunes inside
    assert all(st2["class_lanes"][c] for c in present)


@pytest.fixture(scope="module")
def ion8000():
    """8,000 atoms: beyond one workgroup's LDS in FP32 and FP64 (tiled kernel)."""
    import tempfile
    ion = synth.synthetic_ion(8000, 18.0 * 4 ** (1 / 3), 0.3, 33, 3001)
    path = os.path.join(tempfile.mkdtemp(prefix="cosims_t_"), "ion8000.mfj")
    synth.write_mfj(path, "ion8000", ion)
    p = E.default_params(temp=298.0, traj=1000, iter=1, multipole=0, dispersion_cutoff=0, seed=79)
    return E.Molecule.from_file(path, params=p)


@pytest.mark.parametrize("prec", [E.PREC_MIXED, E.PREC_FP64])
def test_tiled_kernel_tile_size_does_not_change_bits(ion8000, prec):
    """8,000 atoms run the tiled kernel (molecule streamed through LDS in
    tiles sized by the workgroup: 256 threads hold a quarter of a CU's LDS,
    1024 all of it).  Tiles start at multiples of 64, so every lane sums the
    same atoms in the same order: identical results at both sizes."""
    kw = dict(job=8, iters=2, traj=600, lanes=16, keep=True)
    a = ion8000.run(prec, wg_size=256, **kw)
    b = ion8000.run(prec, wg_size=1024, **kw)
    assert a["wg_size"] == 256 and b["wg_size"] == 1024
    assert np.array_equal(a["status"], b["status"])
    assert np.array_equal(a["chi"], b["chi"])
    assert np.array_equal(a["omega"], b["omega"])


def test_large_molecule_beyond_lds(ion8000):
    """Beyond LDS (the tiled kernel for these exact types): FP64 groups match
    the lane kernel, mixed precision matches FP64."""
    kw = dict(job=6, iters=1, traj=1000, keep=True)
    lane = ion8000.run(E.PREC_FP64, lanes=1, **kw)
    grp = ion8000.run(E.PREC_FP64, lanes=16, wg_size=256, **kw)
    mix = ion8000.run(E.PREC_MIXED, lanes=16, wg_size=256, **kw)
    both, d = _paired(grp, lane)
    assert both.sum() > 0.9 * (lane["status"] != 0).sum()
    assert np.median(d) <= 1e-8, np.median(d)
    np.testing.assert_allclose(grp["omega"], lane["omega"], rtol=5e-4)
    assert abs(mix["wmean"] - lane["wmean"]) / lane["wmean"] <= 1e-3


def test_batch_mixes_tiled_and_lds_launches(ion8000, ion128, ion2000):
    """One launch shape, molecules on both sides of the LDS limit, own
    launches: the 8,000-atom job runs the tiled kernel, the others one LDS
    launch each; every result equals the job alone and cosims_run with the
    same lanes."""
    shape = dict(lanes=16, wg_size=256, own_launches=True)
    jobs = [(ion128, dict(job=41, iters=1, traj=1500)), (ion8000, dict(job=42, iters=1, traj=600)),
            (ion2000, dict(job=43, iters=2, traj=800))]
    for prec in (E.PREC_MIXED, E.PREC_FP64):
        mix, st = E.run_batch(jobs, prec=prec, **shape)
        assert st["launches"] == 3
        for k, job in enumerate(jobs):
            (alone,), _ = E.run_batch([job], prec=prec, **shape)
            assert np.array_equal(mix[k]["omega"], alone["omega"]), (prec, k)
            single = job[0].run(prec, job=job[1]["job"], iters=job[1]["iters"], traj=job[1]["traj"], lanes=16,
                                wg_size=256)
            assert np.array_equal(mix[k]["omega"], single["omega"]), (prec, k)
            assert mix[k]["total_steps"] == single["total_steps"]


def test_batch_tiny_jobs_share_a_table_launch(ion128, ion2000):
    """Own launches: jobs below 65,536 trajectories x atoms share one
    table-kernel launch (a launch each would cost more than their work); a
    larger job keeps its own.  Every result equals cosims_run (LDS kernel)
    bit for bit."""
    shape = dict(lanes=8, wg_size=256)
    jobs = [(ion128, dict(job=50 + k, iters=1, traj=300)) for k in range(4)]
    jobs.append((ion2000, dict(job=60, iters=1, traj=1000)))
    mix, st = E.run_batch(jobs, prec=E.PREC_MIXED, own_launches=True, **shape)
    assert st["launches"] == 2
    for k, (m, spec) in enumerate(jobs):
        single = m.run(E.PREC_MIXED, job=spec["job"], iters=1, traj=spec["traj"], **shape)
        assert np.array_equal(mix[k]["omega"], single["omega"]), k
        assert mix[k]["total_steps"] == single["total_steps"], k


@pytest.mark.parametrize("prec", [E.PREC_MIXED, E.PREC_FP64])
def test_batch_shared_launch_equals_own_launches(ion128, ion2000, ion8000, prec):
    """The default shared mode: the jobs of a launch shape within LDS (128
    and 2,000 atoms) in one persistent launch, each workgroup staging the
    molecule of the job it works on; the 8,000-atom job keeps a tiled launch
    of its own.  The results are those of own launches bit for bit."""
    shape = dict(lanes=16, wg_size=256)
    jobs = [(ion128, dict(job=71, iters=1, traj=1500)), (ion2000, dict(job=72, iters=1, traj=1000)),
            (ion8000, dict(job=73, iters=1, traj=400))]
    shared, st = E.run_batch(jobs, prec=prec, **shape)
    assert st["launches"] == 2
    own, st2 = E.run_batch(jobs, prec=prec, own_launches=True, **shape)
    assert st2["launches"] == 3
    for k in range(len(jobs)):
        assert np.array_equal(shared[k]["omega"], own[k]["omega"]), (prec, k)
        assert shared[k]["total_steps"] == own[k]["total_steps"], (prec, k)


@pytest.mark.parametrize("prec", [E.PREC_MIXED, E.PREC_FP64])
def test_batch_shared_lds_restaging(ion128, ion2000, prec):
    """24 jobs of one launch shape, 128 and 2,000 atoms, 300 to 2,531
    trajectories each, in one shared persistent launch: workgroups start on
    jobs in proportion to their work, spend them and restage the molecule of
    the job with the most trajectories left, many times over.  Every job's
    result equals cosims_run of it alone bit for bit."""
    shape = dict(lanes=16, wg_size=512)
    jobs = [(ion2000 if k % 3 == 0 else ion128, dict(job=200 + k, iters=1, traj=300 + 97 * k)) for k in range(24)]
    shared, st = E.run_batch(jobs, prec=prec, **shape)
    assert st["launches"] == 1
    assert st["n_traj"] == sum(spec["traj"] for _, spec in jobs)
    for k, (m, spec) in enumerate(jobs):
        single = m.run(prec, job=spec["job"], iters=1, traj=spec["traj"], **shape)
        assert np.array_equal(shared[k]["omega"], single["omega"]), (prec, k)
        assert shared[k]["total_steps"] == single["total_steps"], (prec, k)


def test_clustered_branch_counts():
    """The 64-lane clustered kernels count their branches: every force
    evaluation examines every cluster (clusters examined = clusters x
    evaluations exactly); far clusters and those with dipole terms are
    subsets; with the default radii (R_d >= R_m) no atom takes the
    charges-only branch; mixed precision and FP64 take the same branches on
    the same initial conditions except on the few chaotic trajectories; the
    reference FLOPs follow the per-branch table."""
    g = golden("acau_charged")
    m = golden_engine_mol(g, [1, 1, 1, 1])
    inf = m.info()
    assert inf["force_type"] == 1 and inf["ncl"] > 0
    kw = dict(job=5, iters=1, traj=4000, lanes=64, wg_size=256)
    runs = {p: m.run(p, **kw) for p in (E.PREC_MIXED, E.PREC_FP64)}
    for p, r in runs.items():
        hdr, far, fdip, exlj, exq, lj = r["cl_counts"]
        evals = sum(r["round_evals"])
        assert hdr == inf["ncl"] * evals, (p, hdr, evals)
        assert 0 < fdip <= far < hdr
        assert exq == 0 and exlj > 0 and lj > 0
        assert r["pair_evals"] == exlj + exq + lj
        steps = sum(r["round_steps"])
        want = (12 * (hdr - far) + 49 * far + 67 * fdip + 62 * exlj + 48 * exq + 22 * lj + 27 * evals
                + 50 * steps + 24 * r["pot_evals"])
        assert abs(r["ref_flops"] - want) <= 1e-9 * want
    a, b = runs[E.PREC_MIXED]["cl_counts"], runs[E.PREC_FP64]["cl_counts"]
    for k in range(6):
        assert abs(a[k] - b[k]) <= 0.01 * max(1, b[k]), (k, a, b)


def test_exact_type_reference_flops(ion2000):
    """Exact type 4: 58 FLOP a pair + 21 an evaluation + 50 a step + 24 a
    potential atom pair; pairs = atoms x evaluations."""
    r = ion2000.run(E.PREC_MIXED, job=3, iters=1, traj=3000, lanes=16, wg_size=256)
    evals, steps = sum(r["round_evals"]), sum(r["round_steps"])
    assert r["pair_evals"] == 2000 * evals
    assert r["pot_evals"] % 2000 == 0 and r["pot_evals"] >= 2000 * r["n_hit"]
    want = 58 * r["pair_evals"] + 21 * evals + 50 * steps + 24 * r["pot_evals"]
    assert abs(r["ref_flops"] - want) <= 1e-9 * want
    assert r["cl_counts"] == [0] * 6
