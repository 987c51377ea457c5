"""CCS request server (MI355X): continuous batching over persistent launches.

A request's result must be exactly cosims_run over the same blocks (iter =
blocks, traj = block_traj, the server's launch shape), and the same whether
it ran alone or shared launches with other requests, whatever the admission
order (bit-identical per-block CCS).  Tolerance: none (equality).
"""
import os
import time

import numpy as np
import pytest

from cosims_amd import engine as E

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

SHAPE = dict(lanes=8, wg_size=256)
BLOCK = 2000


def _mol(name, atoms=None):
    return E.read_molfile(os.path.join(GOLDEN, name), os.path.join(GOLDEN, atoms) if atoms else None)


def _params(seed, **kw):
    kw.setdefault("setup_gpu", 1)
    return E.default_params(temp=298.0, seed=seed, **kw)


REQUESTS = [  # (molecule file, params, job, target relative error)
    ("ion128.mfj", dict(multipole=0, dispersion_cutoff=0), 1, 0.02),
    ("acau.mfj", dict(), 2, 0.03),                                    # charged, clustered type 1
    ("ion2000.mfj", dict(multipole=0, dispersion_cutoff=0), 3, 0.03),
    ("acau.pdb", dict(multipole=0, dispersion_cutoff=0), 4, 0.05),    # neutral, exact type 6
]


def _server(**kw):
    cfg = dict(prec=E.PREC_MIXED, block_traj=BLOCK, min_blocks=3, max_blocks=40, slice_traj=32768, **SHAPE)
    cfg.update(kw)
    return E.Server(**cfg)


def _submit_all(srv, reqs):
    return [srv.submit(_mol(f), _params(100 + j, **p), job=j, target=t) for f, p, j, t in reqs]


@pytest.fixture(scope="module")
def mixed_results():
    with _server() as srv:
        tickets = _submit_all(srv, REQUESTS)
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
    return res, st


def test_server_requests_finish_at_target(mixed_results):
    res, st = mixed_results
    assert st["done"] == len(REQUESTS) and st["failed"] == 0 and st["active"] == 0
    for (f, p, j, t), r in zip(REQUESTS, res):
        assert r["status"] == 0
        assert 3 <= r["blocks"] <= 40
        assert r["rel_err"] <= t or r["blocks"] == 40, (f, r["rel_err"])
        assert r["t_submit_ms"] <= r["t_ready_ms"] <= r["t_admit_ms"] <= r["t_done_ms"]
        assert r["n_traj"] == r["blocks"] * BLOCK
        # setup phases: the boundary marches always, PCM clustering for the clustered request only
        phases = r["setup_boundary_ms"] + r["setup_cluster_ms"] + r["setup_moments_ms"] + r["setup_upload_ms"]
        assert 0 < r["setup_boundary_ms"] and phases <= r["setup_ms"], f
        assert (r["setup_cluster_ms"] > 0) == (p.get("multipole", 1) == 1), f
    assert st["slices"] >= 1 and st["traj"] >= sum(r["n_traj"] for r in res)


def test_server_result_equals_run(mixed_results):
    """The blocks are CCS integrals 0..k-1 of the request's Philox stream."""
    res, _ = mixed_results
    for (f, p, j, t), r in zip(REQUESTS, res):
        mol = _mol(f)
        m = E.Molecule.from_file(os.path.join(GOLDEN, f), params=_params(100 + j, **p))
        run = m.run(E.PREC_MIXED, job=j, iters=r["blocks"], traj=BLOCK, **SHAPE)
        assert np.array_equal(r["omega"], run["omega"]), f
        assert r["wmean"] == run["wmean"] and r["werr"] == run["werr"]
        assert mol["charged"] == bool(m.info()["charged"])


def test_server_alone_equals_mix(mixed_results):
    res, _ = mixed_results
    for k, req in enumerate(REQUESTS):
        with _server() as srv:
            (t,) = _submit_all(srv, [req])
            alone = srv.wait(t)
        assert alone["blocks"] == res[k]["blocks"], req[0]
        assert np.array_equal(alone["omega"], res[k]["omega"]), req[0]


def test_server_policy_does_not_change_results(mixed_results):
    """Equal-share slices instead of shortest-remaining-work-first: other
    slice compositions, the same results."""
    res, _ = mixed_results
    with _server(policy=E.SERVE_EQUAL_SHARE, slice_traj=8192) as srv:
        tickets = _submit_all(srv, REQUESTS)
        out = [srv.wait(t) for t in tickets]
    for r, ref in zip(out, res):
        assert np.array_equal(r["omega"], ref["omega"])


@pytest.mark.parametrize("issue", [dict(issue_fraction=0.5), dict(prior_rel_sd=2.0),
                                   dict(issue_fraction=0.5, prior_rel_sd=3.0)])
def test_server_issue_policy_does_not_change_results(mixed_results, issue):
    """How many blocks a slice hands a request (half or all of the estimate,
    a prior-sized first issue) changes only the blocks past the stopping
    point, never the blocks taken or the result."""
    res, _ = mixed_results
    with _server(**issue) as srv:
        tickets = _submit_all(srv, REQUESTS)
        out = [srv.wait(t) for t in tickets]
        st = srv.stats()
    for r, ref in zip(out, res):
        assert r["blocks"] == ref["blocks"]
        assert np.array_equal(r["omega"], ref["omega"])
    assert st["traj"] >= sum(r["n_traj"] for r in out)


def test_server_rejects_bad_issue_settings():
    for kw in (dict(issue_fraction=1.5), dict(issue_fraction=-0.1), dict(prior_rel_sd=-1.0)):
        with pytest.raises(E.CosimsError, match="invalid argument"):
            _server(**kw)


def _slice_reserve():
    """Device bytes the server sets aside for its slices' trajectory buffers
    (taken out of hbm_budget before any request is admitted)."""
    big = 10 << 30
    with _server(hbm_budget=big) as srv:
        return big - srv.stats()["budget"]


def test_server_budget_must_cover_slice_buffers():
    reserve = _slice_reserve()
    assert reserve >= 3 * (32768 + 65536 + BLOCK) * 100  # three slices in flight
    with pytest.raises(E.CosimsError, match="out of memory"):
        _server(hbm_budget=reserve)


def test_server_budget_serialises_admission(mixed_results):
    """A budget one byte beyond the slice buffers admits one request at a
    time (its bytes come back only once no slice in flight holds it); the
    results do not change."""
    res, _ = mixed_results
    with _server(hbm_budget=_slice_reserve() + 1) as srv:
        tickets = _submit_all(srv, REQUESTS)
        out = [srv.wait(t) for t in tickets]
        # slices issued past the last stopping point still hold that request
        t_end = time.time() + 60
        while srv.stats()["bytes_in_use"] and time.time() < t_end:
            time.sleep(0.01)
        st = srv.stats()
    assert st["done"] == len(REQUESTS)
    assert st["bytes_in_use"] == 0 and st["bytes_peak"] > 0
    spans = sorted((r["t_admit_ms"], r["t_done_ms"]) for r in out)
    for (a0, d0), (a1, d1) in zip(spans, spans[1:]):
        assert a1 >= d0  # never two requests active at once
    for r, ref in zip(out, res):
        assert np.array_equal(r["omega"], ref["omega"])


def test_server_continuous_admission():
    """Requests submitted while others run are admitted at the next slice
    (their admission precedes the earlier requests' completion)."""
    with _server(min_blocks=100, max_blocks=100, slice_traj=16384) as srv:
        first = srv.submit(_mol("ion2000.mfj"), _params(7, multipole=0, dispersion_cutoff=0), job=1, target=1e-9)
        while srv.stats()["active"] == 0:
            time.sleep(0.005)
        # host setup: its boundary marches would otherwise queue behind the running slices
        late = srv.submit(_mol("ion128.mfj"), _params(8, multipole=0, dispersion_cutoff=0, setup_gpu=0), job=2,
                          target=1e-9)
        a, b = srv.wait(first), srv.wait(late)
    assert a["blocks"] == b["blocks"] == 100
    assert b["t_admit_ms"] < a["t_done_ms"]


def test_server_size_rule_equals_run():
    """lanes=0: every request runs with the lanes of its size class; the
    result is cosims_run with those lanes."""
    with _server(lanes=0, wg_size=512) as srv:
        tickets = _submit_all(srv, REQUESTS)
        res = [srv.wait(t) for t in tickets]
    for (f, p, j, t), r in zip(REQUESTS, res):
        m = E.Molecule.from_file(os.path.join(GOLDEN, f), params=_params(100 + j, **p))
        inf = m.info()
        L = E.job_lanes(inf)
        assert L == 64 or inf["force_type"] in (4, 6)  # the clustered types run 64 lanes
        run = m.run(E.PREC_MIXED, job=j, iters=r["blocks"], traj=BLOCK, lanes=L, wg_size=512)
        assert np.array_equal(r["omega"], run["omega"]), (f, L)


def test_server_class_shapes_equal_run():
    """Per-size-class lanes and workgroup sizes: each request runs with its
    class's shape and gives cosims_run's result with that shape."""
    CL = [8, 16, 0, 0, 32]       # class 2, 3: the size rule
    CW = [256, 0, 0, 512, 1024]  # class 1, 2: wg_size
    with _server(lanes=0, wg_size=256, class_lanes=CL, class_wg=CW) as srv:
        tickets = _submit_all(srv, REQUESTS)
        res = [srv.wait(t) for t in tickets]
    for (f, p, j, t), r in zip(REQUESTS, res):
        m = E.Molecule.from_file(os.path.join(GOLDEN, f), params=_params(100 + j, **p))
        inf = m.info()
        k = E.job_class(inf)
        L = E.job_lanes(inf, CL)
        W = max(CW[k] or 256, L)
        assert (r["lanes"], r["wg_size"]) == (L, W), (f, k)
        run = m.run(E.PREC_MIXED, job=j, iters=r["blocks"], traj=BLOCK, lanes=L, wg_size=W)
        assert np.array_equal(r["omega"], run["omega"]), (f, L, W)


def test_server_rejects_bad_class_shapes():
    for kw in (dict(class_lanes=[0, 0, 12, 0, 0]), dict(class_wg=[0, 100, 0, 0, 0]),
               dict(class_wg=[2048, 0, 0, 0, 0])):
        with pytest.raises(E.CosimsError, match="invalid argument"):
            _server(lanes=0, **kw)


def test_server_rejects_bad_requests():
    with _server() as srv:
        bad = dict(xyz=np.zeros((0, 3)), mass=[], sigma=[], epsilon=[], charged=0, sigma_max=3.5)
        with pytest.raises(E.CosimsError, match="invalid argument"):
            srv.submit(bad, _params(1))
        with pytest.raises(E.CosimsError, match="unknown ticket"):
            srv.wait(12345)


def test_server_large_molecule_equals_run():
    """A request beyond one workgroup's LDS (8,000 atoms: the tiled kernel)
    gives cosims_run's result over the same blocks."""
    import tempfile

    from cosims_amd import synth
    ion = synth.synthetic_ion(8000, 18.0 * 4 ** (1 / 3), 0.3, 33, 3001)
    path = os.path.join(tempfile.mkdtemp(prefix="cosims_s_"), "ion8000.mfj")
    synth.write_mfj(path, "ion8000", ion)
    mol = E.read_molfile(path)
    p = dict(multipole=0, dispersion_cutoff=0)
    with _server(max_blocks=6) as srv:
        r = srv.wait(srv.submit(mol, _params(300, **p), job=30, target=0.05))
    assert r["status"] == 0 and 3 <= r["blocks"] <= 6
    m = E.Molecule.from_file(path, params=_params(300, **p))
    run = m.run(E.PREC_MIXED, job=30, iters=r["blocks"], traj=BLOCK, **SHAPE)
    assert np.array_equal(r["omega"], run["omega"])
    assert r["wmean"] == run["wmean"]


def test_server_own_launches_equal_shared(mixed_results):
    """Slices with a launch of their own per request give the results of
    the default shared LDS launches bit for bit."""
    res, _ = mixed_results
    with _server(own_launches=True) as srv:
        out = [srv.wait(t) for t in _submit_all(srv, REQUESTS)]
    for r, ref in zip(out, res):
        assert r["blocks"] == ref["blocks"]
        assert np.array_equal(r["omega"], ref["omega"])
        assert r["wmean"] == ref["wmean"]
