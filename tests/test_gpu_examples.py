"""The reference's recorded example runs, regenerated on the GPU (MI355X).

examples/*/omega_*.txt hold the output of the reference's own runs (1 OpenMP
thread, fixed seeds).  With the reference's initial-condition stream
(COSIMS_RNG_REFERENCE) and the FP64 lane-per-trajectory kernel (the
reference's operation order), every printed CCS integral, its standard error
and the final statistics must come out digit for digit.
"""
import json
import os

import pytest

from cosims_amd import engine as E

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
LOGS = json.load(open(os.path.join(GOLD, "example_logs.json")))


@pytest.mark.parametrize("entry", LOGS, ids=[e["log"].split("/")[-1] for e in LOGS])
def test_example_log_reproduced(entry, tmp_path):
    p = E.default_params(seed=entry["seed"])
    if entry["input"] is not None:
        f = tmp_path / "input.txt"
        f.write_text(entry["input"])
        E.read_params(str(f), p)
    m = E.Molecule.from_file(os.path.join(GOLD, entry["mol"]),
                             os.path.join(GOLD, entry["atoms_dat"]) if entry["atoms_dat"] else None, params=p)
    r = m.run(E.PREC_FP64, lanes=1, rng=E.RNG_REFERENCE, iters=entry["iters"], traj=entry["traj"])
    got = [f"{x:.2f}" for x in r["omega"]]
    got_err = [f"{x:.2f}" for x in r["omega_err"]]
    assert got == entry["omega"], (got, entry["omega"])
    assert got_err == entry["omega_err"]
    fin = entry["final"]
    assert f"{r['mean']:.3f}" == fin["mean"]
    assert f"{r['sd']:.3f}" == fin["sd"]
    assert f"{r['wmean']:.3f}" == fin["wmean"]
    assert f"{r['werr']:.3f}" == fin["werr"]
    assert f"{r['max_ergy']:.4f}" == fin["max_ergy"]
