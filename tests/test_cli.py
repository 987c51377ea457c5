"""The command-line program's argument handling and its setup log, on the CPU.

The flags and error behaviour are fileParams::interpretCMD's
(src/fileParams.cpp:78-201): a bad flag prints "ERROR: Check command line
arguments." and exits with status 0.  With the molecule set up on host
threads (-g 0), everything the log prints before the trajectory phase must
equal the reference's example logs line for line (settings, atom types,
masses, boundary fit, clusters, velocity window); on a machine without a GPU
the run then stops with an engine error.
"""
import os
import subprocess

import pytest

from conftest import CLI, LOGS, run_cli_example

BAD = [["-m"], ["-x", "1"], ["-m", "a.pdb", "-s", "12a"], ["-m", "a.pdb", "-n", "two"], ["-mm", "a.pdb"],
       ["-m", "-a"], ["-m", "a.pdb", "-p", "fp16"], ["-m", "a.pdb", "-r", "mt"]]


@pytest.mark.parametrize("args", BAD, ids=[" ".join(a) for a in BAD])
def test_bad_arguments(args, tmp_path):
    r = subprocess.run([CLI, *args], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert r.stdout == "ERROR: Check command line arguments.\nProgram will now exit...\n"
    assert not list(tmp_path.iterdir())       # nothing written before the flags parse


def test_no_molecule_and_wrong_extension(tmp_path):
    r = subprocess.run([CLI, "-s", "5"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert "ERROR: No molecule file was given" in r.stdout and r.returncode == 0
    (tmp_path / "x.xyz").write_text("1\n")
    r = subprocess.run([CLI, "-m", "x.xyz"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.stdout.endswith("ERROR: Check command line arguments.\nProgram will now exit...\n")


def test_unreadable_molecule(tmp_path):
    r = subprocess.run([CLI, "-m", "missing.mfj", "-s", "1"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0 and "ERROR: Cannot read molecule file." in r.stdout


def test_input_file_notes(tmp_path):
    from conftest import GOLDEN
    import shutil
    shutil.copy(os.path.join(GOLDEN, "acau.pdb"), tmp_path)
    (tmp_path / "in.txt").write_text("name t\nbogus 1\ntemp 300\ntemp 310\nalone\n# comment x\n")
    r = subprocess.run([CLI, "-m", "acau.pdb", "-i", "in.txt", "-s", "3", "-g", "0"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    out = r.stdout
    assert "ERROR: Can not interpret option 'bogus' on \n\tline number 2 in in.txt;" in out
    assert "WARNING: Parameter TEMP already specified at line number 4;" in out
    assert "ERROR: Line number 5 in in.txt is incomplete;" in out
    assert (tmp_path / "omega_t_0.txt").exists()          # name + "_" + int(charge)
    assert "Temperature of the molecule system (Kelvin):" in out and "310.00" in out


@pytest.mark.parametrize("entry", LOGS, ids=[e["log"].split("/")[-1] for e in LOGS])
def test_setup_log_matches_example(entry, tmp_path):
    from golden.logfilter import log_lines
    rc, out, log = run_cli_example(entry, str(tmp_path), ["-g", "0", "-p", "fp64", "-r", "reference", "-l", "1"])
    assert os.path.exists(log), out                       # same omega_<name>.txt name as the reference's
    got = log_lines(open(log).read())
    want = entry["lines"]
    n_setup = next(i for i, l in enumerate(want) if l.startswith((" progress:", " Number of completed")))
    assert got[:n_setup] == want[:n_setup]
