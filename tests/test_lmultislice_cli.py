"""The LMultiSlice front end (multislice_amd/lmultislice over
include/msx/lmultislice.hpp): the XML command of Dispatcher::parseCommand
(Dispatcher.cu:25-159) on the reference's own command files, its error
messages, and the .coo reader (ModelCoo.cpp:20-90).  CPU only: the --parse
and --read-coo modes do no GPU work (the full run is tests/test_gpu_lmultislice.py)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "multislice_amd", "lmultislice")
GOLD = os.path.join(ROOT, "tests", "golden", "lmultislice")


def run(*args, cwd=None):
    assert os.path.exists(EXE), "multislice_amd/lmultislice not built (run __graft_entry__.build())"
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=60, cwd=cwd)


def parsed(path):
    r = run("--parse", path)
    assert r.returncode == 0, r.stderr + r.stdout
    return dict(line.split("=", 1) for line in r.stdout.strip().splitlines())


@pytest.mark.parametrize("name,io,n", [("Fe_10440839_64x64.xml", ("model/Fe_10440839.coo", "result/Fe_10440839_64x64"), 64),
                                       ("Fe_4_4096x4096.xml", ("model/Fe_4.coo", "result/Fe_4_4096x4096"), 4096)])
def test_reference_command_files(name, io, n):
    c = parsed(os.path.join(GOLD, name))
    assert (c["fileNameInput"], c["fileNameOutput"]) == io
    assert int(c["nx"]) == int(c["ny"]) == n
    assert float(c["dpa"]) == 10.0 and float(c["radiuc"]) == 20.0 and int(c["numberslices"]) == 1
    # float fields as the reference stores them (float(atof(...)))
    assert abs(float(c["aperture"]) - 10.37) < 1e-6 and abs(float(c["cs"]) - 1.3) < 1e-6
    assert float(c["defocus"]) == -70.0 and float(c["keV"]) == 200.0


def test_empty_field_and_missing_file(tmp_path):
    with open(os.path.join(GOLD, "Fe_10440839_64x64.xml")) as f:
        text = f.read()
    bad = tmp_path / "bad.xml"
    bad.write_text(text.replace("<ny>64</ny>", "<ny></ny>"))
    r = run("--parse", str(bad))
    assert r.returncode == 255 and "Empty ny field!" in r.stderr
    r = run("--parse", str(tmp_path / "absent.xml"))
    assert r.returncode == 255 and "doesn't exist" in r.stderr
    # a run stops at the model when the model file is absent (no GPU work is reached)
    r = run(os.path.join(GOLD, "Fe_10440839_64x64.xml"), cwd=str(tmp_path))
    assert r.returncode == 255 and "Can not read file model/Fe_10440839.coo" in r.stdout


def test_coo_reader(tmp_path):
    coo = tmp_path / "si.coo"
    coo.write_text("Si two atoms, decimal commas and a fraction\n"
                   "5,431 5,431 10,862 90 90 90 2\n"
                   "0 0 0 14 0 0 1\n"
                   "1/4 0,25 0.125 14 0,5 2 0.75\n")
    r = run("--read-coo", str(coo))
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert [float(v) for v in lines[0].split("=")[1].split()] == pytest.approx([5.431, 5.431, 10.862, 90, 90, 90],
                                                                                 rel=1e-7)
    assert lines[1] == "atoms=2"
    sym, *vals = lines[3].split()
    assert sym == "Si"
    assert [float(v) for v in vals] == pytest.approx([0.25, 0.25, 0.125, 0.5, 2.0, 0.75], rel=1e-7)
    short = tmp_path / "short.coo"
    short.write_text("comment\n5 5 5 90 90 90 3\n0 0 0 14 0 0 1\n")
    assert run("--read-coo", str(short)).returncode == 255
