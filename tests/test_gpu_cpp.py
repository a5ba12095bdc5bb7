"""GPU test of the C++ LMultiSlice mirror (include/msx/lmultislice.hpp): the
compiled driver tests/cpp/lms_check runs ModelPotential / ModelSimulated /
Microscope / CPotentialBuilder through libmsx.so; its outputs are compared with
the oracle on the same Si [001] model."""
import os
import subprocess

import numpy as np
import pytest

from multislice_amd import specimen
from oracle import Oracle as O
from conftest import max_rel, rel_l2

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "lms_check")


def test_cpp_lmultislice_mirror(tmp_path):
    assert os.path.exists(EXE), "tests/cpp/lms_check not built (run __graft_entry__.build())"
    prefix = str(tmp_path / "si")
    r = subprocess.run([EXE, prefix], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "builder=42" in r.stdout
    nx = ny = 128
    nz = 4
    a = 5.431
    pot = np.fromfile(prefix + "_potential.bin", np.float64).reshape(nz, ny, nx)
    img = np.fromfile(prefix + "_image.bin", np.float64).reshape(ny, nx, 2)
    img = (img[..., 0] + 1j * img[..., 1]).T  # -> (ix, iy)
    # the same model through the oracle
    at, cell = specimen.si_001((2, 2, 1))
    V = O.potential(at["Z"], np.stack([at["x"], at["y"], at["z"]], 1), cell, nx, ny, nz, 4.0,
                    table=0, formula=0)
    assert rel_l2(pot.transpose(0, 2, 1), V) < 1e-6  # model coordinates are float32 there
    t = O.transmission(V, cell[0], cell[1], 200.0)
    exitw = O.multislice(t, np.full(nz, a / 4), np.arange(nz), np.ones((nx, ny)), cell[0], cell[1], 200.0)
    ref_img, _ = O.ctf_image(exitw, cell[0], cell[1], 200.0, cs=1.2e7, df=-600.0, aobj=0.02)
    e = rel_l2(img, ref_img)
    m = max_rel(img, ref_img)
    print(f"C++ mirror image vs oracle FP64: rel-L2 {e:.2e}, max-rel {m:.2e}")
    assert e < 1e-4
    assert m <= 1e-4
    with open(prefix + "_pt.mrc", "rb") as f:
        head = np.frombuffer(f.read(1024), np.int32)
    assert list(head[:4]) == [nx, ny, nz, 2]
    assert os.path.getsize(prefix + "_pt.mrc") == 1024 + 4 * nx * ny * nz
