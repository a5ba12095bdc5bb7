"""A full run of the LMultiSlice command-line program (multislice_amd/lmultislice,
Dispatcher::Run, Dispatcher.cu:161-242) on a .coo model and XML command
written here: the potential file <output>_pt.mrc and the image file
<output>.mrc (|psi| of the image wave) against the oracle on the same atoms,
with the LMultiSlice potential formula and table that ModelPotential uses."""
import os
import subprocess

import numpy as np
import pytest

from multislice_amd import specimen
from oracle import Oracle as O
from conftest import max_rel, rel_l2

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "multislice_amd", "lmultislice")

COMMAND = """<?xml version="1.0" ?>
<action>
    <io>
        <fileNameInput>model/si.coo</fileNameInput>
        <fileNameOutput>result/si_128x128</fileNameOutput>
    </io>
    <image>
        <nx>128</nx>
        <ny>128</ny>
       <dpa>11.78</dpa>        <!--- dots per angstrem --->
    </image>
    <slicing>
        <radiuc>3.0</radiuc>
        <numberslices>4</numberslices>
    </slicing>
    <microscope>
        <aperture>20</aperture>
        <cs>1.2</cs>
        <defocus>-60</defocus>     <!---// nm --->
        <keV>200</keV>
    </microscope>
</action>
"""


def read_mrc(path):
    with open(path, "rb") as f:
        raw = f.read()
    nx, ny, nz, mode = np.frombuffer(raw[:16], np.int32)
    assert mode == 2
    return np.frombuffer(raw[1024:], np.float32).reshape(nz, ny, nx)


def test_lmultislice_run_matches_oracle(tmp_path):
    assert os.path.exists(EXE), "multislice_amd/lmultislice not built (run __graft_entry__.build())"
    at, cell = specimen.si_001((2, 2, 1))
    (tmp_path / "model").mkdir()
    (tmp_path / "result").mkdir()
    frac = np.stack([at["x"] / cell[0], at["y"] / cell[1], at["z"] / cell[2]], 1)
    with open(tmp_path / "model" / "si.coo", "w") as f:
        f.write("Si [001] 2x2x1\n")
        f.write(f"{cell[0]:.4f} {cell[1]:.4f} {cell[2]:.4f} 90 90 90 {len(at)}\n")
        for p in frac:
            f.write(f"{p[0]:.6f} {p[1]:.6f} {p[2]:.6f} 14 0 0 1\n")
    (tmp_path / "si.xml").write_text(COMMAND)
    r = subprocess.run([EXE, "si.xml"], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr + r.stdout
    assert "finished successful" in r.stdout
    pot = read_mrc(tmp_path / "result" / "si_128x128_pt.mrc")      # [kz][iy][ix]
    img = read_mrc(tmp_path / "result" / "si_128x128.mrc")[0]      # [iy][ix]
    # the oracle on the model as the program holds it (float32 cell and fractions)
    c32 = np.float32([round(v, 4) for v in cell[:3]]).astype(np.float64)
    f32 = np.float32(np.round(frac, 6)).astype(np.float64)
    xyz = np.stack([f32[:, 0] * np.float32(c32[0]), f32[:, 1] * np.float32(c32[1]),
                    f32[:, 2] * np.float32(c32[2])], 1)
    V = O.potential(at["Z"], xyz, tuple(c32), 128, 128, 4, 3.0, table=0, formula=1)
    e_pot = rel_l2(pot.transpose(0, 2, 1), V)
    t = O.transmission(V, c32[0], c32[1], 200.0)
    w = O.multislice(t, np.full(4, c32[2] / 4), np.arange(4), np.ones((128, 128)), c32[0], c32[1], 200.0)
    img_w, _ = O.ctf_image(w, c32[0], c32[1], 200.0, cs=1.2e7, df=-600.0, aobj=0.020)
    e_img = rel_l2(img.T, np.abs(img_w))
    m_img = max_rel(img.T, np.abs(img_w))
    print(f"lmultislice: potential vs oracle {e_pot:.2e}, |psi| image vs oracle FP64 {e_img:.2e} "
          f"(max-rel {m_img:.2e})")
    assert e_pot < 1e-6  # float32 in the file
    assert e_img < 1e-5
    assert m_img <= 1e-4
