"""The C++ mirror's Microscope and PotentialBuilder surfaces, called with the
reference's signatures (Microscope.h:3-34, PotentialBuilder.h:14-22) by
tests/cpp/lms_check --surface (no GPU: these members are host arithmetic).

Microscope::Es and Microscope::alpha (Microscope.cpp:17-25) are pinned by
closed forms of their special cases rather than by restating them: with
Cs = 0 alpha is -10 pi lambda k^2 defocus and Es is exp(-(pi a k df / 1000)^2)
(lambda cancels); with defocus = 0 alpha is pi Cs lambda^3 k^4 / 2 and Es is
exp(-(pi a Cs lambda^2 k^3 / 1000)^2); alpha is linear in (Cs, defocus), and
the general Es combines the two special arguments before squaring.
"""
import math
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import BOTH_SUITES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "lms_check")


def _run():
    if not os.path.exists(EXE):
        pytest.skip("tests/cpp/lms_check not built (run __graft_entry__.build())")
    r = subprocess.run([EXE, "--surface"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr + r.stdout
    mics, vals, tail = {}, {}, None
    for line in r.stdout.splitlines():
        kv = dict(re.findall(r"(\w+)=([-+0-9.eE]+)", line))
        if line.startswith("mic "):
            mics[int(line.split()[1])] = {k: float(v) for k, v in kv.items()}
        elif line.startswith("val "):
            vals.setdefault(int(line.split()[1]), []).append({k: float(v) for k, v in kv.items()})
        elif line.startswith("fnPotentialBuilder"):
            tail = {k: int(v) for k, v in kv.items()}
    return mics, vals, tail


@BOTH_SUITES
def test_microscope_alpha_and_es_closed_forms(suite):
    mics, vals, _ = _run()
    lam = mics[0]["lambda"]
    # lambda and sigma at 200 kV (closed form of slicelib.cpp:2387-2401, 2795-2806)
    assert lam == pytest.approx(0.025079, abs=1e-6)
    assert mics[0]["sigma"] == pytest.approx(0.728840, abs=1e-6)
    a, cs, df = mics[0]["aperture"], mics[0]["cs"], mics[0]["defocus"]
    assert (mics[1]["cs"], mics[2]["defocus"]) == (0.0, 0.0)
    for full, no_cs, no_df in zip(vals[0], vals[1], vals[2]):
        k = full["k"]
        # Cs = 0: alpha = -10 pi lambda k^2 df, Es = exp(-(pi a k df / 1000)^2)
        assert no_cs["alpha"] == pytest.approx(-10 * math.pi * lam * k * k * df, rel=1e-13, abs=1e-300)
        assert no_cs["Es"] == pytest.approx(math.exp(-(math.pi * a * k * df / 1000) ** 2), rel=1e-12)
        # defocus = 0: alpha = pi Cs lambda^3 k^4 / 2, Es = exp(-(pi a Cs lambda^2 k^3 / 1000)^2)
        assert no_df["alpha"] == pytest.approx(0.5 * math.pi * cs * lam ** 3 * k ** 4, rel=1e-13, abs=1e-300)
        assert no_df["Es"] == pytest.approx(math.exp(-(math.pi * a * cs * lam ** 2 * k ** 3 / 1000) ** 2), rel=1e-12)
        # alpha is linear in (Cs, defocus); Es squares the sum of the two arguments
        assert full["alpha"] == pytest.approx(no_cs["alpha"] + no_df["alpha"], rel=1e-12, abs=1e-12)
        arg = math.sqrt(-math.log(no_cs["Es"])) + math.sqrt(-math.log(no_df["Es"])) if k > 0 else 0.0
        assert full["Es"] == pytest.approx(math.exp(-arg * arg), rel=1e-9, abs=1e-300)
    # k = 0: no phase, full envelope
    assert vals[0][0]["alpha"] == 0.0 and vals[0][0]["Es"] == 1.0


@BOTH_SUITES
def test_microscope_setters_and_defaults(suite):
    mics, vals, _ = _run()
    m = mics[3]  # Microscope(void), then setKeV/setCs/setAperture/setDefocus
    assert (m["keV"], m["cs"], m["aperture"], m["defocus"]) == (300.0, 1.2, 20.0, -60.0)
    assert m["lambda"] == pytest.approx(0.019687, abs=1e-6)
    assert m["sigma"] == pytest.approx(0.652616, abs=1e-6)
    # alpha with negative defocus: both terms positive
    assert all(v["alpha"] >= 0 for v in vals[3])


@BOTH_SUITES
def test_potential_builder_exports_and_empty_constructors(suite):
    _, _, tail = _run()
    # PotentialBuilder.cpp:9-15: nPotentialBuilder starts at 0, fnPotentialBuilder() returns 42
    assert tail["fnPotentialBuilder"] == 42
    assert tail["nPotentialBuilder"] == 1  # the caller incremented the exported variable once
    # ModelPotential(void) / ModelSimulated(void): nothing to compute, the reference's -1
    assert tail["empty_potential"] == -1 and tail["empty_image"] == -1


def _mrc_values(ch):
    # the values tests/cpp/lms_check --mrc writes (5 x 3 x 2 pixels, ch channels)
    i = np.arange(5 * 3 * 2 * ch)
    return (i % 7 - 3) * 37.25 + (i % 5) * 1000.5 * ((i % 3) - 1) + 0.75


@BOTH_SUITES
def test_mrc_modes_round_trip(suite, tmp_path):
    """MRC::save (MRC.cpp:206-345) in every mode of ImageSpace::mrcMode
    (otherEnums.h:21-29): mode word, sizes, cell lengths and angles of the
    model, exactly nx*ny*nz*channels values of the mode's type (unsigned byte,
    short, float, complex short, complex float, unsigned short, RGB bytes),
    integers truncated toward zero and saturated, header statistics of the
    stored values; an unknown mode and a too-small image are refused."""
    if not os.path.exists(EXE):
        pytest.skip("tests/cpp/lms_check not built (run __graft_entry__.build())")
    r = subprocess.run([EXE, "--mrc", str(tmp_path / "t")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr + r.stdout
    types = {0: np.uint8, 1: np.int16, 2: np.float32, 3: np.int16, 4: np.float32, 6: np.uint16, 16: np.uint8}
    chans = {0: 1, 1: 1, 2: 1, 3: 2, 4: 2, 6: 1, 16: 3}
    for mode, dt in types.items():
        raw = (tmp_path / f"t_m{mode}.mrc").read_bytes()
        hi = np.frombuffer(raw[:1024], np.int32)
        hf = np.frombuffer(raw[:1024], np.float32)
        assert tuple(hi[:4]) == (5, 3, 2, mode)
        assert tuple(hi[7:10]) == (5, 3, 2) and tuple(hi[16:19]) == (1, 2, 3)
        assert tuple(hf[10:16]) == (7.5, 6.25, 3.5, 90.0, 95.5, 120.0)
        assert raw[208:212] == b"MAP "
        data = np.frombuffer(raw[1024:], dt)
        assert len(raw) == 1024 + 30 * chans[mode] * np.dtype(dt).itemsize
        v = _mrc_values(chans[mode])
        if np.issubdtype(dt, np.integer):
            info = np.iinfo(dt)
            want = np.trunc(np.clip(v, info.min, info.max)).astype(dt)
        else:
            want = v.astype(np.float32)
        assert np.array_equal(data, want), mode
        q = want.astype(np.float64)
        assert hf[19] == np.float32(q.min()) and hf[20] == np.float32(q.max())
        assert hf[21] == pytest.approx(q.mean(), rel=1e-6)
        assert hf[54] == pytest.approx(q.std(), rel=1e-6)
    assert not (tmp_path / "t_bad.mrc").exists()
