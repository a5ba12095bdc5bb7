"""Musli's STEM image and wave files (msx_musli_stem_write / _wave_write,
after MSProject.cs:697-790 and :522-596).  No C# runtime is here to run the
reference, so the trailer is pinned by .NET's composite-format rules for the
reference's format string (right-aligned widths, general format with the
given significant digits, "n0" for the final 0), worked by hand for these
values; the data part is the raw doubles.  Host code of libmsx.so, run in both the
CPU and the GPU suite."""
import numpy as np

import multislice_amd as msx
from conftest import BOTH_SUITES


@BOTH_SUITES
def test_stem_file(tmp_path, suite):
    img = np.arange(12.0).reshape(3, 4) * 0.125
    p = tmp_path / "a.stem"
    msx.musli_stem_write(p, img, 7.81, 7.8125, 195.25, 300.0)
    raw = p.read_bytes()
    assert len(raw) == 12 * 8 + 160
    assert np.array_equal(np.frombuffer(raw[:96], np.float64).reshape(3, 4), img)
    trailer = raw[96:]
    text = trailer.rstrip(b"\0").decode()
    assert text == ("~~~ " + "        7.81" + " " + "      7.8125" + " " + "      195.25" + " " + "    1.5708" + " "
                    + "    1.5708" + " " + " 1.5707963" + " " + "       300" + " " + "4" + " " + "3" + " "
                    + "    195.25" + " " + "    0" + " ")
    assert set(trailer[len(text):]) == {0}


@BOTH_SUITES
def test_wave_file(tmp_path, suite):
    rng = np.random.default_rng(3)
    w = rng.standard_normal((5, 6)) + 1j * rng.standard_normal((5, 6))
    p = tmp_path / "a.wav"
    msx.musli_wave_write(p, w, 46.86, 23.43, 1.9525, 200.0, gamma=1.2)
    raw = p.read_bytes()
    assert len(raw) == 5 * 6 * 16 + 160
    back = np.frombuffer(raw[:-160], np.float64).view(np.complex128).reshape(5, 6)
    assert np.array_equal(back, w)
    f = raw[-160:].rstrip(b"\0").decode().split()
    assert f[0] == "___" and [float(v) for v in f[1:4]] == [46.86, 23.43, 1.9525]
    assert f[6] == "1.2" and f[7] == "200" and (f[8], f[9]) == ("6", "5") and f[11] == "0"


@BOTH_SUITES
def test_general_format_switches_to_exponent_like_net(tmp_path, suite):
    """G10 of 1.5e-07 and 12345678901 is exponent form with two exponent digits."""
    p = tmp_path / "b.stem"
    msx.musli_stem_write(p, np.zeros((1, 1)), 1.5e-7, 12345678901.0, 1.0, 100.0)
    f = p.read_bytes()[8:].rstrip(b"\0").decode().split()
    assert f[1] == "1.5e-07" and f[2] == "1.23456789e+10"
