"""CPU tests pinning the oracle's partial-coherence image (Kirkland mulslice
partial coherence, mulslice.cpp:539-676) before it is used as the parity
reference of msx_partial_coherence.

Pins:
  * chi with two- and three-fold astigmatism against the compiled
    reference's multipole chi() (slicelib.cpp:258-318), an independent
    parametrisation of the same function (golden vectors, and live when
    oracle/_ref is built);
  * the illumination-angle set against the Gauss circle count of lattice
    points, in exact integer arithmetic;
  * the defocus steps against the closed-form Gaussian weights;
  * the image: one angle and one defocus give the coherent objective-lens
    image; vacuum gives 1; the defocus integral is the weighted mean of
    single-defocus images; a four-angle ring is the mean of coherent images
    of incident waves built from the reference's freqn positions
    (xpos = i a/(n-1)).
"""
import json
import math
import os

import numpy as np
import pytest

from multislice_amd import specimen
from oracle import Oracle as O
from oracle import ref

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _astig_from_multipole(c12a, c12b, c23a, c23b):
    """(dfa2, dfa2phi, dfa3, dfa3phi) with dfa2 sin(2(phi-phi2)) = C12a cos 2phi + C12b sin 2phi
    and dfa3 sin(3(phi-phi3)) = C23a cos 3phi + C23b sin 3phi."""
    return (math.hypot(c12a, c12b), math.atan2(-c12a, c12b) / 2.0,
            math.hypot(c23a, c23b), math.atan2(-c23a, c23b) / 3.0)


def _check_chi_rows(rows):
    kev = 300.0
    lam = O.wavelength(kev)
    for cs, df, c12a, c12b, c23a, c23b, alx, aly, w in rows:
        a2, p2, a3, p3 = _astig_from_multipole(c12a, c12b, c23a, c23b)
        got = O.chi_astig(alx / lam, aly / lam, kev, cs, df, a2, p2, a3, p3)
        want = 2 * np.pi / lam * w
        # the reference holds its coefficients in float
        scale = 2 * np.pi / lam * (0.25 * abs(cs) * (alx * alx + aly * aly) ** 2 + 0.5 * abs(df) * (alx * alx + aly * aly)
                                   + (abs(c12a) + abs(c12b)) * (alx * alx + aly * aly)
                                   + (abs(c23a) + abs(c23b)) * (alx * alx + aly * aly) ** 1.5) + 1e-30
        assert abs(got - want) <= 2e-7 * scale, (cs, df, c12a, c12b, c23a, c23b, alx, aly, got, want)


def test_chi_astig_matches_reference_multipole_golden():
    with open(os.path.join(GOLD, "slicelib_vectors.json")) as f:
        rows = json.load(f)["chi_multi"]
    assert len(rows) == 32
    _check_chi_rows(rows)


def test_chi_astig_matches_reference_multipole_live():
    R = ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(2)
    rows = []
    for _ in range(200):
        cs, df = rng.choice([0.0, 1e4, 1.2e7]), rng.uniform(-900, 900)
        c = rng.uniform(-2000, 2000, 4)
        alx, aly = rng.uniform(-0.03, 0.03, 2)
        rows.append((cs, df, *c, alx, aly, R.ref_chi_multi(cs, df, *c, alx, aly)))
    _check_chi_rows(rows)


def _lattice_count(r2_lo, r2_hi, sx=1, sy=1):
    """Number of integer (i, j) with r2_lo <= sx i^2 + sy j^2 <= r2_hi."""
    n = int(math.isqrt(int(r2_hi / min(sx, sy)))) + 2
    return sum(1 for i in range(-n, n + 1) for j in range(-n, n + 1) if r2_lo <= sx * i * i + sy * j * j <= r2_hi)


@pytest.mark.parametrize("R_lo,R_hi", [(0.0, 3.3), (1.5, 3.3), (0.0, 0.5), (2.05, 5.71)])
def test_illumination_angles_are_lattice_points_in_the_cone(R_lo, R_hi):
    """Angles iq/a inside the cone acmin..acmax: the Gauss circle count
    (cone radii in units of one grid step, squares kept off integers)."""
    nx = ny = 128
    a = 20.0
    kev = 200.0
    lam = O.wavelength(kev)
    step = lam / a
    iq = O.illum_angles(nx, ny, a, a, kev, R_lo * step, R_hi * step)
    assert len(iq) == _lattice_count(R_lo ** 2, R_hi ** 2)
    r2 = (iq.astype(np.int64) ** 2).sum(1)
    assert np.all(r2 <= R_hi ** 2) and np.all(r2 >= R_lo ** 2)
    assert len({tuple(v) for v in iq}) == len(iq)


def test_illumination_angles_rectangular_cell():
    """ax = 2 by: the cone in lattice units is an ellipse i^2 + 4 j^2 <= R^2 (with R in x steps)."""
    kev, ax, by = 300.0, 24.0, 12.0
    lam = O.wavelength(kev)
    iq = O.illum_angles(64, 32, ax, by, kev, 0.0, 4.4 * lam / ax)
    assert len(iq) == _lattice_count(0, 4.4 ** 2, 1, 4)


def test_defocus_steps_closed_form():
    off, w = O.defocus_steps(30.0, 7.0)
    ndf = int(2.5 * 30.0 / 7.0)
    assert len(off) == 2 * ndf + 1 == 21
    i = np.arange(-ndf, ndf + 1)
    assert np.allclose(off, 7.0 * i, rtol=0, atol=1e-12)
    assert np.allclose(w, np.exp(-0.5 * (7.0 * i / 30.0) ** 2), rtol=1e-15, atol=0)
    off, w = O.defocus_steps(0.0, 5.0)
    assert list(off) == [0.0] and list(w) == [1.0]
    with pytest.raises(ValueError):
        O.defocus_steps(10.0, 0.0)


@pytest.fixture(scope="module")
def si_small():
    at, cell = specimen.si_001((2, 2, 1))
    n, nsl = 64, 4
    V = O.potential(at["Z"], np.stack([at["x"], at["y"], at["z"]], 1), cell, n, n, nsl, 3.0)
    t = O.transmission(V, cell[0], cell[1], 200.0)
    return t, np.full(nsl, cell[2] / nsl), np.arange(nsl), cell


LENS = dict(cs=1.3e7, df=700.0, aobj=0.024)


def test_single_angle_single_defocus_is_the_coherent_image(si_small):
    t, cz, layer, cell = si_small
    kev = 200.0
    img, n = O.partial_image(t, cz, layer, cell[0], cell[1], kev, LENS, 0.0, 1e-6, 0.0, 1.0)
    assert n == 1
    w = O.multislice(t, cz, layer, np.ones(t.shape[1:], np.complex128), cell[0], cell[1], kev)
    ref_img = O.ctf_image(w, cell[0], cell[1], kev, LENS["cs"], LENS["df"], LENS["aobj"])[1]
    assert np.abs(img - ref_img).max() < 1e-12 * np.abs(ref_img).max()


def test_vacuum_is_uniform():
    n = 32
    t = np.ones((1, n, n), np.complex128)
    img, nill = O.partial_image(t, [2.0], [0], 10.0, 10.0, 300.0, dict(LENS, dfa2=40.0, dfa3=500.0),
                                0.0, 1e-6, 25.0, 10.0)
    assert nill == 1
    assert np.abs(img - 1.0).max() < 1e-12


def test_defocus_integral_is_weighted_mean(si_small):
    t, cz, layer, cell = si_small
    kev, sig, dd = 200.0, 40.0, 18.0
    lens = dict(LENS, dfa2=25.0, dfa2phi=0.4, dfa3=300.0, dfa3phi=-0.2)
    step = O.wavelength(kev) / cell[0]
    img, nill = O.partial_image(t, cz, layer, cell[0], cell[1], kev, lens, 0.0, 1.05 * step, sig, dd)
    assert nill == 5
    off, w = O.defocus_steps(sig, dd)
    acc = 0.0
    for o, wt in zip(off, w):
        acc = acc + wt * O.partial_image(t, cz, layer, cell[0], cell[1], kev, dict(lens, df=lens["df"] + o),
                                         0.0, 1.05 * step, 0.0, 1.0)[0]
    acc /= w.sum()
    assert np.abs(img - acc).max() < 1e-12 * np.abs(acc).max()


def _xpos(n, a):
    """Positions of freqn (slicelib.cpp:433-450): the compiled reference when
    built, else the golden vector for this (n, a)."""
    R = ref()
    if R is not None:
        ko = np.zeros(n, np.float32)
        ko2 = np.zeros(n, np.float32)
        xo = np.zeros(n, np.float32)
        R.ref_freqn(ko, ko2, xo, n, a)
        return xo.astype(np.float64)
    with open(os.path.join(GOLD, "slicelib_vectors.json")) as f:
        for c in json.load(f)["freqn"]:
            if c["n"] == n and abs(c["a"] - a) < 1e-12:
                return np.array(c["xo"], np.float64)
    pytest.skip(f"no freqn vector for n={n}, a={a}")


def test_four_angle_ring_uses_reference_positions():
    """acmin/acmax picking exactly (+-1, 0), (0, +-1): the image is the mean of
    coherent images of exp(2 pi i (qx xpos + qy ypos)) with xpos from freqn,
    which spans [0, a] over n-1 steps (so the tilted wave is not periodic)."""
    n, a, kev = 256, 21.724, 200.0
    at, _ = specimen.si_001((4, 4, 1))
    nsl = 2
    V = O.potential(at["Z"], np.stack([at["x"], at["y"], at["z"]], 1), (a, a, 5.431), n, n, nsl, 3.0)
    t = O.transmission(V, a, a, kev)
    cz, layer = np.full(nsl, 5.431 / nsl), np.arange(nsl)
    lens = dict(cs=1.2e7, df=-500.0, aobj=0.03, dfa2=30.0, dfa2phi=0.3)
    step = O.wavelength(kev) / a
    img, nill = O.partial_image(t, cz, layer, a, a, kev, lens, 0.5 * step, 1.2 * step, 0.0, 1.0)
    assert nill == 4
    xo = _xpos(n, a)
    acc = np.zeros((n, n))
    for iqx, iqy in [(0, -1), (-1, 0), (1, 0), (0, 1)]:
        w0 = np.exp(2j * np.pi * (iqx / a * xo[:, None] + iqy / a * xo[None, :]))
        w = O.multislice(t, cz, layer, w0, a, a, kev)
        acc += _astig_image(w, a, kev, lens)
    acc /= 4
    err = np.linalg.norm(img - acc) / np.linalg.norm(acc)
    assert err < 1e-5, err  # float positions of the reference against the exact i a/(n-1)
    # the periodic positions i a/n give a clearly different image
    xp = np.arange(n) * a / n
    acc2 = np.zeros((n, n))
    for iqx, iqy in [(0, -1), (-1, 0), (1, 0), (0, 1)]:
        w0 = np.exp(2j * np.pi * (iqx / a * xp[:, None] + iqy / a * xp[None, :]))
        acc2 += _astig_image(O.multislice(t, cz, layer, w0, a, a, kev), a, kev, lens)
    assert np.linalg.norm(img - acc2 / 4) / np.linalg.norm(acc) > 1e-3


def _astig_image(w, a, kev, lens):
    """|ifft(fft(w) exp(-i chi))|^2 with numpy FFTs (exponent conventions of
    the reference: fft is +i unscaled, ifft is -i scaled) and the oracle's chi."""
    n = w.shape[0]
    lam = O.wavelength(kev)
    W = np.fft.ifft2(w) * (n * n)  # exponent +i, unscaled
    k = np.fft.fftfreq(n, d=a / n)
    ko, ko2, _ = O.freqn(n, a)
    chi = np.array([[O.chi_astig(kx, ky, kev, lens["cs"], lens["df"], lens.get("dfa2", 0.0), lens.get("dfa2phi", 0.0),
                                 lens.get("dfa3", 0.0), lens.get("dfa3phi", 0.0)) for ky in k] for kx in k])
    ap = (ko2[:, None] + ko2[None, :]) <= np.float32((lens["aobj"] / lam) ** 2)
    out = np.fft.fft2(W * np.exp(-1j * chi) * ap) / (n * n)  # exponent -i, scaled
    return np.abs(out) ** 2
