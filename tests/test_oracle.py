"""CPU tests pinning the oracle (oracle/msx_oracle.cpp) before it is trusted
as the parity reference of the HIP engine.

Pins, per oracle function:
  * the reference's own slicelib.cpp and FFTW 2.1.5, compiled into oracle/_ref
    (live when built) and as committed golden vectors (tests/golden/);
  * closed forms: lambda and sigma at 100/200/300 kV, scipy's K0/I0, the DFT
    definition, the Hankel-transform identity linking vzatom to featom;
  * invariants and brute force for the potential splat, the propagator,
    the slice loop, the probe and the detectors.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest
from scipy import integrate, special

from oracle import Oracle as O
from oracle import lib, ref, ref_fftw2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _gold(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# ------------------------------------------------------------------ scalars
def test_wavelength_sigma_closed_form():
    # lambda = hc / sqrt(E(2 m0c^2 + E)); sigma = 2 pi (m0c^2+E)/(2m0c^2+E)/(lambda E)
    for kev, lam, sig in [(100, 0.037014, 0.924396), (200, 0.025079, 0.728840), (300, 0.019687, 0.652616)]:
        assert abs(O.wavelength(kev) - lam) < 5e-7
        assert abs(O.sigma(kev) - sig) < 5e-7


def test_scalars_match_reference_golden():
    g = _gold("slicelib_vectors.json")
    for k, v in g["wavelength"]:
        assert O.wavelength(k) == v
    for k, v in g["sigma"]:
        assert O.sigma(k) == v
    L = lib()
    for x, v in g["bessi0"]:
        assert L.orc_bessi0(x) == v
    for x, v in g["bessk0"]:
        assert L.orc_bessk0(x) == v


def test_bessel_against_scipy():
    # Abramowitz & Stegun 9.8.1-9.8.6 polynomial accuracy (|eps| < 2e-7 relative)
    L = lib()
    for x in np.concatenate([np.geomspace(1e-3, 2, 40), np.linspace(2.01, 30, 40)]):
        assert abs(L.orc_bessk0(x) / special.k0(x) - 1) < 3e-7
    for x in np.linspace(0.0, 30, 61):
        assert abs(L.orc_bessi0(x) / special.i0(x) - 1) < 3e-7


def test_vzatom_featom_match_reference_golden():
    g = _gold("slicelib_vectors.json")
    for z, r, v in g["vzatom"]:
        assert O.vzatom(z, r, table=1, formula=0) == v
    for z, k2, v in g["featom"]:
        assert O.featom(z, k2, table=1) == v


def test_vzatom_featom_match_reference_live():
    R = ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(0)
    for z in range(1, 104):
        for r in np.concatenate([rng.random(5) * 5, [1e-12, 0.02]]):
            assert O.vzatom(z, r, 1, 0) == R.ref_vzatom(z, r)
        for k2 in rng.random(3) * 9:
            assert O.featom(z, k2, 1) == R.ref_featom(z, k2)


@pytest.mark.parametrize("Z", [1, 6, 14, 38, 79])
@pytest.mark.parametrize("table", [0, 1])
def test_vzatom_hankel_identity(Z, table):
    """2 pi int_0^inf vz(r) J0(2 pi k r) r dr = (al/2pi) f_e(k^2): the Kirkland
    formula is the projected Fourier pair of featom (pins al, ag = al/2 and the
    exp(-pi^2 r^2/d) Gaussian).  The LMultiSlice variant (exp(-2 pi r^2/d),
    ModelPotential.cu:247-251) does not satisfy it (quirk Q2)."""
    al = 300.8242834
    for k in [0.05, 0.3, 1.0, 2.0]:
        f = lambda r: O.vzatom(Z, r, table, 0) * special.j0(2 * np.pi * k * r) * r
        val = 2 * np.pi * sum(integrate.quad(f, a, b, limit=400)[0]
                              for a, b in [(0, 0.05), (0.05, 0.5), (0.5, 3), (3, 12), (12, 40)])
        expect = al / (2 * np.pi) * O.featom(Z, k * k, table)
        assert abs(val / expect - 1) < 2e-5, (k, val, expect)
    # the LMultiSlice Gaussian form breaks the identity for an element with Gaussian weight
    k = 1.0
    f = lambda r: O.vzatom(Z, r, table, 1) * special.j0(2 * np.pi * k * r) * r
    val = 2 * np.pi * sum(integrate.quad(f, a, b, limit=400)[0]
                          for a, b in [(0, 0.05), (0.05, 0.5), (0.5, 3), (3, 12), (12, 40)])
    assert abs(val / (al / (2 * np.pi) * O.featom(Z, 1.0, table)) - 1) > 1e-3


def test_vzatom_lmultislice_formula_direct():
    """ModelPotential.cu:232-254 evaluated term by term in numpy (scipy K0 gives
    the same to the A&S accuracy)."""
    p = np.array([lib().orc_fparam(0, 14, i) for i in range(12)])
    for r in [0.01, 0.2, 1.1]:
        x = 6.2831853071796 * r
        suml = sum(p[2 * k] * special.k0(x * np.sqrt(p[2 * k + 1])) for k in range(3)) * 300.73079394295
        sumg = sum((p[6 + 2 * k] / p[7 + 2 * k]) * np.exp(-(6.2831853071796 * r * r) / p[7 + 2 * k])
                   for k in range(3)) * 150.36539697148
        assert abs(O.vzatom(14, r, 0, 1) / (suml + sumg) - 1) < 3e-7


def test_freqn_match_reference_golden():
    g = _gold("slicelib_vectors.json")
    for case in g["freqn"]:
        ko, ko2, xo = O.freqn(case["n"], case["a"])
        assert np.array_equal(ko, np.float32(case["ko"]))
        assert np.array_equal(ko2, np.float32(case["ko2"]))
        assert np.array_equal(xo, np.float32(case["xo"]))


def test_chi_closed_form_matches_reference_golden():
    """slicelib chi() with Cs, defocus only: ((al^2 Cs/4) - df/2) al^2; the
    oracle's phase is (2 pi/lambda) times it (chi_k2 in msx_oracle.cpp)."""
    g = _gold("slicelib_vectors.json")
    for cs, df, ax, ay, v in g["chi"]:
        a2 = ax * ax + ay * ay
        # the reference stores Cs, df as float params
        csf, dff = float(np.float32(cs)), float(np.float32(df))
        assert abs(((a2 * csf * 0.25) - 0.5 * dff) * a2 - v) <= 1e-12 * max(1.0, abs(v))


def test_fparams_tables_match_reference_rows():
    g = _gold("fparams_rows.json")["rows"]
    for table, name in [(0, "lmultislice"), (1, "kirkland")]:
        for z, row in g[name].items():
            got = [lib().orc_fparam(table, int(z), i) for i in range(12)]
            assert got == row, (name, z)


# ------------------------------------------------------------------ FFT
def test_fft_matches_dft_and_numpy():
    rng = np.random.default_rng(1)
    for nx, ny in [(6, 10), (8, 8), (9, 12), (5, 7)]:
        x = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
        for s in (+1, -1):
            d = O.dft2(x, s)
            assert np.abs(O.fft2(x, s) - d).max() < 1e-12
            ref_np = np.fft.ifft2(x) * x.size if s > 0 else np.fft.fft2(x)
            assert np.abs(d - ref_np).max() < 1e-11


def test_fft_matches_reference_fftw2_golden():
    g = _gold("slicelib_vectors.json")["fftw2_plus"]
    x = (np.array(g["in_re"]) + 1j * np.array(g["in_im"])).astype(np.complex64)
    want = np.array(g["out_re"]) + 1j * np.array(g["out_im"])
    got = O.fft2(x, +1, f32=True)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 3e-7


def test_fft_matches_reference_fftw2_live():
    F = ref_fftw2()
    if F is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(2)
    for nx, ny, s in [(64, 48, 1), (128, 128, -1), (384, 32, 1)]:
        x = (rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))).astype(np.complex64)
        y = x.copy()
        F.ref_fftw2_2d(y.view(np.float32).reshape(-1), nx, ny, s)
        o = O.fft2(x.astype(np.complex128), s)
        assert np.linalg.norm(y - o) / np.linalg.norm(o) < 3e-7


# ------------------------------------------------------------------ potential
def _brute_potential(Z, xyz, cell, nx, ny, nsl, radius, table, formula):
    a, b, c = cell
    dz = c / nsl
    V = np.zeros((nsl, nx, ny))
    X = np.arange(nx)[:, None] * (a / nx)
    Y = np.arange(ny)[None, :] * (b / ny)
    for z, (x, y, zz) in zip(Z, xyz):
        kz = int(np.floor(zz / dz))
        # nearest periodic image per axis
        dX = x - X
        dX -= a * np.round(dX / a)
        dY = y - Y
        dY -= b * np.round(dY / b)
        r = np.sqrt(dX ** 2 + dY ** 2)
        v = np.vectorize(lambda rr: O.vzatom(z, rr, table, formula))(r)
        V[kz] += np.where(r < radius, v, 0.0)
    return V


def test_potential_brute_force():
    rng = np.random.default_rng(3)
    cell = (7.1, 6.3, 4.0)
    Z = np.array([6, 14, 79, 8, 26])
    xyz = rng.random((5, 3)) * np.array(cell)
    for table, formula in [(0, 0), (1, 0), (0, 1)]:
        V = O.potential(Z, xyz, cell, 24, 20, 2, 2.5, table=table, formula=formula)
        B = _brute_potential(Z, xyz, cell, 24, 20, 2, 2.5, table, formula)
        assert np.abs(V - B).max() <= 1e-9 * np.abs(B).max()


def test_potential_additivity_and_translation():
    rng = np.random.default_rng(4)
    cell = (8.0, 8.0, 2.0)
    Z = np.array([14, 38])
    xyz = np.array([[1.0, 2.0, 0.5], [5.5, 6.25, 0.7]])
    V = O.potential(Z, xyz, cell, 32, 32, 1, 3.0)
    V1 = O.potential(Z[:1], xyz[:1], cell, 32, 32, 1, 3.0)
    V2 = O.potential(Z[1:], xyz[1:], cell, 32, 32, 1, 3.0)
    assert np.abs(V - V1 - V2).max() < 1e-9 * V.max()
    # shifting the atoms by whole pixels rolls the image exactly
    sh = xyz + np.array([3 * 0.25, 5 * 0.25, 0.0])
    Vs = O.potential(Z, sh, cell, 32, 32, 1, 3.0)
    assert np.abs(np.roll(V, (3, 5), axis=(1, 2)) - Vs).max() < 1e-9 * V.max()


# ------------------------------------------------------------------ propagation
def _k2(n, a):
    return O.freqn(n, a)[1]


def test_propagator_semigroup_and_edges():
    n, a = 64, 12.0
    p1x, p1y = O.propagator(n, n, a, a, 200.0, 1.3)
    p2x, p2y = O.propagator(n, n, a, a, 200.0, 2.1)
    p3x, p3y = O.propagator(n, n, a, a, 200.0, 3.4)
    assert np.abs(p1x * p2x - p3x).max() < 1e-12
    assert np.abs(p1y * p2y - p3y).max() < 1e-12
    lam = O.wavelength(200.0)
    k = np.array([(i if i <= n // 2 else i - n) / a for i in range(n)])
    assert np.abs(p1x - np.exp(-1j * np.pi * 1.3 * lam * k * k)).max() < 1e-12


def test_bandlimit_edge_is_strict():
    """slicelib.cpp:755-778 keeps k2 < k2max: a pixel sitting exactly on
    k2max is removed.  With bw = 1/2, n = 64, a = 16: k2max = 1.0 exactly
    and bin (16, 0) has kx2 + ky2 = 1.0 exactly."""
    n, a = 64, 16.0
    k2max = O.k2max(n, n, a, a, 0.5)
    assert k2max == 1.0
    kx2 = _k2(n, a)
    assert kx2[16] == 1.0
    K2 = (kx2[:, None] + kx2[None, :]).astype(np.float32)
    t = np.ones((1, n, n), np.complex128)
    rng = np.random.default_rng(5)
    psi = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    out = O.multislice(t, [0.0], [0], psi, a, a, 200.0, bw=0.5)
    spec = O.fft2(out, +1)
    assert np.abs(spec[K2 >= k2max]).max() < 1e-9
    assert abs(spec[16, 0]) < 1e-9 and abs(spec[15, 0]) > 1e-3
    assert np.abs(spec[K2 < k2max]).min() > 0


def test_vacuum_identity_and_conservation():
    n, a = 64, 10.0
    rng = np.random.default_rng(6)
    psi = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    t = np.ones((1, n, n), np.complex128)
    once = O.multislice(t, [0.0], [0], psi, a, a, 300.0)  # band-limited copy
    fwd = O.multislice(t, [2.5], [0], once, a, a, 300.0)
    back = O.multislice(t, [-2.5], [0], fwd, a, a, 300.0)
    assert np.abs(back - once).max() < 1e-12
    assert abs(np.vdot(fwd, fwd).real / np.vdot(once, once).real - 1) < 1e-12


def test_weak_phase_limit():
    """One thin slice with a tiny potential: psi ~ 1 + i sigma V (band limited)."""
    n, a, kev = 64, 10.0, 200.0
    rng = np.random.default_rng(7)
    V = rng.standard_normal((1, n, n)) * 1e-3
    t = O.transmission(V, a, a, kev, bw=0.0)
    out = O.multislice(t, [0.0], [0], np.ones((n, n)), a, a, kev, bw=10.0)
    s = O.sigma(kev) / 1000.0
    assert np.abs(out - (1 + 1j * s * V[0])).max() < (s * 1e-3 * 5) ** 2 * 5


def test_translation_covariance():
    n, a = 64, 12.8
    rng = np.random.default_rng(8)
    V = rng.random((2, n, n)) * 30
    t = O.transmission(V, a, a, 200.0)
    psi0 = np.ones((n, n))
    out = O.multislice(t, [1.0, 2.0], [0, 1, 0, 1], psi0, a, a, 200.0)
    ts = np.roll(t, (3, -5), axis=(1, 2))
    outs = O.multislice(ts, [1.0, 2.0], [0, 1, 0, 1], psi0, a, a, 200.0)
    assert np.abs(np.roll(out, (3, -5), axis=(0, 1)) - outs).max() < 1e-12


def test_f32_oracle_close_to_f64():
    n, a = 128, 15.0
    rng = np.random.default_rng(9)
    V = rng.random((3, n, n)) * 40
    t64 = O.transmission(V, a, a, 200.0)
    t32 = O.transmission(V.astype(np.float32), a, a, 200.0, f32=True)
    assert np.linalg.norm(t32 - t64) / np.linalg.norm(t64) < 1e-6
    layer = np.arange(30) % 3
    w64 = O.multislice(t64, [1.2, 1.3, 1.4], layer, np.ones((n, n)), a, a, 200.0)
    w32 = O.multislice(t32, [1.2, 1.3, 1.4], layer, np.ones((n, n), np.complex64), a, a, 200.0, f32=True)
    assert np.linalg.norm(w32 - w64) / np.linalg.norm(w64) < 5e-6


# ------------------------------------------------------------------ probe / detector
def test_probe_normalised_and_shift():
    n, a, kev = 64, 12.8, 300.0
    p = O.probe(n, n, a, a, kev, 0.025, cs=1e4, df=50.0, pos=(3.2, 4.0))
    assert abs(np.vdot(p, p).real - 1) < 1e-12
    dx = a / n
    q = O.probe(n, n, a, a, kev, 0.025, cs=1e4, df=50.0, pos=(3.2 + 5 * dx, 4.0 - 2 * dx))
    assert np.abs(np.roll(p, (5, -2), axis=(0, 1)) - q).max() < 1e-12
    # aberration-free probe is centred on its position (intensity maximum)
    c = O.probe(n, n, a, a, kev, 0.025, pos=(10 * dx, 20 * dx))
    assert np.unravel_index(np.argmax(np.abs(c)), c.shape) == (10, 20)


def test_detector_partition_sums_to_total():
    n, a, kev = 64, 12.8, 300.0
    kbw = np.sqrt(O.k2max(n, n, a, a))
    t = np.exp(1j * np.random.default_rng(10).random((2, n, n)))
    dets = [(0.0, 0.3001), (0.3001, 0.9001), (0.9001, kbw + 1), (0.0, kbw + 1)]
    out = O.stem(t, [1.0, 1.0], [0, 1, 0], a, a, kev, 0.02, [(1.0, 2.0), (5.0, 5.0)], dets, [1, 3])
    assert np.allclose(out[:, :3].sum(1), out[:, 3], rtol=1e-12)
    # vacuum-like phase objects conserve intensity up to the band-limit loss
    assert (out[:, 3] <= 1 + 1e-12).all() and (out[:, 3] > 0.5).all()


def test_ctf_image_closed_form():
    """A single plane-wave beam g through the lens picks up exp(-i chi(g))."""
    n, a, kev = 64, 12.8, 200.0
    lam = O.wavelength(kev)
    ix, iy = 3, -2
    x = np.arange(n)[:, None] * (a / n)
    y = np.arange(n)[None, :] * (a / n)
    gx, gy = ix / a, iy / a
    # Kirkland's fft has exponent +i: exp(-2 pi i g.r) lands on bin (ix, iy)
    psi = np.exp(-2j * np.pi * (gx * x + gy * y))
    img_w, img = O.ctf_image(psi, a, a, kev, cs=1.2e7, df=-600.0, aobj=0.02)
    k2 = gx * gx + gy * gy
    chi = np.pi * lam * k2 * (0.5 * 1.2e7 * lam * lam * k2 + 600.0)
    assert np.abs(img_w - psi * np.exp(-1j * chi)).max() < 1e-12
    assert np.abs(img - 1).max() < 1e-12
