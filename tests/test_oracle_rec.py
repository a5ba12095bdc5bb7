"""Pins the oracle's reciprocal-space potential (Muslib Potential_rec_DT_DW,
Pot_rec_DT_DW.cpp:3-132) and its Doyle-Turner table on the CPU.

Reciprocal space -- checks: table rows against the reference bank (fixture); a single atom at the
origin gives the real, radial f_e(|g|/2) exp(-B|g|^2/4)/area at every kept
reflection and 0 on the Nyquist row and column; the translation theorem;
Hermitian symmetry without absorption; the oblique reciprocal metric against
explicit reciprocal basis vectors; absorption as the factor (1 + i ads); the
supercell multiplier against the same structure written as a doubled cell;
and the Fourier sum against the closed-form real-space Gaussian potential
(Curves.cpp:12-21) summed over periodic images.

Real space (Potential_rs_DT_DW_cutpot / _periodic, Pot_RS.cpp:111-486) --
checks: the pixel integral of one atom equals the forward scattering sum_j a_j;
cut and periodic forms agree for boxes inside the cell; the box rule against
explicit pixel-interval tests; and the periodic form against the independent
reciprocal-space Fourier series of the same structure.
"""
import json
import os

import numpy as np
import pytest

from oracle import Oracle as O
from oracle.wrap import lib

HERE = os.path.dirname(os.path.abspath(__file__))


def fe_ref(Z, s, dw=0.0):
    return O.dt_fe(Z, s) * np.exp(-dw * s * s)


def test_table_rows_match_reference_bank():
    rows = json.load(open(os.path.join(HERE, "golden", "dt_rows.json")))["rows"]
    for z, vals in rows.items():
        z = int(z)
        for s in (0.0, 0.3, 1.1):
            expect = sum(vals[2 * j] * np.exp(-vals[2 * j + 1] * s * s) for j in range(4))
            assert O.dt_fe(z, s) == pytest.approx(expect, rel=1e-15, abs=1e-15)


def _kept(n, mult):
    """signed index for each grid index, or None where the coefficient stays 0"""
    out = []
    for i in range(n):
        if i < n // 2:
            out.append(i // mult if i % mult == 0 else None)
        else:
            j = n - i
            out.append(-(j // mult) if (i > n // 2 and j % mult == 0) else None)
    return out


def test_single_atom_at_origin():
    a, b, N, M, dw = 9.0, 7.0, 16, 12, 0.6
    g = O.potential_rec_g([29], [0.0], [0.0], a, b, N, M, dw=[dw])
    hs, ks = _kept(N, 1), _kept(M, 1)
    for row in range(M):
        for col in range(N):
            h, k = hs[col], ks[row]
            if h is None or k is None:
                assert g[row, col] == 0
                continue
            s = 0.5 * np.hypot(h / a, k / b)
            assert g[row, col].real == pytest.approx(fe_ref(29, s, dw) / (a * b), rel=1e-13)
            assert abs(g[row, col].imag) <= 1e-15


def test_translation_theorem_and_hermitian():
    a, b, N, M = 8.0, 8.0, 16, 16
    x0, y0 = 0.3125, 0.7
    g0 = O.potential_rec_g([14], [0.0], [0.0], a, b, N, M, dw=[0.4])
    g1 = O.potential_rec_g([14], [x0], [y0], a, b, N, M, dw=[0.4])
    hs, ks = np.array([h if h is not None else 0 for h in _kept(N, 1)]), \
        np.array([k if k is not None else 0 for k in _kept(M, 1)])
    phase = np.exp(-2j * np.pi * (hs[None, :] * x0 + ks[:, None] * y0))
    np.testing.assert_allclose(g1, g0 * phase, rtol=0, atol=1e-14)
    # V(-g) = conj V(g) without absorption: the real-space potential is real
    gm = g1[(-np.arange(M)) % M][:, (-np.arange(N)) % N]
    np.testing.assert_allclose(gm, np.conj(g1), rtol=0, atol=1e-15)


def test_oblique_metric():
    a, b, gam, N, M = 6.0, 5.0, np.deg2rad(105.0), 12, 12
    g = O.potential_rec_g([8], [0.0], [0.0], a, b, N, M, gam=gam)
    # direct basis a = (a, 0), b = (b cos gam, b sin gam); reciprocal basis = inv(D)^T
    D = np.array([[a, 0.0], [b * np.cos(gam), b * np.sin(gam)]])
    R = np.linalg.inv(D).T
    area = abs(np.linalg.det(D))
    hs, ks = _kept(N, 1), _kept(M, 1)
    for row in range(M):
        for col in range(N):
            h, k = hs[col], ks[row]
            if h is None or k is None:
                continue
            gv = h * R[0] + k * R[1]
            assert g[row, col].real == pytest.approx(fe_ref(8, 0.5 * np.linalg.norm(gv)) / area, rel=1e-12)


def test_absorption_factor():
    a, N = 10.0, 16
    g0 = O.potential_rec_g([22, 22], [0.1, 0.6], [0.2, 0.45], a, a, N, N)
    g1 = O.potential_rec_g([22, 22], [0.1, 0.6], [0.2, 0.45], a, a, N, N, ads=[0.08, 0.08])
    np.testing.assert_allclose(g1, g0 * (1 + 0.08j), rtol=0, atol=1e-15)


def test_supercell_multiplier():
    # nmult = 2 on a cell a: the same as the doubled cell 2a holding two copies
    a, b, N, M = 5.0, 6.0, 32, 16
    x, y = 0.23, 0.61
    g2 = O.potential_rec_g([38], [x], [y], a, b, N, M, nmult=2)
    gd = O.potential_rec_g([38, 38], [x / 2, x / 2 + 0.5], [y, y], 2 * a, b, N, M)
    np.testing.assert_allclose(g2, gd, rtol=0, atol=1e-14)


def test_fourier_sum_equals_real_space_gaussians():
    # with B = 1.5 A^2 the coefficients beyond the kept band are below 1e-9 of
    # the largest, so the truncated Fourier sum equals the periodic sum of the
    # real-space potential 4 pi sum_j a_j/(b_j+B) exp(-4 pi^2 r^2/(b_j+B))
    a, b, N, M, dw = 6.0, 5.0, 96, 80, 1.5
    x0, y0 = 0.37, 0.58
    g = O.potential_rec_g([79], [x0], [y0], a, b, N, M, dw=[dw])
    v = np.fft.ifft2(g) * (N * M)  # sum_g V(g) exp(+2 pi i g.r); rows = y
    assert np.abs(v.imag).max() < 1e-12 * np.abs(v.real).max()
    rng = np.random.default_rng(0)
    for _ in range(12):
        ix, iy = int(rng.integers(N)), int(rng.integers(M))
        rx, ry = ix * a / N, iy * b / M
        tot = 0.0
        for mx in range(-3, 4):
            for my in range(-3, 4):
                d = np.hypot(rx - (x0 + mx) * a, ry - (y0 + my) * b)
                tot += O.dt_rs(79, d, dw)
        assert v.real[iy, ix] == pytest.approx(tot, rel=1e-8, abs=1e-9 * abs(v.real).max())


def test_exported():
    for name in ("orc_dt_fe", "orc_dt_rs", "orc_potential_rec_dt"):
        assert hasattr(lib(), name)


# ---------------------------------------------------------------- real space
def test_rs_integral_is_forward_scattering():
    # integral of 4 pi a/(b+B) exp(-4 pi^2 r^2/(b+B)) over the plane is a: the
    # pixel sum of one atom's potential times the pixel area is sum_j a_j
    a, N = 24.0, 192
    for Z, dw in ((6, 0.5), (79, 1.0)):
        g = O.potential_rs_g([Z], [0.5], [0.5], a, a, N, N, atrad=11.9, periodic=False, dw=[dw])
        tot = g.real.sum() * (a / N) ** 2
        p = [v for v in json.load(open(os.path.join(HERE, "golden", "dt_rows.json")))["rows"][str(Z)]]
        assert tot == pytest.approx(sum(p[2 * j] for j in range(4) if p[2 * j + 1] > 0), rel=2e-5)


def test_rs_cut_equals_periodic_inside_the_cell():
    a, b, N, M = 10.0, 9.0, 64, 48
    rng = np.random.default_rng(1)
    xf, yf = rng.uniform(0.35, 0.65, 5), rng.uniform(0.35, 0.65, 5)
    Z = [6, 8, 14, 38, 79]
    gc = O.potential_rs_g(Z, xf, yf, a, b, N, M, atrad=2.5, periodic=False, dw=np.full(5, 0.4))
    gp = O.potential_rs_g(Z, xf, yf, a, b, N, M, atrad=2.5, periodic=True, dw=np.full(5, 0.4))
    np.testing.assert_array_equal(gc, gp)


def test_rs_box_rule_against_interval_tests():
    # the cut box of one atom written as explicit pixel-interval tests
    a, b, N, M, atrad = 8.0, 7.0, 40, 36, 2.2
    xf, yf = 0.06, 0.93
    g = O.potential_rs_g([22], [xf], [yf], a, b, N, M, atrad=atrad, periodic=False)
    ibb, iee = int((xf - atrad / a) * N), int((xf + atrad / a) * N)   # int() truncates toward zero
    jbb, jee = int((yf - atrad / b) * M), int((yf + atrad / b) * M)
    for j in range(M):
        for i in range(N):
            inside = max(ibb, 0) <= i < min(iee, N) and max(jbb, 0) <= j < min(jee, M)
            if not inside:
                assert g[j, i] == 0
            else:
                x, y = i * a / N - xf * a, j * b / M - yf * b
                assert g[j, i].real == pytest.approx(O.dt_rs(22, np.hypot(x, y)), rel=1e-12)


def test_rs_periodic_equals_reciprocal_space():
    # two independent forms of one periodic potential: Gaussians summed over
    # the nine nearest images (boxes covering the whole cell) and the Fourier
    # series of the Doyle-Turner factors (N = 192: the coefficients beyond the
    # kept band are below exp(-(0.31 + B) 8^2/4) = 6e-13 of the largest)
    a, N, dw = 12.0, 192, 1.5
    rng = np.random.default_rng(2)
    xf, yf = rng.random(4), rng.random(4)
    Z = [6, 14, 38, 79]
    grs = O.potential_rs_g(Z, xf, yf, a, a, N, N, atrad=5.99, periodic=True, dw=np.full(4, dw))
    grec = O.potential_rec_g(Z, xf, yf, a, a, N, N, dw=np.full(4, dw))
    vrec = (np.fft.ifft2(grec) * N * N).real
    err = np.abs(grs.real - vrec).max() / np.abs(vrec).max()
    assert err < 1e-6, err


# ---------------------------------------------------------------- tabulated curves
def test_rec_curve_from_fine_dt_table_matches_dt():
    # a curve table sampled from the Doyle-Turner formula (CreatePots_rec_DT_DW)
    # reproduces the formula up to the linear-interpolation error h^2 f''/8
    a, b, N, M = 8.0, 7.0, 32, 32
    Z, dws = [14, 79], [0.4, 0.7]
    xf, yf = [0.1, 0.55], [0.3, 0.8]
    smax = 0.5 * np.hypot(N / 2 / a, M / 2 / b) * 1.05
    errs = []
    for npts in (4000, 400):
        dds = smax / (npts - 2)
        curves = O.dt_curves_rec(Z, dws, npts, dds, a * b)
        gc = O.potential_rec_curve_g([0, 1], xf, yf, a, b, N, M, curves, dds)
        gd = O.potential_rec_g(Z, xf, yf, a, b, N, M, dw=dws)
        errs.append(np.abs(gc - gd).max() / np.abs(gd).max())
    # second-order convergence: 10x the points, ~100x smaller error
    assert errs[0] < 1e-5 and 40 < errs[1] / errs[0] < 250, errs


def test_rs_curve_from_fine_dt_table_matches_dt():
    a, b, N, M, atrad = 9.0, 8.0, 48, 40, 2.5
    Z, dws = [8, 38], [0.5, 0.3]
    xf, yf = [0.05, 0.6], [0.9, 0.4]
    rmax = atrad * np.sqrt(2) * 1.05
    errs = []
    for npts in (8000, 800):
        dds = rmax / (npts - 2)
        curves = O.dt_curves_rs(Z, dws, npts, dds)
        gc = O.potential_rs_curve_g([0, 1], xf, yf, a, b, N, M, atrad, curves, dds)
        gd = O.potential_rs_g(Z, xf, yf, a, b, N, M, atrad, dw=dws)
        errs.append(np.abs(gc - gd).max() / np.abs(gd).max())
    assert errs[0] < 1e-5 and 40 < errs[1] / errs[0] < 250, errs


def test_curve_past_table_end_contributes_nothing():
    # a table shorter than the reach: the reflections / pixels past the last
    # point get nothing (Pot_rec_curve_DW.cpp:40, Pot_RS.cpp:41)
    a, N = 8.0, 32
    curves = np.ones((1, 5))
    g = O.potential_rec_curve_g([0], [0.0], [0.0], a, a, N, N, curves, dds=0.1)
    s = 0.5 * np.hypot(*np.meshgrid(np.fft.fftfreq(N, a / N), np.fft.fftfreq(N, a / N)))
    assert np.all(g[s / 0.1 >= 4] == 0) and np.all(g[(s / 0.1 < 4) & (s > 0)].real == 1.0)
