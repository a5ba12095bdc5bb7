"""The oracle's oblique-cell slice loop (Musli ApplyPropagator,
Propagator.cpp:55-107, in Kirkland's sign convention), pinned without
restating it:

* plane waves: in vacuum the grid's plane wave of frequency indices (i, j)
  is an eigenvector of one slice, with eigenvalue exp(-i pi lambda dz |g|^2)
  where |g|^2 comes from the reciprocal basis of the oblique cell (a along x,
  b at angle gam), computed here by inverting the real-space basis matrix
  (not from Musli's s formula), and 0 outside the disk |g| < bw min(N/2a, M/2b);
* a right angle reduces it to the orthogonal slice loop (mulslice.cpp:
  703-739), the independent Kirkland restatement;
* the Fresnel semigroup: two vacuum slices dz1, dz2 (tilted beam included)
  equal one slice dz1 + dz2.
"""
import numpy as np
import pytest

from oracle import Oracle as O

KEV = 200.0


def _g2(nx, ny, ax, by, gam):
    """|g|^2 on the grid from the reciprocal basis (rows of inv(B)^T)."""
    B = np.array([[ax, 0.0], [by * np.cos(gam), by * np.sin(gam)]])  # rows: supercell vectors a, b
    R = np.linalg.inv(B).T                                            # rows: a*, b*
    fx = np.fft.fftfreq(nx, 1.0 / nx)
    fy = np.fft.fftfreq(ny, 1.0 / ny)
    gx = fx[:, None, None] * R[0] + fy[None, :, None] * R[1]
    return (gx ** 2).sum(-1), fx, fy


@pytest.mark.parametrize("gam", [np.deg2rad(70.0), np.deg2rad(104.0)])
def test_plane_waves_are_eigenvectors_with_the_oblique_metric(gam):
    nx, ny, ax, by, dz = 32, 24, 9.0, 7.5, 2.3
    g2, fx, fy = _g2(nx, ny, ax, by, gam)
    lam = O.wavelength(KEV)
    kmax = (2.0 / 3.0) * min(nx / (2 * ax), ny / (2 * by))
    t = np.ones((1, nx, ny), np.complex128)
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    rng = np.random.default_rng(0)
    for _ in range(12):
        i, j = int(rng.integers(nx)), int(rng.integers(ny))
        # real-space eigenvector of cfpix::fft (exponent +i) for index (i, j)
        w = np.exp(-2j * np.pi * (i * ix / nx + j * iy / ny))
        out = O.multislice_oblique(t, [dz], [0], w, ax, by, gam, KEV)
        lamb = np.exp(-1j * np.pi * lam * dz * g2[i, j]) if g2[i, j] < kmax ** 2 else 0.0
        assert np.abs(out - lamb * w).max() < 1e-11, (i, j)


def test_right_angle_reduces_to_the_orthogonal_loop():
    nx, ny, ax, by = 48, 32, 10.0, 7.0
    rng = np.random.default_rng(1)
    t = np.exp(1j * rng.random((2, nx, ny)))
    w = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    cz, layer = [1.7, 2.1], [0, 1, 0, 1, 1]
    tilt = (0.004, -0.002)
    a = O.multislice_oblique(t, cz, layer, w, ax, by, np.pi / 2, KEV, tilt)
    b = O.multislice(t, cz, layer, w, ax, by, KEV, tilt)
    assert np.abs(a - b).max() / np.abs(b).max() < 1e-12


@pytest.mark.parametrize("gam", [np.deg2rad(60.0), np.deg2rad(115.0)])
def test_fresnel_semigroup(gam):
    nx, ny, ax, by = 40, 40, 8.0, 8.5
    rng = np.random.default_rng(2)
    w = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    t = np.ones((1, nx, ny), np.complex128)
    tilt = (0.003, 0.006)
    two = O.multislice_oblique(t, [1.3], [0], O.multislice_oblique(t, [2.2], [0], w, ax, by, gam, KEV, tilt),
                               ax, by, gam, KEV, tilt)
    one = O.multislice_oblique(t, [3.5], [0], w, ax, by, gam, KEV, tilt)
    assert np.abs(two - one).max() / np.abs(one).max() < 1e-12
    # the oblique disk differs from the orthogonal one
    orth = O.multislice(t, [3.5], [0], w, ax, by, KEV, tilt)
    assert np.abs(one - orth).max() / np.abs(orth).max() > 1e-3
