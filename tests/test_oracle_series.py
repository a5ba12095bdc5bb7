"""The oracle's thickness series (mulslice.cpp:682-737: beam values and the
integrated intensity after every slice), pinned without restating it:

* beams: the value recorded after slice s equals numpy's inverse FFT
  (exponent +i, times 1/(nx ny): cfpix::fft's transform and the scale of
  mulslice.cpp:721-722) of the oracle's exit wave after s + 1 slices, at
  (h, k); zero outside the band limit;
* the index rule of mulslice.cpp:693-700: a negative index counts from the
  end once, then the index is clamped to the grid;
* intensity: sum |wave|^2 of the same exit waves (Parseval pins it to the
  beams' sum over the whole grid).
"""
import numpy as np

from oracle import Oracle as O

KEV = 200.0


def _case():
    nx, ny, ax, by = 32, 24, 11.0, 8.5
    rng = np.random.default_rng(5)
    t = np.exp(1j * 0.7 * rng.random((2, nx, ny)))
    w = np.exp(1j * 0.3 * rng.random((nx, ny)))
    return nx, ny, ax, by, t, [1.8, 2.2], [0, 1, 1, 0, 1], w


def test_beams_are_the_transform_of_each_exit_wave():
    nx, ny, ax, by, t, cz, layer, w = _case()
    hk = [(0, 0), (1, 0), (0, 2), (-1, -1), (3, -2), (nx // 2, 0)]
    tilt = (0.002, -0.001)
    exitw, beams, inten = O.multislice_series(t, cz, layer, w, ax, by, KEV, hk, tilt)
    k2max = O.k2max(nx, ny, ax, by, 2.0 / 3.0)
    kx2 = (np.fft.fftfreq(nx, ax / nx)) ** 2
    ky2 = (np.fft.fftfreq(ny, by / ny)) ** 2
    for s in range(len(layer)):
        ws = O.multislice(t, cz, layer[: s + 1], w, ax, by, KEV, tilt)
        f = np.fft.ifft2(ws)
        for ib, (h, k) in enumerate(hk):
            h, k = h % nx, k % ny
            assert abs(beams[s, ib] - f[h, k]) < 1e-12 * np.abs(f).max(), (s, h, k)
            if kx2[h] + ky2[k] > 1.01 * k2max:
                assert beams[s, ib] == 0.0
        assert abs(inten[s] - np.sum(np.abs(ws) ** 2)) < 1e-10 * inten[s]
        # Parseval: sum over the whole grid of |beam|^2 * nx ny = sum |wave|^2
        assert abs(np.sum(np.abs(f) ** 2) * nx * ny - inten[s]) < 1e-10 * inten[s]
    assert np.abs(exitw - O.multislice(t, cz, layer, w, ax, by, KEV, tilt)).max() == 0.0


def test_beam_index_rule_wraps_once_then_clamps():
    nx, ny, ax, by, t, cz, layer, w = _case()
    asked = [(-1, -2), (nx + 3, 0), (-nx - 5, 1), (2, ny + 7), (-3, -ny - 1)]
    meant = [(nx - 1, ny - 2), (nx - 1, 0), (0, 1), (2, ny - 1), (nx - 3, 0)]
    _, a, _ = O.multislice_series(t, cz, layer, w, ax, by, KEV, asked)
    _, b, _ = O.multislice_series(t, cz, layer, w, ax, by, KEV, meant)
    assert np.array_equal(a, b)
