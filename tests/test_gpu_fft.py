"""GPU parity of the hand-written Stockham 2D FFT (msx_fft2d) against the
DFT definition (numpy pocketfft, FP64) -- the transform of cfpix::fft/ifft
(cfpix.cpp:115-145): sign +1 = exponent +i unscaled, sign -1 = exponent -i
scaled by 1/(nx*ny)."""
import numpy as np
import pytest

import multislice_amd as msx
from conftest import rel_l2

pytestmark = pytest.mark.gpu

SQUARE = [16, 32, 64, 128, 256, 384, 512, 768, 1024, 2048]


def _ref(x, sign):
    return np.fft.ifft2(x) * x.shape[-1] * x.shape[-2] if sign > 0 else np.fft.fft2(x) / (x.shape[-1] * x.shape[-2])


@pytest.mark.parametrize("n", SQUARE)
@pytest.mark.parametrize("sign", [1, -1])
def test_fft_square_f32(ctx, n, sign):
    rng = np.random.default_rng(n + sign)
    x = (rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))).astype(np.complex64)
    y = x.copy()
    ctx.fft2d(y, sign)
    err = rel_l2(y, _ref(x.astype(np.complex128), sign))
    # FP32 Cooley-Tukey error grows like eps*sqrt(log2 N)
    assert err < 1.5e-6, err


@pytest.mark.parametrize("n", SQUARE)
@pytest.mark.parametrize("sign", [1, -1])
def test_fft_square_f64(ctx, n, sign):
    rng = np.random.default_rng(n - sign)
    x = rng.standard_normal((1, n, n)) + 1j * rng.standard_normal((1, n, n))
    y = x.copy()
    ctx.fft2d(y, sign)
    assert rel_l2(y, _ref(x, sign)) < 1e-14


@pytest.mark.parametrize("nx,ny", [(256, 512), (512, 256), (16, 2048), (2048, 64), (384, 768), (128, 32)])
def test_fft_rect(ctx, nx, ny):
    rng = np.random.default_rng(nx * 7 + ny)
    x = rng.standard_normal((3, nx, ny)) + 1j * rng.standard_normal((3, nx, ny))
    y = x.copy()
    ctx.fft2d(y, +1)
    assert rel_l2(y, _ref(x, +1)) < 1e-14
    ctx.fft2d(y, -1)
    assert rel_l2(y, x) < 1e-14


def test_fft_impulse_and_plane_wave(ctx):
    n = 256
    x = np.zeros((1, n, n), np.complex128)
    x[0, 3, 5] = 1.0
    y = x.copy()
    ctx.fft2d(y, +1)
    kx, ky = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    assert np.abs(y[0] - np.exp(2j * np.pi * (3 * kx + 5 * ky) / n)).max() < 1e-12
    ctx.fft2d(y, -1)
    assert np.abs(y - x).max() < 1e-14


def test_fft_device_buffers(ctx):
    torch = pytest.importorskip("torch")
    n = 512
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((4, n, n)) + 1j * rng.standard_normal((4, n, n))).astype(np.complex64)
    t = torch.from_numpy(x.copy()).cuda()
    ctx.fft2d(t, +1)
    assert rel_l2(t.cpu().numpy(), _ref(x.astype(np.complex128), +1)) < 1.5e-6


def test_fft_rejects_bad_size(ctx):
    x = np.zeros((1, 100, 128), np.complex64)
    with pytest.raises(msx.MsxError) as e:
        ctx.fft2d(x, +1)
    assert e.value.code == -4
