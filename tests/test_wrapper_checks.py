"""The Python wrappers refuse wrongly shaped buffers before calling the C ABI
(whose kernels size every read and write from these shapes).  These run on
the CPU: the context is a stand-in with no handle, so a call that got past
its checks would fail in the C library instead of raising ValueError here."""
import numpy as np
import pytest
import torch

import multislice_amd as msx
from multislice_amd import engine


@pytest.fixture
def ctx():
    c = engine.Context.__new__(engine.Context)
    c.h = None
    c.device = 0
    return c


def _trans(nl=2, n=16):
    return np.ones((nl, n, n), np.complex64)


def test_multislice_rejects(ctx):
    t = _trans()
    psi = np.zeros((1, 16, 16), np.complex64)
    with pytest.raises(ValueError, match="nlayer, nx, ny"):
        ctx.multislice(t[0], [1.0, 1.0], [0, 1], psi, 5.0, 5.0, 200.0)
    with pytest.raises(ValueError, match="layer indices"):
        ctx.multislice(t, [1.0, 1.0], [0, 2], psi, 5.0, 5.0, 200.0)  # layer 2 of 2 images
    with pytest.raises(ValueError, match="layer indices"):
        ctx.multislice(t[:1], [1.0, 1.0], [0, 1], psi, 5.0, 5.0, 200.0)  # more thicknesses than images
    with pytest.raises(ValueError, match="layer indices"):
        ctx.multislice(t, [1.0, 1.0], [-1], psi, 5.0, 5.0, 200.0)
    with pytest.raises(ValueError, match="psi must end in"):
        ctx.multislice(t, [1.0, 1.0], [0, 1], np.zeros((1, 16, 8), np.complex64), 5.0, 5.0, 200.0)
    with pytest.raises(ValueError, match="psi0"):
        ctx.multislice(t, [1.0, 1.0], [0, 1], psi, 5.0, 5.0, 200.0, psi0=np.ones((8, 8), np.complex64))
    with pytest.raises(ValueError, match="cz"):
        ctx.multislice(t, [], [0], psi, 5.0, 5.0, 200.0)


def test_stem_rejects(ctx):
    t = _trans()
    args = ([1.0, 1.0], [0, 1], 5.0, 5.0, 300.0, dict(alpha_max=0.02), np.zeros((3, 2)), [(0.0, 1.0)], [2])
    with pytest.raises(ValueError, match=r"out must be \(1, 1, 3\)"):
        ctx.stem(t, *args, out=np.zeros((1, 1, 2)))
    with pytest.raises(ValueError, match="float64"):
        ctx.stem(t, *args, out=np.zeros((1, 1, 3), np.float32))
    with pytest.raises(ValueError, match="layer indices"):
        ctx.stem(t, [1.0, 1.0], [0, 3], *args[2:])
    with pytest.raises(ValueError, match="positions"):
        ctx.stem(t, *args[:6], np.zeros(3), *args[7:])


def test_transmission_and_potential_reject(ctx):
    V = np.zeros((2, 16, 16), np.float32)
    with pytest.raises(ValueError, match="out must be"):
        ctx.transmission(V, 5.0, 5.0, 200.0, out=np.zeros((1, 16, 16), np.complex64))
    with pytest.raises(ValueError, match="Vabs must be"):
        ctx.transmission(V, 5.0, 5.0, 200.0, Vabs=np.zeros((2, 16, 8), np.float32))
    with pytest.raises(ValueError, match="V must be"):
        ctx.transmission(V[0], 5.0, 5.0, 200.0)
    at = np.zeros(1, engine.ATOM_DTYPE)
    at["Z"] = 6
    with pytest.raises(ValueError, match="out must be"):
        ctx.potential(at, (5.0, 5.0, 5.0), 16, 16, 2, 3.0, out=np.zeros((2, 16, 15), np.float32))
    with pytest.raises(ValueError, match="out must be"):
        ctx.potential_rec(at, (5.0, 5.0, 5.0), 16, 16, 2, out=np.zeros((3, 16, 16), np.float32))
    with pytest.raises(ValueError, match="types"):
        ctx.potential_curve("rec", at, [0, 0], np.zeros((1, 8)), 0.1, (5.0, 5.0, 5.0), 16, 16, 2)


def test_partial_and_batch_reject(ctx):
    t = _trans()
    with pytest.raises(ValueError, match="layer indices"):
        ctx.partial_coherence(t, [1.0], [1], 5.0, 5.0, 200.0, dict(cs=0.0, df=0.0, aobj=0.02), 0.0, 0.001, 0.0, 10.0)
    tt = torch.ones((2, 16, 16), dtype=torch.complex64)
    job = dict(trans=tt, psi=torch.ones((16, 8), dtype=torch.complex64), ax=5.0, by=5.0, kev=200.0, cz=[1.0, 1.0],
               layer=[0, 1])
    with pytest.raises(ValueError, match="psi must be one"):
        ctx.multislice_batch([job])
    job["psi"] = torch.ones((16, 16), dtype=torch.complex64)
    job["layer"] = [0, 5]
    with pytest.raises(ValueError, match="layer indices"):
        ctx.multislice_batch([job])


def test_unet_and_displace_reject(ctx):
    x = torch.zeros((1, 2, 8, 8))
    w = torch.zeros((4, 2 * 9))
    with pytest.raises(ValueError, match="out must be"):
        engine.unet_conv(ctx, 0, 3, torch.zeros((1, 4, 8, 4)), w, x=x)
    with pytest.raises(ValueError, match="bias must be"):
        engine.unet_conv(ctx, 0, 3, torch.zeros((1, 4, 8, 8)), w, x=x, bias=torch.zeros(3))
    with pytest.raises(ValueError, match="w must be"):
        engine.unet_conv(ctx, 0, 3, torch.zeros((1, 4, 8, 8)), torch.zeros((4, 10)), x=x)
    with pytest.raises(ValueError, match="mode 0 needs"):
        engine.unet_conv(ctx, 2, 3, torch.zeros((4, 18)), w, x=x)
    net = engine.UNet.__new__(engine.UNet)
    net.h, net.ctx, net.in_channels = None, ctx, 1
    with pytest.raises(ValueError, match="out must be"):
        net.forward(np.zeros((1, 8, 8), np.float32), out=np.zeros((1, 8, 4), np.float32))
    with pytest.raises(ValueError, match="target"):
        net.train_step(np.zeros((1, 8, 8), np.float32), np.zeros((2, 8, 8), np.float32), 1e-3)
    net.in_channels = 4
    with pytest.raises(ValueError, match="inputs must be"):
        net.forward(np.zeros((1, 3, 8, 8), np.float32))
    at = np.zeros(4, engine.ATOM_DTYPE)
    with pytest.raises(ValueError, match="out must be"):
        msx.displace(at, {6: 0.1}, 1, 0, out=np.zeros(3, engine.ATOM_DTYPE))
