"""Surrogate U-Net on the GPU (msx_unet_*) against the FP64 oracle.

The GPU kernels round every product operand to bfloat16 and accumulate in
FP32.  The oracle's bf16 mode rounds the same operands and sums in FP64.

Single convolutions (every kernel mode, ragged shapes): the two differ only by
FP32 accumulation order, tolerance 2e-5 relative L2.

Whole networks: an activation whose FP32 and FP64 values round to different
bfloat16 neighbours perturbs every later layer, and the perturbation grows
with depth.  Measured on the CPU, a relative perturbation of every conv
output and output gradient of 1e-7 / 1e-6 (FP32 rounding of short / long
sums) moves the bf16 oracle's own output by 5e-3 / 1.1e-2 and its worst
gradient tensor by 8.7e-4 / 2.1e-2 at 4 levels and 32 base channels (2.5e-3 /
6.8e-3 for the gradients of the smaller networks).  The network tolerances
are therefore 2e-3 (up to 3 levels, 8 channels) and 2e-2 (4 levels, 32
channels) on the output, 1e-2 and 4e-2 per gradient tensor.  Against the plain
FP64 oracle (no rounding) they are the bf16 error budget: 5e-2 on the output,
8e-2 on the gradient.
"""
import numpy as np
import pytest

import multislice_amd as msx
from oracle.unet_ref import UNetRef, conv_specs

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def rand_params(levels, base, seed, cin=1):
    rng = np.random.default_rng(seed)
    p = []
    for ci, co, k, _ in conv_specs(levels, base, cin):
        p.append(rng.standard_normal(co * ci * k * k) * np.sqrt(2.0 / (ci * k * k)))
        p.append(rng.standard_normal(co) * 0.1)
    return np.concatenate(p).astype(np.float32)


def images(n, hw, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:hw, 0:hw]
    base = 1.0 + 0.5 * np.cos(2 * np.pi * xx / 8) * np.cos(2 * np.pi * yy / 8)
    return (base[None] + 0.2 * rng.standard_normal((n, hw, hw))).astype(np.float32)


SHAPES = [(1, 8, 2, 16), (2, 8, 2, 16), (3, 8, 2, 32), (4, 32, 1, 64)]
FWD_TOL = {1: 2e-3, 2: 2e-3, 3: 2e-3, 4: 2e-2}
GRAD_TOL = {1: 1e-2, 2: 1e-2, 3: 1e-2, 4: 4e-2}

LAYERS = [(2, 40, 96, 12, 20, 3), (1, 3, 130, 9, 7, 3), (2, 96, 32, 32, 32, 3), (2, 32, 1, 16, 16, 1),
          (3, 5, 70, 8, 8, 1)]


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("relu", [False, True])
@pytest.mark.parametrize("n,cin,cout,h,w,ks", LAYERS)
def test_conv_layer_matches_oracle(ctx, mode, relu, n, cin, cout, h, w, ks):
    import torch
    from oracle.unet_ref import bf16_round, col2im, im2col
    rng = np.random.default_rng(hash((mode, relu, n, cin, cout, h, w, ks)) % 2**32)
    x = rng.standard_normal((n, cin, h, w)).astype(np.float32)
    W = (rng.standard_normal((cout, cin * ks * ks)) / np.sqrt(cin * ks * ks)).astype(np.float32)
    b = rng.standard_normal(cout).astype(np.float32)
    g = rng.standard_normal((n, cout, h, w)).astype(np.float32)
    y = rng.standard_normal((n, cout, h, w)).astype(np.float32)  # mask source: about half <= 0
    cols = bf16_round(im2col(x.astype(np.float64), ks))
    G = g * (y > 0) if relu else g.astype(np.float64)
    Gm = bf16_round(G.reshape(n, cout, h * w))
    if mode == 0:
        ref = (np.matmul(bf16_round(W), cols) + b[None, :, None]).reshape(n, cout, h, w)
        if relu:
            ref = np.maximum(ref, 0)
        out = torch.empty((n, cout, h, w), device="cuda")
    elif mode == 1:
        ref = col2im(np.matmul(bf16_round(W).T, Gm), (n, cin, h, w), ks)
        out = torch.empty((n, cin, h, w), device="cuda")
    else:
        ref = np.einsum("nop,nkp->ok", Gm, cols)
        out = torch.zeros((cout, cin * ks * ks), device="cuda")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    msx.unet_conv(ctx, mode, ks, out, t(W), x=t(x), bias=t(b), g=t(g), y=t(y), relu=relu)
    e = rel(out.cpu().numpy(), ref)
    assert e < 2e-5, e


@pytest.mark.parametrize("levels,base,n,hw", SHAPES)
def test_forward_matches_oracle(ctx, levels, base, n, hw):
    net = msx.UNet(ctx, levels, base, seed=1)
    p = rand_params(levels, base, 7)
    net.set_params(p)
    x = images(n, hw, 1)
    y = net.forward(x)
    z = y.astype(np.float64) - x
    z_bf = UNetRef(levels, base, p.astype(np.float64), bf16=True).forward(x) - x
    z_64 = UNetRef(levels, base, p.astype(np.float64)).forward(x) - x
    e_bf, e_64 = rel(z, z_bf), rel(z, z_64)
    print(f"levels {levels} base {base} {n}x{hw}^2: vs bf16 oracle {e_bf:.2e}, vs fp64 oracle {e_64:.2e}")
    assert e_bf < FWD_TOL[levels]
    assert e_64 < 5e-2
    net.close()


def _slices(levels, base):
    o = 0
    for ci, co, k, _ in conv_specs(levels, base):
        yield slice(o, o + co * ci * k * k)
        o += co * ci * k * k
        yield slice(o, o + co)
        o += co


@pytest.mark.parametrize("levels,base,n,hw", SHAPES)
def test_gradient_matches_oracle(ctx, levels, base, n, hw):
    net = msx.UNet(ctx, levels, base, seed=1)
    p = rand_params(levels, base, 11)
    net.set_params(p)
    x = images(n, hw, 2)
    t = images(n, hw, 3)
    loss, g = net.gradient(x, t)
    l_bf, _, g_bf = UNetRef(levels, base, p.astype(np.float64), bf16=True).loss_and_grad(x, t)
    l_64, _, g_64 = UNetRef(levels, base, p.astype(np.float64)).loss_and_grad(x, t)
    assert abs(loss - l_bf) <= 2e-3 * l_bf
    worst = 0.0
    for s in _slices(levels, base):
        if np.linalg.norm(g_bf[s]) > 1e-6 * np.linalg.norm(g_bf):
            worst = max(worst, rel(g[s], g_bf[s]))
    e_64 = rel(g, g_64)
    print(f"levels {levels} base {base}: loss {loss:.6g} (oracle {l_bf:.6g}); worst tensor vs bf16 oracle "
          f"{worst:.2e}; all vs fp64 oracle {e_64:.2e}")
    assert worst < GRAD_TOL[levels]
    assert e_64 < 8e-2
    net.close()


def test_device_buffers_match_host(ctx):
    import torch
    net = msx.UNet(ctx, 3, 8, seed=5)
    x = images(3, 32, 4)
    t = images(3, 32, 6)
    y_h = net.forward(x)
    y_d = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    np.testing.assert_array_equal(y_h, y_d)
    l_h, g_h = net.gradient(x, t)
    l_d, g_d = net.gradient(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    assert abs(l_h - l_d) <= 1e-12 * l_h
    # weight gradients are summed with float atomics: order may differ
    assert rel(g_d, g_h.astype(np.float64)) < 1e-5
    net.close()


def test_adam_first_step(ctx):
    # Adam's first bias-corrected step is lr * g / (|g| + eps)
    net = msx.UNet(ctx, 2, 8, seed=9)
    p0 = net.params()
    x, t = images(2, 16, 7), images(2, 16, 8)
    _, g = net.gradient(x, t)
    lr = 1e-3
    net.train_step(x, t, lr)
    p1 = net.params()
    big = np.abs(g) > 1e-2 * np.abs(g).max()
    expect = p0 - lr * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(p1[big], expect[big], rtol=0, atol=1e-6)
    net.close()


def test_training_reduces_loss(ctx):
    # learn to remove a fixed high-frequency pattern from smooth images
    net = msx.UNet(ctx, 3, 8, seed=3)
    rng = np.random.default_rng(0)
    yy, xx = np.mgrid[0:32, 0:32]
    clean = np.stack([1 + 0.3 * np.sin(2 * np.pi * (a * xx + b * yy) / 32)
                      for a, b in rng.integers(1, 4, (8, 2))]).astype(np.float32)
    noisy = (clean + 0.1 * np.cos(np.pi * xx) * np.cos(np.pi * yy)).astype(np.float32)
    first = net.train_step(noisy, clean, 2e-3)
    for _ in range(150):
        last = net.train_step(noisy, clean, 2e-3)
    print(f"loss {first:.3e} -> {last:.3e}")
    assert last < 0.1 * first
    net.close()


def test_rejects_bad_shapes(ctx):
    net = msx.UNet(ctx, 3, 8)
    with pytest.raises(msx.MsxError):
        net.forward(np.zeros((1, 30, 32), np.float32))  # 30 is not a multiple of 4
    with pytest.raises(msx.MsxError):
        msx.UNet(ctx, 7, 8)
    net.close()


@pytest.mark.parametrize("levels,base,cin", [(3, 8, 4), (4, 32, 4)])
def test_multichannel_input_matches_oracle(ctx, levels, base, cin):
    """Several input channels (one image per frozen-phonon configuration):
    forward and gradient against the oracle, the residual being the channel
    mean; a one-channel input shape is refused."""
    import torch
    net = msx.UNet(ctx, levels, base, seed=1, in_channels=cin)
    p = rand_params(levels, base, 13, cin)
    assert net.nparam == p.size
    net.set_params(p)
    hw = 32 if levels < 4 else 64
    x = np.stack([images(2, hw, 20 + c) for c in range(cin)], axis=1)
    t = images(2, hw, 30)
    y = net.forward(x)
    ref = UNetRef(levels, base, p.astype(np.float64), bf16=True, cin=cin)
    base_img = x.astype(np.float64).mean(axis=1)
    e_fwd = rel(y.astype(np.float64) - base_img, ref.forward(x) - base_img)
    loss, g = net.gradient(x, t)
    l_bf, _, g_bf = ref.loss_and_grad(x, t)
    e_g = rel(g, g_bf)
    print(f"levels {levels} base {base} in_channels {cin}: forward vs bf16 oracle {e_fwd:.2e}, "
          f"loss {loss:.6g} (oracle {l_bf:.6g}), gradient {e_g:.2e}")
    assert e_fwd < FWD_TOL[levels]
    assert abs(loss - l_bf) <= 2e-3 * l_bf
    assert e_g < GRAD_TOL[levels]
    yd = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    np.testing.assert_array_equal(y, yd)
    with pytest.raises(ValueError, match="inputs must be"):
        net.forward(x[:, 0])
    net.close()
