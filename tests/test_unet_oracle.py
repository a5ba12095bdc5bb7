"""Pins the FP64 U-Net oracle (oracle/unet_ref.py) on the CPU.

- bfloat16 rounding against torch's float32 -> bfloat16 cast (round to
  nearest even) and hand-checked ties;
- the im2col convolution against the nested-loop definition;
- the analytic gradient of the loss against central finite differences, for
  one-, two- and three-level networks (pooling, upsampling, concatenation and
  the residual head all on the path), with one input channel and with
  several (the residual is then the channel mean);
- a network whose input channels are copies of one image equals the
  one-channel network whose first-layer weights are their sums.
"""
import numpy as np
import pytest
import torch

from oracle.unet_ref import UNetRef, bf16_round, conv_direct, conv_specs, im2col, param_count


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-6, 6, 10000),
                        [0.0, -0.0, 1.0, -2.5, 1e-30, 3.0e38]])
    ref = torch.from_numpy(v.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(bf16_round(v), ref)


def test_bf16_round_ties_to_even():
    # spacing of bfloat16 at 1.0 is 2^-7
    assert bf16_round(1 + 2.0 ** -8) == 1.0                      # tie -> even (mantissa 0)
    assert bf16_round(1 + 3 * 2.0 ** -8) == 1 + 2.0 ** -6        # tie -> even (mantissa 2)
    assert bf16_round(1 + 2.0 ** -8 + 2.0 ** -20) == 1 + 2.0 ** -7  # above the tie rounds up


@pytest.mark.parametrize("ks", [1, 3])
def test_im2col_conv_matches_definition(ks):
    rng = np.random.default_rng(ks)
    x = rng.standard_normal((2, 3, 5, 4))
    w = rng.standard_normal((4, 3 * ks * ks))
    b = rng.standard_normal(4)
    y = (np.matmul(w, im2col(x, ks)) + b[None, :, None]).reshape(2, 4, 5, 4)
    np.testing.assert_allclose(y, conv_direct(x, w, b, ks), rtol=1e-12, atol=1e-12)


def test_param_layout():
    # levels 2, base 4: enc (1->4, 4->4), (4->8, 8->8); dec (12->4, 4->4); head 4->1
    assert conv_specs(2, 4) == [(1, 4, 3, True), (4, 4, 3, True), (4, 8, 3, True), (8, 8, 3, True),
                                (12, 4, 3, True), (4, 4, 3, True), (4, 1, 1, False)]
    assert param_count(2, 4) == (36 + 4) + (144 + 4) + (288 + 8) + (576 + 8) + (432 + 4) + (144 + 4) + (4 + 1)
    # the production shape: 4 levels, 32 base channels, conv by conv
    sizes = [1 * 32 * 9 + 32, 32 * 32 * 9 + 32, 32 * 64 * 9 + 64, 64 * 64 * 9 + 64, 64 * 128 * 9 + 128,
             128 * 128 * 9 + 128, 128 * 256 * 9 + 256, 256 * 256 * 9 + 256,
             384 * 128 * 9 + 128, 128 * 128 * 9 + 128, 192 * 64 * 9 + 64, 64 * 64 * 9 + 64,
             96 * 32 * 9 + 32, 32 * 32 * 9 + 32, 32 + 1]
    assert param_count(4, 32) == sum(sizes) == 1_946_305


def _rand_params(levels, base, rng, cin=1):
    p = []
    for ci, co, k, _ in conv_specs(levels, base, cin):
        p.append(rng.standard_normal(co * ci * k * k) * np.sqrt(2.0 / (ci * k * k)))
        p.append(rng.standard_normal(co) * 0.1)
    return np.concatenate(p)


@pytest.mark.parametrize("levels,base,hw,cin", [(1, 3, 6, 1), (2, 2, 8, 1), (3, 2, 8, 1), (2, 2, 8, 3)])
def test_gradient_matches_finite_differences(levels, base, hw, cin):
    rng = np.random.default_rng(10 * levels + base + cin)
    p = _rand_params(levels, base, rng, cin)
    x = rng.standard_normal((2, hw, hw) if cin == 1 else (2, cin, hw, hw))
    t = rng.standard_normal((2, hw, hw))
    loss, _, g = UNetRef(levels, base, p, cin=cin).loss_and_grad(x, t)
    idx = rng.choice(p.size, 40, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = UNetRef(levels, base, pp, cin=cin).loss_and_grad(x, t)[0]
        lm = UNetRef(levels, base, pm, cin=cin).loss_and_grad(x, t)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_untrained_identity_head():
    # a zero head makes the network the identity (residual connection)
    levels, base = 2, 2
    rng = np.random.default_rng(3)
    p = _rand_params(levels, base, rng)
    ci, co, k, _ = conv_specs(levels, base)[-1]
    p[-(co * ci * k * k + co):] = 0.0
    x = rng.standard_normal((1, 8, 8))
    np.testing.assert_array_equal(UNetRef(levels, base, p).forward(x), x)


def test_copied_channels_equal_one_channel_network():
    levels, base, cin = 2, 3, 4
    rng = np.random.default_rng(3)
    p4 = _rand_params(levels, base, rng, cin)
    w0 = base * cin * 9
    # one-channel parameters: first-layer weights summed over the input channels
    p1 = np.concatenate([p4[:w0].reshape(base, cin, 9).sum(axis=1).ravel(), p4[w0:]])
    assert p1.size == param_count(levels, base)
    x = rng.standard_normal((2, 8, 8))
    y4 = UNetRef(levels, base, p4, cin=cin).forward(np.repeat(x[:, None], cin, axis=1))
    y1 = UNetRef(levels, base, p1).forward(x)
    np.testing.assert_allclose(y4, y1, rtol=1e-12, atol=1e-12)
