"""What the trained surrogate gains over the few-configuration mean alone.

A scaled-down version of the surrogate workload, small enough for the test
suite: 128 random Si [001] specimens 5-10 nm thick (random tilt <= 10 mrad,
random scan window), 128^2 grid, 32x32 ADF scans (40 mrad to the band limit).
A 3-level, 16-channel U-Net is trained on the GPU by the library's kernels on
engine-generated pairs (4-config mean -> mean of 8 further configs, 2000 Adam
steps) and measured on 12 held-out specimens against plain 32-config
multislice.  With per-config noise sigma, the 4-config mean alone errs by
0.47 sigma and an ideal estimator of the infinite mean by 0.17 sigma (the
32-config mean's own scatter).  Measured: the surrogate removes about a third
of the input's error (ratio 0.66); the test requires a ratio below 0.85.
"""
import numpy as np
import pytest

import multislice_amd as msx
from multislice_amd import surrogate as S

pytestmark = pytest.mark.gpu


def test_surrogate_beats_few_config_mean(ctx):
    setup = S.PairSetup(kinds=("si",), thickness=(50.0, 100.0), nconf=32)
    xtr, ttr, _ = S.training_pairs(ctx, 128, setup, seed=11)
    xev, tev, meta = S.converged_pairs(ctx, 12, setup, seed=12)
    net = msx.UNet(ctx, 3, 16, seed=0)
    xn, tn = S.normalise(xtr, ttr)
    hist = S.train(net, xn, tn, steps=2000, batch=16, lr=1e-3)
    e_s, e_x = S.evaluate(net, xev, tev, setup)
    print(f"loss {hist[:20].mean():.4g} -> {hist[-50:].mean():.4g}; held-out rel-RMS vs 32-config: "
          f"surrogate {e_s.mean():.4f}, 4-config mean {e_x.mean():.4f}")
    for m, a, b in zip(meta, e_s, e_x):
        print(f"  {m['kind']:7s} {m['thickness']:6.1f} A  surrogate {a:.4f}  input {b:.4f}")
    assert e_s.mean() < 0.85 * e_x.mean()
    net.close()
