"""What the trained surrogate gains over the few-configuration mean alone.

A scaled-down version of the surrogate workload, small enough for the test
suite: 384 random specimens of all three kinds (SrTiO3 [001], Si [001], Au
clusters on amorphous carbon), 5-15 nm thick, random tilt <= 10 mrad and scan
window, 128^2 grid, 32x32 ADF scans (40 mrad to the band limit).  A 3-level,
16-channel U-Net takes the 4 configurations' images as its input channels and
is trained on the GPU by the library's kernels on engine-generated pairs
(target: the mean of 16 further configurations, 6000 Adam steps of 32
images); it is measured on 18 held-out specimens against plain 32-config
multislice, run as it is deployed (the mean of its passes over the 8
symmetries of the square).  The test requires the surrogate's mean error
below 0.85 of the 4-config mean's, and below the 4-config mean's for every
specimen kind.  Training is deterministic (the weight gradients' split
reductions are summed in a fixed order), so the figures repeat run to run.
"""
import numpy as np
import pytest

import multislice_amd as msx
from multislice_amd import surrogate as S

pytestmark = pytest.mark.gpu


def test_surrogate_beats_few_config_mean(ctx):
    setup = S.PairSetup(thickness=(50.0, 150.0), nconf=32, detectors=((0.040, None),))
    xtr, ttr, _ = S.training_pairs(ctx, 384, setup, seed=11)
    xev, tev, meta = S.converged_pairs(ctx, 18, setup, seed=12)
    net = msx.UNet(ctx, 3, 16, seed=0, in_channels=setup.ninput)
    xn, tn = S.normalise(xtr[:, 0], ttr[:, 0])
    hist = S.train(net, xn, tn, steps=6000, batch=32, lr=1e-3)
    e_s, e_x = S.evaluate(net, xev[:, 0], tev[:, 0], setup, tta=True)
    print(f"loss {hist[:20].mean():.4g} -> {hist[-50:].mean():.4g}; held-out rel-RMS vs 32-config: "
          f"surrogate {e_s.mean():.4f}, 4-config mean {e_x.mean():.4f}")
    for m, a, b in zip(meta, e_s, e_x):
        print(f"  {m['kind']:7s} {m['thickness']:6.1f} A  surrogate {a:.4f}  input {b:.4f}")
    kinds = sorted(set(m["kind"] for m in meta))
    assert len(kinds) == 3
    assert e_s.mean() < 0.85 * e_x.mean()
    for k in kinds:
        sel = np.array([m["kind"] == k for m in meta])
        print(f"  {k}: surrogate {e_s[sel].mean():.4f}, input {e_x[sel].mean():.4f}")
        assert e_s[sel].mean() < e_x[sel].mean(), k
    net.close()
