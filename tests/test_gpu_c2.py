"""The large-HRTEM configuration at its full grid: SrTiO3 [001], 52x52 cells
(203.06 A), 2048^2 complex64, 300 kV, frozen phonons averaged after the
objective lens (Cs 1.2 mm, defocus -60 nm, 20 mrad aperture).

Against the oracle (FP64 transmission, slice loop and objective-lens image,
fed by the GPU's FP64 potential, itself checked against the oracle at smaller
sizes), as deep as the oracle's CPU time allows: a slab of 1 cell (2 slices,
13,520 atoms) with 2 frozen-phonon configurations, and a slab of 5 cells (10
slices) with one.  The full configuration (50 cells, 100 slices, 676,000
atoms, 16 configurations) runs through the FP32 and the FP64 engine, and the
FP32 image's error against the FP64 one is stated and bounded.  Bars: 1e-5
relative L2 and 1e-4 max-rel (max |a - b| / max |b|) for FP32 images, 1e-10
for the FP64 engine.
"""
import numpy as np
import pytest

import multislice_amd as msx
from multislice_amd import hrtem, specimen
from oracle import Oracle as O
from conftest import max_rel

pytestmark = pytest.mark.gpu

KEV, CS, DF, AOBJ = 300.0, 1.2e7, -600.0, 0.020


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _slab_vs_oracle(ctx, ncell, nconf, seed):
    import torch
    at0, cell = specimen.srtio3_001((52, 52, ncell))
    N, ns = 2048, 2 * ncell
    cz = np.full(ns, cell[2] / ns)
    layer = np.arange(ns)
    img32 = np.zeros((N, N))
    img64 = np.zeros((N, N))
    ref = np.zeros((N, N))
    for c in range(nconf):
        at = msx.displace(at0, specimen.SRTIO3_U, seed=seed, config=c)
        V = torch.empty((ns, N, N), dtype=torch.float64, device="cuda")
        ctx.potential(at, cell, N, N, ns, 3.0, prec=msx.F64, out=V)
        for prec, dt, ct, acc in ((msx.F32, torch.float32, torch.complex64, img32),
                                  (msx.F64, torch.float64, torch.complex128, img64)):
            t = ctx.transmission(V.to(dt).contiguous(), cell[0], cell[1], KEV)
            psi = torch.ones((N, N), dtype=ct, device="cuda")
            ctx.multislice(t, cz, layer, psi, cell[0], cell[1], KEV)
            acc += ctx.ctf_image(psi, cell[0], cell[1], KEV, CS, DF, AOBJ).cpu().numpy() / nconf
            del t, psi
        Vh = V.cpu().numpy()
        del V
        tr = O.transmission(Vh, cell[0], cell[1], KEV)
        w = O.multislice(tr, cz, layer, np.ones((N, N), np.complex128), cell[0], cell[1], KEV)
        ref += O.ctf_image(w, cell[0], cell[1], KEV, CS, DF, AOBJ)[1] / nconf
    e32, e64 = rel(img32, ref), rel(img64, ref)
    m32, m64 = max_rel(img32, ref), max_rel(img64, ref)
    print(f"C2 2048^2, {ns} slices, {nconf} configs vs oracle FP64: FP32 rel-L2 {e32:.2e} max-rel {m32:.2e}; "
          f"FP64 rel-L2 {e64:.2e} max-rel {m64:.2e}")
    assert e32 < 1e-5 and m32 <= 1e-4
    assert e64 < 1e-10 and m64 <= 1e-10


def test_c2_averaged_ctf_images_match_oracle(ctx):
    _slab_vs_oracle(ctx, 1, 2, 77)


@pytest.mark.timeout(600)
def test_c2_ten_slices_match_oracle(ctx):
    _slab_vs_oracle(ctx, 5, 1, 78)


@pytest.mark.timeout(600)
def test_c2_full_workload_fp32_vs_fp64(ctx):
    """C2 itself: 52x52x50 cells, 2048^2, 100 slices, 16 configurations (seed
    and path of `python -m multislice_amd.hrtem c2`), through the FP32 and the
    FP64 engine; the FP32 image against the FP64 image."""
    import torch
    at, cell = specimen.srtio3_001((52, 52, 50))
    args = (ctx, at, cell, 2048, 100, KEV, specimen.SRTIO3_U, 16, 2024, CS, DF, AOBJ)
    img32, _ = hrtem.frozen_phonon_image(*args, prec=msx.F32)
    img32 = img32.cpu().numpy()
    torch.cuda.empty_cache()
    img64, _ = hrtem.frozen_phonon_image(*args, prec=msx.F64)
    img64 = img64.cpu().numpy()
    torch.cuda.empty_cache()
    e, m = rel(img32, img64), max_rel(img32, img64)
    q = np.quantile(np.abs(img32 - img64) / np.abs(img64), [0.5, 0.999])
    print(f"C2 full (100 slices, 16 configs, 2048^2): FP32 vs FP64 engine rel-L2 {e:.2e}, max-rel {m:.2e}, "
          f"pointwise relative median {q[0]:.1e} / 99.9th percentile {q[1]:.1e}; image mean {img64.mean():.6f}")
    assert e < 1e-5
    assert m <= 1e-4
