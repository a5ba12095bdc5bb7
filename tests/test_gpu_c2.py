"""The large-HRTEM configuration at its full grid: SrTiO3 [001], 52x52 cells
(203.06 A), 2048^2 complex64, 300 kV, frozen phonons averaged after the
objective lens (Cs 1.2 mm, defocus -60 nm, 20 mrad aperture).

A slab of 1 cell (2 slices of 1.9525 A, 13,520 atoms) and 2 frozen-phonon configurations
keeps the oracle's FP64 run to seconds: the GPU's potential (checked against
the oracle at smaller sizes) feeds both the engine and the oracle's FP64
transmission, slice loop and objective-lens image; the averaged images agree
to 1e-5 relative L2 (FP32 engine) and 1e-10 (FP64 engine).
"""
import numpy as np
import pytest

import multislice_amd as msx
from multislice_amd import specimen
from oracle import Oracle as O

pytestmark = pytest.mark.gpu

KEV, CS, DF, AOBJ = 300.0, 1.2e7, -600.0, 0.020


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_c2_averaged_ctf_images_match_oracle(ctx):
    import torch
    at0, cell = specimen.srtio3_001((52, 52, 1))
    N, ns = 2048, 2
    cz = np.full(ns, cell[2] / ns)
    layer = np.arange(ns)
    img32 = np.zeros((N, N))
    img64 = np.zeros((N, N))
    ref = np.zeros((N, N))
    nconf = 2
    for c in range(nconf):
        at = msx.displace(at0, specimen.SRTIO3_U, seed=77, config=c)
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
    print(f"C2 2048^2, {ns} slices, {nconf} configs: FP32 {e32:.2e}, FP64 {e64:.2e}")
    assert e32 < 1e-5
    assert e64 < 1e-10
