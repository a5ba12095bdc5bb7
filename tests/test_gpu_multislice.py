"""GPU parity of the transmission build and the slice loop against the CPU
oracle (restatement of Kirkland mulslice: mulslice.cpp:414-421, 515-522,
703-739; slicelib.cpp:744-782).

Bar: FP32 engine exit waves within 1e-5 relative L2 of the oracle's FP32
(Kirkland float) run on identical inputs; the FP64 engine within 1e-11 of
the FP64 oracle.  Each FP32 result's error against the FP64 oracle is
printed (pytest -s) and bounded.
"""
import numpy as np
import pytest

import multislice_amd as msx
from multislice_amd import specimen
from oracle import Oracle as O
from conftest import max_rel, rel_l2

pytestmark = pytest.mark.gpu

KEV = 200.0


@pytest.fixture(scope="module")
def si_layers():
    """Config 1 specimen: Si [001], 4x4 cells laterally, one cell (4 slices of a/4) of
    potential, 256^2; the 80-slice stack repeats the 4 layers 20 times."""
    at, cell = specimen.si_001((4, 4, 1))
    V = O.potential(at["Z"], np.stack([at["x"], at["y"], at["z"]], 1), cell, 256, 256, 4, 5.0,
                    table=msx.TABLE_LMULTISLICE, formula=msx.FORMULA_KIRKLAND)
    return at, cell, V


def test_potential_matches_oracle(ctx, si_layers):
    at, cell, V = si_layers
    V64 = ctx.potential(at, cell, 256, 256, 4, 5.0, table=msx.TABLE_LMULTISLICE,
                        formula=msx.FORMULA_KIRKLAND, prec=msx.F64)
    assert rel_l2(V64, V) < 1e-13
    V32 = ctx.potential(at, cell, 256, 256, 4, 5.0, table=msx.TABLE_LMULTISLICE,
                        formula=msx.FORMULA_KIRKLAND, prec=msx.F32)
    assert rel_l2(V32, V) < 1e-7


@pytest.mark.parametrize("bw", [2.0 / 3.0, 0.0])
def test_transmission_matches_oracle(ctx, si_layers, bw):
    at, cell, V = si_layers
    t32 = ctx.transmission(V.astype(np.float32), cell[0], cell[1], KEV, bw)
    o32 = O.transmission(V.astype(np.float32), cell[0], cell[1], KEV, bw, f32=True)
    o64 = O.transmission(V, cell[0], cell[1], KEV, bw)
    e32 = rel_l2(t32, o32)
    print(f"transmission bw={bw:.3f}: FP32 vs oracle FP32 {e32:.2e}, vs FP64 {rel_l2(t32, o64):.2e}")
    assert e32 < 2e-6
    t64 = ctx.transmission(V, cell[0], cell[1], KEV, bw)
    assert rel_l2(t64, o64) < 1e-13


def _stack():
    a = 5.431
    cz = np.full(4, a / 4)
    layer = np.tile(np.arange(4), 20)  # 80 slices = 20 x (abcd)
    return cz, layer


def test_c1_exit_wave_fp32_parity(ctx, si_layers):
    at, cell, V = si_layers
    cz, layer = _stack()
    o32t = O.transmission(V.astype(np.float32), cell[0], cell[1], KEV, f32=True)
    o64t = O.transmission(V, cell[0], cell[1], KEV)
    psi0 = np.ones((256, 256), np.complex64)
    ref32, inten = O.multislice(o32t, cz, layer, psi0, cell[0], cell[1], KEV, f32=True, want_intensity=True)
    ref64 = O.multislice(o64t, cz, layer, psi0.astype(np.complex128), cell[0], cell[1], KEV)
    # engine on the same (oracle FP32) transmission functions
    psi = np.zeros((1, 256, 256), np.complex64)
    ctx.multislice(o32t, cz, layer, psi, cell[0], cell[1], KEV, psi0=psi0)
    e = rel_l2(psi[0], ref32)
    e64 = rel_l2(psi[0], ref64)
    eo = rel_l2(ref32, ref64)
    print(f"C1 exit wave: FP32 engine vs oracle FP32 {e:.2e}; vs oracle FP64 {e64:.2e}; "
          f"oracle FP32 vs FP64 {eo:.2e}")
    assert e < 1e-5
    assert e64 < 1e-5
    # intensity monitor of mulslice.cpp:727-737 decreases only by band-limit loss
    inten = inten / (256 * 256)  # plane wave of unit amplitude: sum |psi|^2 = nx*ny
    assert 0.9 < inten[-1] <= 1.0 + 1e-5


def test_c1_exit_wave_fp64_parity(ctx, si_layers):
    at, cell, V = si_layers
    cz, layer = _stack()
    o64t = O.transmission(V, cell[0], cell[1], KEV)
    psi0 = np.ones((256, 256), np.complex128)
    ref64 = O.multislice(o64t, cz, layer, psi0, cell[0], cell[1], KEV)
    psi = np.zeros((1, 256, 256), np.complex128)
    ctx.multislice(o64t, cz, layer, psi, cell[0], cell[1], KEV, psi0=psi0)
    assert rel_l2(psi[0], ref64) < 1e-11


def test_c1_end_to_end_engine(ctx, si_layers):
    """atoms -> potential -> transmission -> 80 slices entirely on the GPU."""
    at, cell, V = si_layers
    cz, layer = _stack()
    V32 = ctx.potential(at, cell, 256, 256, 4, 5.0, table=msx.TABLE_LMULTISLICE,
                        formula=msx.FORMULA_KIRKLAND, prec=msx.F32)
    t32 = ctx.transmission(V32, cell[0], cell[1], KEV)
    psi = np.zeros((1, 256, 256), np.complex64)
    ctx.multislice(t32, cz, layer, psi, cell[0], cell[1], KEV, psi0=np.ones((256, 256), np.complex64))
    ref32 = O.multislice(O.transmission(V.astype(np.float32), cell[0], cell[1], KEV, f32=True), cz, layer,
                         np.ones((256, 256), np.complex64), cell[0], cell[1], KEV, f32=True)
    ref64 = O.multislice(O.transmission(V, cell[0], cell[1], KEV), cz, layer,
                         np.ones((256, 256), np.complex128), cell[0], cell[1], KEV)
    e, e64 = rel_l2(psi[0], ref32), rel_l2(psi[0], ref64)
    print(f"C1 end-to-end: vs oracle FP32 {e:.2e}, vs oracle FP64 {e64:.2e}")
    assert e < 1e-5 and e64 < 1e-5


@pytest.mark.parametrize("nx,ny", [(64, 64), (128, 256), (384, 384), (512, 128)])
def test_random_stack_tilt_batch(ctx, nx, ny):
    """Random weak/strong phase layers, crystal tilt, several incident waves."""
    rng = np.random.default_rng(nx + 3 * ny)
    ax, by = nx * 0.11, ny * 0.09
    V = rng.random((3, nx, ny)) * 40.0
    t64 = O.transmission(V, ax, by, 300.0)
    cz = np.array([1.5, 2.0, 0.7])
    layer = rng.integers(0, 3, 17)
    w0 = rng.standard_normal((2, nx, ny)) + 1j * rng.standard_normal((2, nx, ny))
    tilt = (0.004, -0.003)
    ref = [O.multislice(t64, cz, layer, w0[i], ax, by, 300.0, tilt=tilt) for i in range(2)]
    psi = w0.copy()
    ctx.multislice(t64, cz, layer, psi, ax, by, 300.0, tilt=tilt)
    assert rel_l2(psi, np.stack(ref)) < 1e-11
    psi32 = w0.astype(np.complex64)
    ctx.multislice(t64.astype(np.complex64), cz, layer, psi32, ax, by, 300.0, tilt=tilt)
    assert rel_l2(psi32, np.stack(ref)) < 1e-5


def test_vacuum_is_identity(ctx):
    n = 256
    t = np.ones((1, n, n), np.complex128)
    rng = np.random.default_rng(1)
    # a band-limited wave is unchanged by vacuum propagation of total thickness 0 mod phase
    w = rng.standard_normal((1, n, n)) + 1j * rng.standard_normal((1, n, n))
    Wk = np.fft.ifft2(w[0])
    kx2, ky2 = O.freqn(n, 20.0)[1], O.freqn(n, 20.0)[1]
    mask = (kx2[:, None] + ky2[None, :]) < O.k2max(n, n, 20.0, 20.0)
    w = np.fft.fft2(Wk * mask)[None]
    # propagate +dz then -dz: identity
    psi = w.copy()
    ctx.multislice(t, [3.0], [0], psi, 20.0, 20.0, 200.0)
    ctx.multislice(t, [-3.0], [0], psi, 20.0, 20.0, 200.0)
    assert rel_l2(psi, w) < 1e-12


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ctf_image_matches_oracle(ctx, si_layers, prec):
    """Objective-lens image (mulslice.cpp:612-656 at one defocus; Kirkland chi)."""
    at, cell, V = si_layers
    cz, layer = _stack()
    o64t = O.transmission(V, cell[0], cell[1], KEV)
    exitw = O.multislice(o64t, cz, layer, np.ones((256, 256)), cell[0], cell[1], KEV)
    ref_w, ref_i = O.ctf_image(exitw, cell[0], cell[1], KEV, cs=1.2e7, df=-600.0, aobj=0.02)
    src = exitw[None] if prec == "f64" else exitw[None].astype(np.complex64)
    inten, w = ctx.ctf_image(src, cell[0], cell[1], KEV, 1.2e7, -600.0, 0.02, want_wave=True)
    tol = 1e-12 if prec == "f64" else 2e-6
    assert rel_l2(w[0], ref_w) < tol
    assert rel_l2(inten[0], ref_i) < 2 * tol
    mr = max_rel(inten[0], ref_i)
    print(f"CTF image {prec}: rel-L2 {rel_l2(inten[0], ref_i):.2e}, max-rel {mr:.2e}")
    assert mr <= (1e-11 if prec == "f64" else 1e-4)
