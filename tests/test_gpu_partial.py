"""GPU parity of msx_partial_coherence (Kirkland mulslice partial coherence,
mulslice.cpp:539-676) against the oracle's FP64 restatement, which
tests/test_oracle_partial.py pins.

Bar: FP32 images within 1e-5 relative L2 and 1e-4 max relative error of the
FP64 oracle; the FP64 engine within 1e-10."""
import numpy as np
import pytest

import multislice_amd as msx
from multislice_amd import specimen
from oracle import Oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

KEV = 200.0
LENS = dict(cs=1.3e7, df=650.0, aobj=0.022, dfa2=35.0, dfa2phi=0.35, dfa3=420.0, dfa3phi=-0.6)


def _stack(nx, ny, cells):
    at, cell = specimen.srtio3_001(cells)
    nsl = 2 * cells[2]
    V = O.potential(at["Z"], np.stack([at["x"], at["y"], at["z"]], 1), cell, nx, ny, nsl, 3.0)
    t = O.transmission(V, cell[0], cell[1], KEV)
    return t, np.full(nsl, cell[2] / nsl), np.arange(nsl), cell


@pytest.fixture(scope="module")
def small():
    t, cz, layer, cell = _stack(128, 128, (3, 3, 2))
    step = O.wavelength(KEV) / cell[0]
    args = (cz, layer, cell[0], cell[1], KEV, LENS, 0.0, 2.3 * step, 30.0, 15.0)
    ref, nill = O.partial_image(t, *args)
    return t, args, ref, nill


@pytest.mark.parametrize("prec", [msx.F64, msx.F32])
def test_partial_coherence_matches_oracle(ctx, small, prec):
    t, args, ref, nill = small
    tt = t if prec == msx.F64 else t.astype(np.complex64)
    got, na, nd = ctx.partial_coherence(tt, *args, batch=4)
    assert na == nill == 21 and nd == 11
    e2 = rel_l2(got, ref)
    emax = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"partial coherence 128^2 {'F64' if prec == msx.F64 else 'F32'}: rel-L2 {e2:.2e}, max rel {emax:.2e}")
    if prec == msx.F64:
        assert e2 < 1e-10
    else:
        assert e2 < 1e-5 and emax < 1e-4


def test_partial_coherence_rectangular_grid(ctx):
    t, cz, layer, cell = _stack(256, 128, (6, 3, 1))
    step = O.wavelength(KEV) / cell[1]
    args = (cz, layer, cell[0], cell[1], KEV, dict(LENS, dfa3=0.0), 0.5 * step, 1.2 * step, 20.0, 10.0)
    ref, nill = O.partial_image(t, *args)
    got, na, nd = ctx.partial_coherence(t.astype(np.complex64), *args)
    assert na == nill and nd == 11
    e2 = rel_l2(got, ref)
    print(f"partial coherence 256x128 F32 ({na} angles): rel-L2 {e2:.2e}")
    assert e2 < 1e-5


def test_batch_size_does_not_change_the_image(ctx, small):
    t, args, ref, nill = small
    t32 = t.astype(np.complex64)
    a = ctx.partial_coherence(t32, *args, batch=1)[0]
    b = ctx.partial_coherence(t32, *args, batch=0)[0]
    assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()


def test_one_angle_one_defocus_is_the_coherent_image(ctx, small):
    """The mixed-domain exit of the slice loop and its scaling against the
    real-space path: msx_multislice then msx_ctf_image."""
    import torch
    t, args, ref, nill = small
    cz, layer, ax, by = args[0], args[1], args[2], args[3]
    lens = dict(cs=LENS["cs"], df=LENS["df"], aobj=LENS["aobj"])
    td = torch.from_numpy(t.astype(np.complex64)).cuda()
    got, na, nd = ctx.partial_coherence(td, cz, layer, ax, by, KEV, lens, 0.0, 1e-6, 0.0, 1.0)
    assert na == 1 and nd == 1 and got.is_cuda
    psi = torch.ones(t.shape[1:], dtype=torch.complex64, device="cuda")
    ctx.multislice(td, cz, layer, psi, ax, by, KEV)
    coh = ctx.ctf_image(psi, ax, by, KEV, lens["cs"], lens["df"], lens["aobj"]).double()
    assert rel_l2(got.cpu().numpy(), coh.cpu().numpy()) < 1e-6


def test_graph_replay_is_identical(ctx, small):
    import torch
    t, args, ref, nill = small
    td = torch.from_numpy(t.astype(np.complex64)).cuda()
    ctx.set_graphs(True)
    try:
        a = ctx.partial_coherence(td, *args, batch=8)[0].cpu().numpy()
        h0 = ctx.graph_stats()
        b = ctx.partial_coherence(td, *args, batch=8)[0].cpu().numpy()
        h1 = ctx.graph_stats()
    finally:
        ctx.set_graphs(False)
    assert np.array_equal(a, b)
    assert h1[1] > h0[1]  # replays
    assert rel_l2(a, ref) < 1e-5


def test_rejects(ctx, small):
    t, args, ref, nill = small
    cz, layer, ax, by = args[0], args[1], args[2], args[3]
    t32 = t.astype(np.complex64)
    step = O.wavelength(KEV) / ax
    with pytest.raises(msx.MsxError):  # acmax < acmin
        ctx.partial_coherence(t32, cz, layer, ax, by, KEV, LENS, 2 * step, step, 0.0, 1.0)
    with pytest.raises(msx.MsxError):  # a ring with no grid angle
        ctx.partial_coherence(t32, cz, layer, ax, by, KEV, LENS, 1.1 * step, 1.3 * step, 0.0, 1.0)
    with pytest.raises(msx.MsxError):  # defocus spread without a step
        ctx.partial_coherence(t32, cz, layer, ax, by, KEV, LENS, 0.0, step, 10.0, 0.0)
