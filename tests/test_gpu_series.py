"""Thickness series (msx_multislice_series; mulslice.cpp:682-737): beam values
and integrated intensity after every slice and exit waves kept at
intermediate thicknesses, against the oracle's FP64 loop (whose beams are
pinned on the CPU to numpy's transform of its exit waves), on a rectangular
radix-2 x radix-3 grid with a tilted beam, in-band, out-of-band, negative and
clamped beam indices; host and device buffers; an oblique cell."""
import ctypes as C

import numpy as np
import pytest

import multislice_amd as msx
from oracle import Oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

KEV = 200.0
NX, NY, AX, BY = 256, 384, 25.6, 36.0
TILT = (0.002, -0.003)


def _case(seed=7):
    rng = np.random.default_rng(seed)
    t = np.exp(1j * 0.6 * rng.random((3, NX, NY)))
    cz = rng.uniform(1.6, 2.4, 3)
    layer = np.array([0, 1, 2, 1, 0, 2, 1], np.int32)
    return t, cz, layer


# (0, 0), low orders, negative indices, an index clamped at the grid edge,
# and beams outside the 2/3 band limit (exactly zero)
BEAMS = [(0, 0), (1, 0), (0, 1), (-1, -1), (3, -2), (NX + 9, 0), (NX // 2, 0), (100, 150)]
AT = [1, 4, 7]


def _oracle(t, cz, layer, w):
    ex, beams, inten = O.multislice_series(t, cz, layer, w, AX, BY, KEV, BEAMS, TILT)
    waves = [O.multislice(t, cz, layer[:a], w, AX, BY, KEV, TILT) for a in AT]
    return ex, beams, inten / (NX * NY), waves


@pytest.mark.parametrize("prec", [msx.F64, msx.F32])
def test_series_matches_oracle(ctx, prec):
    t, cz, layer = _case()
    w0 = np.exp(1j * 0.2 * np.random.default_rng(1).random((NX, NY)))
    ex, beams_o, inten_o, waves_o = _oracle(t, cz, layer, w0)
    dt = np.complex128 if prec == msx.F64 else np.complex64
    tt = t.astype(dt)
    psi = np.zeros((1, NX, NY), dt)
    waves, beams, inten = ctx.multislice_series(tt, cz, layer, psi, AX, BY, KEV, at=AT, beams=BEAMS, intensity=True,
                                                tilt=TILT, psi0=w0.astype(dt))
    # the exit wave is msx_multislice's, bit for bit, and the last kept wave is it
    ref = np.zeros_like(psi)
    ctx.multislice(tt, cz, layer, ref, AX, BY, KEV, tilt=TILT, psi0=w0.astype(dt))
    assert np.array_equal(psi, ref)
    assert np.array_equal(waves[-1], psi)
    tol = 1e-11 if prec == msx.F64 else 1e-5
    ew = max(rel_l2(waves[i, 0], waves_o[i]) for i in range(len(AT)))
    bscale = np.abs(beams_o).max()
    eb = np.abs(beams[:, 0, :] - beams_o).max() / bscale
    ei = np.abs(inten[:, 0] - inten_o).max() / inten_o.max()
    print(f"series {NX}x{NY}, {len(layer)} slices, {'F64' if prec == msx.F64 else 'F32'}: kept waves rel-L2 {ew:.2e}, "
          f"beams max err / max beam {eb:.2e}, intensity max rel {ei:.2e}")
    assert ew < tol and eb < tol and ei < tol
    # out of the band limit the beam is exactly zero, as the reference's
    # propagate leaves it
    zero = np.abs(beams_o).max(axis=0) == 0.0
    assert zero.any() and not zero.all()
    assert np.all(beams[:, 0, zero] == 0)


def test_series_device_buffers_match_host(ctx):
    torch = pytest.importorskip("torch")
    t, cz, layer = _case(11)
    rng = np.random.default_rng(2)
    w = (np.exp(1j * rng.random((2, NX, NY)))).astype(np.complex64)
    tt = t.astype(np.complex64)
    ph = w.copy()
    wh, bh, ih = ctx.multislice_series(tt, cz, layer, ph, AX, BY, KEV, at=[2, 5], beams=BEAMS, intensity=True,
                                       tilt=TILT)
    pd = torch.from_numpy(w).cuda()
    wd, bd, idd = ctx.multislice_series(torch.from_numpy(tt).cuda(), cz, layer, pd, AX, BY, KEV, at=[2, 5],
                                        beams=BEAMS, intensity=True, tilt=TILT)
    assert np.array_equal(pd.cpu().numpy(), ph)
    assert np.array_equal(wd.cpu().numpy(), wh)
    assert np.array_equal(bd.cpu().numpy(), bh)
    assert np.array_equal(idd.cpu().numpy(), ih)
    # two different incident waves: two different series
    assert not np.array_equal(bh[:, 0], bh[:, 1])


def test_series_oblique_cell(ctx):
    t, cz, layer = _case(3)
    gam = np.deg2rad(75.0)
    w0 = np.ones((NX, NY), np.complex128)
    psi = np.zeros((1, NX, NY), np.complex128)
    hk = [(0, 0), (1, 1), (-2, 3)]
    waves, beams, inten = ctx.multislice_series(t, cz, layer, psi, AX, BY, KEV, at=[3], beams=hk, intensity=True,
                                                psi0=w0, gam=gam)
    for s in (2, len(layer) - 1):
        ws = O.multislice_oblique(t, cz, layer[: s + 1], w0, AX, BY, gam, KEV)
        f = np.fft.ifft2(ws)
        fb = np.array([f[h % NX, k % NY] for h, k in hk])
        assert np.abs(beams[s, 0] - fb).max() < 1e-11 * np.abs(fb).max()
        assert abs(inten[s, 0] - np.sum(np.abs(ws) ** 2) / (NX * NY)) < 1e-11
        if s == 2:
            assert rel_l2(waves[0, 0], ws) < 1e-11


def test_series_rejects_bad_records(ctx):
    from multislice_amd.engine import Series, _check
    t = np.ones((1, 64, 64), np.complex64)
    st, keep = msx.Context._stack(64, 64, 6.4, 6.4, KEV, [2.0], [0, 0, 0])
    psi = np.ones((1, 64, 64), np.complex64)
    waves = np.zeros((2, 1, 64, 64), np.complex64)
    for at, msg in (([2, 2], "ascend"), ([0], "ascend"), ([4], "ascend")):
        a = np.array(at, np.int32)
        se = Series(a.ctypes.data, len(a), waves.ctypes.data, None, 0, None, None)
        with pytest.raises(msx.MsxError, match=msg):
            _check(msx.load().msx_multislice_series(ctx.h, C.byref(st), C.c_void_p(t.ctypes.data), None,
                                                    C.c_void_p(psi.ctypes.data), 1, msx.F32, msx.HOST, C.byref(se)))
    hk = np.zeros(2, np.int32)
    se = Series(None, 0, None, hk.ctypes.data, 1, None, None)
    with pytest.raises(msx.MsxError, match="needs hk and beams"):
        _check(msx.load().msx_multislice_series(ctx.h, C.byref(st), C.c_void_p(t.ctypes.data), None,
                                                C.c_void_p(psi.ctypes.data), 1, msx.F32, msx.HOST, C.byref(se)))
    with pytest.raises(ValueError, match="ascend"):
        ctx.multislice_series(t, [2.0], [0, 0, 0], psi, 6.4, 6.4, KEV, at=[3, 1])
