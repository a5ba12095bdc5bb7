"""Oblique-cell propagation (msx_stack.gam; Musli ApplyPropagator,
Propagator.cpp:55-107, in this engine's sign convention) on the GPU against
the oracle's FP64 restatement (pinned on the CPU by the oblique reciprocal
metric, the right-angle reduction and the Fresnel semigroup): random phase
stacks, tilted beam, radix-2, radix-3 and rectangular grids, the 2/3 band
limit and Musli's full Nyquist disk (bw = 1)."""
import numpy as np
import pytest

import multislice_amd as msx
from oracle import Oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

KEV = 300.0


@pytest.mark.parametrize("nx,ny,gam,bw", [(256, 256, np.deg2rad(70.0), 2.0 / 3.0), (384, 256, np.deg2rad(104.0), 2.0 / 3.0),
                                          (256, 512, np.deg2rad(60.0), 1.0)])
def test_oblique_multislice_matches_oracle(ctx, nx, ny, gam, bw):
    rng = np.random.default_rng(nx + ny)
    ax, by = 0.1 * nx, 0.1 * ny
    nl, ns = 3, 9
    t = np.exp(1j * 0.8 * rng.random((nl, nx, ny)))
    cz, layer = rng.uniform(1.5, 2.5, nl), rng.integers(0, nl, ns)
    tilt = (0.003, -0.005)
    ref = O.multislice_oblique(t, cz, layer, np.ones((nx, ny), np.complex128), ax, by, gam, KEV, tilt, bw)
    for prec, tt in ((msx.F64, t), (msx.F32, t.astype(np.complex64))):
        psi = np.zeros((1, nx, ny), tt.dtype)
        ctx.multislice(tt, cz, layer, psi, ax, by, KEV, tilt=tilt, bw=bw, psi0=np.ones((nx, ny), tt.dtype), gam=gam)
        e = rel_l2(psi[0], ref)
        print(f"oblique {nx}x{ny} gam {np.rad2deg(gam):.0f} deg bw {bw:.3f} {'F64' if prec == msx.F64 else 'F32'}: "
              f"rel-L2 vs oracle {e:.2e}")
        assert e < (1e-11 if prec == msx.F64 else 1e-5)


def test_right_angle_is_the_orthogonal_path(ctx):
    rng = np.random.default_rng(3)
    t = np.exp(1j * rng.random((2, 128, 128))).astype(np.complex64)
    a = np.zeros((1, 128, 128), np.complex64)
    b = np.zeros_like(a)
    one = np.ones((128, 128), np.complex64)
    ctx.multislice(t, [2.0, 1.5], [0, 1, 0], a, 12.8, 12.8, KEV, psi0=one)
    ctx.multislice(t, [2.0, 1.5], [0, 1, 0], b, 12.8, 12.8, KEV, psi0=one, gam=np.pi / 2)
    assert np.array_equal(a, b)


def test_oblique_refused_where_unsupported(ctx):
    t = np.ones((1, 64, 64), np.complex64)
    with pytest.raises(msx.MsxError, match="oblique"):
        st, keep = msx.Context._stack(64, 64, 6.4, 6.4, KEV, [2.0], [0], gam=1.0)
        import ctypes as C
        pr = msx.engine.Probe(0.02, 0.0, 0.0, 0.0)
        pos = np.zeros(2)
        det = (msx.engine.Detector * 1)(msx.engine.Detector(0.0, 1.0))
        ck = np.array([1], np.int32)
        out = np.zeros(1)
        msx.engine._check(msx.load().msx_stem(ctx.h, C.byref(st), C.c_void_p(t.ctypes.data), C.byref(pr),
                                              C.c_void_p(pos.ctypes.data), 1, det, 1, C.c_void_p(ck.ctypes.data), 1,
                                              1.0, 0, C.c_void_p(out.ctypes.data), 0, msx.F32, msx.HOST))
    with pytest.raises(msx.MsxError, match="cell angle"):
        ctx.multislice(t, [2.0], [0], np.zeros((1, 64, 64), np.complex64), 6.4, 6.4, KEV, gam=4.0)
