"""FFT plans the autotuner chooses between, beyond the work-group shapes:

* the alternative FFT plan of each sweep (512: rows 16 x 16 x 2 on 32
  threads per line, columns 4 x 4 x 4 x 8, against the built-in 8 x 8 x 8
  on 64 threads) is another factorisation of the same transform: the
  slice loop and STEM agree with the oracle's FP64 loop to the FP32 bars of
  the built-in plan, in every combination of row and column plan;
* the row sweep's second plan at 256, 384, 768, 1024 and 2048 (other thread
  counts per line and radix sequences) meets the same bar against the
  oracle's FP64 loop as the built-in plan;
* msx_set_tuning refuses a plan the grid does not have.
"""
import numpy as np
import pytest

import multislice_amd as msx
from oracle import Oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

KEV = 300.0
N = 512
A = 0.09152 * N  # C3's pixel size


def _phase_stack(seed=4, nl=3):
    rng = np.random.default_rng(seed)
    t = np.exp(1j * 0.5 * rng.random((nl, N, N)))
    return t, rng.uniform(1.8, 2.1, nl)


def _shapes(**kw):
    t = dict(row_wpb=8, col_wpb=-1, batch=64, ms_per_slice=0.0, col_tile=8, row_plan=0, col_plan=0)
    t.update(kw)
    return t


def _stem(ctx, t, cz, layer, pos, tuning):
    ctx.set_tuning(N, prec=msx.F32, tuning=tuning)
    try:
        return ctx.stem(t.astype(np.complex64), cz, layer, A, A, KEV, dict(alpha_max=0.021),
                        pos, [(0.0, 0.5), (1.0, 3.6)], [len(layer)], batch=tuning["batch"])
    finally:
        ctx.set_tuning(N, prec=msx.F32, tuning=None)


@pytest.mark.parametrize("rp,cp", [(1, 1), (1, 0), (0, 1)])
def test_alternative_plans_match_oracle(ctx, rp, cp):
    t, cz = _phase_stack(7)
    layer = [0, 1, 2, 1, 0]
    w0 = np.ones((N, N), np.complex128)
    ref = O.multislice(t, cz, layer, w0, A, A, KEV)
    res = {}
    for plan in ((0, 0), (rp, cp)):
        ctx.set_tuning(N, prec=msx.F32, tuning=_shapes(row_plan=plan[0], col_plan=plan[1]))
        try:
            psi = np.zeros((1, N, N), np.complex64)
            ctx.multislice(t.astype(np.complex64), cz, layer, psi, A, A, KEV, psi0=w0.astype(np.complex64))
        finally:
            ctx.set_tuning(N, prec=msx.F32, tuning=None)
        res[plan] = rel_l2(psi[0], ref)
    print(f"plans row {rp} col {cp}: rel-L2 vs oracle {res[(rp, cp)]:.2e} (built-in plans {res[(0, 0)]:.2e})")
    assert res[(rp, cp)] < 1e-6 and res[(rp, cp)] < 2.0 * res[(0, 0)]
    # STEM through the alternative column plan (probe generation, detector
    # epilogue) against the built-in one and the oracle's FP64 STEM
    pos = np.array([[1.3, 2.2], [20.0, 31.5], [45.0, 7.0]])
    a = _stem(ctx, t, cz, layer, pos, _shapes(row_plan=rp, col_plan=cp))
    b = _stem(ctx, t, cz, layer, pos, _shapes())
    o = O.stem(t, cz, layer, A, A, KEV, 0.021, pos, [(0.0, 0.5), (1.0, 3.6)], [len(layer)], nthreads=8)
    ea, eb = np.abs(a - o).max() / np.abs(o).max(), np.abs(b - o).max() / np.abs(o).max()
    print(f"STEM plans row {rp} col {cp}: max rel vs oracle {ea:.2e} (built-in plans {eb:.2e})")
    assert ea < 1e-5 and eb < 1e-5


@pytest.mark.parametrize("n,nsl", [(256, 5), (384, 5), (768, 4), (1024, 3), (2048, 2)])
def test_alternative_row_plans_match_oracle(ctx, n, nsl):
    rng = np.random.default_rng(n)
    t = np.exp(1j * 0.5 * rng.random((2, n, n)))
    cz = rng.uniform(1.8, 2.1, 2)
    layer = [s % 2 for s in range(nsl)]
    a = 0.09 * n
    w0 = np.ones((n, n), np.complex128)
    ref = O.multislice(t, cz, layer, w0, a, a, KEV)
    res = {}
    for rp in (0, 1):
        ctx.set_tuning(n, prec=msx.F32, tuning=dict(row_wpb=0, col_wpb=0, batch=0, row_plan=rp))
        try:
            psi = np.zeros((2, n, n), np.complex64)
            ctx.multislice(t.astype(np.complex64), cz, layer, psi, a, a, KEV, psi0=w0.astype(np.complex64))
        finally:
            ctx.set_tuning(n, prec=msx.F32, tuning=None)
        res[rp] = max(rel_l2(psi[0], ref), rel_l2(psi[1], ref))
    print(f"{n}^2 row plans 0 / 1: rel-L2 vs oracle {res[0]:.2e} / {res[1]:.2e}")
    assert res[1] < 1e-6 and res[1] < 2.0 * res[0] + 1e-7


def test_set_tuning_refuses_what_the_grid_lacks(ctx):
    with pytest.raises(msx.MsxError, match="FFT plans"):
        ctx.set_tuning(N, prec=msx.F32, tuning=_shapes(row_plan=2))
    with pytest.raises(msx.MsxError, match="FFT plans"):
        ctx.set_tuning(384, prec=msx.F32, tuning=_shapes(col_plan=1, col_tile=0))
    ctx.set_tuning(N, prec=msx.F32, tuning=None)
