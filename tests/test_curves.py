"""Curve tables of the tabulated-curve potentials (msx_create_pots_rec_dt /
msx_create_pots_rs_dt, Musli CreatePots_rec_DT_DW / CreatePots_rs_DT_DW,
Muslib/Curves.cpp:45-81) against the oracle's Doyle-Turner amplitudes
(scat_ampl_rec/rs_DT_DW restated, pinned in tests/test_oracle_rec.py by the
forward-scattering integral and the Fourier pair), sampled where the
reference samples them: s accumulated by repeated addition of dds.  Host
code of libmsx.so, run in both suites."""
import numpy as np
import pytest

import multislice_amd as msx
from conftest import BOTH_SUITES
from oracle import Oracle as O


def _grid(npoints, dds):
    s, out = 0.0, []
    for _ in range(npoints):
        out.append(s)
        s += dds
    return np.array(out)


@BOTH_SUITES
def test_rec_curves(suite):
    types, dw, npts, dds, vol = [8, 22, 38, 79], [0.3, 0.45, 0.6, 0.0], 600, 0.005, 1.0 / 15.249
    got = msx.create_pots_rec_dt(types, dw, npts, dds, vol)
    s = _grid(npts, dds)
    want = np.stack([[O.dt_fe(z, v) for v in s] * np.exp(-w * s * s) * vol for z, w in zip(types, dw)])
    assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max()
    # the oracle's own table builder (s = i dds, divided by the area) agrees to the accumulation of s
    assert np.abs(got - O.dt_curves_rec(types, dw, npts, dds, 15.249)).max() <= 1e-12 * np.abs(want).max()


@BOTH_SUITES
def test_rs_curves(suite):
    types, dw, npts, dds = [6, 14, 79], [0.5, 0.0, 0.2], 500, 0.01
    got = msx.create_pots_rs_dt(types, dw, npts, dds)
    s = _grid(npts, dds)
    want = np.stack([[O.dt_rs(z, v, w) for v in s] for z, w in zip(types, dw)])
    assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max()
    assert np.abs(got - O.dt_curves_rs(types, dw, npts, dds)).max() <= 1e-12 * np.abs(want).max()


@BOTH_SUITES
def test_curve_builders_reject(suite):
    with pytest.raises(msx.MsxError):
        msx.create_pots_rec_dt([93], [0.0], 10, 0.1, 1.0)
    with pytest.raises(msx.MsxError):
        msx.create_pots_rs_dt([6], [0.0], 10, 0.0)
    # row 0 of the bank is empty: a zero curve, as the reference reads it
    assert not msx.create_pots_rs_dt([0], [0.0], 5, 0.1).any()
