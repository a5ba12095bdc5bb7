"""Frozen-phonon displacements (msx_displace): the Philox-4x32-10 block
function against the published known-answer vectors (Salmon et al. 2011,
Random123 kat_vectors), the Gaussian moments, and reproducibility per
(seed, configuration, atom)."""
import numpy as np
import pytest

import multislice_amd as msx
from conftest import BOTH_SUITES
from multislice_amd import specimen


@pytest.mark.parametrize("ctr,key,want", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
@BOTH_SUITES
def test_philox_known_answers(ctr, key, want, suite):
    assert list(msx.philox4x32_10(ctr, key)) == want


def _atoms(n, Z=38):
    a = np.zeros(n, msx.ATOM_DTYPE)
    a["Z"] = Z
    a["occ"] = 1
    a["x"] = np.arange(n) * 0.01
    return a


@BOTH_SUITES
def test_displacement_moments(suite):
    n = 200000
    at = _atoms(n)
    d = msx.displace(at, {38: 0.088}, seed=7, config=3)
    dx = np.stack([d["x"] - at["x"], d["y"] - at["y"], d["z"] - at["z"]], 1)
    assert np.abs(dx.mean(0)).max() < 4 * 0.088 / np.sqrt(n)
    assert np.abs(dx.std(0) / 0.088 - 1).max() < 0.01
    assert abs(np.corrcoef(dx[:, 0], dx[:, 1])[0, 1]) < 0.01
    # 4th moment of a Gaussian: 3 sigma^4
    assert abs((dx[:, 2] ** 4).mean() / 0.088 ** 4 / 3 - 1) < 0.03


@BOTH_SUITES
def test_displacement_reproducible_and_per_element(suite):
    at, _ = specimen.srtio3_001((3, 3, 2))
    a = msx.displace(at, specimen.SRTIO3_U, seed=11, config=0)
    b = msx.displace(at, specimen.SRTIO3_U, seed=11, config=0)
    c = msx.displace(at, specimen.SRTIO3_U, seed=11, config=1)
    assert np.array_equal(a, b)
    assert not np.array_equal(a["x"], c["x"])
    # an element with u = 0 does not move; the order of other atoms does not matter
    u = dict(specimen.SRTIO3_U)
    u[8] = 0.0
    d = msx.displace(at, u, seed=11, config=0)
    o = at["Z"] == 8
    assert np.array_equal(d["x"][o], at["x"][o]) and np.array_equal(d["x"][~o], a["x"][~o])
