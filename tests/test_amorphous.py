"""The amorphous-specimen generator of the uneven-job workload
(msx_amorphous: random sequential addition, host code of libmsx.so) and the
uneven-job specimen built on it.  CPU only."""
import numpy as np
import pytest

import multislice_amd as msx
from multislice_amd import hrtem


def _min_image_dist(at, box):
    p = np.stack([at["x"], at["y"], at["z"]], 1)
    d = p[:, None, :] - p[None, :, :]
    d -= box * np.round(d / box)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    return r.min()


def test_density_spacing_and_box():
    box = np.array([18.0, 21.0, 26.0])
    at = msx.amorphous(box, seed=4)
    target = round(2.0 / 12.011 * 6.02214076e23 * 1e-24 * box.prod())
    assert len(at) == target  # 2.0 g/cm^3 is far below the RSA jamming density
    assert _min_image_dist(at, box) >= 1.4
    for f, L in zip("xyz", box):
        assert at[f].min() >= 0 and at[f].max() < L
    assert np.all(at["Z"] == 6) and np.all(at["occ"] == 1)


def test_deterministic_in_seed():
    a = msx.amorphous((15.0, 15.0, 15.0), seed=9)
    b = msx.amorphous((15.0, 15.0, 15.0), seed=9)
    c = msx.amorphous((15.0, 15.0, 15.0), seed=10)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)


def test_thin_box_uses_all_pairs():
    """Fewer than three cells along an axis: the generator tests every atom
    along it, so the spacing holds across the periodic boundary."""
    box = np.array([3.5, 40.0, 2.0])
    at = msx.amorphous(box, seed=1)
    assert len(at) > 10
    assert _min_image_dist(at, box) >= 1.4


def test_stops_at_the_candidate_limit():
    """An impossible density (hard spheres of 1.4 A beyond jamming) places
    fewer atoms, never closer than the spacing."""
    box = np.array([10.0, 10.0, 10.0])
    at = msx.amorphous(box, density_g_cm3=12.0, seed=2, max_tries=5)
    assert 0 < len(at) < round(12.0 / 12.011 * 0.602214076 * 1000)
    assert _min_image_dist(at, box) >= 1.4


def test_rejects_bad_input():
    with pytest.raises(msx.MsxError):
        msx.amorphous((0.0, 10.0, 10.0))
    with pytest.raises(msx.MsxError):
        msx.amorphous((10.0, 10.0, 10.0), Z=0, mass=1.0)


def test_uneven_specimen_gold_cluster():
    rng = np.random.default_rng(3)
    at, cell = hrtem.uneven_specimen(rng, 256, 30, seed=7, max_au=50)
    box = np.array(cell)
    assert np.allclose(box, (25.6, 25.6, 60.0))
    au = at[at["Z"] == 79]
    c = at[at["Z"] == 6]
    assert len(au) == 41 and len(c) > 3000  # 41 Au atoms drawn with this seed
    pa = np.stack([au["x"], au["y"], au["z"]], 1)
    pc = np.stack([c["x"], c["y"], c["z"]], 1)
    d = pc[:, None, :] - pa[None, :, :]
    d -= box * np.round(d / box)
    assert np.sqrt((d ** 2).sum(-1)).min() >= 1.4
    # fcc nearest-neighbour distance a/sqrt(2) inside the cluster
    assert abs(_min_image_dist(au, box) - 4.078 / np.sqrt(2)) < 1e-9
    for f, L in zip("xyz", box):
        assert at[f].min() >= 0 and at[f].max() < L
