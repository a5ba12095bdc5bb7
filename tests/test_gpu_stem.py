"""GPU parity of the batched STEM path (probe generation, slice loop, annular
detector epilogue) against the oracle's restatement of Musli ConvergedBeam /
SaveSTEM (MSProject.cs:458-514, 792-842) in Kirkland's conventions, plus
size-independent properties at the full workload grid."""
import numpy as np
import pytest

import multislice_amd as msx
from multislice_amd import specimen
from oracle import Oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

KEV = 300.0


def _srtio3_small(n=128, cells=(3, 3, 2)):
    at, cell = specimen.srtio3_001(cells)
    nsl = 2 * cells[2]
    V = O.potential(at["Z"], np.stack([at["x"], at["y"], at["z"]], 1), cell, n, n, nsl, 3.0)
    return at, cell, V, nsl


@pytest.fixture(scope="module")
def small():
    at, cell, V, nsl = _srtio3_small()
    t64 = O.transmission(V, cell[0], cell[1], KEV)
    return cell, V, t64, nsl


def test_probe_generation_matches_oracle(ctx, small):
    """One vacuum slice of zero thickness: the detector over all k returns the
    probe norm; the probe itself is checked through the exit of a 1-slice run."""
    cell, V, t64, nsl = small
    n = t64.shape[1]
    lam = msx.wavelength(KEV)
    kmax = np.sqrt(O.k2max(n, n, cell[0], cell[1]))
    out = ctx.stem(np.ones((1, n, n), np.complex128), [0.0], [0], cell[0], cell[1], KEV,
                   dict(alpha_max=0.021, cs=1e4, df=30.0), [(1.3, 2.2), (0.0, 0.0)],
                   [(0.0, 10.0), (0.0, 0.2 * kmax)], [1])
    assert abs(out[0, 0, 0] - 1.0) < 1e-12  # whole probe inside the 2/3 band limit
    assert abs(out[0, 0, 1] - 1.0) < 1e-12


@pytest.mark.parametrize("prec", [msx.F64, msx.F32])
def test_stem_matches_oracle(ctx, small, prec):
    cell, V, t64, nsl = small
    lam = msx.wavelength(KEV)
    rng = np.random.default_rng(3)
    pos = rng.random((7, 2)) * np.array(cell[:2])
    det = [(0.0, 0.010 / lam), (0.030 / lam, 0.060 / lam), (0.0, 10.0)]
    ck = [1, nsl // 2, nsl]
    layer = np.arange(nsl)
    cz = np.full(nsl, 3.905 / 2)
    probe = dict(alpha_max=0.021, cs=1e4, df=-20.0)
    ref = O.stem(t64, cz, layer, cell[0], cell[1], KEV, 0.021, pos, det, ck, cs=1e4, df=-20.0)
    t = t64 if prec == msx.F64 else t64.astype(np.complex64)
    got = ctx.stem(t, cz, layer, cell[0], cell[1], KEV, probe, pos, det, ck, batch=3)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"STEM {'F64' if prec == msx.F64 else 'F32'}: max rel err vs oracle FP64 {err:.2e}")
    assert err < (1e-11 if prec == msx.F64 else 1e-5)


@pytest.mark.parametrize("nx,ny,cells", [(384, 384, (3, 3, 2)), (256, 512, (3, 6, 2))])
def test_stem_radix3_and_rectangular_grids(ctx, nx, ny, cells):
    """The radix-3 length (384) and a rectangular grid take other FFT plans and
    column-group counts in the detector epilogue; FP32 against the FP64 oracle."""
    at, cell = specimen.srtio3_001(cells)
    nsl = 2 * cells[2]
    V = O.potential(at["Z"], np.stack([at["x"], at["y"], at["z"]], 1), cell, nx, ny, nsl, 3.0)
    t64 = O.transmission(V, cell[0], cell[1], KEV)
    lam = msx.wavelength(KEV)
    pos = np.random.default_rng(11).random((5, 2)) * np.array(cell[:2])
    det = [(0.0, 0.012 / lam), (0.025 / lam, 0.080 / lam)]
    cz, layer = np.full(nsl, cell[2] / nsl), np.arange(nsl)
    ref = O.stem(t64, cz, layer, cell[0], cell[1], KEV, 0.021, pos, det, [nsl], cs=1e4, df=10.0)
    got = ctx.stem(t64.astype(np.complex64), cz, layer, cell[0], cell[1], KEV,
                   dict(alpha_max=0.021, cs=1e4, df=10.0), pos, det, [nsl], batch=2)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    per_det = (np.abs(got - ref).max(axis=2) / np.abs(ref).max(axis=2)).max()  # the weak ADF ring on its own scale
    print(f"STEM {nx}x{ny} F32: max rel err vs oracle FP64 {err:.2e} (per detector {per_det:.2e})")
    assert err < 1e-5
    assert per_det < 1e-4


def test_stem_accumulate_and_weight(ctx, small):
    cell, V, t64, nsl = small
    pos = [(1.0, 2.0), (3.0, 1.5)]
    det = [(0.0, 1.0)]
    args = (t64, np.full(nsl, 1.9525), np.arange(nsl), cell[0], cell[1], KEV, dict(alpha_max=0.02), pos, det, [nsl])
    a = ctx.stem(*args)
    b = ctx.stem(*args, weight=0.25, out=a.copy())
    assert np.allclose(b, 1.25 * a, rtol=1e-13, atol=0)


def test_detector_partition_full_grid(ctx):
    """C3 grid (512^2 over 46.86 A): BF + ring + outer ring tiling [0, k_BW]
    sum to the total intensity inside the band limit (= 1 in vacuum)."""
    n, a = 512, 12 * 3.905
    kbw = np.sqrt(O.k2max(n, n, a, a))
    t = np.ones((1, n, n), np.complex64)
    out = ctx.stem(t, [1.9525], [0], a, a, KEV, dict(alpha_max=0.021, cs=1e4), [(23.0, 23.0), (20.1, 19.7)],
                   [(0.0, 0.5001), (0.5001, 1.2001), (1.2001, kbw + 1), (0.0, kbw + 1)], [1], batch=2)
    parts = out[0, :3].sum(0)
    assert np.allclose(parts, out[0, 3], rtol=2e-6)
    assert np.allclose(out[0, 3], 1.0, rtol=2e-6)


def test_probe_shift_is_translation(ctx, small):
    """Shifting the probe and the specimen by one pixel leaves the detector
    signal unchanged (periodic supercell)."""
    cell, V, t64, nsl = small
    n = t64.shape[1]
    dx = cell[0] / n
    cz, layer = np.full(nsl, 1.9525), np.arange(nsl)
    det = [(0.5, 2.5)]
    a = ctx.stem(t64, cz, layer, cell[0], cell[1], KEV, dict(alpha_max=0.02), [(5.0, 6.0)], det, [nsl])
    t_shift = np.roll(t64, 3, axis=1)
    b = ctx.stem(t_shift, cz, layer, cell[0], cell[1], KEV, dict(alpha_max=0.02), [(5.0 + 3 * dx, 6.0)], det, [nsl])
    assert abs(a - b).max() / abs(a).max() < 1e-11


def test_repeat_calls_are_independent(ctx, small):
    """A call must not depend on scratch state left by an earlier call
    (detector partials of skipped out-of-band column groups included)."""
    cell, V, t64, nsl = small
    n = t64.shape[1]
    args = dict(cz=np.full(nsl, 1.9525), layer=np.arange(nsl))
    det = [(0.0, 0.6), (0.6, 3.0)]
    pos = [(2.0, 3.0), (4.0, 1.0), (6.5, 6.5)]
    a1 = ctx.stem(t64, args["cz"], args["layer"], cell[0], cell[1], KEV, dict(alpha_max=0.02), pos, det, [1, nsl])
    rng = np.random.default_rng(9)
    other = np.exp(1j * rng.random((nsl, n, n)) * 3)
    ctx.stem(other, args["cz"], args["layer"], cell[0], cell[1], KEV, dict(alpha_max=0.03), pos[::-1], det, [nsl])
    a2 = ctx.stem(t64, args["cz"], args["layer"], cell[0], cell[1], KEV, dict(alpha_max=0.02), pos, det, [1, nsl])
    assert np.array_equal(a1, a2)


def _uniform53(hi, lo):
    """(u + 0.5) 2^-53 from the top 53 bits of two 32-bit words (msx_philox.h)."""
    return ((((int(hi) << 32) | int(lo)) >> 11) + 0.5) * 2.0 ** -53


def _device_displace(ctx, at, u, seed, config):
    torch = pytest.importorskip("torch")
    dev_in = torch.from_numpy(at.view(np.uint8).copy()).cuda()
    return msx.displace(dev_in, u, seed=seed, config=config, ctx=ctx).cpu().numpy().view(msx.ATOM_DTYPE)


def test_device_displacement_known_answer(ctx):
    """Atom 0 of configuration 0 with seed 0 draws Philox-4x32-10 of counter
    (0, 0, 0, 0) and key (0, 0): the first published known-answer vector
    (Salmon et al. 2011, Random123 kat_vectors), 6627e8d5 e169c58d bc57ac4c
    9b00dbd8.  Its Box-Muller pair is the x and y offset of that atom (in units
    of u); z uses the next counter (0, 0, 0, 1)."""
    at = np.zeros(3, msx.ATOM_DTYPE)
    at["Z"], at["occ"] = 38, 1.0
    d = _device_displace(ctx, at, {38: 1.0}, 0, 0)
    u1, u2 = _uniform53(0x6627e8d5, 0xe169c58d), _uniform53(0xbc57ac4c, 0x9b00dbd8)
    r = np.sqrt(-2.0 * np.log(u1))
    assert abs(d["x"][0] - r * np.cos(2 * np.pi * u2)) < 1e-14
    assert abs(d["y"][0] - r * np.sin(2 * np.pi * u2)) < 1e-14
    w = msx.philox4x32_10([0, 0, 0, 1], [0, 0])
    rz = np.sqrt(-2.0 * np.log(_uniform53(w[0], w[1])))
    assert abs(d["z"][0] - rz * np.cos(2 * np.pi * _uniform53(w[2], w[3]))) < 1e-14


def test_device_displacement_against_block_function(ctx):
    """Every atom's offsets from the device equal Box-Muller of the host
    Philox block function (pinned by the known-answer vectors) at counter
    (atom, 0, configuration, 0 / 1) and key (seed lo, seed hi)."""
    n, seed, config = 512, 0x1234_5678_9ABC, 7
    at = np.zeros(n, msx.ATOM_DTYPE)
    at["Z"], at["occ"] = 8, 1.0
    at["x"] = np.arange(n) * 0.37
    d = _device_displace(ctx, at, {8: 0.095}, seed, config)
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for i in range(0, n, 7):
        a = msx.philox4x32_10([i, 0, config, 0], key)
        b = msx.philox4x32_10([i, 0, config, 1], key)
        ra, rb = np.sqrt(-2 * np.log(_uniform53(a[0], a[1]))), np.sqrt(-2 * np.log(_uniform53(b[0], b[1])))
        assert abs(d["x"][i] - at["x"][i] - 0.095 * ra * np.cos(2 * np.pi * _uniform53(a[2], a[3]))) < 1e-13
        assert abs(d["y"][i] - 0.095 * ra * np.sin(2 * np.pi * _uniform53(a[2], a[3]))) < 1e-14
        assert abs(d["z"][i] - 0.095 * rb * np.cos(2 * np.pi * _uniform53(b[2], b[3]))) < 1e-14


def test_device_displacement_moments(ctx):
    """Mean 0 and RMS u per axis, independent axes, Gaussian 4th moment, on
    200,000 atoms displaced on the GPU."""
    n, u = 200000, 0.088
    at = np.zeros(n, msx.ATOM_DTYPE)
    at["Z"], at["occ"] = 38, 1.0
    d = _device_displace(ctx, at, {38: u}, 7, 3)
    dx = np.stack([d["x"], d["y"], d["z"]], 1)
    assert np.abs(dx.mean(0)).max() < 4 * u / np.sqrt(n)
    assert np.abs(dx.std(0) / u - 1).max() < 0.01
    assert abs(np.corrcoef(dx[:, 0], dx[:, 1])[0, 1]) < 0.01
    assert abs((dx[:, 2] ** 4).mean() / u ** 4 / 3 - 1) < 0.03


def test_displacement_host_equals_device(ctx):
    torch = pytest.importorskip("torch")
    at, _ = specimen.srtio3_001((6, 6, 10))
    host = msx.displace(at, specimen.SRTIO3_U, seed=99, config=5)
    dev_in = torch.from_numpy(at.view(np.uint8).copy()).cuda()
    dev = msx.displace(dev_in, specimen.SRTIO3_U, seed=99, config=5, ctx=ctx).cpu().numpy().view(msx.ATOM_DTYPE)
    for f in ("x", "y", "z"):
        assert np.abs(dev[f] - host[f]).max() < 1e-13
    assert np.array_equal(dev["Z"], host["Z"])


def test_stem_largest_grid(ctx):
    """The largest supported grid (2048^2: column tiles of 2 columns, the
    longest FFT plan) through the detector epilogue, against the oracle."""
    n, a = 2048, 4 * 3.905 * 2
    rng = np.random.default_rng(20)
    t = np.exp(1j * 0.3 * rng.random((2, n, n)))  # a weak random phase object
    cz, layer = np.full(2, 1.9525), np.arange(2)
    lam = msx.wavelength(KEV)
    det = [(0.0, 0.012 / lam), (0.03 / lam, 0.09 / lam)]
    pos = [(3.1, 4.7), (12.0, 9.5)]
    ref = O.stem(t, cz, layer, a, a, KEV, 0.021, pos, det, [2], cs=1e4, df=5.0)
    got = ctx.stem(t.astype(np.complex64), cz, layer, a, a, KEV, dict(alpha_max=0.021, cs=1e4, df=5.0), pos, det,
                   [2])
    err = np.abs(got - ref).max() / np.abs(ref).max()
    per_det = (np.abs(got - ref).max(axis=2) / np.abs(ref).max(axis=2)).max()
    print(f"STEM 2048^2 F32: max rel err vs oracle FP64 {err:.2e} (per detector {per_det:.2e})")
    assert err < 1e-5
    assert per_det < 1e-4


def test_stem_rejects(ctx, small):
    cell, V, t64, nsl = small
    cz, layer = np.full(nsl, 1.9525), np.arange(nsl)
    args = (t64, cz, layer, cell[0], cell[1], KEV, dict(alpha_max=0.02))
    with pytest.raises(msx.MsxError):  # no probe
        ctx.stem(*args, np.zeros((0, 2)), [(0.0, 1.0)], [nsl])
    with pytest.raises(msx.MsxError):  # no detector
        ctx.stem(*args, [(1.0, 1.0)], np.zeros((0, 2)), [nsl])
    with pytest.raises(msx.MsxError):  # checkpoints must increase within 1..nslice
        ctx.stem(*args, [(1.0, 1.0)], [(0.0, 1.0)], [nsl, 1])
    with pytest.raises(msx.MsxError):
        ctx.stem(*args, [(1.0, 1.0)], [(0.0, 1.0)], [nsl + 1])
    with pytest.raises(msx.MsxError):  # more detectors than the epilogue holds
        ctx.stem(*args, [(1.0, 1.0)], [(0.0, 1.0)] * 9, [nsl])
