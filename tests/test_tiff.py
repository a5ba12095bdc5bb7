"""Kirkland floating-point TIFF files (msx_tiff_write / msx_tiff_read,
replacing floatTIFF::write / read of floatTIFF.cpp): the reference's own
floatTIFF, compiled into oracle/_ref, pins the format -- files written by both
agree byte for byte except the 20-byte date, and each side reads the other's
files -- plus a committed file written by the reference.  Host code of
libmsx.so: the tests that need no compiled reference run in both the CPU and
the GPU suite."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import multislice_amd as msx
from conftest import BOTH_SUITES
from oracle import ref

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tiff")


def ours_write(path, pix, npix, params, rmin, rmax, imin, imax, dx, dy):
    pix = np.ascontiguousarray(pix, np.float32)
    prm = np.ascontiguousarray(params, np.float32)
    return msx.load().msx_tiff_write(os.fsencode(path), C.c_void_p(pix.ctypes.data), pix.shape[1], pix.shape[0],
                                     npix, C.c_void_p(prm.ctypes.data), len(prm), rmin, rmax, imin, imax, dx, dy)


def ours_read(path):
    nx, ny, npix = C.c_int32(), C.c_int32(), C.c_int32()
    prm = np.zeros(512, np.float32)
    assert msx.load().msx_tiff_read(os.fsencode(path), None, 0, C.byref(nx), C.byref(ny), C.byref(npix), None) == 0
    pix = np.zeros((ny.value, nx.value), np.float32)
    rc = msx.load().msx_tiff_read(os.fsencode(path), C.c_void_p(pix.ctypes.data), pix.size, C.byref(nx), C.byref(ny),
                                  C.byref(npix), C.c_void_p(prm.ctypes.data))
    return rc, pix, npix.value, prm


def date_offset(raw):
    """Offset of the DateTime bytes: the value of tag 306 in the third directory."""
    ifd = int.from_bytes(raw[4:8], "little")
    for _ in range(3):
        n = int.from_bytes(raw[ifd:ifd + 2], "little")
        for k in range(n):
            e = raw[ifd + 2 + 12 * k: ifd + 14 + 12 * k]
            if int.from_bytes(e[:2], "little") == 306:
                return int.from_bytes(e[8:12], "little")
        ifd = int.from_bytes(raw[ifd + 2 + 12 * n: ifd + 6 + 12 * n], "little")
    return None


CASES = [  # (ny, nx, npix, dx, dy)
    (6, 8, 1, 0.0, 0.0), (33, 16, 2, 0.2, 0.2), (300, 200, 1, 0.1, 0.25), (257, 512, 2, 0.3, 0.1), (2, 2, 1, 1.0, 1.0),
]


@pytest.mark.parametrize("ny,nx,npix,dx,dy", CASES)
def test_bytes_match_reference_writer(tmp_path, ny, nx, npix, dx, dy):
    R = ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(nx * 7 + ny)
    pix = (rng.standard_normal((ny, nx)) * 3 + 1).astype(np.float32)
    prm = rng.standard_normal(512).astype(np.float32)
    lo, hi = float(pix.min()), float(pix.max())
    args = (npix, prm, lo, hi, lo - 0.5, hi + 0.25, dx, dy)
    a, b = tmp_path / "ours.tif", tmp_path / "ref.tif"
    assert ours_write(a, pix, *args) == 0
    assert R.ref_tiff_write(os.fsencode(b), pix.ravel(), nx, ny, npix, prm, 512, lo, hi, lo - 0.5, hi + 0.25,
                            dx, dy) == 1
    ra, rb = a.read_bytes(), b.read_bytes()
    assert len(ra) == len(rb)
    d = date_offset(ra)
    assert d is not None and d == date_offset(rb)
    diff = [i for i in range(len(ra)) if ra[i] != rb[i]]
    assert all(d <= i < d + 20 for i in diff), diff[:10]
    # cross reads: the reference reads ours, we read the reference's
    buf = np.zeros(nx * ny, np.float32)
    rp = np.zeros(512, np.float32)
    n1, n2, n3 = C.c_int(), C.c_int(), C.c_int()
    assert R.ref_tiff_read(os.fsencode(a), buf, buf.size, C.byref(n1), C.byref(n2), C.byref(n3), rp) == 1
    assert (n1.value, n2.value, n3.value) == (nx, ny, npix)
    assert np.array_equal(buf.reshape(ny, nx), pix)
    rc, got, gn, gp = ours_read(b)
    assert rc == 0 and gn == npix and np.array_equal(got, pix)
    assert gp[0] == npix and np.array_equal(gp[1:], prm[1:])


@BOTH_SUITES
def test_reads_committed_reference_file(suite):
    """A 12 x 5 complex layer written by the reference's floatTIFF (date
    included), with its expected contents."""
    with open(os.path.join(GOLD, "ref_12x5_complex.json")) as f:
        exp = json.load(f)
    rc, pix, npix, prm = ours_read(os.path.join(GOLD, "ref_12x5_complex.tif"))
    assert rc == 0 and npix == 2 and pix.shape == (5, 24)
    assert np.array_equal(pix, np.array(exp["pix"], np.float32))
    assert np.array_equal(prm[:8], np.array(exp["params"][:8], np.float32))
    img, p = msx.tiff_read(os.path.join(GOLD, "ref_12x5_complex.tif"))
    assert img.shape == (5, 12) and np.array_equal(img.real, pix[:, :12]) and np.array_equal(img.imag, pix[:, 12:])


@BOTH_SUITES
def test_python_round_trip_and_errors(tmp_path, suite):
    rng = np.random.default_rng(1)
    z = (rng.standard_normal((40, 24)) + 1j * rng.standard_normal((40, 24))).astype(np.complex64)
    msx.tiff_write(tmp_path / "w.tif", z, params=np.arange(512, dtype=np.float32), dx=0.1, dy=0.1)
    back, p = msx.tiff_read(tmp_path / "w.tif")
    assert np.array_equal(back, z) and p[0] == 2 and p[5] == 5
    pix = np.ones((4, 4), np.float32)
    assert ours_write(tmp_path / "flat.tif", pix, 1, np.zeros(512), 1.0, 1.0, 0, 0, 0, 0) == -1  # rmax == rmin
    assert ours_write(tmp_path / "bad.tif", pix, 3, np.zeros(512), 0.0, 1.0, 0, 0, 0, 0) == -2   # npix
    (tmp_path / "junk.tif").write_bytes(b"not a tiff")
    with pytest.raises(msx.MsxError):
        msx.tiff_read(tmp_path / "junk.tif")
