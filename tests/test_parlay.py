"""Kirkland's layer stacking sequence (msx_parlay, replacing parlay of
slicelib.cpp:558-713): the reference's own parser, compiled into oracle/_ref,
pins every code and index -- as committed golden vectors and, when the build
is present, live on random sequences.  Host code of libmsx.so: the golden
vectors run in both the CPU and the GPU suite."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import multislice_amd as msx
from conftest import BOTH_SUITES
from oracle import ref

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "parlay_vectors.json")


def ours(s, lmax, nsmax):
    out = np.zeros(max(nsmax, 1), np.int32)
    n = C.c_int32(0)
    rc = msx.load().msx_parlay(s.encode(), C.c_void_p(out.ctypes.data), nsmax, lmax, C.byref(n))
    return rc, out[:n.value].tolist()


@BOTH_SUITES
def test_golden_vectors(suite):
    with open(GOLD) as f:
        cases = json.load(f)["cases"]
    assert len(cases) == 20
    for c in cases:
        assert ours(c["sequence"], c["lmax"], c["nsmax"]) == (c["code"], c["layers"]), c["sequence"]


def test_random_sequences_against_reference():
    R = ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(8)
    alphabet = list("abcXYZ") + ["(", ")", " ", "\t", "2", "3", "10", "0", "-"]
    for _ in range(3000):
        s = "".join(rng.choice(alphabet, rng.integers(0, 14)))
        lmax = int(rng.choice([0, 2, 25, 51]))
        nsmax = int(rng.choice([5, 40, 400]))
        out = np.zeros(nsmax, np.int32)
        n = C.c_int(0)
        rc = R.ref_parlay(s.encode(), out, nsmax, lmax, C.byref(n))
        assert ours(s, lmax, nsmax) == (rc, out[:n.value].tolist()), repr(s)


@BOTH_SUITES
def test_python_wrapper(suite):
    assert msx.parlay("2(abc)d").tolist() == [0, 1, 2, 0, 1, 2, 3]
    assert len(msx.parlay("50(ab)")) == 100
    with pytest.raises(msx.MsxError) as e:
        msx.parlay("3(ab")
    assert e.value.code == -8
    with pytest.raises(msx.MsxError):
        msx.parlay("abc", nlayers=2)
