"""Write golden vectors of Kirkland's stacking-sequence parser from the
reference's own slicelib.cpp (parlay, built into oracle/_ref by
oracle/build_ref.sh): sequences, lmax, nsmax and the expected code and
layer indices.  Output: tests/golden/parlay_vectors.json.

    python tests/golden/make_parlay_vectors.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import ref  # noqa: E402

CASES = [
    ("2(abc)d", 25, 100), ("5( 2(abc) 3(cba) )", 25, 1000), ("a\tb c", 25, 10), ("ab ", 25, 10),
    ("10(ab)", 1, 100), ("10(ab)", 0, 100), ("3(a", 25, 10), ("a)", 25, 10), ("2 x", 25, 10),
    ("a-b", 25, 10), ("Zz", 51, 10), ("Z", 50, 10), ("0(ab)c", 25, 10), ("1(ab)c", 25, 10),
    ("4(abcd)", 25, 10), ("12 (a2(bc))", 25, 200), ("", 25, 10), ("()", 25, 10), ("2(3(4(a)))", 0, 100),
    ("abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ", 51, 100),
]


def run(R, s, lmax, nsmax):
    out = np.zeros(max(nsmax, 1), np.int32)
    n = C.c_int(0)
    rc = R.ref_parlay(s.encode(), out, nsmax, lmax, C.byref(n))
    return rc, out[:n.value].tolist()


def main():
    R = ref()
    if R is None:
        raise SystemExit("oracle/_ref not built (bash oracle/build_ref.sh)")
    rows = []
    for s, lmax, nsmax in CASES:
        rc, layers = run(R, s, lmax, nsmax)
        rows.append({"sequence": s, "lmax": lmax, "nsmax": nsmax, "code": rc, "layers": layers})
    with open(os.path.join(HERE, "parlay_vectors.json"), "w") as f:
        json.dump({"source": "ImageSimulation/multislice/slicelib.cpp parlay, compiled by oracle/build_ref.sh",
                   "cases": rows}, f, indent=0)
    print("wrote parlay_vectors.json")


if __name__ == "__main__":
    main()
