"""Write a small complex float TIFF with the reference's own floatTIFF
(compiled into oracle/_ref by oracle/build_ref.sh) and its expected contents.
Output: tests/golden/tiff/ref_12x5_complex.tif and .json.

    python tests/golden/make_tiff_fixture.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import ref  # noqa: E402


def main():
    R = ref()
    if R is None:
        raise SystemExit("oracle/_ref not built (bash oracle/build_ref.sh)")
    rng = np.random.default_rng(12)
    pix = rng.standard_normal((5, 24)).astype(np.float32)
    prm = np.zeros(512, np.float32)
    prm[1:8] = [200.0, 0.0251, 10.862, 10.862, 1.3578, 3.0, 0.6666667]
    out = os.path.join(HERE, "tiff")
    os.makedirs(out, exist_ok=True)
    rc = R.ref_tiff_write(os.fsencode(os.path.join(out, "ref_12x5_complex.tif")), pix.ravel(), 24, 5, 2, prm, 512,
                          float(pix[:, :12].min()), float(pix[:, :12].max()), float(pix[:, 12:].min()),
                          float(pix[:, 12:].max()), 0.2, 0.2)
    assert rc == 1, rc
    prm[0] = 2.0
    with open(os.path.join(out, "ref_12x5_complex.json"), "w") as f:
        json.dump({"source": "ImageSimulation/multislice/floatTIFF.cpp write(), compiled by oracle/build_ref.sh",
                   "pix": pix.tolist(), "params": prm[:8].tolist()}, f)
    print("wrote tiff/ref_12x5_complex.tif and .json")


if __name__ == "__main__":
    main()
