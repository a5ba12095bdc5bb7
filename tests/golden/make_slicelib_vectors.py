"""Write golden vectors from the reference's own slicelib.cpp (built into
oracle/_ref by oracle/build_ref.sh) and its vendored FFTW 2.1.5.

Output: tests/golden/slicelib_vectors.json (inputs and expected outputs
only).  Run in a checkout where oracle/_ref has been built:

    python tests/golden/make_slicelib_vectors.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import ref, ref_fftw2  # noqa: E402


def main():
    R = ref()
    F = ref_fftw2()
    if R is None or F is None:
        raise SystemExit("oracle/_ref not built (bash oracle/build_ref.sh)")
    out = {"source": "ImageSimulation/multislice/slicelib.cpp and ImageSimulation/fftw-2.1.5 "
                     "compiled by oracle/build_ref.sh"}
    kev = [60.0, 100.0, 120.0, 200.0, 300.0, 400.0]
    out["wavelength"] = [[k, R.ref_wavelength(k)] for k in kev]
    out["sigma"] = [[k, R.ref_sigma(k)] for k in kev]
    xs = [1e-4, 0.01, 0.3, 1.0, 1.999, 2.0, 2.001, 3.75, 3.76, 7.5, 20.0]
    out["bessi0"] = [[x, R.ref_bessi0(x)] for x in xs]
    out["bessk0"] = [[x, R.ref_bessk0(x)] for x in xs]
    Zs = [1, 6, 8, 14, 22, 26, 38, 79, 92, 103]
    rs = [0.0, 1e-6, 0.01, 0.05, 0.1, 0.3, 0.7, 1.5, 3.0, 6.0]
    out["vzatom"] = [[z, r, R.ref_vzatom(z, r)] for z in Zs for r in rs]
    k2s = [0.0, 0.01, 0.25, 1.0, 4.0, 16.0]
    out["featom"] = [[z, k2, R.ref_featom(z, k2)] for z in Zs for k2 in k2s]
    fq = []
    for n, a in [(256, 21.724), (257, 20.3), (384, 31.5), (16, 1.0)]:
        ko = np.zeros(n, np.float32)
        ko2 = np.zeros(n, np.float32)
        xo = np.zeros(n, np.float32)
        R.ref_freqn(ko, ko2, xo, n, a)
        fq.append({"n": n, "a": a, "ko": ko.tolist(), "ko2": ko2.tolist(), "xo": xo.tolist()})
    out["freqn"] = fq
    out["chi"] = [[cs, df, ax, ay, R.ref_chi_csdf(cs, df, ax, ay)]
                  for cs in (0.0, 1.2e7, 1e4) for df in (0.0, -600.0, 300.0)
                  for ax, ay in ((0.01, 0.0), (0.003, -0.004), (0.021, 0.001))]
    out["chi_multi"] = [[cs, df, c12a, c12b, c23a, c23b, ax, ay,
                         R.ref_chi_multi(cs, df, c12a, c12b, c23a, c23b, ax, ay)]
                        for cs, df in ((1.2e7, -600.0), (0.0, 0.0))
                        for c12a, c12b, c23a, c23b in ((0.0, 0.0, 0.0, 0.0), (-35.3, 48.5, 0.0, 0.0),
                                                       (0.0, 0.0, 812.0, -1410.0), (12.0, -7.0, -300.0, 95.0))
                        for ax, ay in ((0.01, 0.0), (0.003, -0.004), (-0.012, 0.007), (0.0, -0.015))]
    seed = C.c_ulong(12345)
    out["rangauss"] = {"seed": 12345, "values": [R.ref_rangauss(C.byref(seed)) for _ in range(16)]}
    rng = np.random.default_rng(7)
    x = (rng.standard_normal((8, 12)) + 1j * rng.standard_normal((8, 12))).astype(np.complex64)
    fw = x.copy()
    F.ref_fftw2_2d(fw.view(np.float32).reshape(-1), 8, 12, +1)
    out["fftw2_plus"] = {"nx": 8, "ny": 12, "in_re": x.real.tolist(), "in_im": x.imag.tolist(),
                         "out_re": fw.real.tolist(), "out_im": fw.imag.tolist()}
    with open(os.path.join(HERE, "slicelib_vectors.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote slicelib_vectors.json")


if __name__ == "__main__":
    main()
