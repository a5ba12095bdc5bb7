"""Time the reference's CPU slice loop: Kirkland's coherent loop
(mulslice.cpp:703-739) over the reference's own FFTW 2.1.5 (float), built
into oracle/_ref by build_ref.sh (ref_fftw2_loop), beside the oracle's FP32
loop (its textbook mixed-radix FFT, the proxy timed on machines without the
reference tree).  TEST INFRASTRUCTURE / MEASUREMENT ONLY.

    python -m oracle.time_ref_loop [--threads 8]

Workloads: the plain-HRTEM case (Si [001] 4x4x20 cells, 256^2, 80 slices of
a/4, 200 kV, one plane wave) and single probes of the batched-STEM case
(SrTiO3 [001] 12x12x50 cells, 512^2, 100 slices, 300 kV, 21 mrad, Cs 1 um),
one probe per thread.  Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import Oracle as O  # noqa: E402
from oracle.wrap import ref_fftw2 as fftw2  # noqa: E402


def _tables(nx, ny, ax, by, kev, cz):
    kx2 = O.freqn(nx, ax)[1]
    ky2 = O.freqn(ny, by)[1]
    px, py = [], []
    for d in cz:
        a, b = O.propagator(nx, ny, ax, by, kev, d)
        px.append(a)
        py.append(b)
    px, py = np.array(px).astype(np.complex64), np.array(py).astype(np.complex64)
    return kx2, ky2, px, py, np.float32(O.k2max(nx, ny, ax, by))


def run_ref(trans, cz, layer, waves, ax, by, kev, nthreads, measure=False):
    """(exit waves, planning seconds, loop seconds)"""
    L = fftw2()
    if L is None or not hasattr(L, "ref_fftw2_loop"):
        raise SystemExit("oracle/_ref/libfftw2f_ref.so not built (bash oracle/build_ref.sh)")
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L.ref_fftw2_loop.restype = C.c_int
    L.ref_fftw2_loop.argtypes = [fp, C.c_int, ip, C.c_int, C.c_int, C.c_int, fp, fp, fp, fp, fp, fp, C.c_float, fp,
                                 C.c_int, C.c_int, C.c_void_p, C.c_int, np.ctypeslib.ndpointer(np.float64)]
    nl, nx, ny = trans.shape
    kx2, ky2, px, py, k2max = _tables(nx, ny, ax, by, kev, cz)
    t = np.ascontiguousarray(trans, np.complex64).view(np.float32).reshape(-1)
    w = np.ascontiguousarray(waves, np.complex64).copy()
    lay = np.ascontiguousarray(layer, np.int32)
    f = lambda a: np.ascontiguousarray(a, np.float32).reshape(-1)  # noqa: E731
    sec = np.zeros(2)
    rc = L.ref_fftw2_loop(t, nl, lay, len(lay), nx, ny, f(px.real), f(px.imag), f(py.real), f(py.imag), kx2, ky2,
                          k2max, w.view(np.float32).reshape(-1), len(w), nthreads, None, int(measure), sec)
    if rc != 0:
        raise RuntimeError("ref_fftw2_loop failed")
    return w, sec[0], sec[1]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args(argv)
    from multislice_amd import specimen
    out = {"cpu": os.uname().machine, "cores_available": os.cpu_count()}
    # plain HRTEM: Si [001] 4x4x20, 256^2, 80 slices, 200 kV, one plane wave
    # (the 80 slices are the 4 atomic planes of one cell, 20 times over)
    n = 256
    at1, cell1 = specimen.si_001((4, 4, 1))
    V = O.potential(at1["Z"], np.stack([at1["x"], at1["y"], at1["z"]], 1), cell1, n, n, 4, 4.0)
    t = O.transmission(V.astype(np.float32), cell1[0], cell1[1], 200.0, f32=True)
    cz = np.full(4, 5.431 / 4)
    layer = np.tile(np.arange(4), 20)
    w0 = np.ones((1, n, n), np.complex64)
    ref = O.multislice(t, cz, layer, w0[0], cell1[0], cell1[1], 200.0, f32=True)
    w, tp, dt = run_ref(t, cz, layer, w0, cell1[0], cell1[1], 200.0, 1)
    _, tpm, dtm = run_ref(t, cz, layer, w0, cell1[0], cell1[1], 200.0, 1, measure=True)
    t0 = time.perf_counter()
    O.multislice(t, cz, layer, w0[0], cell1[0], cell1[1], 200.0, f32=True)
    dto = time.perf_counter() - t0
    out["hrtem_si_256_80slices"] = {
        "fftw2_estimate_ms_per_slice_1thread": round(1e3 * dt / 80, 3),
        "fftw2_measure_ms_per_slice_1thread": round(1e3 * dtm / 80, 3), "fftw2_measure_planning_s": round(tpm, 1),
        "oracle_ms_per_slice_1thread": round(1e3 * dto / 80, 3),
        "exit_wave_rel_l2_fftw2_vs_oracle_f32": float(np.linalg.norm(w[0] - ref) / np.linalg.norm(ref))}
    # batched STEM, single probes: SrTiO3 12x12x50 (two layers), 512^2, 100 slices, 300 kV
    at, cell = specimen.srtio3_001((12, 12, 1))
    n = 512
    V = O.potential(at["Z"], np.stack([at["x"], at["y"], at["z"]], 1), cell, n, n, 2, 3.0)
    t = O.transmission(V.astype(np.float32), cell[0], cell[1], 300.0, f32=True)
    cz = np.full(2, 3.905 / 2)
    layer = np.arange(100) % 2
    res = {}
    for nt in sorted({1, a.threads}):
        probes = np.stack([O.probe(n, n, cell[0], cell[1], 300.0, 0.021, cs=1e4, pos=(20.0 + 0.1 * k, 20.5))
                           for k in range(nt)]).astype(np.complex64)
        w, tp, dt = run_ref(t, cz, layer, probes, cell[0], cell[1], 300.0, nt)
        res[f"fftw2_estimate_probe_slices_per_s_{nt}threads"] = round(nt * 100 / dt, 2)
        if nt == 1:
            t0 = time.perf_counter()
            o = O.multislice(t, cz, layer, probes[0], cell[0], cell[1], 300.0, f32=True)
            res["oracle_probe_slices_per_s_1thread"] = round(100 / (time.perf_counter() - t0), 2)
            res["exit_wave_rel_l2_fftw2_vs_oracle_f32"] = float(np.linalg.norm(w[0] - o) / np.linalg.norm(o))
    out["stem_srtio3_512_100slices_single_probes"] = res
    print(json.dumps(out))


if __name__ == "__main__":
    main()
