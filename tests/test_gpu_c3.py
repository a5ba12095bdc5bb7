"""Parity of the batched-STEM benchmark workload itself, at its full size and
with its launch shapes: SrTiO3 [001] 12x12x50 cells (46.86 A, 36,000 atoms,
frozen-phonon configuration 0 of the benchmark's seed), 512^2, 100 slices of
1.9525 A, 300 kV, 21 mrad, Cs 1 um, BF 0-10 mrad and ADF 70-200 mrad (the
ADF band clipped at the 2/3 band limit, 71.7 mrad, on this grid), launch
shapes 8 waves per row work-group / balanced column schedule / 128 probes
per launch / 8-column tiles, HIP graphs on.  136 probes of the benchmark's
scan make one call run a full 128-probe batch and a partial one.

The engine's FP32 chain from atoms (FP32 potential from the radial tables,
FP32 transmission, FP32 STEM) is compared with the oracle's FP64 chain
(oracle potential, transmission and STEM of Musli ConvergedBeam / SaveSTEM,
MSProject.cs:458-514,792-842, MSUseOnceCalculator.cs:125-277) on 8 probes
drawn from both batches; every probe of the FP32 run is compared with the
FP64 engine on the same stack, and the FP64 engine with the oracle.

Bars: the FP64 engine within 1e-11 of the oracle; FP32 detector images
within 1e-4 max-rel (max |got - ref| / max |ref| per detector, the image
bar); and, because FP32 rounding at this depth (100 slices of a strongly
scattering crystal) reaches 1e-5, the FP32 engine's error against FP64 no
larger than that of the reference's own FP32 loop (the oracle's statement-
for-statement Kirkland float restatement) on the same probes and waves."""
import numpy as np
import pytest
import torch

import multislice_amd as msx
from multislice_amd import specimen
from oracle import Oracle as O

pytestmark = pytest.mark.gpu

A, CELLS, N, NSL, KEV, ALPHA, CS = 3.905, (12, 12, 50), 512, 100, 300.0, 0.021, 1.0e4
SCAN = 128
PICK = [0, 45, 91, 127, 128, 131, 133, 135]  # probes of both launch batches
TUNING = dict(row_wpb=8, col_wpb=-1, batch=128, col_tile=8)


def scan_positions(n):
    x0, step = 5 * A, 2 * A / SCAN
    ix, iy = np.meshgrid(np.arange(SCAN), np.arange(SCAN), indexing="ij")
    return np.stack([x0 + ix.ravel() * step, x0 + iy.ravel() * step], 1)[:n]


def detectors():
    lam = msx.wavelength(KEV)
    return [(0.0, 0.010 / lam), (0.070 / lam, 0.200 / lam)]


def max_rel(got, ref):
    """per detector: max |got - ref| / max |ref| over the probes (last checkpoint)"""
    g, r = np.asarray(got)[-1], np.asarray(ref)[-1]
    return np.abs(g - r).max(axis=1) / np.abs(r).max(axis=1)


@pytest.fixture(scope="module")
def c3():
    at0, cell = specimen.srtio3_001(CELLS)
    at = msx.displace(at0, specimen.SRTIO3_U, seed=1234, config=0)
    xyz = np.stack([at["x"], at["y"], at["z"]], 1)
    V64 = O.potential(at["Z"], xyz, cell, N, N, NSL, 3.0)
    t64 = O.transmission(V64, cell[0], cell[1], KEV)
    pos = scan_positions(136)
    cz, layer = np.full(NSL, A / 2), np.arange(NSL)
    ref = O.stem(t64, cz, layer, cell[0], cell[1], KEV, ALPHA, pos[PICK], detectors(), [NSL], cs=CS, nthreads=16)
    ref32 = O.stem(t64.astype(np.complex64), cz, layer, cell[0], cell[1], KEV, ALPHA, pos[PICK], detectors(), [NSL],
                   cs=CS, nthreads=16, f32=True)
    return at, cell, t64, pos, ref, ref32


@pytest.fixture(scope="module")
def bench_ctx():
    c = msx.Context(0)
    c.set_tuning(N, N, msx.F32, TUNING)
    c.set_graphs(True)
    yield c
    c.close()


def test_c3_fp32_chain_from_atoms_matches_oracle(bench_ctx, c3):
    at, cell, t64, pos, ref, ref32 = c3
    cz, layer = np.full(NSL, A / 2), np.arange(NSL)
    V = torch.empty((NSL, N, N), dtype=torch.float32, device="cuda")
    bench_ctx.potential(at, cell, N, N, NSL, 3.0, table=msx.TABLE_LMULTISLICE, formula=msx.FORMULA_KIRKLAND,
                        prec=msx.F32, out=V)
    t = bench_ctx.transmission(V, cell[0], cell[1], KEV)
    out = [bench_ctx.stem(t, cz, layer, cell[0], cell[1], KEV, dict(alpha_max=ALPHA, cs=CS), pos, detectors(),
                          [NSL]) for _ in range(2)]
    caps, replays = bench_ctx.graph_stats()
    assert caps >= 1 and replays >= 1  # the graph path ran (the second call replays)
    assert np.array_equal(out[0], out[1])
    mr = max_rel(out[1][:, :, PICK], ref)
    print(f"C3 FP32 chain from atoms vs oracle FP64 (8 probes): BF max-rel {mr[0]:.2e}, ADF max-rel {mr[1]:.2e}")
    assert mr.max() <= 1e-4


def test_c3_stack_fp32_vs_fp64_engine_and_oracle(bench_ctx, c3):
    at, cell, t64, pos, ref, ref32 = c3
    cz, layer = np.full(NSL, A / 2), np.arange(NSL)
    t32 = torch.from_numpy(t64.astype(np.complex64)).cuda()
    got32 = bench_ctx.stem(t32, cz, layer, cell[0], cell[1], KEV, dict(alpha_max=ALPHA, cs=CS), pos, detectors(),
                           [NSL])
    del t32
    td = torch.from_numpy(t64).cuda()
    got64 = bench_ctx.stem(td, cz, layer, cell[0], cell[1], KEV, dict(alpha_max=ALPHA, cs=CS), pos, detectors(),
                           [NSL])
    del td
    e64 = max_rel(got64[:, :, PICK], ref)
    e32 = max_rel(got32[:, :, PICK], ref)
    eref = max_rel(ref32, ref)
    eall = max_rel(got32, got64)
    print(f"C3 FP64 engine vs oracle: max-rel {e64.max():.2e}; FP32 engine vs oracle FP64: BF {e32[0]:.2e} "
          f"ADF {e32[1]:.2e} (the reference's FP32 loop: BF {eref[0]:.2e} ADF {eref[1]:.2e}); FP32 vs FP64 engine, "
          f"all 136 probes: BF {eall[0]:.2e} ADF {eall[1]:.2e}")
    assert e64.max() <= 1e-11
    assert e32.max() <= 1e-4 and eall.max() <= 1e-4
    assert (e32 <= 1.1 * eref).all()


def test_c3_exit_wave_fp32_error(c3):
    """One probe's exit wave through the 100 slices (msx_multislice): the
    FP32 engine against the FP64 oracle, beside the reference FP32 loop's."""
    at, cell, t64, pos, ref, ref32 = c3
    cz, layer = np.full(NSL, A / 2), np.arange(NSL)
    probe = O.probe(N, N, cell[0], cell[1], KEV, ALPHA, cs=CS, pos=tuple(pos[PICK[1]]))
    w64 = O.multislice(t64, cz, layer, probe, cell[0], cell[1], KEV)
    w32o = O.multislice(t64.astype(np.complex64), cz, layer, probe.astype(np.complex64), cell[0], cell[1], KEV,
                        f32=True)
    ctx = msx.Context(0)
    w = probe[None].astype(np.complex64)
    ctx.multislice(t64.astype(np.complex64), cz, layer, w, cell[0], cell[1], KEV)
    ctx.close()
    e = np.linalg.norm(w[0] - w64) / np.linalg.norm(w64)
    eo = np.linalg.norm(w32o - w64) / np.linalg.norm(w64)
    print(f"C3 exit wave (100 slices): FP32 engine rel-L2 {e:.2e} vs FP64 oracle; the reference's FP32 loop {eo:.2e}")
    assert e <= 1.1 * eo
    assert e <= 1.5e-5
