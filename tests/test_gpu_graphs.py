"""Runtime: slice loops replayed from cached HIP graphs (msx_set_graphs) give the same
bits as direct launches: every replay is checked against a direct run, with
direct work on another context between replays; a changed argument (another
transmission stack) captures a new graph instead of replaying a stale one.
The autotuner's launch shapes change only how waves are grouped into
work-groups: results stay bit-identical (a different STEM batch regroups the
detector sums, so it is compared to 1e-12)."""
import numpy as np
import pytest

import multislice_amd as msx
from multislice_amd import specimen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def si_stack():
    import torch
    ctx = msx.Context(0)
    at, cell = specimen.si_001((4, 4, 5))
    V = torch.empty((20, 256, 256), dtype=torch.float32, device="cuda")
    ctx.potential(at, cell, 256, 256, 20, 3.0, prec=msx.F32, out=V)
    t = ctx.transmission(V, cell[0], cell[1], 200.0)
    V2 = V * 1.1
    t2 = ctx.transmission(V2, cell[0], cell[1], 200.0)
    torch.cuda.synchronize()
    yield ctx, t, t2, cell
    ctx.close()


def test_multislice_graph_replays_match_direct(si_stack):
    import torch
    ctx_d, t, t2, cell = si_stack
    g = msx.Context(0)
    g.set_graphs(True)
    cz = np.full(20, cell[2] / 20)
    layer = np.arange(20)
    outs = []
    # one wave buffer each: the graph key holds the wave pointer
    psi_g = torch.empty((256, 256), dtype=torch.complex64, device="cuda")
    psi_d = torch.empty((256, 256), dtype=torch.complex64, device="cuda")
    for rep in range(4):
        tt = t if rep != 2 else t2
        psi_g.fill_(1)
        g.multislice(tt, cz, layer, psi_g, cell[0], cell[1], 200.0)
        psi_d.fill_(1)
        ctx_d.multislice(tt, cz, layer, psi_d, cell[0], cell[1], 200.0)  # direct launches in between
        assert torch.equal(psi_g, psi_d), f"replay {rep} differs"
        outs.append(psi_g.cpu())
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[3]) and not torch.equal(outs[0], outs[2])
    captures, replays = g.graph_stats()
    assert captures == 2 and replays == 2, (captures, replays)
    g.close()


def test_stem_graph_replays_match_direct(si_stack):
    ctx_d, t, _, cell = si_stack
    g = msx.Context(0)
    g.set_graphs(True)
    cz = np.full(20, cell[2] / 20)
    layer = np.arange(20)
    lam = msx.wavelength(200.0)
    det = [(0.0, 0.01 / lam), (0.04 / lam, 0.09 / lam)]
    probe = dict(alpha_max=0.02, cs=1e4)
    rng = np.random.default_rng(0)
    for rep in range(3):
        pos = rng.uniform(0, cell[0], (64, 2))  # same buffer sizes, new positions each call
        a = g.stem(t, cz, layer, cell[0], cell[1], 200.0, probe, pos, det, [10, 20], batch=64)
        b = ctx_d.stem(t, cz, layer, cell[0], cell[1], 200.0, probe, pos, det, [10, 20], batch=64)
        np.testing.assert_array_equal(a, b)
    captures, replays = g.graph_stats()
    assert captures == 1 and replays == 2, (captures, replays)
    g.close()


def test_stem_batches_of_one_call_share_a_graph(si_stack):
    """The launch batches of one call write their detector sums to one
    scratch block (copied into place afterwards), so every full batch replays
    the same graph; the short last batch captures its own."""
    ctx_d, t, _, cell = si_stack
    g = msx.Context(0)
    g.set_graphs(True)
    cz = np.full(20, cell[2] / 20)
    layer = np.arange(20)
    lam = msx.wavelength(200.0)
    det = [(0.0, 0.01 / lam), (0.04 / lam, 0.09 / lam)]
    probe = dict(alpha_max=0.02, cs=1e4)
    pos = np.random.default_rng(5).uniform(0, cell[0], (4 * 32 + 5, 2))
    a = g.stem(t, cz, layer, cell[0], cell[1], 200.0, probe, pos, det, [7, 20], batch=32)
    b = ctx_d.stem(t, cz, layer, cell[0], cell[1], 200.0, probe, pos, det, [7, 20], batch=32)
    np.testing.assert_array_equal(a, b)
    one = ctx_d.stem(t, cz, layer, cell[0], cell[1], 200.0, probe, pos[37:38], det, [7, 20], batch=1)
    np.testing.assert_array_equal(a[:, :, 37:38], one)
    captures, replays = g.graph_stats()
    assert captures == 2 and replays == 3, (captures, replays)
    g.close()


def test_autotune_keeps_results_and_reports_shapes(si_stack):
    ctx_d, t, _, cell = si_stack
    c = msx.Context(0)
    cz = np.full(20, cell[2] / 20)
    layer = np.arange(20)
    lam = msx.wavelength(200.0)
    det = [(0.0, 0.01 / lam), (0.04 / lam, 0.09 / lam)]
    probe = dict(alpha_max=0.02, cs=1e4)
    pos = np.random.default_rng(1).uniform(0, cell[0], (100, 2))
    ref = c.stem(t, cz, layer, cell[0], cell[1], 200.0, probe, pos, det, [20], batch=50)
    best = c.autotune(256, prec=msx.F32, max_batch=256)
    print("tuned 256^2 FP32:", best)
    assert best["row_wpb"] in (1, 2, 4, 8) and best["col_wpb"] in (1, 2, 4, 8, -1, -2)
    assert 16 <= best["batch"] <= 256 and best["ms_per_slice"] > 0
    assert best["col_tile"] in (16, 32)  # the built-in 16-column tile at 256^2 or its alternative
    assert best["row_plan"] in (0, 1) and best["col_plan"] == 0  # 256^2 has a second row plan only
    got = c.stem(t, cz, layer, cell[0], cell[1], 200.0, probe, pos, det, [20], batch=50)
    # another FFT plan is another factorisation (other rounding); launch
    # shapes alone leave every bit as it was
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
    c.set_tuning(256, prec=msx.F32, tuning=dict(best, row_plan=0, col_plan=0))
    np.testing.assert_array_equal(c.stem(t, cz, layer, cell[0], cell[1], 200.0, probe, pos, det, [20], batch=50), ref)
    # an explicit shape, and the tuned default batch
    c.set_tuning(256, prec=msx.F32, tuning=dict(row_wpb=1, col_wpb=8, batch=32))
    np.testing.assert_array_equal(c.stem(t, cz, layer, cell[0], cell[1], 200.0, probe, pos, det, [20], batch=50), ref)
    np.testing.assert_allclose(c.stem(t, cz, layer, cell[0], cell[1], 200.0, probe, pos, det, [20]), ref,
                               rtol=1e-12, atol=0)
    c.set_tuning(256, prec=msx.F32, tuning=None)
    c.close()


@pytest.mark.parametrize("n,tiles", [(256, (16, 32)), (512, (8, 4))])
def test_column_tile_choice_keeps_results(n, tiles):
    """Both column tile widths of a grid give the same exit waves (each column
    is transformed by the same code, which compiles to the same arithmetic in
    both tile instantiations, so bit for bit) and STEM
    detector sums equal to FP32 rounding (the two instantiations may fuse
    products into FMAs differently, and their FP64 partial sums are grouped
    differently)."""
    import torch
    c = msx.Context(0)
    rng = np.random.default_rng(n)
    nl = 3
    t = np.exp(1j * rng.random((nl, n, n)) * 2).astype(np.complex64)
    td = torch.from_numpy(t).cuda()
    cz, layer = np.full(nl, 2.0), np.arange(7) % nl
    a = 0.1 * n
    lam = msx.wavelength(300.0)
    det = [(0.0, 0.01 / lam), (0.02 / lam, 0.06 / lam)]
    pos = rng.uniform(0, a, (20, 2))
    waves, stems = [], []
    for tile in tiles:
        c.set_tuning(n, prec=msx.F32, tuning=dict(row_wpb=0, col_wpb=0, batch=8, col_tile=tile))
        psi = torch.ones((3, n, n), dtype=torch.complex64, device="cuda")
        c.multislice(td, cz, layer, psi, a, a, 300.0)
        waves.append(psi.cpu().numpy())
        stems.append(c.stem(td, cz, layer, a, a, 300.0, dict(alpha_max=0.02), pos, det, [4, 7], batch=8))
    np.testing.assert_array_equal(waves[0], waves[1])
    np.testing.assert_allclose(stems[0], stems[1], rtol=1e-6, atol=0)
    with pytest.raises(msx.MsxError):
        c.set_tuning(n, prec=msx.F32, tuning=dict(row_wpb=0, col_wpb=0, batch=8, col_tile=3))
    c.close()


@pytest.mark.parametrize("nx,ny,tile,prec", [(512, 512, 8, msx.F32), (256, 256, 32, msx.F32),
                                             (256, 16, 32, msx.F32), (512, 256, 0, msx.F64)])
def test_balanced_column_schedule_keeps_results(nx, ny, tile, prec):
    """The column sweep's balanced schedule (col_wpb < 0: one work-group per
    resident slot, (column group, wave) units dealt out in wave-major order)
    runs every column through the same code as the chunked schedule: exit
    waves, partial-coherence-style CTF sweeps and STEM detector sums (whose
    per-group partials it writes, zero for the groups outside the band) are
    bit-identical.  256 x 16 with 32-column tiles has lanes past the image
    edge, whose stores must be dropped."""
    import torch
    c = msx.Context(0)
    rng = np.random.default_rng(nx + ny)
    nl = 3
    cdt, ndt = (torch.complex64, np.complex64) if prec == msx.F32 else (torch.complex128, np.complex128)
    t = np.exp(1j * rng.random((nl, nx, ny)) * 2).astype(ndt)
    td = torch.from_numpy(t).cuda()
    cz, layer = np.full(nl, 2.0), np.arange(7) % nl
    a, b = 0.1 * nx, 0.1 * ny
    lam = msx.wavelength(300.0)
    det = [(0.0, 0.01 / lam), (0.02 / lam, 0.06 / lam)]
    pos = rng.uniform(0, min(a, b), (37, 2))
    psi0 = torch.from_numpy((rng.random((5, nx, ny)) + 1j * rng.random((5, nx, ny))).astype(ndt)).cuda()
    out = {}
    for cw in (4, 1, -1, -2):
        c.set_tuning(nx, ny, prec=prec, tuning=dict(row_wpb=0, col_wpb=cw, batch=16, col_tile=tile))
        psi = psi0.clone()
        c.multislice(td, cz, layer, psi, a, b, 300.0)
        stem = c.stem(td, cz, layer, a, b, 300.0, dict(alpha_max=0.02), pos, det, [3, 7], batch=16)
        out[cw] = (psi.cpu().numpy(), stem)
        print(f"{nx}x{ny} tile {tile} col_wpb {cw}: |psi| sum {np.abs(out[cw][0]).sum():.6e}", flush=True)
    for cw in (1, -1, -2):
        np.testing.assert_array_equal(out[cw][0], out[4][0])
        np.testing.assert_array_equal(out[cw][1], out[4][1])
    assert np.isfinite(out[4][0]).all() and np.isfinite(out[4][1]).all() and (out[4][1][:, 0] > 0).all()
    c.close()
