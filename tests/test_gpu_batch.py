"""Uneven HRTEM jobs batched into shared launches (msx_multislice_batch).

Jobs of mixed grids (including radix-3 sizes), voltages and slice counts:
every exit wave must equal, bit for bit, the same job run alone through
msx_multislice (itself checked against the oracle), and one job is checked
against the oracle directly.  The per-job FP32 error against an FP64 run of
the same batch is reported and bounded (1e-5 relative L2).
"""
import numpy as np
import pytest

import multislice_amd as msx
from conftest import max_rel, rel_l2
from oracle import Oracle as O

pytestmark = pytest.mark.gpu


def make_jobs(rng, njobs, sizes, dtype):
    import torch
    jobs = []
    for k in range(njobs):
        n = int(rng.choice(sizes))
        L = 0.1 * n
        kev = float(rng.choice([200.0, 300.0]))
        ns = int(rng.integers(3, 25))
        nl = min(ns, 4)
        V = np.abs(rng.standard_normal((nl, n, n))).cumsum(axis=1) % 25.0  # smooth-ish positive potentials
        Vt = torch.from_numpy(V.astype(np.float64 if dtype == "f64" else np.float32)).cuda()
        jobs.append(dict(V=Vt, n=n, ax=L, by=L, kev=kev, cz=np.full(nl, 2.0), layer=np.arange(ns) % nl))
    return jobs


def run_batch(ctx, jobs, cdtype):
    import torch
    out = []
    for j in jobs:
        j["trans"] = ctx.transmission(j["V"], j["ax"], j["by"], j["kev"])
        j["psi"] = torch.ones((j["n"], j["n"]), dtype=cdtype, device="cuda")
        out.append(j)
    ctx.multislice_batch(out)
    return [j["psi"].cpu().numpy() for j in out]


@pytest.mark.parametrize("rp,cp", [(1, 0), (0, 1), (1, 1)])
def test_batch_equals_single_jobs_on_tuned_plans(ctx, rp, cp):
    """With a tuning that selects the alternative FFT plans of a grid, the
    batched jobs (per-wave propagators: the work-queue column sweep) still
    equal each job run alone bit for bit."""
    import torch
    rng = np.random.default_rng(9)
    jobs = make_jobs(rng, 4, [512], "f32")
    jobs[0]["kev"], jobs[1]["kev"] = 200.0, 300.0  # two propagators: the work-queue sweep
    ctx.set_tuning(512, prec=msx.F32, tuning=dict(row_wpb=8, col_wpb=-1, batch=0, col_tile=8, row_plan=rp, col_plan=cp))
    try:
        got = run_batch(ctx, jobs, torch.complex64)
        for j, g in zip(jobs, got):
            psi = torch.ones((j["n"], j["n"]), dtype=torch.complex64, device="cuda")
            ctx.multislice(j["trans"], j["cz"], j["layer"], psi, j["ax"], j["by"], j["kev"])
            np.testing.assert_array_equal(g, psi.cpu().numpy())
    finally:
        ctx.set_tuning(512, prec=msx.F32, tuning=None)


def test_batch_equals_single_jobs(ctx):
    import torch
    rng = np.random.default_rng(4)
    jobs = make_jobs(rng, 14, [64, 128, 384], "f32")
    got = run_batch(ctx, jobs, torch.complex64)
    for j, g in zip(jobs, got):
        psi = torch.ones((j["n"], j["n"]), dtype=torch.complex64, device="cuda")
        ctx.multislice(j["trans"], j["cz"], j["layer"], psi, j["ax"], j["by"], j["kev"])
        np.testing.assert_array_equal(g, psi.cpu().numpy())
    # one job against the oracle (FP32 restatement)
    j = jobs[0]
    ref = O.multislice(j["trans"].cpu().numpy(), j["cz"], j["layer"], np.ones((j["n"], j["n"]), np.complex64),
                       j["ax"], j["by"], j["kev"], f32=True)
    ref = ref[0] if isinstance(ref, tuple) else ref
    assert np.linalg.norm(got[0] - ref) / np.linalg.norm(ref) < 1e-5


def test_batch_mixed_optics_on_one_grid(ctx):
    """Jobs on one grid with different voltages, tilts, slice thicknesses and
    slice counts share each slice step's launches (per-wave propagator
    tables, the column sweep's work counter): each exit wave equals its
    msx_multislice run bit for bit, in FP32 and FP64."""
    import torch
    rng = np.random.default_rng(12)
    for cdt, rdt in ((torch.complex64, np.float32), (torch.complex128, np.float64)):
        jobs = []
        for k in range(9):
            n, L = 256, 25.6
            ns = int(rng.integers(4, 30))
            nl = 3
            V = (np.abs(rng.standard_normal((nl, n, n))).cumsum(axis=2) % 20.0).astype(rdt)
            kev = float(rng.choice([100.0, 200.0, 300.0]))
            cz = rng.uniform(1.5, 2.5, nl)
            tilt = tuple(rng.uniform(-0.01, 0.01, 2)) if k % 3 else (0.0, 0.0)
            t = ctx.transmission(torch.from_numpy(V).cuda(), L, L, kev)
            jobs.append(dict(trans=t, ax=L, by=L, kev=kev, cz=cz, layer=rng.integers(0, nl, ns), tilt=tilt,
                             psi=torch.ones((n, n), dtype=cdt, device="cuda")))
        ctx.multislice_batch(jobs)
        for j in jobs:
            psi = torch.ones_like(j["psi"])
            ctx.multislice(j["trans"], j["cz"], j["layer"], psi, j["ax"], j["by"], j["kev"], tilt=j["tilt"])
            assert torch.equal(j["psi"], psi)


def test_batch_fp32_vs_fp64_per_job(ctx):
    import torch
    jobs32 = make_jobs(np.random.default_rng(7), 10, [128, 256, 384], "f32")
    jobs64 = make_jobs(np.random.default_rng(7), 10, [128, 256, 384], "f64")
    g32 = run_batch(ctx, jobs32, torch.complex64)
    g64 = run_batch(ctx, jobs64, torch.complex128)
    errs = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(g32, g64)]
    print("per-job FP32 vs FP64 rel-L2:", ", ".join(f"{e:.1e}" for e in errs))
    assert max(errs) < 1e-5


def test_amorphous_au_jobs_match_oracle(ctx):
    """Uneven jobs with their real specimens (amorphous carbon + an Au
    cluster, built by msx_amorphous, potentials and transmissions on the
    GPU): batched FP32 and FP64 exit waves against the oracle's FP64 pipeline
    (CPU potential, transmission and slice loop) on the same atoms."""
    import torch
    from multislice_amd import hrtem
    kw = dict(sizes=(128, 256), seed=21, max_au=12, slices=(8, 24))
    jobs32, specs = hrtem.uneven_jobs(ctx, np.random.default_rng(21), 3, **kw)
    jobs64, specs64 = hrtem.uneven_jobs(ctx, np.random.default_rng(21), 3, prec=msx.F64, **kw)
    outs = []
    for jobs, cdt in ((jobs32, torch.complex64), (jobs64, torch.complex128)):
        for j in jobs:
            j["psi"] = torch.ones((j["n"], j["n"]), dtype=cdt, device="cuda")
        ctx.multislice_batch(jobs)
        outs.append([j["psi"].cpu().numpy() for j in jobs])
    for j, (at, cell), (at2, _), a, b in zip(jobs32, specs, specs64, *outs):
        assert np.array_equal(at, at2) and np.any(at["Z"] == 6)
        n, ns = j["n"], len(j["layer"])
        V = O.potential(at["Z"], np.stack([at["x"], at["y"], at["z"]], 1), cell, n, n, ns, 3.0)
        t = O.transmission(V, cell[0], cell[1], j["kev"])
        ref = O.multislice(t, j["cz"], j["layer"], np.ones((n, n), np.complex128), cell[0], cell[1], j["kev"])
        e32, e64 = rel_l2(a, ref), rel_l2(b, ref)
        print(f"a-C + {int(np.sum(at['Z'] == 79))} Au, {n}^2, {ns} slices, {j['kev']:.0f} kV: "
              f"FP32 {e32:.1e}, FP64 {e64:.1e}")
        assert e32 < 1e-5
        assert e64 < 1e-10


@pytest.mark.timeout(600)
def test_large_grid_jobs_fp32_vs_fp64_and_oracle(ctx):
    """The uneven-job configuration's large grids with its slice range: jobs
    of 512^2, 768^2 and 1024^2 with 20-200 slices of 2 A (amorphous carbon +
    Au, 200/300 kV), batched, FP32 against the FP64 engine per job (rel-L2
    1e-5, max-rel 1e-4), bit-identical to the same job run alone; the 768^2
    job (the radix-3 grid) against the oracle's FP64 pipeline."""
    import torch
    from multislice_amd import hrtem
    kw = dict(sizes=(512, 768, 1024), seed=33, max_au=50, slices=(20, 200))
    jobs32, specs = hrtem.uneven_jobs(ctx, np.random.default_rng(33), 5, **kw)
    jobs64, _ = hrtem.uneven_jobs(ctx, np.random.default_rng(33), 5, prec=msx.F64, **kw)
    assert {j["n"] for j in jobs32} == {512, 768, 1024}  # seed 33 draws 1024, 512, 768, 1024, 512
    outs = []
    for jobs, cdt in ((jobs32, torch.complex64), (jobs64, torch.complex128)):
        for j in jobs:
            j["psi"] = torch.ones((j["n"], j["n"]), dtype=cdt, device="cuda")
        ctx.multislice_batch(jobs)
        outs.append([j["psi"].cpu().numpy() for j in jobs])
    for j, a, b in zip(jobs32, *outs):
        e, m = rel_l2(a, b), max_rel(a, b)
        print(f"{j['n']}^2, {len(j['layer'])} slices, {j['kev']:.0f} kV: FP32 vs FP64 rel-L2 {e:.1e} max-rel {m:.1e}")
        assert e < 1e-5 and m <= 1e-4
        psi = torch.ones((j["n"], j["n"]), dtype=torch.complex64, device="cuda")
        ctx.multislice(j["trans"], j["cz"], j["layer"], psi, j["ax"], j["by"], j["kev"])
        assert np.array_equal(psi.cpu().numpy(), a)
    for j, (at, cell), b in zip(jobs64, specs, outs[1]):
        if j["n"] != 768:
            continue
        n, ns = j["n"], len(j["layer"])
        V = O.potential(at["Z"], np.stack([at["x"], at["y"], at["z"]], 1), cell, n, n, ns, 3.0)
        t = O.transmission(V, cell[0], cell[1], j["kev"])
        ref = O.multislice(t, j["cz"], j["layer"], np.ones((n, n), np.complex128), cell[0], cell[1], j["kev"])
        e64 = rel_l2(b, ref)
        print(f"768^2 job, {ns} slices: FP64 engine vs oracle rel-L2 {e64:.1e}")
        assert e64 < 1e-10
        break


def test_batch_rejects_host_buffers(ctx):
    with pytest.raises(msx.MsxError):
        st, _ = msx.Context._stack(64, 64, 6.4, 6.4, 200.0, [2.0], [0])
        arr = (msx.engine.Job * 1)(msx.engine.Job(None, None, None, None))
        msx.engine._check(msx.load().msx_multislice_batch(ctx.h, arr, 1, msx.F32, msx.HOST))
