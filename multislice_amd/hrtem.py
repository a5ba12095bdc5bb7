"""HRTEM with frozen phonons, and batches of uneven HRTEM jobs.

`frozen_phonon_image` is the large-HRTEM path of the reference (LMultiSlice
imageCalculation per configuration, ModelSimulated.cu:21-192, with Kirkland's
objective lens, mulslice.cpp:612-656): every configuration's atoms are
displaced on the GPU (msx_displace), binned and splatted into its potential
there (msx_potential_device_atoms) and turned into its transmission stack;
all configurations advance through one batched slice loop
(msx_multislice_batch: one wave per configuration, each with its own
transmission), and the objective-lens images are averaged.

    python -m multislice_amd.hrtem c1      # Si 4x4x20, 256^2, 80 slices, one plane wave (ms/slice)
    python -m multislice_amd.hrtem c2      # SrTiO3 52x52x50, 2048^2, 100 slices, 16 configurations
    python -m multislice_amd.hrtem c4      # 256 uneven jobs, batched vs one by one
    python -m multislice_amd.hrtem pc      # partial-coherence images, Si 256^2 and 512^2, 80 slices

print one JSON line of timings each.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import engine, specimen


def _sync():
    import torch
    torch.cuda.synchronize()
    return time.perf_counter()


def frozen_phonon_image(ctx, atoms, cell, n, nslices, kev, u_rms, nconf, seed, cs, df, aobj, radius=3.0,
                        prec=engine.F32):
    """Mean objective-lens image (n x n, torch float64 on the GPU) over nconf
    frozen-phonon configurations, and the seconds spent per stage."""
    import torch
    rdt, cdt = (torch.float64, torch.complex128) if prec == engine.F64 else (torch.float32, torch.complex64)
    cz = np.full(nslices, cell[2] / nslices)
    layer = np.arange(nslices)
    atd = torch.from_numpy(np.ascontiguousarray(atoms, engine.ATOM_DTYPE).view(np.uint8).copy()).cuda()
    atc = torch.empty_like(atd)
    V = torch.empty((nslices, n, n), dtype=rdt, device="cuda")
    tm = {}
    t0 = _sync()
    trans = []
    for c in range(nconf):
        at = engine.displace(atd, u_rms, seed, c, ctx=ctx, out=atc)
        ctx.potential(at, cell, n, n, nslices, radius, prec=prec, out=V)
        trans.append(ctx.transmission(V, cell[0], cell[1], kev))
    t1 = _sync()
    tm["displace_potential_transmission"] = t1 - t0
    psis = [torch.ones((n, n), dtype=cdt, device="cuda") for _ in range(nconf)]
    t1 = _sync()
    ctx.multislice_batch([dict(trans=t, psi=p, ax=cell[0], by=cell[1], kev=kev, cz=cz, layer=layer)
                          for t, p in zip(trans, psis)])
    t2 = _sync()
    tm["multislice_batched"] = t2 - t1
    img = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    for p in psis:
        img += ctx.ctf_image(p, cell[0], cell[1], kev, cs, df, aobj).double()
    img /= nconf
    tm["ctf_average"] = _sync() - t2
    return img, tm


def c1(ctx, reps=20):
    """The small-HRTEM configuration: Si [001] 4x4x20 cells (2560 atoms at
    exact sites), 256^2, 80 slices of a/4, 200 kV, one plane wave: the
    potential and transmission build, then the slice loop of one wave
    (launch-latency bound), timed with direct launches and replayed from a
    cached HIP graph (mean of reps runs after a warm-up)."""
    import torch
    at, cell = specimen.si_001((4, 4, 20))
    n, ns, kev = 256, 80, 200.0
    V = torch.empty((ns, n, n), dtype=torch.float32, device="cuda")
    t0 = _sync()
    ctx.potential(at, cell, n, n, ns, 3.0, prec=engine.F32, out=V)
    t = ctx.transmission(V, cell[0], cell[1], kev)
    t_build = _sync() - t0
    cz, layer = np.full(ns, cell[2] / ns), np.arange(ns)
    psi = torch.empty((n, n), dtype=torch.complex64, device="cuda")
    out = {}
    for graphs in (False, True):
        ctx.set_graphs(graphs)
        psi.fill_(1)
        ctx.multislice(t, cz, layer, psi, cell[0], cell[1], kev)
        t0 = _sync()
        for _ in range(reps):
            psi.fill_(1)
            ctx.multislice(t, cz, layer, psi, cell[0], cell[1], kev)
        dt = (_sync() - t0) / reps
        out["graphs" if graphs else "direct"] = {"ms_per_run": round(1e3 * dt, 3),
                                                 "ms_per_slice": round(1e3 * dt / ns, 4)}
    ctx.set_graphs(False)
    return {"workload": "C1 Si [001] 4x4x20 cells (2560 atoms), 256^2, 80 slices of 1.35775 A, 200 kV, one "
                        "plane wave", "build_ms": round(1e3 * t_build, 2), "slice_loop": out,
            "mean_exit_intensity": float((psi.abs() ** 2).mean())}


def c2(ctx):
    """The large-HRTEM configuration: SrTiO3 [001] 52x52x50 cells, 2048^2,
    100 slices of 1.9525 A, 300 kV, 16 configurations, Cs 1.2 mm, defocus
    -60 nm, 20 mrad aperture (one warm-up run, then a timed run)."""
    at, cell = specimen.srtio3_001((52, 52, 50))
    args = (ctx, at, cell, 2048, 100, 300.0, specimen.SRTIO3_U, 16, 2024, 1.2e7, -600.0, 0.020)
    frozen_phonon_image(*args)
    img, tm = frozen_phonon_image(*args)
    ms = tm["multislice_batched"]
    return {"workload": "C2 frozen-phonon HRTEM, SrTiO3 52x52x50 (676,000 atoms), 2048^2, 100 slices, 16 configs",
            "seconds": {k: round(v, 4) for k, v in tm.items()}, "total_s": round(sum(tm.values()), 3),
            "ms_per_wave_slice": round(1e3 * ms / (16 * 100), 4), "image_mean": float(img.mean())}


def uneven_specimen(rng, n, nslices, seed, max_au=50):
    """Specimen of one uneven job: amorphous carbon (2.0 g/cm^3, random
    sequential addition with 1.4 A minimum spacing, msx_amorphous) filling a
    periodic 0.1n x 0.1n x 2*nslices A box, plus 0..max_au Au atoms of an fcc
    cluster (a = 4.078 A) at a random centre; carbon atoms within 1.4 A of a
    gold atom are removed.  Returns (atoms, cell)."""
    L, c = 0.1 * n, 2.0 * nslices
    box = np.array([L, L, c])
    at = engine.amorphous((L, L, c), seed=seed)
    nau = int(rng.integers(0, max_au + 1))
    if nau:
        au = specimen.au_cluster(nau, rng.random(3) * box)
        pa = np.stack([au["x"], au["y"], au["z"]], 1) % box
        au["x"], au["y"], au["z"] = pa[:, 0], pa[:, 1], pa[:, 2]
        pc = np.stack([at["x"], at["y"], at["z"]], 1)
        keep = np.ones(len(at), bool)
        for q in pa:
            d = pc - q
            d -= box * np.round(d / box)
            keep &= (d * d).sum(1) >= 1.4 * 1.4
        at = np.concatenate([at[keep], au])
    return at, (L, L, c)


def uneven_jobs(ctx, rng, njobs, sizes=(256, 384, 512, 768, 1024), seed=5, prec=engine.F32, max_au=50,
                slices=(20, 200)):
    """The uneven-job set: grid n from `sizes`, `slices` (20-200) slices of 2 A, 200 or
    300 kV, each job its own amorphous-carbon + Au specimen (uneven_specimen);
    potentials (radius 3 A) and transmission stacks are built on the GPU, one
    layer per slice.  Returns job dicts (trans on the GPU) and their atoms."""
    import torch
    rdt = torch.float64 if prec == engine.F64 else torch.float32
    jobs, specs = [], []
    for k in range(njobs):
        n = int(rng.choice(sizes))
        ns = int(rng.integers(slices[0], slices[1] + 1))
        kev = float(rng.choice([200.0, 300.0]))
        at, cell = uneven_specimen(rng, n, ns, seed * 1000003 + k, max_au)
        V = torch.empty((ns, n, n), dtype=rdt, device="cuda")
        ctx.potential(at, cell, n, n, ns, 3.0, prec=prec, out=V)
        trans = ctx.transmission(V, cell[0], cell[1], kev)
        del V
        jobs.append(dict(n=n, trans=trans, ax=cell[0], by=cell[1], kev=kev, cz=np.full(ns, 2.0), layer=np.arange(ns)))
        specs.append((at, cell))
    return jobs, specs


def c4(ctx, njobs=256, seed=5, f64_every=16):
    """Uneven jobs: one msx_multislice_batch call against one msx_multislice
    call per job (same jobs, same results), after a warm-up of each; then the
    FP32 exit waves of every f64_every-th job against an FP64 run of the same
    job (FP64 potential, transmission and slice loop)."""
    import torch
    t0 = _sync()
    jobs, specs = uneven_jobs(ctx, np.random.default_rng(seed), njobs, seed=seed)
    t_build = _sync() - t0
    wave_slices = sum(len(j["layer"]) for j in jobs)
    natoms = sum(len(a) for a, _ in specs)

    def batched():
        for j in jobs:
            j["psi"] = torch.ones((j["n"], j["n"]), dtype=torch.complex64, device="cuda")
        t0 = _sync()
        ctx.multislice_batch(jobs)
        return _sync() - t0

    def one_by_one():
        t0 = _sync()
        for j in jobs:
            psi = torch.ones((j["n"], j["n"]), dtype=torch.complex64, device="cuda")
            ctx.multislice(j["trans"], j["cz"], j["layer"], psi, j["ax"], j["by"], j["kev"])
        return _sync() - t0

    batched(), one_by_one()
    tb, ts = batched(), one_by_one()
    errs = []
    for k in range(0, njobs, f64_every):
        j = jobs[k]
        at, cell = specs[k]
        n, ns = j["n"], len(j["layer"])
        V = torch.empty((ns, n, n), dtype=torch.float64, device="cuda")
        ctx.potential(at, cell, n, n, ns, 3.0, prec=engine.F64, out=V)
        t64 = ctx.transmission(V, cell[0], cell[1], j["kev"])
        del V
        psi = torch.ones((n, n), dtype=torch.complex128, device="cuda")
        ctx.multislice(t64, j["cz"], j["layer"], psi, cell[0], cell[1], j["kev"])
        del t64
        errs.append(float(torch.linalg.vector_norm(j["psi"].to(torch.complex128) - psi) / torch.linalg.vector_norm(psi)))
    return {"workload": f"C4 {njobs} uneven HRTEM jobs: amorphous C (2.0 g/cm^3, 1.4 A spacing) + 0-50 Au, "
                        f"n in 256..1024 at 0.1 A/px, 20-200 slices of 2 A, 200/300 kV",
            "atoms": natoms, "wave_slices": wave_slices, "build_s": round(t_build, 2), "batched_s": round(tb, 3),
            "one_by_one_s": round(ts, 3), "speedup": round(ts / tb, 2),
            "wave_slices_per_s_batched": round(wave_slices / tb),
            "fp32_vs_fp64_rel_l2": {"jobs": len(errs), "max": max(errs), "median": float(np.median(errs))}}


def _pc_case(ctx, cells, n, reps):
    import torch
    at, cell = specimen.si_001(cells)
    ns, kev = 4 * cells[2], 200.0
    V = torch.empty((ns, n, n), dtype=torch.float32, device="cuda")
    ctx.potential(at, cell, n, n, ns, 3.0, out=V)
    t = ctx.transmission(V, cell[0], cell[1], kev)
    del V
    lens = dict(cs=1.3e7, df=700.0, aobj=0.024, dfa2=30.0, dfa2phi=0.5)
    args = (t, np.full(ns, cell[2] / ns), np.arange(ns), cell[0], cell[1], kev, lens, 0.0, 2.5e-3, 40.0, 10.0)
    img, na, nd = ctx.partial_coherence(*args)
    t0 = _sync()
    for _ in range(reps):
        img, na, nd = ctx.partial_coherence(*args)
    dt = (_sync() - t0) / reps
    ctx.profile(True)
    ctx.partial_coherence(*args)
    prof = {}
    for name, kind in (("row_sweeps", engine.PROF_ROW), ("column_sweeps", engine.PROF_COL)):
        nl, ms, by = ctx.profile_read(kind)
        prof[name] = {"launches": nl, "ms": round(ms, 3), "GB_per_s": round(by / ms / 1e6, 1) if ms else None}
    ctx.profile(False)
    return {"workload": f"Si [001] {cells[0]}x{cells[1]}x{cells[2]}, {n}^2, {ns} slices, 200 kV, "
                        f"{na} illumination angles x {nd} defocus steps",
            "ms": round(1e3 * dt, 3), "ms_per_angle": round(1e3 * dt / na, 4),
            "wave_slices_per_s": round(na * ns / dt), "lens_images_per_s": round(na * nd / dt),
            "image_mean": float(img.mean()), "kernels": prof}


def pc(ctx, reps=5):
    """Partial-coherence images (msx_partial_coherence) of Si [001] slabs
    (slices of a/4, 200 kV): illumination cone 0-2.5 mrad, defocus spread
    40 A sampled every 10 A (21 steps), Cs 1.3 mm, defocus 700 A, 2-fold
    astigmatism 30 A, 24 mrad aperture.  The 4x4x20 slab on 256^2 (13 grid
    angles) and the 8x8x20 slab on 512^2 (more angles on the finer angular
    grid); one warm-up, then the mean of `reps` timed calls, and the sweep
    kernel times from the profiler on a separate call."""
    return {"cases": [_pc_case(ctx, (4, 4, 20), 256, reps), _pc_case(ctx, (8, 8, 20), 512, reps)]}


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m multislice_amd.hrtem")
    ap.add_argument("workload", choices=["c1", "c2", "c4", "pc"])
    a = ap.parse_args(argv)
    ctx = engine.Context(0)
    run = {"c1": c1, "c2": c2, "c4": c4, "pc": pc}[a.workload]
    print(json.dumps(run(ctx)), flush=True)


if __name__ == "__main__":
    main()
    sys.exit(0)
