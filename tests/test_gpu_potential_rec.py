"""Reciprocal-space (msx_potential_rec) and real-space (msx_potential_rs)
Doyle-Turner potentials and the absorptive
transmission (msx_transmission_abs) on the GPU against the oracle.

The oracle gives the Fourier coefficients of each slice (Potential_rec_DT_DW
restated, FP64); numpy's FP64 inverse FFT turns them into V(r).  FP64 engine
results agree to 1e-11 relative L2 (FP64 FFT rounding); FP32 outputs to 1e-6
(float storage of the FP64 result).
"""
import numpy as np
import pytest

import multislice_amd as msx
from multislice_amd.specimen import atoms_array
from oracle import Oracle as O

pytestmark = pytest.mark.gpu

KVA2 = 300.8242834 / (2 * np.pi)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def oracle_rec(at, a, b, c, gam, nmult, mmult, nx, ny, nslices, dw, ads):
    dz = c / nslices
    V = np.zeros((nslices, nx, ny))
    Va = np.zeros((nslices, nx, ny))
    kz = np.floor(at["z"] / dz).astype(int)
    xf = (at["x"] - at["y"] * np.cos(gam) / np.sin(gam)) / a
    yf = at["y"] / (b * np.sin(gam))
    for s in range(nslices):
        m = kz == s
        if not m.any():
            continue
        g = O.potential_rec_g(at["Z"][m], xf[m] % 1.0, yf[m] % 1.0, a, b, nx, ny, gam=gam, nmult=nmult,
                              mmult=mmult, dw=dw[m], ads=ads[m], occ=at["occ"][m].astype(np.float64))
        v = np.fft.ifft2(g) * (nx * ny)  # [iy][ix]
        V[s], Va[s] = KVA2 * v.real.T, KVA2 * v.imag.T
    return V, Va


def random_atoms(rng, n, a, b, c, zs):
    xyz = rng.random((n, 3)) * [a, b, c]
    at = atoms_array(rng.choice(zs, n), xyz)
    at["occ"] = rng.uniform(0.5, 1.0, n)
    return at


CASES = [
    # a, b, c, gam(deg), nmult, mmult, nx, ny, nslices, natoms
    (9.0, 9.0, 6.0, 90.0, 1, 1, 64, 64, 3, 40),
    (5.431, 5.431, 5.431, 90.0, 2, 2, 64, 64, 2, 24),
    (7.0, 6.0, 4.0, 100.0, 1, 1, 48 * 8, 64, 2, 30),
    (11.0, 8.0, 4.0, 90.0, 1, 2, 128, 128, 1, 120),
]


@pytest.mark.parametrize("a,b,c,gam,nmult,mmult,nx,ny,nslices,natoms", CASES)
def test_potential_rec_matches_oracle(ctx, a, b, c, gam, nmult, mmult, nx, ny, nslices, natoms):
    rng = np.random.default_rng(natoms)
    gam = np.deg2rad(gam)
    at = random_atoms(rng, natoms, a, b, c, [6, 8, 14, 38, 79])
    dw = rng.choice([0.3, 0.6], natoms)
    ads = rng.choice([0.0, 0.05], natoms)
    Vr, Var = oracle_rec(at, a, b, c, gam, nmult, mmult, nx, ny, nslices, dw, ads)
    V, Va = ctx.potential_rec(at, (a, b, c), nx, ny, nslices, dw=dw, ads=ads, gamma=gam, nmult=nmult, mmult=mmult,
                              prec=msx.F64, absorptive=True)
    e, ea = rel(V, Vr), rel(Va, Var)
    print(f"{nx}x{ny} gam {np.rad2deg(gam):.0f} mult {nmult},{mmult}: V {e:.2e}, Vabs {ea:.2e}")
    assert e < 1e-11 and ea < 1e-11
    V32 = ctx.potential_rec(at, (a, b, c), nx, ny, nslices, dw=dw, ads=ads, gamma=gam, nmult=nmult, mmult=mmult,
                            prec=msx.F32)
    assert rel(V32, Vr) < 1e-6


def test_potential_rec_device_buffers(ctx):
    import torch
    rng = np.random.default_rng(5)
    at = random_atoms(rng, 30, 8.0, 8.0, 4.0, [14, 8])
    Vh = ctx.potential_rec(at, (8.0, 8.0, 4.0), 64, 64, 2, dw=0.5, prec=msx.F32)
    Vd = torch.empty((2, 64, 64), dtype=torch.float32, device="cuda")
    ctx.potential_rec(at, (8.0, 8.0, 4.0), 64, 64, 2, dw=0.5, prec=msx.F32, out=Vd)
    np.testing.assert_array_equal(Vd.cpu().numpy(), Vh)


def test_transmission_with_absorption(ctx):
    rng = np.random.default_rng(9)
    a, kev, bw = 8.0, 200.0, 2.0 / 3.0
    at = random_atoms(rng, 20, a, a, 2.0, [14, 79])
    V, Va = ctx.potential_rec(at, (a, a, 2.0), 64, 64, 1, dw=0.5, ads=0.1, prec=msx.F64, absorptive=True)
    t = ctx.transmission(V, a, a, kev, bw=bw, Vabs=Va)
    sig = msx.sigma(kev) / 1000.0
    raw = np.exp(1j * sig * V[0]) * np.exp(-sig * Va[0])
    k2 = np.add.outer(O.freqn(64, a)[1], O.freqn(64, a)[1])
    T = np.fft.fft2(raw)
    T[k2 >= O.k2max(64, 64, a, a, bw)] = 0
    ref = np.fft.ifft2(T)
    assert rel(t[0], ref) < 1e-12
    assert np.mean(np.abs(raw) ** 2) < 1.0  # absorption removes intensity
    # without Vabs: the plain transmission path
    t0 = ctx.transmission(V, a, a, kev, bw=bw)
    np.testing.assert_allclose(t0, O.transmission(V, a, a, kev, bw), rtol=0, atol=1e-12)


def test_potential_rec_rejects(ctx):
    at = atoms_array([93], [[1.0, 1.0, 1.0]])
    with pytest.raises(msx.MsxError):
        ctx.potential_rec(at, (5.0, 5.0, 2.0), 32, 32, 1)  # Z outside the Doyle-Turner table
    with pytest.raises(msx.MsxError):
        ctx.potential_rec(atoms_array([6], [[1.0, 1.0, 1.0]]), (5.0, 5.0, 2.0), 32, 32, 1, gamma=0.0)


def oracle_rs(at, a, b, c, gam, atrad, periodic, nx, ny, nslices, dw, ads):
    dz = c / nslices
    V = np.zeros((nslices, nx, ny))
    Va = np.zeros((nslices, nx, ny))
    kz = np.floor(at["z"] / dz).astype(int)
    xf = (at["x"] - at["y"] * np.cos(gam) / np.sin(gam)) / a
    yf = at["y"] / (b * np.sin(gam))
    for s in range(nslices):
        m = kz == s
        if not m.any():
            continue
        g = O.potential_rs_g(at["Z"][m], xf[m] % 1.0, yf[m] % 1.0, a, b, nx, ny, atrad, periodic=periodic, gam=gam,
                             dw=dw[m], ads=ads[m], occ=at["occ"][m].astype(np.float64))
        V[s], Va[s] = KVA2 * g.real.T, KVA2 * g.imag.T
    return V, Va


RS_CASES = [
    # a, b, c, gam(deg), atrad, periodic, nx, ny, nslices, natoms
    (9.0, 9.0, 6.0, 90.0, 2.5, True, 64, 64, 3, 40),
    (9.0, 8.0, 4.0, 90.0, 3.0, False, 80, 72, 2, 30),
    (7.0, 6.0, 4.0, 100.0, 2.0, True, 56, 48, 2, 30),
    (12.0, 12.0, 2.0, 90.0, 5.9, True, 100, 100, 1, 6),
]


@pytest.mark.parametrize("a,b,c,gam,atrad,periodic,nx,ny,nslices,natoms", RS_CASES)
def test_potential_rs_matches_oracle(ctx, a, b, c, gam, atrad, periodic, nx, ny, nslices, natoms):
    rng = np.random.default_rng(natoms + nx)
    gam = np.deg2rad(gam)
    at = random_atoms(rng, natoms, a, b, c, [6, 8, 14, 38, 79])
    dw = rng.choice([0.3, 0.6], natoms)
    ads = rng.choice([0.0, 0.05], natoms)
    Vr, Var = oracle_rs(at, a, b, c, gam, atrad, periodic, nx, ny, nslices, dw, ads)
    V, Va = ctx.potential_rs(at, (a, b, c), nx, ny, nslices, atrad, periodic=periodic, dw=dw, ads=ads, gamma=gam,
                             prec=msx.F64, absorptive=True)
    e, ea = rel(V, Vr), rel(Va, Var)
    print(f"{nx}x{ny} atrad {atrad} periodic {periodic}: V {e:.2e}, Vabs {ea:.2e}")
    assert e < 1e-12 and ea < 1e-12
    V32 = ctx.potential_rs(at, (a, b, c), nx, ny, nslices, atrad, periodic=periodic, dw=dw, ads=ads, gamma=gam)
    assert rel(V32, Vr) < 1e-6


def test_potential_rs_periodic_matches_rec(ctx):
    # the two Muslib potentials of one periodic structure agree (boxes cover
    # the cell, B large enough for the Fourier series to converge on the grid)
    rng = np.random.default_rng(3)
    a = 12.0
    at = random_atoms(rng, 8, a, a, 2.0, [6, 14, 79])
    Vrs = ctx.potential_rs(at, (a, a, 2.0), 256, 256, 1, 5.99, dw=1.5, prec=msx.F64)
    Vrec = ctx.potential_rec(at, (a, a, 2.0), 256, 256, 1, dw=1.5, prec=msx.F64)
    assert np.abs(Vrs - Vrec).max() < 1e-6 * np.abs(Vrec).max()


@pytest.mark.parametrize("kind", ["rec", "rs"])
def test_curve_potentials_match_oracle(ctx, kind):
    rng = np.random.default_rng(17 if kind == "rec" else 18)
    a, b, c, gam, nx, ny, ns = 9.0, 8.0, 4.0, np.deg2rad(96.0), 64, 64, 2
    at = random_atoms(rng, 30, a, b, c, [6])
    types = rng.integers(0, 3, 30).astype(np.int32)
    Zs, dws = [8, 38, 79], [0.5, 0.3, 0.6]
    ads = rng.choice([0.0, 0.07], 30)
    if kind == "rec":
        npts = 300
        dds = 0.5 * np.hypot(nx / 2 / a, ny / 2 / b) / 250  # the table ends inside the band: both branches run
        curves = O.dt_curves_rec(Zs, dws, npts, dds, a * b * np.sin(gam))
    else:
        npts, atrad = 400, 2.5
        dds = atrad * 1.2 / 350
        curves = O.dt_curves_rs(Zs, dws, npts, dds)
    dz = c / ns
    kz = np.floor(at["z"] / dz).astype(int)
    xf = (at["x"] - at["y"] * np.cos(gam) / np.sin(gam)) / a
    yf = at["y"] / (b * np.sin(gam))
    Vr, Var = np.zeros((ns, nx, ny)), np.zeros((ns, nx, ny))
    for s in range(ns):
        m = kz == s
        occ = at["occ"][m].astype(np.float64)
        if kind == "rec":
            g = O.potential_rec_curve_g(types[m], xf[m] % 1, yf[m] % 1, a, b, nx, ny, curves, dds, gam=gam,
                                        nmult=2, ads=ads[m], occ=occ)
            v = np.fft.ifft2(g) * (nx * ny)
        else:
            v = O.potential_rs_curve_g(types[m], xf[m] % 1, yf[m] % 1, a, b, nx, ny, atrad, curves, dds,
                                       periodic=True, gam=gam, ads=ads[m], occ=occ)
        Vr[s], Var[s] = KVA2 * v.real.T, KVA2 * v.imag.T
    kw = dict(nmult=2) if kind == "rec" else dict(atrad=atrad, periodic=True)
    V, Va = ctx.potential_curve(kind, at, types, curves, dds, (a, b, c), nx, ny, ns, ads=ads, gamma=gam,
                                prec=msx.F64, absorptive=True, **kw)
    e, ea = rel(V, Vr), rel(Va, Var)
    print(f"{kind} curves: V {e:.2e}, Vabs {ea:.2e}")
    assert e < 1e-11 and ea < 1e-11
    with pytest.raises(msx.MsxError):  # a type outside the table
        ctx.potential_curve(kind, at, np.full(30, 3), curves, dds, (a, b, c), nx, ny, ns, **kw)
