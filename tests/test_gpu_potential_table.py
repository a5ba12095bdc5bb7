"""The FP32 splat evaluates V_Z(r^2) from per-species radial tables (cubic
interpolation on octave-uniform grids, ModelPotential.cu:232-254 /
slicelib.cpp:2666-2699 evaluated in FP64 at the grid points); the FP64 splat
sums the exact formula.  On specimens with several species, atoms exactly on
pixel sites (r = 0, clamped to 1e-10) and radii that span many octaves, the
FP32 potential must stay within FP32 resolution of the FP64 one, and the FP64
one within 1e-13 of the oracle."""
import numpy as np
import pytest

import multislice_amd as msx
from oracle import Oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def specimen(rng, n, cell, zs):
    at = np.zeros(n, dtype=msx.ATOM_DTYPE)
    at["Z"] = rng.choice(zs, n)
    at["occ"] = rng.uniform(0.5, 1.0, n)
    at["x"] = rng.uniform(0, cell[0], n)
    at["y"] = rng.uniform(0, cell[1], n)
    at["z"] = rng.uniform(0, cell[2], n)
    # a few atoms exactly on pixel sites
    at["x"][:4] = np.arange(4) * cell[0] / 64
    at["y"][:4] = np.arange(4) * cell[1] / 64
    return at


@pytest.mark.parametrize("radius,table,formula", [(3.0, 0, msx.FORMULA_KIRKLAND), (20.0, 1, msx.FORMULA_KIRKLAND),
                                                  (3.0, 1, 1), (20.0, 0, 1)])
def test_fp32_table_splat_within_fp32_resolution(radius, table, formula):
    ctx = msx.Context(0)
    rng = np.random.default_rng(int(radius) + 7 * table + 13 * formula)
    cell = (12.3, 9.7, 4.0)
    at = specimen(rng, 60, cell, [1, 6, 8, 14, 22, 38, 79, 92])
    kw = dict(table=table, formula=formula)
    V64 = ctx.potential(at, cell, 64, 64, 2, radius, prec=msx.F64, **kw)
    V32 = ctx.potential(at, cell, 64, 64, 2, radius, prec=msx.F32, **kw)
    e = rel_l2(V32, V64)
    big = np.abs(V64) > 1e-3 * np.abs(V64).max()
    emax = float(np.max(np.abs(V32[big] - V64[big]) / np.abs(V64[big])))
    print(f"radius {radius} table {table} formula {formula}: FP32 (table) vs FP64 (exact) "
          f"rel-L2 {e:.2e}, max rel {emax:.2e}")
    assert e < 1e-7 and emax < 3e-7
    # the FP64 splat against the oracle on the same atoms
    xyz = np.stack([at["x"], at["y"], at["z"]], 1)
    Vo = O.potential(at["Z"], xyz, cell, 64, 64, 2, radius, table=table, formula=formula, occ=at["occ"])
    assert rel_l2(V64, Vo) < 1e-13
    ctx.close()


@pytest.mark.parametrize("cell,nx,ny,radius", [((120.0, 90.0, 4.0), 24, 20, 1.5), ((47.0, 150.0, 4.0), 10, 40, 1.2)])
def test_coarse_grid_tiles_reach_many_bins(cell, nx, ny, radius):
    """Coarse pixels and a small cutoff: a 16 x 32 pixel tile then reaches
    more than 32 bin columns (the bins are at least the cutoff wide), so the
    splat walks its bins in several column groups, and along the other axis
    the tile covers the whole period (every bin row, no tile filter).  Both
    precisions against the oracle's per-pixel sum over every atom."""
    ctx = msx.Context(0)
    rng = np.random.default_rng(nx + ny)
    at = specimen(rng, 300, cell, [6, 8, 38, 79])
    V64 = ctx.potential(at, cell, nx, ny, 2, radius, prec=msx.F64)
    V32 = ctx.potential(at, cell, nx, ny, 2, radius, prec=msx.F32)
    xyz = np.stack([at["x"], at["y"], at["z"]], 1)
    Vo = O.potential(at["Z"], xyz, cell, nx, ny, 2, radius, occ=at["occ"])
    assert np.abs(Vo).max() > 0
    e64, e32 = rel_l2(V64, Vo), rel_l2(V32, Vo)
    print(f"cell {cell[:2]} grid {nx}x{ny} radius {radius}: FP64 {e64:.1e}, FP32 {e32:.1e} rel-L2 vs oracle")
    assert e64 < 1e-13 and e32 < 1e-7
    ctx.close()


def test_potential_edge_cases():
    """No atoms, or only atoms outside the slab: a zero potential in both
    precisions (no species, so no radial tables); an element outside 1..103
    is refused with the atom's index.  Calls in a row reuse the context's
    binning buffers (a large call, then a small one, then the large one
    again gives the same potential)."""
    ctx = msx.Context(0)
    cell = (10.0, 10.0, 4.0)
    for prec in (msx.F32, msx.F64):
        V = ctx.potential(np.zeros(0, dtype=msx.ATOM_DTYPE), cell, 32, 32, 2, 3.0, prec=prec)
        assert V.shape == (2, 32, 32) and not V.any()
        out = np.zeros(3, dtype=msx.ATOM_DTYPE)
        out["Z"], out["occ"] = 14, 1.0
        out["z"] = [-0.5, 4.0, 7.0]  # below the slab, at its top face, above it
        V = ctx.potential(out, cell, 32, 32, 2, 3.0, prec=prec)
        assert not V.any()
    bad = np.zeros(2, dtype=msx.ATOM_DTYPE)
    bad["Z"] = [14, 104]
    with pytest.raises(msx.MsxError) as e:
        ctx.potential(bad, cell, 32, 32, 2, 3.0)
    assert "atom 1" in str(e.value)
    rng = np.random.default_rng(3)
    big = specimen(rng, 400, cell, [8, 22, 38])
    small = specimen(rng, 5, cell, [79])
    V1 = ctx.potential(big, cell, 64, 64, 2, 3.0)
    ctx.potential(small, cell, 64, 64, 2, 3.0)
    V2 = ctx.potential(big, cell, 64, 64, 2, 3.0)
    np.testing.assert_array_equal(V1, V2)
    ctx.close()


def _device_atoms(at):
    import torch
    return torch.from_numpy(np.ascontiguousarray(at).view(np.uint8).copy()).cuda()


@pytest.mark.parametrize("radius", [3.0, 1.1])
def test_device_binning_equals_host_binning(radius):
    """msx_potential_device_atoms bins on the GPU (cell key per atom, stable
    radix sort, cell starts by lower bound): the potential must be identical,
    bit for bit, to msx_potential's host binning in both precisions, with atoms
    on slice boundaries, on the cell edges, at negative and far-away
    coordinates, outside the slab, and several atoms in one cell in input
    order."""
    ctx = msx.Context(0)
    rng = np.random.default_rng(11 + int(radius))
    cell = (12.3, 9.7, 4.0)
    at = specimen(rng, 900, cell, [1, 6, 8, 14, 22, 38, 79])
    nsl = 5
    dz = cell[2] / nsl
    at["z"][10:20] = np.arange(10) % nsl * dz          # on slice boundaries
    at["x"][20:25] = [0.0, cell[0], -cell[0], 3 * cell[0] + 0.25, -1e-300]
    at["y"][25:30] = [cell[1], -0.5, 1e7, -1e7, np.nextafter(cell[1], 0)]
    at["z"][30:33] = [-1e-12, cell[2], 1e9]           # outside the slab
    at[40:60] = at[40]                                # one cell, many atoms
    at["occ"][40:60] = np.linspace(0.1, 1.0, 20)
    for prec in (msx.F32, msx.F64):
        kw = dict(prec=prec, table=0, formula=msx.FORMULA_KIRKLAND)
        Vh = ctx.potential(at, cell, 64, 48, nsl, radius, **kw)
        Vd = ctx.potential(_device_atoms(at), cell, 64, 48, nsl, radius, **kw)
        np.testing.assert_array_equal(Vd, Vh)
        # device output buffer too
        import torch
        out = torch.empty((nsl, 64, 48), dtype=torch.float32 if prec == msx.F32 else torch.float64, device="cuda")
        ctx.potential(_device_atoms(at), cell, 64, 48, nsl, radius, out=out, **kw)
        np.testing.assert_array_equal(out.cpu().numpy(), Vh)
    ctx.close()


def test_device_binning_edge_cases():
    """No atoms, all atoms outside the slab, invalid atoms (the first one is
    named, as by the host binning), and calls in a row of different sizes."""
    ctx = msx.Context(0)
    cell = (10.0, 10.0, 4.0)
    V = ctx.potential(_device_atoms(np.zeros(0, dtype=msx.ATOM_DTYPE)), cell, 32, 32, 2, 3.0)
    assert V.shape == (2, 32, 32) and not V.any()
    out = np.zeros(3, dtype=msx.ATOM_DTYPE)
    out["Z"], out["occ"] = 14, 1.0
    out["z"] = [-0.5, 4.0, 7.0]
    assert not ctx.potential(_device_atoms(out), cell, 32, 32, 2, 3.0).any()
    rng = np.random.default_rng(5)
    good = specimen(rng, 50, cell, [8, 14])
    for field, val, msg in (("Z", 104, "atom 7: Z = 104"), ("Z", 0, "atom 7: Z = 0"),
                            ("x", np.nan, "atom 7: coordinate not finite"),
                            ("z", np.inf, "atom 7: coordinate not finite")):
        bad = good.copy()
        bad[field][7] = val
        bad["Z"][30] = 200                             # a later invalid atom is not the one named
        for atoms in (bad, _device_atoms(bad)):
            with pytest.raises(msx.MsxError) as e:
                ctx.potential(atoms, cell, 32, 32, 2, 3.0)
            assert msg in str(e.value)
    big = specimen(rng, 3000, cell, [8, 22, 38])
    small = specimen(rng, 5, cell, [79])
    V1 = ctx.potential(_device_atoms(big), cell, 64, 64, 2, 3.0)
    Vs = ctx.potential(_device_atoms(small), cell, 64, 64, 2, 3.0)
    V2 = ctx.potential(_device_atoms(big), cell, 64, 64, 2, 3.0)
    np.testing.assert_array_equal(V1, V2)
    np.testing.assert_array_equal(Vs, ctx.potential(small, cell, 64, 64, 2, 3.0))
    np.testing.assert_array_equal(V1, ctx.potential(big, cell, 64, 64, 2, 3.0))
    ctx.close()


def test_device_atoms_pointer_checked():
    """msx_potential_device_atoms refuses, before any kernel runs, atoms in
    host memory and a record count that runs past the end of the device
    allocation holding them."""
    import ctypes as C
    import torch
    from multislice_amd import engine
    ctx = msx.Context(0)
    cell = (10.0, 10.0, 4.0)
    good = specimen(np.random.default_rng(9), 50, cell, [8, 14])
    V = torch.empty((2, 32, 32), dtype=torch.float32, device="cuda")
    fn = msx.load().msx_potential_device_atoms
    geo = (10.0, 10.0, 4.0, 32, 32, 2, 3.0, msx.TABLE_LMULTISLICE, msx.FORMULA_LMULTISLICE, C.c_void_p(V.data_ptr()),
           msx.F32, msx.DEVICE)
    host = np.ascontiguousarray(good)
    with pytest.raises(msx.MsxError) as e:
        engine._check(fn(ctx.h, C.c_void_p(host.ctypes.data), len(host), *geo))
    assert "not device memory" in str(e.value)
    dev = _device_atoms(good)
    seg = torch.cuda.memory_snapshot()
    seg_bytes = max(s["total_size"] for s in seg if s["address"] <= dev.data_ptr() < s["address"] + s["total_size"])
    too_many = seg_bytes // msx.ATOM_DTYPE.itemsize + 1
    with pytest.raises(msx.MsxError) as e:
        engine._check(fn(ctx.h, C.c_void_p(dev.data_ptr()), too_many, *geo))
    assert "past the end" in str(e.value)
    engine._check(fn(ctx.h, C.c_void_p(dev.data_ptr()), len(good), *geo))
    np.testing.assert_array_equal(V.cpu().numpy(), ctx.potential(good, cell, 32, 32, 2, 3.0,
                                                                 table=msx.TABLE_LMULTISLICE,
                                                                 formula=msx.FORMULA_LMULTISLICE))
    ctx.close()
