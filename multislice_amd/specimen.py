"""Synthetic specimens for the multislice workloads (host side, numpy).

Atoms are returned as a structured array with the fields of msx_atom
(Z, occ, x, y, z in Angstrom).  Crystal planes are placed at the centre of
the slice that holds them (z + dz/2), so the slice rule of
ModelPotential.cu:63-66 (kz*dz <= z < (kz+1)*dz) never meets an atom on a
slice boundary.  Lattice constants and the Debye-Waller RMS values are the
workload parameters, not fitted values.
"""
import numpy as np

from .engine import ATOM_DTYPE

# fractional positions in the conventional cubic cell
SI_DIAMOND = [(0, 0, 0), (0, .5, .5), (.5, 0, .5), (.5, .5, 0),
              (.25, .25, .25), (.25, .75, .75), (.75, .25, .75), (.75, .75, .25)]
SRTIO3 = [(38, (0, 0, 0)), (22, (.5, .5, .5)), (8, (.5, .5, 0)), (8, (.5, 0, .5)), (8, (0, .5, .5))]


def atoms_array(Z, xyz, occ=None):
    a = np.zeros(len(Z), ATOM_DTYPE)
    a["Z"] = Z
    a["occ"] = 1.0 if occ is None else occ
    xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
    a["x"], a["y"], a["z"] = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    return a


def crystal(basis, a, ncell, dz):
    """Periodic crystal: basis = [(Z, (fx, fy, fz))], cubic lattice constant a,
    ncell = (nx, ny, nz) cells; planes shifted by dz/2 along z."""
    nxc, nyc, nzc = ncell
    Zs, pos = [], []
    for Z, (fx, fy, fz) in basis:
        ii, jj, kk = np.meshgrid(np.arange(nxc), np.arange(nyc), np.arange(nzc), indexing="ij")
        Zs.append(np.full(ii.size, Z, np.int32))
        pos.append(np.stack([(ii.ravel() + fx) * a, (jj.ravel() + fy) * a,
                             (kk.ravel() + fz) * a + 0.5 * dz], 1))
    at = atoms_array(np.concatenate(Zs), np.concatenate(pos))
    order = np.lexsort((at["y"], at["x"], at["z"]))
    return at[order], (nxc * a, nyc * a, nzc * a)


def si_001(ncell=(4, 4, 20), a=5.431):
    """Si diamond along [001]; slice thickness a/4 (one atom plane per slice)."""
    return crystal([(14, p) for p in SI_DIAMOND], a, ncell, a / 4)


def srtio3_001(ncell=(12, 12, 50), a=3.905):
    """SrTiO3 perovskite along [001]; slice thickness a/2 (SrO / TiO2 planes)."""
    return crystal(SRTIO3, a, ncell, a / 2)


def frozen_phonon(atoms, u_rms, seed, config):
    """Isotropic Gaussian displacements (x, y, z) with per-element RMS u_rms
    {Z: u} (Angstrom, per axis), reproducible per (seed, config, atom):
    msx_displace (Philox-4x32-10, identical on host and GPU)."""
    from .engine import displace
    return displace(atoms, u_rms, seed, config)


SRTIO3_U = {38: 0.088, 22: 0.075, 8: 0.095}  # Angstrom, per axis


def amorphous_carbon(cell, density_g_cm3=2.0, min_dist=1.4, seed=0, max_tries=200):
    """Random sequential addition of C atoms at the given mass density with a
    minimum spacing (cell in Angstrom)."""
    a, b, c = cell
    n_target = int(round(density_g_cm3 / 12.011 * 6.02214076e23 * 1e-24 * a * b * c))
    rng = np.random.default_rng(seed)
    cellsz = min_dist
    nb = [max(1, int(s // cellsz)) for s in cell]
    grid = {}
    pts = []

    def key(p):
        return tuple(int(p[i] / cell[i] * nb[i]) % nb[i] for i in range(3))

    for _ in range(n_target * max_tries):
        if len(pts) >= n_target:
            break
        p = rng.random(3) * np.array(cell)
        k = key(p)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    kk = ((k[0] + dx) % nb[0], (k[1] + dy) % nb[1], (k[2] + dz) % nb[2])
                    for q in grid.get(kk, ()):
                        d = p - q
                        d -= np.round(d / np.array(cell)) * np.array(cell)
                        if d @ d < min_dist * min_dist:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            pts.append(p)
            grid.setdefault(k, []).append(p)
    pts = np.array(pts).reshape(-1, 3)
    return atoms_array(np.full(len(pts), 6, np.int32), pts)


def au_cluster(n, center, a=4.078):
    """The n fcc Au sites (a = 4.078 A) nearest to `center`."""
    r = int(np.ceil((n / 4) ** (1 / 3))) + 2
    fcc = np.array([(0, 0, 0), (.5, .5, 0), (.5, 0, .5), (0, .5, .5)])
    g = np.array([(i, j, k) for i in range(-r, r) for j in range(-r, r) for k in range(-r, r)], float)
    sites = (g[:, None, :] + fcc[None, :, :]).reshape(-1, 3) * a
    sites = sites[np.argsort((sites ** 2).sum(1), kind="stable")][:n]
    return atoms_array(np.full(n, 79, np.int32), sites + np.asarray(center))
