"""Molecular integration grids: atom-centred radial x Lebedev grids with the
Becke partition (Treutler-Ahlrichs radii adjustment).

Mirrors occ::numint::generate_atom_grid (src/numint/grid_utils.cpp:576-758)
and MolecularGrid::get_partitioned_atom_grid / populate_molecular_grid_points
(src/numint/molecular_grid.cpp:46-131).  Radial rules are the library's
occxc_radial_grid (grid_utils.cpp:337-572); angular rules are Lebedev-Laikov
grids from scipy.integrate.lebedev_rule (weights summing to 4 pi, which
absorbs the 4 pi factor the reference applies at grid_utils.cpp:679); the
Becke partition runs on the GPU (occxc_becke_partition).
"""
from dataclasses import dataclass

import numpy as np

from . import _lib

# Bragg radii in Angstrom for H..Ne (grid_utils.cpp:22-41)
BRAGG_RADII_ANGSTROM = [0.35, 1.40, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50, 1.50]
ANGSTROM_TO_BOHR = 1.0 / 0.52917721067

# Lebedev point counts -> algebraic order accepted by scipy
LEBEDEV_ORDER = {6: 3, 14: 5, 26: 7, 38: 9, 50: 11, 74: 13, 86: 15, 110: 17, 146: 19,
                 170: 21, 194: 23, 230: 25, 266: 27, 302: 29, 350: 31, 434: 35, 590: 41,
                 770: 47, 974: 53, 1202: 59}

RADIAL_METHODS = {"euler_maclaurin": 0, "treutler_ahlrichs": 1, "becke": 2,
                  "gauss_chebyshev": 3, "mura_knowles": 4}


def bragg_radius(z):
    """Bohr (get_atomic_radius, grid_utils.cpp:82-90)."""
    return BRAGG_RADII_ANGSTROM[int(z) - 1] * ANGSTROM_TO_BOHR


def lebedev(n):
    """(points [n,3], weights [n], summing to 4 pi)."""
    from scipy.integrate import lebedev_rule
    if n not in LEBEDEV_ORDER:
        raise ValueError(f"no Lebedev rule with {n} points")
    x, w = lebedev_rule(LEBEDEV_ORDER[n])
    return np.ascontiguousarray(x.T), np.ascontiguousarray(w)


def radial_grid(method, n, param):
    """Raw radial rule of grid_utils.cpp (no r^2 factor except where the
    reference's rule itself carries it, as Euler-Maclaurin does)."""
    m = RADIAL_METHODS[method] if isinstance(method, str) else int(method)
    r = np.zeros(n)
    w = np.zeros(n)
    cnt = np.zeros(1, dtype=np.int32)
    _lib.check(_lib.lib().occxc_radial_grid(m, n, float(param), _lib.dptr(r), _lib.dptr(w),
                                            _lib.iptr(cnt)))
    return r[:cnt[0]].copy(), w[:cnt[0]].copy()


@dataclass
class GridSettings:
    """Fixed-size atom grids as used by the benchmark workloads: n radial
    shells (Euler-Maclaurin with radial_points = n + 1, alpha = 2 R_Bragg,
    grid_utils.cpp:733-737) times one Lebedev rule, no pruning, no reduced
    H/He grid, Becke partition with Treutler-Ahlrichs adjustment."""
    radial_method: str = "euler_maclaurin"
    radial_points: int = 76
    angular_points: int = 302
    ta_adjustment: bool = True


def atom_grid(z, settings):
    """Unpartitioned grid of one atom at the origin: (points [n,3], weights)."""
    rb = bragg_radius(z)
    if settings.radial_method == "euler_maclaurin":
        r, wr = radial_grid("euler_maclaurin", settings.radial_points, 2.0 * rb)
    elif settings.radial_method == "becke":
        r, wr = radial_grid("becke", settings.radial_points, rb)
        wr = wr * r * r
    elif settings.radial_method == "treutler_ahlrichs":
        r, wr = radial_grid("treutler_ahlrichs", settings.radial_points, 0.0)
        r, wr = r * rb, wr * rb * (r * rb) ** 2
    else:
        raise ValueError(f"radial method {settings.radial_method} not supported for atom grids")
    ang, wa = lebedev(settings.angular_points)
    pts = (r[:, None, None] * ang[None, :, :]).reshape(-1, 3)
    wts = (wr[:, None] * wa[None, :]).reshape(-1)
    return pts, wts


class MolecularGrid:
    def __init__(self, atoms, settings=None):
        self.atoms = list(atoms)
        self.settings = settings or GridSettings()

    def unpartitioned(self):
        """(points [N,3], weights [N], atom index [N]) before partitioning."""
        cache = {}
        pts, wts, own = [], [], []
        for ia, (z, x, y, w) in enumerate(self.atoms):
            if z not in cache:
                cache[z] = atom_grid(z, self.settings)
            p, q = cache[z]
            pts.append(p + np.array([x, y, w]))
            wts.append(q.copy())
            own.append(np.full(len(q), ia, dtype=np.int32))
        return np.concatenate(pts), np.concatenate(wts), np.concatenate(own)

    def atom_arrays(self):
        xyz = np.array([[a[1], a[2], a[3]] for a in self.atoms], dtype=np.float64)
        radii = np.array([bragg_radius(a[0]) for a in self.atoms], dtype=np.float64)
        return xyz, radii

    def build(self, device=0):
        """Partitioned molecular grid, partition evaluated on the GPU."""
        pts, wts, own = self.unpartitioned()
        xyz, radii = self.atom_arrays()
        pts = _lib.f64(pts)
        wts = _lib.f64(wts)
        xyz = _lib.f64(xyz)
        own = _lib.i32(own)
        _lib.check(_lib.lib().occxc_becke_partition(
            len(self.atoms), _lib.dptr(xyz), _lib.dptr(radii), int(self.settings.ta_adjustment),
            len(wts), _lib.dptr(pts), _lib.iptr(own), _lib.dptr(wts), int(device)))
        return pts, wts
