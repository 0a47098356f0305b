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
import sys
from dataclasses import dataclass

import numpy as np

from . import _lib

# Bragg radii in Angstrom for H..Ne (grid_utils.cpp:22-41)
BRAGG_RADII_ANGSTROM = [0.35, 1.40, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50, 1.50]
ANGSTROM_TO_BOHR = 1.0 / 0.52917721067

# Lebedev point counts -> algebraic order accepted by scipy
LEBEDEV_ORDER = {6: 3, 14: 5, 26: 7, 38: 9, 50: 11, 74: 13, 86: 15, 110: 17, 146: 19,
                 170: 21, 194: 23, 230: 25, 266: 27, 302: 29, 350: 31, 434: 35, 590: 41,
                 770: 47, 974: 53, 1202: 59, 1454: 65, 1730: 71, 2030: 77, 2354: 83,
                 2702: 89, 3074: 95, 3470: 101, 3890: 107, 4334: 113, 4802: 119, 5294: 125,
                 5810: 131}

# Lebedev grid levels (point counts) in increasing order (grid_utils.cpp:15-18)
LEBEDEV_LEVELS = [1, 6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194, 230, 266, 302, 350, 434,
                  590, 770, 974, 1202, 1454, 1730, 2030, 2354, 2702, 3074, 3470, 3890, 4334,
                  4802, 5294, 5810]

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


def level_at_or_above(n):
    """nearest_grid_level_at_or_above (grid_utils.cpp:43-53)."""
    for m in LEBEDEV_LEVELS:
        if m >= n:
            return m
    raise ValueError("more than 5810 angular points requested")


def level_below(n):
    """nearest_grid_level_below (grid_utils.cpp:55-68)."""
    if n <= LEBEDEV_LEVELS[0]:
        return LEBEDEV_LEVELS[0]
    for m in reversed(LEBEDEV_LEVELS):
        if m < n:
            return m
    return LEBEDEV_LEVELS[0]


def prune_nwchem(z, max_angular, radii):
    """prune_nwchem_scheme (grid_utils.cpp:234-274): angular points per radial
    shell from the shell's radius in units of the Bragg radius."""
    n = len(radii)
    if max_angular < 50:
        return [max_angular] * n
    if max_angular == 50:
        levels = [5, 6, 6, 6, 5]
    else:
        i = 6
        while i < len(LEBEDEV_LEVELS) and LEBEDEV_LEVELS[i] != max_angular:
            i += 1
        levels = [5, 6, i - 1, i, i - 1]
    if z <= 2:
        bounds = [0.25, 0.5, 1.0, 4.5]
    elif z <= 10:
        bounds = [0.16666667, 0.5, 0.9, 3.5]
    else:
        bounds = [0.1, 0.4, 0.8, 2.5]
    rb = bragg_radius(z)
    out = []
    for r in radii:
        scale = r / rb
        place = next((k for k, b in enumerate(bounds) if b > scale), len(bounds))
        out.append(LEBEDEV_LEVELS[levels[place]])
    return out


@dataclass
class OccGridSettings:
    """OCC's molecular-grid settings (GridSettings, include/occ/numint/
    grid_settings.h:45-95) with its default radial rule, LMG, whose extent
    comes from the basis (MolecularGrid(basis, settings), molecular_grid.cpp:
    11-25, 183-256): NWChem pruning, reduced H/He grid, Becke partition with
    the Treutler-Ahlrichs adjustment.  The defaults are those of DFT's grid;
    nlc() gives the settings NonLocalCorrelationFunctional::set_integration_grid
    uses (nonlocal_correlation.h:36-38: {110, 50, 50, 1e-7, false}, where the
    fifth value lands in int_acc, so the reduced H/He grid stays on)."""
    max_angular_points: int = 302
    min_angular_points: int = 110
    radial_precision: float = 1e-12
    reduced_first_row_element_grid: bool = True
    treutler_alrichs_adjustment: bool = True
    pruning: str = "nwchem"

    @classmethod
    def nlc(cls):
        return cls(max_angular_points=110, min_angular_points=50, radial_precision=1e-7)

    @property
    def ta_adjustment(self):
        return self.treutler_alrichs_adjustment


def basis_exponents(basis, atom):
    """(alpha_max, l_max, alpha_min[0..l_max]) of one atom as
    MolecularGrid::initialize_from_basis records them (molecular_grid.cpp:
    206-228): l_max is the basis' largest l; an l the atom lacks keeps DBL_MAX."""
    l_max = basis.l_max()
    amin = [sys.float_info.max] * (l_max + 1)
    amax = 0.0
    for sh, ia in zip(basis.shells, basis.shell_atom):
        if ia != atom:
            continue
        e = np.asarray(sh.exponents, dtype=np.float64)
        amin[sh.l] = min(float(e.min()), amin[sh.l])
        amax = max(amax, float(e.max()))
    return amax, l_max, amin


def lmg_radial_grid(z, precision, alpha_max, l_max, alpha_min):
    """LMG radial rule (occxc_lmg_radial_grid): points and weights (with r^2)."""
    am = _lib.f64(alpha_min)
    n = np.zeros(1, dtype=np.int32)
    L = _lib.lib()
    _lib.check(L.occxc_lmg_radial_grid(float(precision), float(alpha_max), int(l_max),
                                       _lib.dptr(am), bragg_radius(z), 0, None, None,
                                       _lib.iptr(n)))
    r, w = np.zeros(int(n[0])), np.zeros(int(n[0]))
    _lib.check(L.occxc_lmg_radial_grid(float(precision), float(alpha_max), int(l_max),
                                       _lib.dptr(am), bragg_radius(z), len(r), _lib.dptr(r),
                                       _lib.dptr(w), _lib.iptr(n)))
    return r, w


def occ_atom_grid(z, settings, alpha_max, l_max, alpha_min):
    """generate_atom_grid with the LMG rule (grid_utils.cpp:576-758): radial
    shells outermost, each shell's Lebedev points in the rule's order."""
    r, wr = lmg_radial_grid(z, settings.radial_precision, alpha_max, l_max, alpha_min)
    max_ang = level_at_or_above(settings.max_angular_points)
    if settings.reduced_first_row_element_grid and z < 3:
        max_ang = level_below(max_ang)
    if settings.pruning == "nwchem":
        nang = prune_nwchem(z, max_ang, r)
    else:
        nang = [max_ang] * len(r)
    pts, wts = [], []
    for ri, wi, na in zip(r, wr, nang):
        ang, wa = lebedev(level_at_or_above(na))
        pts.append(ri * ang)
        wts.append(wi * wa)
    return np.concatenate(pts), np.concatenate(wts)


class MolecularGrid:
    """Atom grids of every atom (in atom order), partitioned by Becke cells.
    With OccGridSettings the atom grids are OCC's default ones and need the
    basis (for the LMG extent); with GridSettings they are fixed radial x
    Lebedev products."""

    def __init__(self, atoms, settings=None, basis=None):
        self.atoms = list(atoms)
        self.settings = settings or GridSettings()
        self.basis = basis
        if isinstance(self.settings, OccGridSettings) and basis is None:
            raise ValueError("OCC-style grids need the basis (LMG radial extent)")

    def _atom_grid(self, ia, z):
        if isinstance(self.settings, OccGridSettings):
            return occ_atom_grid(z, self.settings, *basis_exponents(self.basis, ia))
        return atom_grid(z, self.settings)

    def unpartitioned(self):
        """(points [N,3], weights [N], atom index [N]) before partitioning."""
        cache = {}
        pts, wts, own = [], [], []
        for ia, (z, x, y, w) in enumerate(self.atoms):
            if z not in cache:
                cache[z] = self._atom_grid(ia, z)
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
