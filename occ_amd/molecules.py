"""Seeded molecule generators for the benchmark and test workloads.

Geometries are returned as lists of (Z, x, y, z) with coordinates in Bohr
(OCC's Molecule returns atoms in Bohr, src/core/molecule.cpp:115-122; the
conversion factor is occ::units::ANGSTROM_TO_BOHR, include/occ/core/units.h:7-8).

- water(): O-H 0.9572 A, H-O-H 104.52 deg, O at the origin, H in the xz plane.
- water_cluster(n, seed): waters on a 3.0 A simple-cubic lattice, filled in
  x-fastest order over the smallest cube holding n sites; each O jittered by
  U(-0.3, 0.3) A per axis; each molecule rotated by a uniform random rotation
  (unit quaternion from four normal deviates).
- alkane(n): all-trans C_nH_(2n+2), C-C 1.54 A, C-H 1.09 A, tetrahedral angles.
- batch_set(count, seed): the mixed batch of the batched-launch workload:
  count/2 water clusters with n ~ U{1..16} and count/2 alkanes with
  n ~ U{1..24}, each with 6-31G or def2-SVP drawn with equal odds, shuffled.
- sized_molecule(natoms, rng): a water cluster or an alkane (equal odds) with
  about natoms atoms, for request streams with log-uniform molecule sizes.
"""
import math

import numpy as np

BOHR_TO_ANGSTROM = 0.52917721067
ANGSTROM_TO_BOHR = 1.0 / BOHR_TO_ANGSTROM
TETRAHEDRAL = math.acos(-1.0 / 3.0)


def _to_bohr(atoms):
    return [(z, x * ANGSTROM_TO_BOHR, y * ANGSTROM_TO_BOHR, w * ANGSTROM_TO_BOHR)
            for z, x, y, w in atoms]


def water_angstrom():
    r, half = 0.9572, math.radians(104.52 / 2.0)
    return [(8, 0.0, 0.0, 0.0),
            (1, r * math.sin(half), 0.0, r * math.cos(half)),
            (1, -r * math.sin(half), 0.0, r * math.cos(half))]


def water():
    return _to_bohr(water_angstrom())


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(n, seed=0, spacing=3.0, jitter=0.3):
    rng = np.random.default_rng(seed)
    k = 1
    while k ** 3 < n:
        k += 1
    mono = np.array([a[1:] for a in water_angstrom()])
    zs = [a[0] for a in water_angstrom()]
    atoms = []
    for i in range(n):
        site = np.array([i % k, (i // k) % k, i // (k * k)], dtype=float) * spacing
        centre = site + rng.uniform(-jitter, jitter, size=3)
        R = random_rotation(rng)
        for z, p in zip(zs, mono):
            q = centre + R @ p
            atoms.append((z, float(q[0]), float(q[1]), float(q[2])))
    return _to_bohr(atoms)


def alkane_angstrom(n):
    if n < 1:
        raise ValueError("alkane needs n >= 1")
    rcc, rch = 1.54, 1.09
    if n == 1:
        t = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / math.sqrt(3.0)
        return [(6, 0.0, 0.0, 0.0)] + [(1, *(rch * v)) for v in t]
    half = TETRAHEDRAL / 2.0
    # zig-zag backbone in the xy plane, extended by one virtual carbon at each end
    pos = [np.array([i * rcc * math.sin(half), (i % 2) * rcc * math.cos(half), 0.0])
           for i in range(-1, n + 1)]
    atoms = [(6, *pos[i + 1]) for i in range(n)]
    for i in range(n):
        c, prev, nxt = pos[i + 1], pos[i], pos[i + 2]
        b = (c - prev) + (c - nxt)
        b /= np.linalg.norm(b)
        nrm = np.cross(prev - c, nxt - c)
        nrm /= np.linalg.norm(nrm)
        for sgn in (1.0, -1.0):
            d = b * math.cos(half) + sgn * nrm * math.sin(half)
            atoms.append((1, *(c + rch * d)))
        if i == 0:
            d = (prev - c) / np.linalg.norm(prev - c)
            atoms.append((1, *(c + rch * d)))
        if i == n - 1:
            d = (nxt - c) / np.linalg.norm(nxt - c)
            atoms.append((1, *(c + rch * d)))
    return [(z, float(x), float(y), float(w)) for z, x, y, w in atoms]


def alkane(n):
    return _to_bohr(alkane_angstrom(n))


def nuclear_charge(atoms):
    return sum(int(a[0]) for a in atoms)


def batch_set(count=512, seed=0):
    """[(name, atoms, basis_name)] for the batched-launch workload."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        basis = "6-31g" if rng.integers(2) == 0 else "def2-svp"
        if i < count // 2:
            n = int(rng.integers(1, 17))
            out.append((f"(H2O){n}", water_cluster(n, seed=int(rng.integers(1 << 31))), basis))
        else:
            n = int(rng.integers(1, 25))
            out.append((f"C{n}H{2 * n + 2}", alkane(n), basis))
    order = rng.permutation(count)
    return [out[i] for i in order]


def sized_molecule(natoms, rng):
    """(name, atoms) of a water cluster (3n atoms) or an alkane (3n + 2 atoms)
    with about natoms atoms."""
    if rng.integers(2) == 0:
        n = max(1, int(round(natoms / 3.0)))
        return f"(H2O){n}", water_cluster(n, seed=int(rng.integers(1 << 31)))
    n = max(1, int(round((natoms - 2) / 3.0)))
    return f"C{n}H{2 * n + 2}", alkane(n)
