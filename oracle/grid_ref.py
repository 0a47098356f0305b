"""Becke partition restated in numpy (CPU).

TEST INFRASTRUCTURE ONLY: builds grids for CPU tests and checks the GPU
partition kernel.  Follows calculate_atomic_grid_weights_impl with the
Treutler-Ahlrichs adjustment (src/numint/grid_utils.cpp:125-200) and the
normalisation of MolecularGrid::get_partitioned_atom_grid
(src/numint/molecular_grid.cpp:70-78).
"""
import numpy as np


def becke_weights(points, atom_xyz, radii, ta=True):
    """Cell functions P[npts, natoms] in the reference's pair order."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    A = np.asarray(atom_xyz, dtype=np.float64).reshape(-1, 3)
    n_at = len(A)
    d = np.sqrt(((pts[:, None, :] - A[None, :, :]) ** 2).sum(-1))
    R = np.sqrt(((A[:, None, :] - A[None, :, :]) ** 2).sum(-1))
    P = np.ones((len(pts), n_at))
    for i in range(n_at):
        for j in range(i):
            w = (d[:, i] - d[:, j]) / R[i, j]
            ri, rj = radii[i], radii[j]
            if abs(ri - rj) > 1e-14:
                xi = np.sqrt(ri / rj) if ta else ri / rj
                u = (xi - 1) / (xi + 1)
                a = u / (u * u - 1.0)
                w = w + a * (1 - w * w)
            for _ in range(3):
                w = (3 - w * w) * w * 0.5
            zero = np.abs(1.0 - w) < 1e-14
            P[zero, i] = 0.0
            keep = ~zero
            P[keep, i] *= 0.5 * (1.0 - w[keep])
            P[keep, j] *= 0.5 * (1.0 + w[keep])
    return P


def partition(points, weights, atom_of_point, atom_xyz, radii, ta=True, chunk=200000):
    out = np.array(weights, dtype=np.float64, copy=True)
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    own = np.asarray(atom_of_point)
    for lo in range(0, len(out), chunk):
        hi = min(len(out), lo + chunk)
        P = becke_weights(pts[lo:hi], atom_xyz, radii, ta)
        out[lo:hi] *= P[np.arange(hi - lo), own[lo:hi]] / P.sum(1)
    return out


def molecular_grid(mgrid):
    """(points, weights) of an occ_amd.grid.MolecularGrid, partitioned on the CPU."""
    pts, wts, own = mgrid.unpartitioned()
    xyz, radii = mgrid.atom_arrays()
    return pts, partition(pts, wts, own, xyz, radii, mgrid.settings.ta_adjustment)
