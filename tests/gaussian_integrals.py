"""Closed-form overlap integrals of contracted Gaussian AOs (test helper).

Cartesian overlaps by McMurchie-Davidson Hermite expansion coefficients
(Helgaker, Jorgensen, Olsen, Molecular Electronic-Structure Theory, 9.5),
then the CCA spherical transform and common_fac(l) scaling the AO values
carry.  Used to check the oracle's normalisation (unit self-overlap) and the
density path (sum_g w rho = tr(2 D S)).
"""
import math
from functools import lru_cache

import numpy as np

import oracle


def _E(i, j, t, Xab, a, b):
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return 0.0
    if i == j == t == 0:
        return math.exp(-q * Xab * Xab)
    if j == 0:
        return (_E(i - 1, j, t - 1, Xab, a, b) / (2 * p) - q * Xab / a * _E(i - 1, j, t, Xab, a, b)
                + (t + 1) * _E(i - 1, j, t + 1, Xab, a, b))
    return (_E(i, j - 1, t - 1, Xab, a, b) / (2 * p) + q * Xab / b * _E(i, j - 1, t, Xab, a, b)
            + (t + 1) * _E(i, j - 1, t + 1, Xab, a, b))


def _cart(l):
    return [(a, b, l - a - b) for a in range(l, -1, -1) for b in range(l - a, -1, -1)]


@lru_cache(maxsize=None)
def _T(l, pure):
    if pure and l >= 2:
        return oracle.cart_to_sph(l)
    if pure and l == 1:
        return np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]], dtype=float)
    n = (l + 1) * (l + 2) // 2
    return np.eye(n)


def shell_overlap(sa, sb):
    """Overlap block between two occ_amd.basis.Shell objects."""
    A, B = sa.origin, sb.origin
    ca, cb = _cart(sa.l), _cart(sb.l)
    Sc = np.zeros((len(ca), len(cb)))
    for ea, da in zip(sa.exponents, sa.coefficients):
        for eb, db in zip(sb.exponents, sb.coefficients):
            p = ea + eb
            pref = da * db * (math.pi / p) ** 1.5
            for u, (ax, ay, az) in enumerate(ca):
                for v, (bx, by, bz) in enumerate(cb):
                    Sc[u, v] += pref * (_E(ax, bx, 0, A[0] - B[0], ea, eb) *
                                        _E(ay, by, 0, A[1] - B[1], ea, eb) *
                                        _E(az, bz, 0, A[2] - B[2], ea, eb))
    fa = oracle.common_fac(sa.l, sa.pure)
    fb = oracle.common_fac(sb.l, sb.pure)
    return fa * fb * _T(sa.l, sa.pure) @ Sc @ _T(sb.l, sb.pure).T


def overlap_matrix(basis):
    fb = basis.first_bf()
    S = np.zeros((basis.nbf, basis.nbf))
    for i, si in enumerate(basis.shells):
        for j, sj in enumerate(basis.shells):
            S[fb[i]:fb[i] + si.size(), fb[j]:fb[j] + sj.size()] = shell_overlap(si, sj)
    return S
