"""Grid construction and the spatial hierarchy.

- radial rules: the reference's goldens (tests/dft_tests.cpp:217-309, 1e-5);
- Lebedev rules: sizes (dft_tests.cpp:196-200) and exactness on spherical
  polynomials;
- closed form: the hydrogen 1s density integrates to 1 on an atom grid;
- Becke partition: cell fractions sum to 1 at every point;
- hierarchy: the invariants of dft_tests.cpp:1287-1462 and identity of the
  library's host hierarchy with the oracle's.
"""
import numpy as np
import pytest

import grid_ref
import oracle
from occ_amd import molecules
from occ_amd.dft import grid_hierarchy
from occ_amd.grid import GridSettings, MolecularGrid, atom_grid, lebedev, radial_grid


def test_radial_goldens():
    r, w = radial_grid("becke", 3, 0.6614041435977716)
    np.testing.assert_allclose(r, [9.21217133, 0.66140414, 0.04748668], atol=1e-5)
    np.testing.assert_allclose(w, [77.17570606, 1.3852416, 0.39782349], atol=1e-5)
    r, w = radial_grid("gauss_chebyshev", 3, 0.0)
    np.testing.assert_allclose(r, [8.66025404e-01, 6.123234e-17, -8.66025404e-01], atol=1e-5)
    np.testing.assert_allclose(w, [1.04719755] * 3, atol=1e-5)
    r, w = radial_grid("mura_knowles", 3, 1)
    np.testing.assert_allclose(r, [0.02412997, 0.69436324, 4.49497829], atol=1e-5)
    np.testing.assert_allclose(w, [0.14511628, 1.48571429, 8.57142857], atol=1e-5)
    r, w = radial_grid("treutler_ahlrichs", 3, 0.0)
    np.testing.assert_allclose(r, [0.10934791, 1, 3.82014324], atol=1e-5)
    np.testing.assert_allclose(w, [0.34905607, 1.60432893, 4.51614622], atol=1e-5)
    r, w = radial_grid("euler_maclaurin", 10, 1.2)
    assert len(r) == 9
    np.testing.assert_allclose(r, [0.0148148, 0.0750000, 0.2204082, 0.5333333, 1.2000000,
                                   2.7000000, 6.5333333, 19.2000000, 97.2000000], atol=1e-5)
    np.testing.assert_allclose(w, [0.7225637E-05, 0.5273438E-03, 0.1019750E-01, 0.1264198E+00,
                                   0.1382400E+01, 0.1640250E+02, 0.2655921E+03, 0.8847360E+04,
                                   0.2040733E+07], rtol=1e-5)


def test_lebedev_sizes_and_exactness():
    for n, order in [(110, 17), (194, 23), (302, 29)]:
        x, w = lebedev(n)
        assert x.shape == (n, 3)
        assert w.sum() == pytest.approx(4 * np.pi, rel=1e-14)
        # integral of z^4 over the sphere = 4 pi / 5; x^2 y^2 z^2 -> 4 pi / 105
        assert np.sum(w * x[:, 2] ** 4) == pytest.approx(4 * np.pi / 5, rel=1e-13)
        assert np.sum(w * (x[:, 0] * x[:, 1] * x[:, 2]) ** 2) == pytest.approx(
            4 * np.pi / 105, rel=1e-12)


def test_hydrogen_density_integrates_to_one():
    pts, w = atom_grid(1, GridSettings(radial_points=76, angular_points=302))
    r = np.linalg.norm(pts, axis=1)
    assert np.sum(w * np.exp(-2 * r) / np.pi) == pytest.approx(1.0, abs=1e-7)
    assert len(w) == 75 * 302


def test_becke_fractions_sum_to_one(rng):
    atoms = molecules.water_cluster(2, seed=2)
    xyz = np.array([a[1:] for a in atoms])
    from occ_amd.grid import bragg_radius
    radii = np.array([bragg_radius(a[0]) for a in atoms])
    pts = xyz[0] + rng.normal(size=(2000, 3)) * 2.0
    P = grid_ref.becke_weights(pts, xyz, radii)
    frac = P / P.sum(1, keepdims=True)
    np.testing.assert_allclose(frac.sum(1), 1.0, atol=1e-14)
    assert np.all(frac >= 0)


def test_partitioned_water_grid_integrates_density_volume():
    atoms = molecules.water()
    g = MolecularGrid(atoms, GridSettings(radial_points=76, angular_points=302))
    pts, w = grid_ref.molecular_grid(g)
    assert len(w) == 67950
    # sum over atoms of a normalised Gaussian at each nucleus -> natoms
    f = sum(np.exp(-np.sum((pts - np.array(a[1:])) ** 2, 1)) / np.pi ** 1.5 for a in atoms)
    assert np.sum(w * f) == pytest.approx(3.0, abs=1e-8)


def _check_hierarchy(h, n):
    assert sorted(h["perm"].tolist()) == list(range(n))
    offs, cnts = h["offset"], h["count"]
    assert offs[0] == 0 and offs[-1] + cnts[-1] == n
    assert np.all(offs[1:] == offs[:-1] + cnts[:-1])
    assert np.all(cnts <= 256)
    if n >= 64:
        assert np.all(cnts >= 32)
    for k in range(len(offs)):
        p = h["points"][offs[k]:offs[k] + cnts[k]]
        d = np.linalg.norm(p - h["center"][k], axis=1)
        assert d.max() <= h["radius"][k] * (1 + 1e-12) + 1e-14


@pytest.mark.parametrize("n", [1, 50, 300, 5003])
def test_hierarchy_invariants_and_identity(rng, n):
    pts = rng.normal(size=(n, 3)) * 4.0
    w = rng.uniform(size=n)
    a = grid_hierarchy(pts, w)
    b = oracle.build_hierarchy(pts, w)
    _check_hierarchy(a, n)
    _check_hierarchy(b, n)
    np.testing.assert_array_equal(a["perm"], b["perm"])
    np.testing.assert_array_equal(a["offset"], b["offset"])
    np.testing.assert_array_equal(a["count"], b["count"])
    np.testing.assert_allclose(a["center"], b["center"], rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(a["radius"], b["radius"], rtol=1e-15)
    np.testing.assert_array_equal(a["points"], pts[a["perm"]])


# --- OCC's default atom grids: LMG radial rule, NWChem pruning, reduced H/He
# grid (grid_utils.cpp:234-274, 397-515, 576-758).  What pins them: the LMG
# rule integrates the Gaussian radial moments it is built for to its
# precision; every angular count is a Lebedev level, the largest is the
# requested (or, for H/He, the next lower) level; and the VV10 known answer
# of tests/test_oracle_vv10.py is met on the NLC grid at the reference's own
# tolerance.

def _lmg_case(name, atoms, ia, prec):
    from occ_amd import AOBasis
    from occ_amd.grid import basis_exponents, lmg_radial_grid
    b = AOBasis.load(atoms, name, pure=True)
    amax, lmax, amin = basis_exponents(b, ia)
    r, w = lmg_radial_grid(atoms[ia][0], prec, amax, lmax, amin)
    return r, w, amax, amin


@pytest.mark.parametrize("ia", [0, 1])
def test_lmg_integrates_gaussian_moments(ia):
    from math import gamma
    atoms = molecules.water()
    r, w, amax, amin = _lmg_case("def2-svp", atoms, ia, 1e-12)
    assert np.all(np.diff(r) > 0) and np.all(w > 0)
    for l, a_lo in enumerate(amin):
        if a_lo > 1e300:
            continue
        for a in np.geomspace(a_lo, amax, 7):
            exact = gamma(l + 1.5) / (2 * a ** (l + 1.5))
            num = np.sum(w * r ** (2 * l) * np.exp(-a * r * r))
            assert num == pytest.approx(exact, rel=1e-9)


def test_lmg_precision_controls_size():
    atoms = molecules.water()
    r12, _, _, _ = _lmg_case("def2-svp", atoms, 0, 1e-12)
    r7, _, _, _ = _lmg_case("def2-svp", atoms, 0, 1e-7)
    assert len(r7) < len(r12)


def test_nwchem_pruning_and_reduced_first_row():
    from occ_amd import AOBasis
    from occ_amd.grid import (LEBEDEV_LEVELS, OccGridSettings, basis_exponents,
                              lmg_radial_grid, prune_nwchem, level_below)
    atoms = molecules.water()
    b = AOBasis.load(atoms, "def2-svp", pure=True)
    for ia, top in ((0, 302), (1, level_below(302))):
        z = atoms[ia][0]
        r, _ = lmg_radial_grid(z, 1e-12, *basis_exponents(b, ia))
        nang = prune_nwchem(z, top, r)
        assert set(nang) <= set(LEBEDEV_LEVELS) and max(nang) == top
        assert nang[0] == 50 and nang[-1] < top  # innermost and outermost shells pruned
    assert level_below(302) == 266 and level_below(110) == 86
    assert OccGridSettings.nlc().max_angular_points == 110


def test_occ_grid_integrates_density_volume():
    """The OCC default grid of water integrates a normalised Gaussian
    density per atom to 1 (partitioned, numpy Becke)."""
    from occ_amd import AOBasis
    from occ_amd.grid import OccGridSettings
    atoms = molecules.water()
    b = AOBasis.load(atoms, "def2-svp", pure=True)
    g = MolecularGrid(atoms, OccGridSettings(), basis=b)
    pts, w = grid_ref.molecular_grid(g)
    for (_, x, y, z) in atoms:
        d2 = ((pts - np.array([x, y, z])) ** 2).sum(1)
        val = np.sum(w * (2.0 / np.pi) ** 1.5 * np.exp(-2.0 * d2))
        assert val == pytest.approx(1.0, abs=1e-6)  # pruned grid: quadrature error ~1e-7
