"""Pins the oracle's XC build by what the mathematics fixes.

- screened == dense (the reference's own invariant, tests/dft_tests.cpp:
  904-931, restated for Exc / Vxc / N_elec);
- N_elec = sum_g w rho = tr(2 D S) with S from closed-form overlaps (to the
  grid's quadrature error);
- dExc/dD: (Exc(D + eps Delta) - Exc(D - eps Delta)) / (2 eps) = 2 tr(Vxc Delta)
  for symmetric Delta;
- brute force on a tiny case: rho, grad rho and Vxc summed point by point in
  Python from the oracle's AO values;
- only the upper triangle of D is read (density.cpp:14,38).
"""
import numpy as np
import pytest

import grid_ref
import oracle
from conftest import random_density
from gaussian_integrals import overlap_matrix
from occ_amd import AOBasis, molecules
from occ_amd.grid import GridSettings, MolecularGrid

METHODS = {"svwn5": [1, 7], "pbe": [101, 130], "blyp": [106, 131]}


@pytest.fixture(scope="module")
def water_case():
    atoms = molecules.water()
    basis = AOBasis.load(atoms, "6-31g", pure=False)
    g = MolecularGrid(atoms, GridSettings(radial_points=76, angular_points=302))
    pts, w = grid_ref.molecular_grid(g)
    return basis, oracle.OracleBasis.from_flat(basis.flat()), pts, w


@pytest.fixture(scope="module")
def dimer_case():
    atoms = molecules.water_cluster(2, seed=1)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    g = MolecularGrid(atoms, GridSettings(radial_points=41, angular_points=146))
    pts, w = grid_ref.molecular_grid(g)
    return basis, oracle.OracleBasis.from_flat(basis.flat()), pts, w


@pytest.mark.parametrize("method", list(METHODS))
def test_screened_equals_dense(dimer_case, method):
    basis, ob, pts, w = dimer_case
    D = random_density(basis.nbf, 10, 3)
    K1, e1, n1 = oracle.xc_build(ob, pts, w, METHODS[method], D, screened=True)
    K2, e2, n2 = oracle.xc_build(ob, pts, w, METHODS[method], D, screened=False)
    assert abs(e1 - e2) < 1e-8
    assert np.abs(K1 - K2).max() < 1e-8
    assert abs(n1 - n2) < 1e-8


def test_nelec_equals_trace_2DS(water_case):
    basis, ob, pts, w = water_case
    S = overlap_matrix(basis)
    D = random_density(basis.nbf, 5, 4)
    _, _, nel = oracle.xc_build(ob, pts, w, METHODS["svwn5"], D)
    assert nel == pytest.approx(np.trace(2 * D @ S), abs=2e-6)


@pytest.mark.parametrize("method", list(METHODS))
def test_vxc_is_the_derivative_of_exc(water_case, method):
    basis, ob, pts, w = water_case
    D = random_density(basis.nbf, 5, 5)
    r = np.random.default_rng(9)
    Delta = r.normal(size=D.shape) * 0.01
    Delta = Delta + Delta.T
    K, _, _ = oracle.xc_build(ob, pts, w, METHODS[method], D)
    eps = 1e-4
    _, ep, _ = oracle.xc_build(ob, pts, w, METHODS[method], D + eps * Delta)
    _, em, _ = oracle.xc_build(ob, pts, w, METHODS[method], D - eps * Delta)
    fd = (ep - em) / (2 * eps)
    assert fd == pytest.approx(2 * np.sum(K * Delta), rel=1e-6, abs=1e-9)


def test_vxc_symmetric(water_case):
    basis, ob, pts, w = water_case
    D = random_density(basis.nbf, 5, 6)
    for ids in METHODS.values():
        K, _, _ = oracle.xc_build(ob, pts, w, ids, D)
        np.testing.assert_allclose(K, K.T, atol=1e-13)


def test_only_upper_triangle_read(water_case):
    basis, ob, pts, w = water_case
    D = random_density(basis.nbf, 5, 7)
    Dup = np.triu(D)
    Dup = Dup + np.tril(np.random.default_rng(1).normal(size=D.shape), -1)
    K1, e1, n1 = oracle.xc_build(ob, pts, w, METHODS["pbe"], D, nthreads=1)
    K2, e2, n2 = oracle.xc_build(ob, pts, w, METHODS["pbe"], Dup, nthreads=1)
    assert e1 == e2 and n1 == n2 and np.array_equal(K1, K2)


def test_brute_force_tiny():
    atoms = molecules.water()
    basis = AOBasis.load(atoms, "sto-3g", pure=False)
    ob = oracle.OracleBasis.from_flat(basis.flat())
    rng = np.random.default_rng(11)
    pts = rng.normal(size=(300, 3)) * 1.5
    w = rng.uniform(0.01, 0.1, size=300)
    D = random_density(basis.nbf, 5, 8)
    ids = METHODS["pbe"]
    # AO values point by point
    fb = basis.first_bf()
    phi = np.zeros((4, basis.nbf, len(w)))
    for s, sh in enumerate(basis.shells):
        phi[:, fb[s]:fb[s] + sh.size()] = oracle.eval_shell(sh.l, sh.pure, sh.exponents,
                                                             sh.coefficients, sh.origin, pts)
    D2 = 2 * D
    rho = np.einsum("mp,mn,np->p", phi[0], D2, phi[0])
    grad = np.stack([2 * np.einsum("mp,mn,np->p", phi[k], D2, phi[0]) for k in (1, 2, 3)])
    sigma = (grad ** 2).sum(0)
    exc, vr, vs = np.zeros_like(rho), np.zeros_like(rho), np.zeros_like(rho)
    for f in ids:
        a, b, c = oracle.functional_eval(f, rho, sigma)
        exc += a
        vr += b
        vs += c
    E = np.sum(w * rho * exc)
    A = 0.5 * w * vr * phi[0] + 2 * w * vs * np.einsum("kp,kmp->mp", grad, phi[1:])
    Kt = phi[0] @ A.T
    V = Kt + Kt.T
    K, e, n = oracle.xc_build(ob, pts, w, ids, D, screened=False, density_threshold=0.0)
    assert e == pytest.approx(E, rel=1e-12)
    assert n == pytest.approx(np.sum(w * rho), rel=1e-12)
    np.testing.assert_allclose(K, V, rtol=1e-10, atol=1e-12)
