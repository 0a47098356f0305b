"""Pins the oracle's basis normalisation and collocation.

- normalisation: every spherical AO (and every s/p AO) has unit self-overlap
  (closed-form Gaussian integrals), and the oracle's C normalisation equals
  the package's numpy one;
- common_fac: the literal constants of gto.cpp:7-16;
- collocation + CCA transform: the reference's own gau2grid compiled from its
  sources into oracle/_ref (when present) and the committed fixture written
  from it (tests/golden/gau2grid_collocation.npz);
- gradients: central finite differences of the values.
"""
import numpy as np
import pytest

import oracle
from gaussian_integrals import overlap_matrix, shell_overlap
from occ_amd import AOBasis, molecules
from occ_amd.basis import normalize_contraction

GOLDEN = __import__("pathlib").Path(__file__).parent / "golden" / "gau2grid_collocation.npz"


@pytest.mark.parametrize("name", ["6-31g", "def2-svp", "def2-tzvp", "sto-3g", "3-21g"])
def test_spherical_ao_unit_self_overlap(name):
    basis = AOBasis.load(molecules.water(), name, pure=True)
    for sh in basis.shells:
        S = shell_overlap(sh, sh)
        np.testing.assert_allclose(np.diag(S), 1.0, atol=1e-12)
        np.testing.assert_allclose(S, np.eye(sh.size()), atol=1e-12)


def test_cartesian_sp_unit_and_d_unnormalised():
    basis = AOBasis.load(molecules.water(), "def2-svp", pure=False)
    for sh in basis.shells:
        d = np.diag(shell_overlap(sh, sh))
        if sh.l <= 1:
            np.testing.assert_allclose(d, 1.0, atol=1e-12)
        else:
            # common_fac = 1 for Cartesian l >= 2 (gto.cpp:13): xx has <xx|xx> = 3 <xy|xy>
            np.testing.assert_allclose(d[0] / d[1], 3.0, rtol=1e-12)


def test_oracle_normalisation_equals_package(rng):
    for l in range(5):
        exps = 10 ** rng.uniform(-1, 3, 4)
        raw = rng.normal(size=4)
        np.testing.assert_allclose(oracle.normalize_shell(l, exps, raw),
                                   normalize_contraction(l, exps, raw), rtol=1e-13)


def test_common_fac_constants():
    assert oracle.common_fac(0, 1) == 0.282094791773878143
    assert oracle.common_fac(1, 0) == 0.488602511902919921
    assert oracle.common_fac(2, 1) == pytest.approx(np.sqrt(5 / (4 * np.pi)), rel=1e-15)
    assert oracle.common_fac(3, 0) == 1.0


def test_cart_to_sph_l2_literals():
    # gau2grid_transform.c:85-114: sqrt3 xy, sqrt3 yz, zz - (xx+yy)/2, sqrt3 xz, sqrt3/2 (xx-yy)
    s3 = 1.7320508075688772
    T = oracle.cart_to_sph(2)
    ref = np.array([[0, s3, 0, 0, 0, 0], [0, 0, 0, 0, s3, 0], [-0.5, 0, 0, -0.5, 0, 1],
                    [0, 0, s3, 0, 0, 0], [s3 / 2, 0, 0, -s3 / 2, 0, 0]])
    np.testing.assert_allclose(T, ref, atol=1e-15)


def _collocation_cases():
    r = np.random.default_rng(7)
    cases = []
    for l in range(5):
        for pure in (0, 1):
            for nprim in (1, 3, 6):
                exps = 10 ** r.uniform(-1.5, 2, nprim)
                coefs = r.normal(size=nprim)
                cen = r.normal(size=3)
                # centre, mid-range and tail points
                xyz = np.concatenate([cen[None, :] + 1e-9, cen + r.normal(size=(12, 3)),
                                      cen + r.normal(size=(6, 3)) * 6.0])
                cases.append((l, pure, exps, coefs, cen, xyz))
    return cases


def test_collocation_matches_gau2grid_library():
    if not oracle.gau2grid_available():
        pytest.skip("reference gau2grid not built here (no OCC sources)")
    for l, pure, exps, coefs, cen, xyz in _collocation_cases():
        mine = oracle.eval_shell(l, pure, exps, coefs, cen, xyz)
        ref = oracle.gau2grid_deriv1(l, pure, exps, coefs, cen, xyz) * oracle.common_fac(l, pure)
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(mine - ref).max() <= 1e-13 * scale


def test_collocation_matches_golden_fixture():
    z = np.load(GOLDEN)
    for k in range(int(z["ncase"])):
        l, pure = int(z[f"l{k}"]), int(z[f"pure{k}"])
        mine = oracle.eval_shell(l, pure, z[f"exps{k}"], z[f"coefs{k}"], z[f"center{k}"],
                                 z[f"xyz{k}"])
        ref = z[f"phi{k}"] * oracle.common_fac(l, pure)
        assert np.abs(mine - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1e-300)


def test_hessian_matches_gau2grid_library():
    if not oracle.gau2grid_available():
        pytest.skip("reference gau2grid not built here (no OCC sources)")
    for l, pure, exps, coefs, cen, xyz in _collocation_cases():
        mine = oracle.eval_shell2(l, pure, exps, coefs, cen, xyz)
        ref = oracle.gau2grid_deriv2(l, pure, exps, coefs, cen, xyz) * oracle.common_fac(l, pure)
        assert np.abs(mine - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1e-300)


def test_hessian_matches_golden_fixture():
    """phi_xx..phi_zz of gg_collocation_deriv2 (gau2grid_deriv2.c), written by
    tests/golden/make_gau2grid_fixture.py."""
    z = np.load(GOLDEN)
    for k in range(int(z["ncase"])):
        l, pure = int(z[f"l{k}"]), int(z[f"pure{k}"])
        mine = oracle.eval_shell2(l, pure, z[f"exps{k}"], z[f"coefs{k}"], z[f"center{k}"],
                                  z[f"xyz{k}"])
        fac = oracle.common_fac(l, pure)
        ref1 = z[f"phi{k}"] * fac
        ref2 = z[f"hess{k}"] * fac
        assert np.abs(mine[:4] - ref1).max() <= 1e-13 * max(np.abs(ref1).max(), 1e-300)
        assert np.abs(mine[4:] - ref2).max() <= 1e-13 * max(np.abs(ref2).max(), 1e-300)


def test_hessian_finite_difference(rng):
    """d_j of the gradient component i equals the (i, j) Hessian entry."""
    pairs = {(0, 0): 4, (0, 1): 5, (0, 2): 6, (1, 1): 7, (1, 2): 8, (2, 2): 9}
    for l in range(5):
        for pure in (0, 1):
            exps = np.array([0.8, 2.5])
            coefs = np.array([0.6, 0.4])
            cen = np.array([0.1, -0.2, 0.3])
            xyz = rng.normal(size=(20, 3))
            v = oracle.eval_shell2(l, pure, exps, coefs, cen, xyz)
            np.testing.assert_allclose(v[:4], oracle.eval_shell(l, pure, exps, coefs, cen, xyz),
                                       rtol=1e-14, atol=1e-15)
            h = 1e-5
            for (i, j), q in pairs.items():
                e = np.zeros(3)
                e[j] = h
                fd = (oracle.eval_shell(l, pure, exps, coefs, cen, xyz + e)[1 + i] -
                      oracle.eval_shell(l, pure, exps, coefs, cen, xyz - e)[1 + i]) / (2 * h)
                np.testing.assert_allclose(v[q], fd, atol=1e-7)


def test_collocation_gradient_finite_difference(rng):
    for l in range(5):
        exps = np.array([0.8, 2.5])
        coefs = np.array([0.6, 0.4])
        cen = np.zeros(3)
        xyz = rng.normal(size=(20, 3))
        v = oracle.eval_shell(l, 1, exps, coefs, cen, xyz)
        h = 1e-5
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fd = (oracle.eval_shell(l, 1, exps, coefs, cen, xyz + e, 0)[0] -
                  oracle.eval_shell(l, 1, exps, coefs, cen, xyz - e, 0)[0]) / (2 * h)
            np.testing.assert_allclose(v[1 + k], fd, atol=1e-8)


def test_overlap_matrix_symmetric_positive():
    basis = AOBasis.load(molecules.water(), "6-31g", pure=False)
    S = overlap_matrix(basis)
    np.testing.assert_allclose(S, S.T, atol=1e-14)
    assert np.linalg.eigvalsh(S).min() > 0
