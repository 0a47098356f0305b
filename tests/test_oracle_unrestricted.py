"""Pins the oracle's unrestricted (UKS) path: the polarised functionals and
compute_K_dft<d, Unrestricted> (include/occ/dft/dft.h:190-388,
src/dft/dft_kernels.cpp:8-70, src/gto/density.cpp:18-30,45-70,
src/dft/xc_potential_matrix.cpp:61-109).

- closed-shell limit: every polarised functional at (rho/2, rho/2, sigma/4,
  sigma/4, sigma/4) gives the unpolarised exc, vrho_a = vrho_b = vrho and
  (v_aa + v_ab + v_bb)/4 = vsigma, and the unrestricted build with
  D_a = D_b = D/2 gives the restricted Exc and V_a = V_b = Vxc;
- fully polarised Dirac exchange: e = -(3/4)(6/pi)^(1/3) rho^(4/3);
- PW92 at z = 1 is the ferromagnetic fit and its second z-derivative at
  z = 0 is the spin stiffness (Perdew, Wang, PRB 45, 13244, Table I);
- complex-step derivatives against central differences;
- dExc/dD_s(Delta) = 2 tr(V_s Delta) for each spin on an open-shell density,
  and screened == dense.
libxc itself is absent here: between the closed-shell and fully polarised
limits the spin interpolations (VWN5 / PW92 f(zeta), PBE's phi(zeta), the
general LYP form) are pinned by the limits above, the PW92 spin stiffness and
the derivative checks, not by values computed with libxc."""
import numpy as np
import pytest

import grid_ref
import oracle
from conftest import random_density
from occ_amd import AOBasis, molecules
from occ_amd.grid import GridSettings, MolecularGrid

IDS = [1, 7, 12, 13, 101, 106, 130, 131]
METHODS = {"svwn5": [1, 7], "pbe": [101, 130], "blyp": [106, 131]}


@pytest.mark.parametrize("fid", IDS)
def test_polarised_closed_shell_limit(rng, fid):
    rho = 10 ** rng.uniform(-6, 1, 300)
    sig = 10 ** rng.uniform(-4, 1, 300) * rho ** (8 / 3)
    e_u, v_u, s_u = oracle.functional_eval(fid, rho, sig)
    e_p, v_p, s_p = oracle.functional_eval_pol(fid, np.stack([rho / 2, rho / 2], 1),
                                               np.stack([sig / 4] * 3, 1))
    np.testing.assert_allclose(e_p, e_u, rtol=1e-13)
    np.testing.assert_allclose(v_p[:, 0], v_u, rtol=1e-13)
    np.testing.assert_allclose(v_p[:, 1], v_u, rtol=1e-13)
    if fid > 100:
        np.testing.assert_allclose(s_p.sum(1) / 4, s_u, rtol=1e-12)


def test_fully_polarised_dirac_exchange(rng):
    rho = 10 ** rng.uniform(-4, 2, 50)
    for r in rho:
        e = oracle.energy_density_pol(1, r, 0.0)
        assert e == pytest.approx(-0.75 * (6 / np.pi) ** (1 / 3) * r ** (4 / 3), rel=1e-13)


def _pw_G(rs, A, a1, b1, b2, b3, b4):
    den = 2 * A * (b1 * rs ** 0.5 + b2 * rs + b3 * rs ** 1.5 + b4 * rs ** 2)
    return -2 * A * (1 + a1 * rs) * np.log1p(1 / den)


@pytest.mark.parametrize("rs", [0.5, 1.0, 3.0, 8.0])
def test_pw92_endpoints(rs):
    rho = 3 / (4 * np.pi * rs ** 3)
    # z = 1: the ferromagnetic fit (lda_c_pw_mod constants)
    eps1 = oracle.energy_density_pol(13, rho, 0.0) / rho
    assert eps1 == pytest.approx(_pw_G(rs, 0.01554535, 0.20548, 14.1189, 6.1977, 3.3662,
                                       0.62517), rel=1e-12)
    # second z-derivative at z = 0: the spin stiffness alpha_c = -G(alpha fit)
    h = 1e-3
    e = [oracle.energy_density_pol(13, rho * (1 + z) / 2, rho * (1 - z) / 2) / rho
         for z in (h, 0.0, -h)]
    d2 = (e[0] - 2 * e[1] + e[2]) / h ** 2
    alpha = -_pw_G(rs, 0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671)
    assert d2 == pytest.approx(alpha, rel=2e-5)


@pytest.mark.parametrize("fid", IDS)
def test_polarised_derivatives_finite_difference(rng, fid):
    for _ in range(20):
        ra, rb = 10 ** rng.uniform(-3, 0.5, 2)
        ga, gb = rng.normal(size=3) * ra ** (4 / 3), rng.normal(size=3) * rb ** (4 / 3)
        x = np.array([ra, rb, ga @ ga, ga @ gb, gb @ gb])
        _, v, s = oracle.functional_eval_pol(fid, x[None, :2], x[None, 2:])
        grad = np.concatenate([v[0], s[0]])
        nv = 5 if fid > 100 else 2
        e0 = abs(oracle.energy_density_pol(fid, *x))
        for k in range(nv):
            # fourth-order central difference; its rounding noise is
            # ~1e-16 |e| / h, which sets the absolute tolerance
            h = 1e-3 * abs(x[k])

            def f(t):
                y = x.copy()
                y[k] += t
                return oracle.energy_density_pol(fid, *y)
            fd = (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)
            assert grad[k] == pytest.approx(fd, rel=1e-7, abs=1e-12 * e0 / h)


@pytest.fixture(scope="module")
def dimer():
    atoms = molecules.water_cluster(2, seed=1)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    g = MolecularGrid(atoms, GridSettings(radial_points=41, angular_points=146))
    pts, w = grid_ref.molecular_grid(g)
    return basis, oracle.OracleBasis.from_flat(basis.flat()), pts, w


@pytest.mark.parametrize("method", list(METHODS))
def test_unrestricted_closed_shell_equals_restricted(dimer, method):
    basis, ob, pts, w = dimer
    n = basis.nbf
    D = random_density(n, 10, 3)
    K, e, nel = oracle.xc_build(ob, pts, w, METHODS[method], D)
    Ku, eu, (na, nb) = oracle.xc_build_unrestricted(ob, pts, w, METHODS[method],
                                                    np.vstack([D / 2, D / 2]))
    assert eu == pytest.approx(e, abs=1e-12)
    assert np.abs(Ku[:n] - K).max() < 1e-12 and np.abs(Ku[n:] - K).max() < 1e-12
    assert na == pytest.approx(nel / 2, abs=1e-12) and nb == pytest.approx(nel / 2, abs=1e-12)


@pytest.mark.parametrize("method", list(METHODS))
def test_unrestricted_potential_is_the_derivative(dimer, method):
    basis, ob, pts, w = dimer
    n = basis.nbf
    Du = np.vstack([random_density(n, 10, 5) / 2, random_density(n, 9, 6) / 2])
    ids = METHODS[method]
    Ku, eu, _ = oracle.xc_build_unrestricted(ob, pts, w, ids, Du)
    Ks, es, _ = oracle.xc_build_unrestricted(ob, pts, w, ids, Du, screened=False)
    assert abs(eu - es) < 1e-8 and np.abs(Ku - Ks).max() < 1e-8
    r = np.random.default_rng(9)
    for sp in (0, 1):
        Delta = r.normal(size=(n, n)) * 0.01
        Delta = Delta + Delta.T
        eps = 1e-4
        Dp, Dm = Du.copy(), Du.copy()
        Dp[sp * n:(sp + 1) * n] += eps * Delta
        Dm[sp * n:(sp + 1) * n] -= eps * Delta
        fd = (oracle.xc_build_unrestricted(ob, pts, w, ids, Dp)[1] -
              oracle.xc_build_unrestricted(ob, pts, w, ids, Dm)[1]) / (2 * eps)
        assert fd == pytest.approx(2 * np.sum(Ku[sp * n:(sp + 1) * n] * Delta), rel=1e-7,
                                   abs=1e-9)


@pytest.fixture(scope="module")
def pol_shim(tmp_path_factory):
    """The polarised functional header the GPU kernels use, compiled for the
    host."""
    import ctypes
    import pathlib
    import subprocess
    root = pathlib.Path(__file__).resolve().parent.parent
    d = tmp_path_factory.mktemp("polshim")
    src = d / "shim.cpp"
    src.write_text('#include "xc_functionals_pol.hpp"\n'
                   'extern "C" int shim(int id, const double *x, double *out) {\n'
                   '  out[0] = 0; for (int k = 1; k < 6; ++k) out[k] = 0;\n'
                   '  return occxc::xc_point_pol(id, x[0], x[1], x[2], x[3], x[4], out[0],\n'
                   '                             out + 1, out + 3) ? 0 : -1; }\n')
    so = d / "shim.so"
    subprocess.run(["g++", "-O2", "-fPIC", "-shared", f"-I{root / 'occ_amd' / 'csrc'}",
                    str(src), "-o", str(so)], check=True)
    L = ctypes.CDLL(str(so))
    L.shim.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                       ctypes.POINTER(ctypes.c_double)]
    return L


@pytest.mark.parametrize("fid", IDS)
def test_gpu_polarised_header_matches_oracle(pol_shim, fid, rng):
    import ctypes
    out = (ctypes.c_double * 6)()
    for _ in range(300):
        # spin polarisation up to rho_a / rho_b = 1e4 either way; beyond that
        # 1 - zeta is formed by cancellation and the two implementations
        # round differently (both are exact forms of the same expression)
        ra = 10 ** rng.uniform(-8, 2)
        rb = ra * 10 ** rng.uniform(-4, 4)
        ga, gb = rng.normal(size=3) * ra ** (4 / 3), rng.normal(size=3) * rb ** (4 / 3)
        x = np.array([ra, rb, ga @ ga, ga @ gb, gb @ gb]) if fid > 100 else \
            np.array([ra, rb, 0.0, 0.0, 0.0])
        assert pol_shim.shim(fid, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out) == 0
        oe, ov, os_ = oracle.functional_eval_pol(fid, x[None, :2], x[None, 2:])
        scale = abs(oe[0] * (ra + rb))
        assert out[0] == pytest.approx(oe[0] * (ra + rb), rel=1e-11)
        for k in range(2):
            assert out[1 + k] == pytest.approx(ov[0, k], rel=1e-9, abs=1e-14 * scale / (ra + rb))
        if fid > 100:
            for k in range(3):
                assert out[3 + k] == pytest.approx(os_[0, k], rel=1e-9,
                                                   abs=1e-12 * abs(os_[0]).max())


@pytest.mark.parametrize("fid", IDS)
def test_gpu_polarised_header_one_spin_empty(pol_shim, fid, rng):
    """rho_b = 0 exactly (no beta electrons): finite, and equal to the oracle."""
    import ctypes
    out = (ctypes.c_double * 6)()
    for _ in range(100):
        ra = 10 ** rng.uniform(-6, 2)
        ga = rng.normal(size=3) * ra ** (4 / 3)
        x = np.array([ra, 0.0, ga @ ga if fid > 100 else 0.0, 0.0, 0.0])
        assert pol_shim.shim(fid, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out) == 0
        assert np.all(np.isfinite(np.array(out[:6])))
        oe, ov, os_ = oracle.functional_eval_pol(fid, x[None, :2], x[None, 2:])
        assert out[0] == pytest.approx(oe[0] * ra, rel=1e-11)
        assert out[1] == pytest.approx(ov[0, 0], rel=1e-9)
        assert out[2] == pytest.approx(ov[0, 1], rel=1e-8, abs=1e-12 * abs(ov[0, 0]))
