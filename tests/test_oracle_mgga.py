"""Pins the oracle's meta-GGA path (TPSS exchange and correlation and the
tau terms of the XC build).

What pins what:
- TPSS exchange and correlation energy densities and the correlation's
  vrho / vsigma / vtau at four points: the reference's golden
  (tests/dft_tests.cpp:138-190, tolerance 1e-6 there; matched here to 1e-8);
- derivatives away from tau = tau_W: complex-step against central finite
  differences in rho, sigma and tau;
- uniform-gas limits: at sigma = 0 and tau = tau_unif TPSS exchange is Slater
  exchange (F_x = 1), and at sigma = 0 TPSS correlation is PW92(mod)
  correlation (z = 0 removes every tau term);
- the tau term of the build: dExc/dD = 2 tr(Vxc Delta) for symmetric Delta,
  and a brute-force point-by-point build of a tiny case.
The b = 0.40 parameter of TPSS exchange enters only where alpha is neither
0 nor 1; there it is pinned by the paper alone.
"""
import numpy as np
import pytest

import grid_ref
import oracle
from conftest import random_density
from occ_amd import AOBasis, molecules
from occ_amd.grid import GridSettings, MolecularGrid

TPSS = [202, 231]

# tests/dft_tests.cpp:141-147: rows (rho, grad rho, laplacian, tau)
GOLDEN_IN = np.array([
    [0.06019401627252133, -0.12747174847145987, -0.0, 0.03731311973522952,
     0.09472176570942804, 0.036634279981580156],
    [0.06019401627252133, -0.0, -0.12747174847145987, 0.03731311973522952,
     0.09472176570942804, 0.036634279981580156],
    [0.30171849914696935, -0.0, -0.0, 0.26070301592021455, -2.594981941107205,
     0.02815789498408782],
    [0.3550246251878335, -0.0, -0.0, 0.15685208520532756, -6.307746108179051,
     0.008662278222335566]])


def golden_inputs():
    rho = GOLDEN_IN[:, 0]
    sigma = (GOLDEN_IN[:, 1:4] ** 2).sum(1)
    return rho, sigma, GOLDEN_IN[:, 5]


def test_tpss_reference_golden():
    rho, sigma, tau = golden_inputs()
    ex = oracle.functional_eval_mgga(202, rho, sigma, tau)
    np.testing.assert_allclose(ex[0], [-0.33175882, -0.33175882, -0.56154382, -0.59257418],
                               atol=1e-8)
    exc, vr, vs, vt = oracle.functional_eval_mgga(231, rho, sigma, tau)
    np.testing.assert_allclose(exc, [-0.02276044, -0.02276044, -0.04168478, -0.04268936],
                               atol=1e-8)
    np.testing.assert_allclose(vr, [-5.57565777e-02, -5.57565777e-02, -8.00593613e-02,
                                    -8.53775165e-02], atol=1e-8)
    np.testing.assert_allclose(vs, [3.91761083e-02, 3.91761083e-02, 1.47458233e-01,
                                    5.51608074e-01], atol=1e-8)
    np.testing.assert_allclose(vt, [1.10283011e-03, 1.10283011e-03, -3.50002323e-01,
                                    -1.56487879e+00], atol=1e-8)


def random_mgga_points(rng, n):
    rho = 10 ** rng.uniform(-4, 1, n)
    sigma = 10 ** rng.uniform(-3, 1, n) * rho ** (8 / 3)
    tau_w = sigma / (8 * rho)
    tau = tau_w * 10 ** rng.uniform(0.05, 2.0, n)  # tau_W / tau in (0.01, 0.9)
    return rho, sigma, tau


@pytest.mark.parametrize("fid", TPSS)
def test_complex_step_matches_finite_difference(fid, rng):
    rho, sigma, tau = random_mgga_points(rng, 40)
    _, vr, vs, vt = oracle.functional_eval_mgga(fid, rho, sigma, tau)
    for i in range(len(rho)):
        x = [rho[i], sigma[i], tau[i]]
        for k, v in enumerate((vr[i], vs[i], vt[i])):
            h = 1e-5 * x[k]
            def e(d):
                y = list(x)
                y[k] += d
                return oracle.energy_density_mgga(fid, *y)
            fd = (8 * (e(h) - e(-h)) - (e(2 * h) - e(-2 * h))) / (12 * h)
            scale = abs(oracle.energy_density_mgga(fid, *x)) / x[k]
            assert v == pytest.approx(fd, rel=1e-6, abs=1e-9 * scale)


def test_uniform_gas_limits(rng):
    rho = 10 ** rng.uniform(-4, 1, 30)
    tau_unif = 0.3 * (3 * np.pi ** 2) ** (2 / 3) * rho ** (5 / 3)
    zero = np.full_like(rho, 1e-300)
    ex = oracle.functional_eval_mgga(202, rho, zero, tau_unif)[0]
    slater = oracle.functional_eval(1, rho)[0]
    np.testing.assert_allclose(ex, slater, rtol=1e-12)
    ec = oracle.functional_eval_mgga(231, rho, zero, tau_unif)[0]
    pw = oracle.functional_eval(13, rho)[0]
    np.testing.assert_allclose(ec, pw, rtol=1e-10)


def test_fermi_hole_curvature_limit():
    """sigma above 8 rho tau is evaluated as sigma = 8 rho tau."""
    r, t = 0.2, 0.01
    s_max = 8 * r * t
    for fid in TPSS:
        a = oracle.functional_eval_mgga(fid, [r], [s_max], [t])
        b = oracle.functional_eval_mgga(fid, [r], [2 * s_max], [t])
        for u, v in zip(a, b):
            assert u[0] == v[0]


@pytest.fixture(scope="module")
def water_case():
    atoms = molecules.water()
    basis = AOBasis.load(atoms, "6-31g", pure=False)
    g = MolecularGrid(atoms, GridSettings(radial_points=60, angular_points=194))
    pts, w = grid_ref.molecular_grid(g)
    return basis, oracle.OracleBasis.from_flat(basis.flat()), pts, w


def test_vxc_is_the_derivative_of_exc(water_case):
    basis, ob, pts, w = water_case
    D = random_density(basis.nbf, 5, 5)
    r = np.random.default_rng(19)
    Delta = r.normal(size=D.shape) * 0.01
    Delta = Delta + Delta.T
    K, _, _ = oracle.xc_build(ob, pts, w, TPSS, D)
    eps = 1e-4
    _, ep, _ = oracle.xc_build(ob, pts, w, TPSS, D + eps * Delta)
    _, em, _ = oracle.xc_build(ob, pts, w, TPSS, D - eps * Delta)
    fd = (ep - em) / (2 * eps)
    assert fd == pytest.approx(2 * np.sum(K * Delta), rel=1e-6, abs=1e-9)
    np.testing.assert_allclose(K, K.T, atol=1e-13)


def test_screened_equals_dense(water_case):
    basis, ob, pts, w = water_case
    D = random_density(basis.nbf, 5, 3)
    K1, e1, n1 = oracle.xc_build(ob, pts, w, TPSS, D, screened=True)
    K2, e2, n2 = oracle.xc_build(ob, pts, w, TPSS, D, screened=False)
    assert abs(e1 - e2) < 1e-8 and abs(n1 - n2) < 1e-8
    assert np.abs(K1 - K2).max() < 1e-8


def test_brute_force_tiny():
    atoms = molecules.water()
    basis = AOBasis.load(atoms, "sto-3g", pure=False)
    ob = oracle.OracleBasis.from_flat(basis.flat())
    rng = np.random.default_rng(12)
    pts = rng.normal(size=(300, 3)) * 1.5
    w = rng.uniform(0.01, 0.1, size=300)
    D = random_density(basis.nbf, 5, 8)
    fb = basis.first_bf()
    phi = np.zeros((4, basis.nbf, len(w)))
    for s, sh in enumerate(basis.shells):
        phi[:, fb[s]:fb[s] + sh.size()] = oracle.eval_shell(sh.l, sh.pure, sh.exponents,
                                                             sh.coefficients, sh.origin, pts)
    D2 = 2 * D
    rho = np.einsum("mp,mn,np->p", phi[0], D2, phi[0])
    grad = np.stack([2 * np.einsum("mp,mn,np->p", phi[k], D2, phi[0]) for k in (1, 2, 3)])
    tau = 0.5 * sum(np.einsum("mp,mn,np->p", phi[k], D2, phi[k]) for k in (1, 2, 3))
    sigma = (grad ** 2).sum(0)
    exc, vr, vs, vt = (np.zeros_like(rho) for _ in range(4))
    for f in TPSS:
        a, b, c, d = oracle.functional_eval_mgga(f, rho, sigma, tau)
        exc += a
        vr += b
        vs += c
        vt += d
    E = np.sum(w * rho * exc)
    A = 0.5 * (w * vr) * phi[0] + 2 * (w * vs) * np.einsum("kp,kmp->mp", grad, phi[1:])
    K = phi[0] @ A.T
    K = K + K.T
    K += sum(phi[k] @ (0.5 * w * vt * phi[k]).T for k in (1, 2, 3))
    Ko, Eo, No = oracle.xc_build(ob, pts, w, TPSS, D, density_threshold=0.0, screened=False)
    assert Eo == pytest.approx(E, rel=1e-12, abs=1e-14)
    assert No == pytest.approx(np.sum(w * rho), rel=1e-12)
    np.testing.assert_allclose(Ko, K, atol=1e-12)


def test_gradient_and_unrestricted_reject_mgga(water_case):
    basis, ob, pts, w = water_case
    D = random_density(basis.nbf, 5, 3)
    with pytest.raises(ValueError):
        oracle.xc_build_unrestricted(ob, pts, w, TPSS, np.vstack([D, D]))


@pytest.fixture(scope="module")
def mgga_shim(tmp_path_factory):
    """The GPU kernels' meta-GGA functional header compiled for the host."""
    import ctypes
    import pathlib
    import subprocess
    root = pathlib.Path(__file__).resolve().parent.parent
    d = tmp_path_factory.mktemp("mgga_shim")
    src = d / "shim.cpp"
    src.write_text('#include "xc_functionals_mgga.hpp"\n'
                   'extern "C" int shim(int id, double r, double s, double t, double *e, '
                   'double *v, double *w, double *x) {\n'
                   '  return occxc::xc_point_mgga(id, r, s, t, *e, *v, *w, *x) ? 0 : -1; }\n')
    so = d / "shim.so"
    subprocess.run(["g++", "-O2", "-fPIC", "-shared", f"-I{root / 'occ_amd' / 'csrc'}",
                    str(src), "-o", str(so)], check=True)
    L = ctypes.CDLL(str(so))
    L.shim.argtypes = [ctypes.c_int] + [ctypes.c_double] * 3 + \
        [ctypes.POINTER(ctypes.c_double)] * 4
    return L


@pytest.mark.parametrize("fid", TPSS)
def test_gpu_mgga_header_matches_oracle(mgga_shim, fid, rng):
    import ctypes
    out = [ctypes.c_double() for _ in range(4)]
    rho, sigma, tau = random_mgga_points(rng, 200)
    r0, s0, t0 = golden_inputs()
    rho, sigma, tau = (np.concatenate([a, b]) for a, b in ((rho, r0), (sigma, s0), (tau, t0)))
    e0, v0, w0, x0 = oracle.functional_eval_mgga(fid, rho, sigma, tau)
    for i in range(len(rho)):
        assert mgga_shim.shim(fid, rho[i], sigma[i], tau[i], *(ctypes.byref(q) for q in out)) == 0
        e, v, w, x = (q.value for q in out)
        assert e == pytest.approx(e0[i] * rho[i], rel=1e-12)
        assert v == pytest.approx(v0[i], rel=1e-10, abs=1e-14)
        assert w == pytest.approx(w0[i], rel=1e-10, abs=1e-14 * abs(w0[i]) + 1e-300)
        assert x == pytest.approx(x0[i], rel=1e-10, abs=1e-14)
