"""Pins the oracle's functionals, and checks the GPU kernels' analytic
derivatives (the __host__ __device__ header compiled for the host) against
them.

What pins what:
- gga_x_pbe exc at four points: the reference's golden (tests/dft_tests.cpp:
  71-111, tolerance 1e-6 there; matched here to 5e-9);
- vrho / vsigma: complex-step derivatives of the oracle's energy density
  checked against central finite differences;
- limits: PBE-x and B88 -> Slater at sigma = 0; PBE-c -> PW92(mod) at sigma = 0;
- H-atom exchange energies with the exact density n = exp(-2r)/pi, fully
  spin polarised (E_x[n, 0] = E_x^unpol[2n] / 2): LSDA -0.2680 (closed form),
  B88 -0.3098 (Becke, PRA 38, 3098 (1988), Table I), PBE -0.3059 (Perdew,
  Burke, Ernzerhof, PRL 77, 3865 (1996));
- the correlation functionals (VWN5, PW92, PBE-c, LYP) by their papers'
  limits and identities in tests/test_oracle_functional_pins.py, and, on a
  molecule with OCC's default grid, by the reference's whole-path known
  answers (tests/isolated_xc_tests.cpp) in tests/test_oracle_known_answers.py.
"""
import ctypes
import pathlib
import subprocess

import numpy as np
import pytest

import oracle

ROOT = pathlib.Path(__file__).resolve().parent.parent
IDS = [1, 7, 12, 13, 101, 106, 130, 131]


def test_pbe_x_reference_golden():
    rho = np.array([3.009700173159170558e-02, 3.009700173159170558e-02,
                    1.508591677559790178e-01, 1.775122853194434636e-01])
    g = np.array([[-6.373586084157208120e-02, 0.0, 1.865655330995227498e-02],
                  [0.0, -6.373586084157208120e-02, 1.865655330995227498e-02],
                  [0.0, 0.0, 1.303514966003518905e-01],
                  [0.0, 0.0, 7.842601108050306635e-02]])
    sigma = (g * g).sum(1)
    exc, _, _ = oracle.functional_eval(101, rho, sigma)
    np.testing.assert_allclose(exc, [-0.27851489, -0.27851489, -0.39899553, -0.41654061],
                               atol=5e-9)


@pytest.mark.parametrize("fid", IDS)
def test_complex_step_matches_finite_difference(fid, rng):
    for _ in range(60):
        r = 10 ** rng.uniform(-3, 1)
        s = 10 ** rng.uniform(-1, 1) * r ** (8 / 3)
        _, vr, vs = oracle.functional_eval(fid, [r], [s])
        # 4th-order central difference
        def d(f, x, h):
            return (-f(x + 2 * h) + 8 * f(x + h) - 8 * f(x - h) + f(x - 2 * h)) / (12 * h)
        fd_r = d(lambda x: oracle.energy_density(fid, x, s), r, 1e-4 * r)
        assert abs(fd_r - vr[0]) <= 1e-7 * abs(vr[0]) + 1e-12
        if fid > 100:
            fd_s = d(lambda x: oracle.energy_density(fid, r, x), s, 1e-3 * s)
            assert abs(fd_s - vs[0]) <= 1e-6 * abs(vs[0]) + 1e-12


def test_gga_limits(rng):
    rho = 10 ** rng.uniform(-6, 2, 200)
    zero = np.zeros_like(rho)
    ex, vx, _ = oracle.functional_eval(1, rho)
    for fid in (101, 106):
        e, v, _ = oracle.functional_eval(fid, rho, zero)
        np.testing.assert_allclose(e, ex, rtol=1e-12)
        np.testing.assert_allclose(v, vx, rtol=1e-12)
    ec, vc, _ = oracle.functional_eval(13, rho)
    e, v, _ = oracle.functional_eval(130, rho, zero)
    np.testing.assert_allclose(e, ec, rtol=1e-10)
    np.testing.assert_allclose(v, vc, rtol=1e-9)


def _hydrogen_exchange(fid):
    # Gauss-Laguerre-like radial quadrature on r in (0, 40]
    x, w = np.polynomial.legendre.leggauss(400)
    t = 0.5 * (x + 1.0)
    r = 40.0 * t ** 3
    wr = 40.0 * 3 * t ** 2 * 0.5 * w
    n = np.exp(-2 * r) / np.pi
    rho = 2 * n                      # unpolarised density 2n
    sigma = (2 * 2 * n) ** 2         # |grad(2n)|^2 = (4n)^2
    e = np.array([oracle.energy_density(fid, a, b) for a, b in zip(rho, sigma)])
    return 0.5 * np.sum(4 * np.pi * r * r * e * wr)


def test_hydrogen_atom_exchange_energies():
    assert _hydrogen_exchange(1) == pytest.approx(-0.26803, abs=2e-5)
    assert _hydrogen_exchange(106) == pytest.approx(-0.3098, abs=5e-4)
    assert _hydrogen_exchange(101) == pytest.approx(-0.3059, abs=5e-4)


def test_below_threshold_contributes_nothing():
    for fid in IDS:
        e, v, s = oracle.functional_eval(fid, [1e-16, 0.0], [1e-30, 0.0])
        assert np.all(e == 0) and np.all(v == 0) and np.all(s == 0)


@pytest.fixture(scope="module")
def host_shim(tmp_path_factory):
    d = tmp_path_factory.mktemp("shim")
    src = d / "shim.cpp"
    src.write_text('#include "xc_functionals.hpp"\n'
                   'extern "C" int shim(int id, double r, double s, double *e, double *v, '
                   'double *w) { return occxc::xc_point(id, r, s, *e, *v, *w) ? 0 : -1; }\n')
    so = d / "shim.so"
    subprocess.run(["g++", "-O2", "-fPIC", "-shared", f"-I{ROOT / 'occ_amd' / 'csrc'}",
                    str(src), "-o", str(so)], check=True)
    L = ctypes.CDLL(str(so))
    L.shim.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double] + \
        [ctypes.POINTER(ctypes.c_double)] * 3
    return L


@pytest.mark.parametrize("fid", IDS)
def test_gpu_functional_header_matches_oracle(host_shim, fid, rng):
    """The analytic derivatives the GPU kernels use, compiled for the host."""
    e, v, w = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    for _ in range(300):
        r = 10 ** rng.uniform(-10, 3)
        s = 10 ** rng.uniform(-12, 2) * r ** (8 / 3)
        assert host_shim.shim(fid, r, s, ctypes.byref(e), ctypes.byref(v), ctypes.byref(w)) == 0
        oe, ov, os_ = oracle.functional_eval(fid, [r], [s])
        assert e.value == pytest.approx(oe[0] * r, rel=1e-11)
        assert v.value == pytest.approx(ov[0], rel=1e-10)
        if fid > 100:
            assert w.value == pytest.approx(os_[0], rel=1e-10)
