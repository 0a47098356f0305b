"""GPU parity of the unrestricted (UKS) XC build (occxc_build_unrestricted,
compute_K_dft<d, Unrestricted>, include/occ/dft/dft.h:190-388) and of the
polarised functionals against the CPU oracle, through the C ABI.
Tolerances (stated per result, as for the restricted build):
|dExc| <= 1e-6 Ha, max|dVxc| <= 1e-7 (both spin blocks), |dN_s| <= 1e-8 for
the FP64 path."""
import ctypes

import numpy as np
import pytest

import oracle
from conftest import random_density
from occ_amd import AOBasis, _lib, molecules
from occ_amd.dft import DFT, DensityFunctional, XCBatch, XCContext, functional_ids
from occ_amd.grid import GridSettings, MolecularGrid

pytestmark = pytest.mark.gpu

TOL_E, TOL_V, TOL_N = 1e-6, 1e-7, 1e-8
NAMES = {1: "lda_x", 7: "lda_c_vwn", 12: "lda_c_pw", 13: "lda_c_pw_mod", 101: "gga_x_pbe",
         106: "gga_x_b88", 130: "gga_c_pbe", 131: "gga_c_lyp"}


def _grid(atoms, nrad=31, nang=110):
    g = MolecularGrid(atoms, GridSettings(radial_points=nrad, angular_points=nang))
    return g.build(device=0)


def _open_shell(nbf, na, nb, seed):
    """alpha over beta blocks, each C_s C_s^T / 2 (orb.h:68-82)"""
    return np.vstack([random_density(nbf, na, seed) / 2, random_density(nbf, nb, seed + 1) / 2])


@pytest.mark.parametrize("fid", sorted(NAMES))
def test_polarised_functional_matches_oracle(rng, fid):
    n = 3000
    ra = 10 ** rng.uniform(-10, 3, n)
    rb = ra * 10 ** rng.uniform(-4, 4, n)
    ga = rng.normal(size=(n, 3)) * (ra ** (4 / 3))[:, None] * 10 ** rng.uniform(-3, 1, (n, 1))
    gb = rng.normal(size=(n, 3)) * (rb ** (4 / 3))[:, None] * 10 ** rng.uniform(-3, 1, (n, 1))
    rho = np.stack([ra, rb], 1)
    sig = np.stack([(ga * ga).sum(1), (ga * gb).sum(1), (gb * gb).sum(1)], 1)
    exc, vr, vs = DensityFunctional(NAMES[fid]).evaluate_polarized(rho, sig)
    e0, v0, s0 = oracle.functional_eval_pol(fid, rho, sig)
    np.testing.assert_allclose(exc, e0, rtol=1e-11, atol=1e-300)
    scale = np.abs(e0)[:, None]
    np.testing.assert_allclose(vr, v0, rtol=1e-9, atol=0)
    if fid > 100:
        np.testing.assert_allclose(vs, s0, rtol=1e-9,
                                   atol=1e-12 * np.abs(s0).max(axis=1, keepdims=True).max())
    assert scale.shape[0] == n


def _case(atoms, basis_name, method, pure, seed, precision="fp64", screen=True):
    basis = AOBasis.load(atoms, basis_name, pure=pure)
    pts, wts = _grid(atoms)
    ne = molecules.nuclear_charge(atoms)
    Du = _open_shell(basis.nbf, (ne + 1) // 2, (ne - 1) // 2, seed)
    print(f"uks case {basis_name} {method} nbf={basis.nbf} npts={len(wts)} {precision}",
          flush=True)
    dft = DFT(method, basis, pts, wts, precision=precision, screen=screen)
    V = dft.compute_vxc(Du)
    ob = oracle.OracleBasis.from_flat(basis.flat())
    K, exc, (na, nb) = oracle.xc_build_unrestricted(ob, pts, wts, dft.ids, Du, screened=screen)
    ga, gb = dft.total_density()
    return (abs(dft.exchange_correlation_energy() - exc), float(np.abs(V - K).max()),
            max(abs(ga - na), abs(gb - nb)))


@pytest.mark.parametrize("method", ["svwn5", "pbe", "blyp", "tpss"])
def test_uks_water_631g(method):
    de, dv, dn = _case(molecules.water(), "6-31g", method, False, 31)
    print(f"UKS water/6-31G {method}: |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}")
    assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N


@pytest.mark.parametrize("method", ["pbe", "blyp", "svwn5", "tpss"])
def test_uks_water_dimer_def2svp(method):
    de, dv, dn = _case(molecules.water_cluster(2, seed=1), "def2-svp", method, True, 33)
    print(f"UKS (H2O)2/def2-SVP {method}: |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}")
    assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N


def test_uks_dense_path_and_fp32_cache():
    de, dv, dn = _case(molecules.water(), "6-31g", "pbe", False, 35, screen=False)
    assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N
    de, dv, dn = _case(molecules.water_cluster(2, seed=1), "def2-svp", "blyp", True, 37,
                       precision="fp32cache")
    print(f"UKS (H2O)2/def2-SVP blyp fp32cache: |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}")
    assert de <= TOL_E and dv <= 1e-6 and dn <= 1e-6


def test_uks_closed_shell_equals_restricted_on_gpu():
    atoms = molecules.water_cluster(2, seed=1)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = _grid(atoms)
    D = random_density(basis.nbf, 10, 39)
    n = basis.nbf
    ctx = XCContext(basis, pts, wts, functional_ids("pbe"))
    V, e, nel = ctx.build(D)
    Vu, eu, (na, nb) = ctx.build_unrestricted(np.vstack([D / 2, D / 2]))
    assert abs(eu - e) <= 1e-11
    assert np.abs(Vu[:n] - V).max() <= 1e-11 and np.abs(Vu[n:] - V).max() <= 1e-11
    assert abs(na - nel / 2) <= 1e-11 and abs(nb - nel / 2) <= 1e-11
    # the device-resident entry point gives the host path's numbers
    from occ_amd.device import DeviceArray
    Du = _open_shell(n, 10, 9, 43)
    Vh, eh, (nah, nbh) = ctx.build_unrestricted(Du)
    dD, dV, dr = DeviceArray(2 * n * n), DeviceArray(2 * n * n), DeviceArray(3)
    dD.upload(np.asfortranarray(Du))
    ctx.build_unrestricted_device(dD.ptr, dV.ptr, dr.ptr)
    ctx.synchronize()
    assert np.array_equal(dr.download(), np.array([eh, nah, nbh]))
    assert np.abs(dV.download((2 * n, n)) - Vh).max() <= 1e-14
    # the restricted build is unaffected by the unrestricted one in between
    V2, e2, nel2 = ctx.build(D)
    assert e2 == e and nel2 == nel and np.abs(V2 - V).max() <= 1e-14


def test_uks_rejections():
    atoms = molecules.water()
    basis = AOBasis.load(atoms, "6-31g", pure=False)
    pts, wts = _grid(atoms, 11, 26)
    Du = _open_shell(basis.nbf, 5, 4, 41)
    L = _lib.lib()
    V = np.zeros((2 * basis.nbf, basis.nbf), order="F")
    e, nel = ctypes.c_double(), np.zeros(2)
    Df = np.asfortranarray(Du)
    ctx = XCContext(basis, pts, wts, functional_ids("svwn5"), precision="fp32")
    st = L.occxc_build_unrestricted(ctx.handle, Df.ctypes.data_as(_lib._dp),
                                    V.ctypes.data_as(_lib._dp), ctypes.byref(e), _lib.dptr(nel))
    assert st == 2 and b"FP64 or FP32-cache" in L.occxc_last_error_message()
    batch = XCBatch([basis, basis], [(pts, wts), (pts, wts)], functional_ids("svwn5"))
    st = L.occxc_build_unrestricted(batch.handle, Df.ctypes.data_as(_lib._dp),
                                    V.ctypes.data_as(_lib._dp), ctypes.byref(e), _lib.dptr(nel))
    assert st == 1
    bad = Df.copy()
    bad[0, 0] = np.inf
    ctx = XCContext(basis, pts, wts, functional_ids("svwn5"))
    st = L.occxc_build_unrestricted(ctx.handle, bad.ctypes.data_as(_lib._dp),
                                    V.ctypes.data_as(_lib._dp), ctypes.byref(e), _lib.dptr(nel))
    assert st == 1


@pytest.mark.parametrize("method", ["svwn5", "pbe", "blyp", "tpss"])
def test_uks_one_electron_beta_density_zero(method):
    """A one-electron system (hydrogen atom, D_beta = 0): rho_beta vanishes
    everywhere and zeta = 1; the build stays finite and matches the oracle."""
    atoms = [(1, 0.0, 0.0, 0.0)]
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = _grid(atoms, 41, 146)
    n = basis.nbf
    Du = np.vstack([random_density(n, 1, 81) / 2, np.zeros((n, n))])
    dft = DFT(method, basis, pts, wts)
    V = dft.compute_vxc(Du)
    ob = oracle.OracleBasis.from_flat(basis.flat())
    K, exc, (na, nb) = oracle.xc_build_unrestricted(ob, pts, wts, dft.ids, Du)
    ga, gb = dft.total_density()
    print(f"H atom UKS {method}: Exc={dft.exchange_correlation_energy():.10f} "
          f"|dExc|={abs(dft.exchange_correlation_energy() - exc):.2e} "
          f"max|dV|={np.abs(V - K).max():.2e} N=({ga:.6f}, {gb:.1e})")
    assert np.all(np.isfinite(V))
    assert abs(dft.exchange_correlation_energy() - exc) <= TOL_E
    assert np.abs(V - K).max() <= TOL_V
    assert abs(ga - na) <= TOL_N and gb == 0.0 and nb == 0.0


def test_uks_tpss_fp32_cache_and_closed_shell_limit():
    """Unrestricted meta-GGA (compute_K_dft<2, Unrestricted>) with the FP32
    AO cache, and D_a = D_b = D/2 against the restricted TPSS build."""
    de, dv, dn = _case(molecules.water_cluster(2, seed=1), "def2-svp", "tpss", True, 39,
                       precision="fp32cache")
    print(f"UKS (H2O)2/def2-SVP tpss fp32cache: |dExc|={de:.2e} max|dVxc|={dv:.2e} "
          f"|dN|={dn:.2e}")
    assert de <= TOL_E and dv <= 1e-6 and dn <= 1e-6
    atoms = molecules.water()
    basis = AOBasis.load(atoms, "6-31g", pure=False)
    pts, wts = _grid(atoms)
    D = random_density(basis.nbf, 5, 41)
    dft = DFT("tpss", basis, pts, wts)
    V = dft.compute_vxc(D)
    e = dft.exchange_correlation_energy()
    Vu = dft.compute_vxc(np.vstack([D, D]) / 2)
    eu = dft.exchange_correlation_energy()
    n = basis.nbf
    assert abs(e - eu) < 1e-10
    assert np.abs(Vu[:n] - V).max() < 1e-10 and np.abs(Vu[n:] - V).max() < 1e-10
