"""GPU parity of the restricted meta-GGA (TPSS) XC build
(compute_K_dft<2, Restricted>, include/occ/dft/dft.h:190-388, with
evaluate_density<2, R>, src/gto/density.cpp:75-102, and
xc_potential_matrix<Restricted, 2>, src/dft/xc_potential_matrix.cpp:44-58)
and of the TPSS functionals against the CPU oracle, through the C ABI.
Tolerances as for the restricted build: |dExc| <= 1e-6 Ha,
max|dVxc| <= 1e-7, |dN| <= 1e-8 in FP64."""
import numpy as np
import pytest

import oracle
from conftest import random_density
from occ_amd import AOBasis, molecules
from occ_amd.dft import DFT, DensityFunctional, XCBatch, XCContext
from occ_amd.grid import GridSettings, MolecularGrid
from test_oracle_mgga import golden_inputs

pytestmark = pytest.mark.gpu

TOL_E, TOL_V, TOL_N = 1e-6, 1e-7, 1e-8
TPSS = [202, 231]


def _grid(atoms, nrad=31, nang=110):
    g = MolecularGrid(atoms, GridSettings(radial_points=nrad, angular_points=nang))
    return g.build(device=0)


@pytest.mark.parametrize("name,fid", [("mgga_x_tpss", 202), ("mgga_c_tpss", 231)])
def test_tpss_matches_oracle_and_reference_golden(rng, name, fid):
    rho, sigma, tau = golden_inputs()
    exc, vr, vs, vt = DensityFunctional(name).evaluate_mgga(rho, sigma, tau)
    e0, r0, s0, t0 = oracle.functional_eval_mgga(fid, rho, sigma, tau)
    for a, b in ((exc, e0), (vr, r0), (vs, s0), (vt, t0)):
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-14)
    n = 4000
    r = 10 ** rng.uniform(-8, 3, n)
    s = 10 ** rng.uniform(-6, 2, n) * r ** (8 / 3)
    t = s / (8 * r) * 10 ** rng.uniform(-0.3, 3, n)  # includes tau < tau_W (limited)
    exc, vr, vs, vt = DensityFunctional(name).evaluate_mgga(r, s, t)
    e0, r0, s0, t0 = oracle.functional_eval_mgga(fid, r, s, t)
    np.testing.assert_allclose(exc, e0, rtol=1e-11, atol=1e-300)
    np.testing.assert_allclose(vr, r0, rtol=1e-9, atol=1e-14 * np.abs(r0).max())
    np.testing.assert_allclose(vs, s0, rtol=1e-9, atol=1e-14 * np.abs(s0).max())
    np.testing.assert_allclose(vt, t0, rtol=1e-9, atol=1e-14 * np.abs(t0).max())


def _case(atoms, basis_name, seed, pure, precision="fp64", screen=True, nrad=31, nang=110,
          method="tpss"):
    basis = AOBasis.load(atoms, basis_name, pure=pure)
    pts, wts = _grid(atoms, nrad, nang)
    nocc = max(1, molecules.nuclear_charge(atoms) // 2)
    D = random_density(basis.nbf, min(nocc, basis.nbf), seed)
    print(f"mgga case {basis_name} {method} nbf={basis.nbf} npts={len(wts)} {precision} "
          f"screen={screen}", flush=True)
    dft = DFT(method, basis, pts, wts, precision=precision, screen=screen)
    V = dft.compute_vxc(D)
    ob = oracle.OracleBasis.from_flat(basis.flat())
    K, exc, nel = oracle.xc_build(ob, pts, wts, dft.ids, D, screened=screen)
    return (abs(dft.exchange_correlation_energy() - exc), float(np.abs(V - K).max()),
            abs(dft.total_density() - nel))


def test_tpss_water_631g():
    de, dv, dn = _case(molecules.water(), "6-31g", 11, False)
    print(f"TPSS water/6-31G: |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}")
    assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N


def test_tpss_water_dimer_def2svp():
    de, dv, dn = _case(molecules.water_cluster(2, seed=1), "def2-svp", 12, True)
    print(f"TPSS (H2O)2/def2-SVP: |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}")
    assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N


def test_tpss_def2tzvp_f_shells_and_dense_path():
    de, dv, dn = _case(molecules.water(), "def2-tzvp", 13, True, nrad=25, nang=86)
    assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N
    de, dv, dn = _case(molecules.water(), "6-31g", 14, False, screen=False)
    assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N


def test_mixed_gga_and_mgga_components():
    """A GGA component evaluated in a meta-GGA build (tau unused)."""
    de, dv, dn = _case(molecules.water(), "6-31g", 15, False, method=["mgga_x_tpss",
                                                                       "gga_c_pbe"])
    assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N


def test_tpss_fp32_cache():
    de, dv, dn = _case(molecules.water_cluster(2, seed=1), "def2-svp", 16, True,
                       precision="fp32cache")
    print(f"TPSS fp32cache: |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}")
    assert de <= 1e-5 and dv <= 1e-6 and dn <= 1e-5


def test_tpss_batch_matches_single():
    mols = [molecules.water(), molecules.alkane(2)]
    bases = [AOBasis.load(m, "6-31g", pure=False) for m in mols]
    grids = [_grid(m, 21, 50) for m in mols]
    Ds = [random_density(b.nbf, 4, 50 + i) for i, b in enumerate(bases)]
    batch = XCBatch(bases, grids, TPSS)
    Vs, excs, nels = batch.build(Ds)
    for i, (b, (pts, w)) in enumerate(zip(bases, grids)):
        dft = DFT("tpss", b, pts, w)
        V = dft.compute_vxc(Ds[i])
        assert abs(excs[i] - dft.exchange_correlation_energy()) <= 1e-11
        assert abs(nels[i] - dft.total_density()) <= 1e-11
        assert np.abs(Vs[i] - V).max() <= 1e-11
    batch.close()


def test_tpss_graph_replay_matches_direct_builds():
    from occ_amd.device import DeviceArray, synchronize
    atoms = molecules.water_cluster(2, seed=3)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = _grid(atoms)
    n = basis.nbf
    Ds = [random_density(n, 10, 70 + k) for k in range(2)]
    ctx = XCContext(basis, pts, wts, TPSS)
    dD, dV, dr = DeviceArray(n * n), DeviceArray(n * n), DeviceArray(2)
    direct = []
    for D in Ds:
        dD.upload(D)
        ctx.build_device(dD.ptr, dV.ptr, dr.ptr)
        ctx.synchronize()
        direct.append((dV.download((n, n)), dr.download()))
    ctx.set_graph(True)
    for rep in range(3):
        for k, D in enumerate(Ds):
            dD.upload(D)
            ctx.build_device(dD.ptr, dV.ptr, dr.ptr)
            ctx.synchronize()
            synchronize()
            V, r = dV.download((n, n)), dr.download()
            assert np.abs(V - direct[k][0]).max() <= 1e-13
            assert abs(r[0] - direct[k][1][0]) <= 1e-12 and abs(r[1] - direct[k][1][1]) <= 1e-12
    ctx.close()


def test_mgga_rejections():
    atoms = molecules.water()
    basis = AOBasis.load(atoms, "6-31g", pure=False)
    pts, wts = _grid(atoms, 21, 50)
    with pytest.raises(Exception):
        XCContext(basis, pts, wts, TPSS, precision="fp32")


@pytest.mark.parametrize("name,fid", [("mgga_x_tpss", 202), ("mgga_c_tpss", 231)])
def test_polarised_tpss_matches_oracle(rng, name, fid):
    """occxc_functional_evaluate_mgga_polarized against the oracle's
    polarised TPSS on open-shell points (spin ratios up to 1e3, tau above
    tau_W, a tenth of them with tau below tau_W, which the limits raise)."""
    from test_oracle_mgga import random_mgga_points_pol
    rho, sig, tau = random_mgga_points_pol(rng, 4000)
    tau[:400] *= 0.3
    exc, vr, vs, vt = DensityFunctional(name).evaluate_mgga_polarized(rho, sig, tau)
    e0, r0, s0, t0 = oracle.functional_eval_mgga_pol(fid, rho, sig, tau)
    np.testing.assert_allclose(exc, e0, rtol=1e-11, atol=1e-300)
    np.testing.assert_allclose(vr, r0, rtol=1e-9, atol=1e-13 * np.abs(r0).max())
    np.testing.assert_allclose(vs, s0, rtol=1e-9, atol=1e-13 * np.abs(s0).max())
    np.testing.assert_allclose(vt, t0, rtol=1e-9, atol=1e-13 * np.abs(t0).max())
