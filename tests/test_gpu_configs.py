"""GPU parity at the benchmark configurations' real sizes, through the C ABI,
against the CPU oracle (the restatement of compute_K_dft, dft.h:190-388).
Tolerances for the FP64 path: |dExc| <= 1e-6 Ha, max|dVxc| <= 1e-7,
|dN_elec| <= 1e-8.

- H2O / 6-31G / SVWN5 on its exact grid (75 radial Euler-Maclaurin x 302
  Lebedev per atom, 67,950 points);
- (H2O)32 / def2-SVP / PBE on 75 x 302 (2,174,400 points, 768 AOs): FP64
  gated, the reduced-precision modes' errors printed;
- a system whose leaves are wider than 2,048 AOs ((H2O)90 / def2-SVP on a
  coarse grid), where the tiled kernels spread one leaf over many
  workgroups, in the tiled and the per-leaf kernels;
- B88+LYP in the FP32 and split-BF16 modes on a small cluster (gated at
  1e-6 there; on the large clusters these modes miss 1e-6, which the
  benchmark reports per mode).
"""
import numpy as np
import pytest

import oracle
from conftest import random_density
from occ_amd import AOBasis, molecules
from occ_amd.dft import DFT
from occ_amd.grid import GridSettings, MolecularGrid

pytestmark = pytest.mark.gpu

TOL_E, TOL_V, TOL_N = 1e-6, 1e-7, 1e-8


def _compare(dft, D, pts, wts):
    V = dft.compute_vxc(D)
    ob = oracle.OracleBasis.from_flat(dft.basis.flat())
    K, exc, nel = oracle.xc_build(ob, pts, wts, dft.ids, D)
    de = abs(dft.exchange_correlation_energy() - exc)
    dv = float(np.abs(V - K).max())
    dn = abs(dft.total_density() - nel)
    return de, dv, dn


def test_water_631g_svwn5_exact_grid():
    atoms = molecules.water()
    basis = AOBasis.load(atoms, "6-31g", pure=False)
    pts, wts = MolecularGrid(atoms, GridSettings(radial_points=76,
                                                 angular_points=302)).build(device=0)
    assert len(wts) == 67950
    D = random_density(basis.nbf, 5, 101)
    de, dv, dn = _compare(DFT("svwn5", basis, pts, wts), D, pts, wts)
    print(f"H2O/6-31G SVWN5 67,950 points: |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}")
    assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N


def test_h2o32_def2svp_pbe_full_size():
    atoms = molecules.water_cluster(32, seed=0)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = MolecularGrid(atoms, GridSettings(radial_points=76,
                                                 angular_points=302)).build(device=0)
    assert len(wts) == 2174400 and basis.nbf == 768
    rng = np.random.default_rng(0)
    C, _ = np.linalg.qr(rng.normal(size=(basis.nbf, 160)))
    D = np.asfortranarray(C @ C.T)
    ob = oracle.OracleBasis.from_flat(basis.flat())
    K, exc, nel = oracle.xc_build(ob, pts, wts, [101, 130], D)
    for precision in ("fp64", "fp32cache", "fp32", "bf16x3"):
        dft = DFT("pbe", basis, pts, wts, precision=precision)
        V = dft.compute_vxc(D)
        de = abs(dft.exchange_correlation_energy() - exc)
        dv = float(np.abs(V - K).max())
        dn = abs(dft.total_density() - nel)
        print(f"(H2O)32 PBE {precision}: |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}",
              flush=True)
        dft.ctx.close()
        if precision == "fp64":
            assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N


@pytest.mark.parametrize("tiled", ["1", "0"])
def test_leaves_wider_than_2048_aos(monkeypatch, tiled):
    monkeypatch.setenv("OCCXC_TILED", tiled)
    atoms = molecules.water_cluster(90, seed=3)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = MolecularGrid(atoms, GridSettings(radial_points=11,
                                                 angular_points=26)).build(device=0)
    D = random_density(basis.nbf, 40, 53)
    dft = DFT("pbe", basis, pts, wts)
    info = dft.ctx.info()
    print(f"(H2O)90 nbf {basis.nbf}, {len(wts)} points, max m_b {info['max_local_nbf']}, "
          f"tiled={tiled}", flush=True)
    assert info["max_local_nbf"] > 2048
    de, dv, dn = _compare(dft, D, pts, wts)
    print(f"  |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}")
    assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "fp32cache"])
def test_blyp_reduced_precision_against_oracle(precision):
    atoms = molecules.water_cluster(4, seed=2)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = MolecularGrid(atoms, GridSettings(radial_points=31,
                                                 angular_points=110)).build(device=0)
    D = random_density(basis.nbf, 20, 61)
    de, dv, dn = _compare(DFT("blyp", basis, pts, wts, precision=precision), D, pts, wts)
    print(f"(H2O)4 BLYP {precision}: |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}")
    assert de <= TOL_E and dv <= 1e-6 and dn <= 1e-5
