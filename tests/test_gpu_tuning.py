"""The launch settings the autotuner chooses between (largest AO block per
work item, persistent density / Vxc launches) change only the launch shape:
every setting gives the oracle's result (|dExc| <= 1e-6, max|dVxc| <= 1e-7,
|dN| <= 1e-8) and the settings agree with each other to FP64 rounding, for
one molecule and for a batch context; the committed tuning table's choice is
one of them."""
import numpy as np
import pytest

import oracle
from conftest import random_density
from occ_amd import AOBasis, molecules
from occ_amd.dft import XCBatch, XCContext, functional_ids
from occ_amd.grid import GridSettings, MolecularGrid

pytestmark = pytest.mark.gpu

IDS = [101, 130]
SETTINGS = [(0, -1)] + [(t, p) for t in (32, 64, 96, 128) for p in (0, 1, 2, 3)]


def _grid(atoms, nrad=31, nang=110):
    return MolecularGrid(atoms, GridSettings(radial_points=nrad, angular_points=nang)).build(device=0)


def test_settings_agree_on_one_molecule():
    atoms = molecules.water_cluster(6, seed=9)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = _grid(atoms)
    D = random_density(basis.nbf, 30, 91)
    K, exc, nel = oracle.xc_build(oracle.OracleBasis.from_flat(basis.flat()), pts, wts, IDS, D)
    out = {}
    for tile, pers in SETTINGS:
        ctx = XCContext(basis, pts, wts, IDS, tile_ao=tile, persistent=pers)
        V, e, n = ctx.build(D)
        ctx.close()
        out[tile, pers] = (V, e, n)
        assert abs(e - exc) <= 1e-6 and abs(n - nel) <= 1e-8 and np.abs(V - K).max() <= 1e-7
    V0, e0, n0 = out[0, -1]
    for key, (V, e, n) in out.items():
        assert np.abs(V - V0).max() <= 1e-12 and abs(e - e0) <= 1e-12 and abs(n - n0) <= 1e-12, key


def test_settings_agree_on_a_batch():
    mols = [molecules.water(), molecules.alkane(5), molecules.water_cluster(3, seed=4)]
    bases = [AOBasis.load(m, "def2-svp", pure=True) for m in mols]
    grids = [_grid(m, 21, 50) for m in mols]
    Ds = [random_density(b.nbf, 8, 120 + i) for i, b in enumerate(bases)]
    refs = [oracle.xc_build(oracle.OracleBasis.from_flat(b.flat()), p, w, IDS, D)
            for b, (p, w), D in zip(bases, grids, Ds)]
    first = None
    for tile, pers in SETTINGS:
        batch = XCBatch(bases, grids, IDS, tile_ao=tile, persistent=pers)
        Vs, excs, nels = batch.build(Ds)
        batch.close()
        for i, (K, exc, nel) in enumerate(refs):
            assert abs(excs[i] - exc) <= 1e-6 and abs(nels[i] - nel) <= 1e-8
            assert np.abs(Vs[i] - K).max() <= 1e-7
        if first is None:
            first = (Vs, excs, nels)
        for i in range(len(mols)):
            assert np.abs(Vs[i] - first[0][i]).max() <= 1e-12, (tile, pers, i)
            assert abs(excs[i] - first[1][i]) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32cache", "fp32"])
def test_density_staging_forms_agree(monkeypatch, precision):
    """The density items staged by direct-to-LDS loads (the default) and by
    registers (OCCXC_DENSITY_DMA=0) run the same MFMA k order, so they agree
    to FP64 rounding (the Vxc atomics' order); both meet the oracle's
    tolerances."""
    atoms = molecules.water_cluster(6, seed=9)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = _grid(atoms)
    D = random_density(basis.nbf, 30, 95)
    K, exc, nel = oracle.xc_build(oracle.OracleBasis.from_flat(basis.flat()), pts, wts, IDS, D)
    out = []
    for dma in ("1", "0"):
        monkeypatch.setenv("OCCXC_DENSITY_DMA", dma)
        for tile in (64, 128):
            ctx = XCContext(basis, pts, wts, IDS, precision=precision, tile_ao=tile)
            V, e, n = ctx.build(D)
            ctx.close()
            assert abs(e - exc) <= 1e-6 and abs(n - nel) <= 1e-8 and np.abs(V - K).max() <= 1e-7
            out.append((dma, tile, V, e, n))
    for dma, tile, V, e, n in out[2:]:
        _, _, V0, e0, n0 = out[0 if tile == 64 else 1]
        assert np.abs(V - V0).max() <= 1e-12 and abs(e - e0) <= 1e-12 and abs(n - n0) <= 1e-12, tile


@pytest.mark.parametrize("deriv_ids", ["gga", "lda"])
def test_fp32_vxc_forms_agree(monkeypatch, deriv_ids):
    """FP32 mode's Vxc stage in its two forms: A = sum_q c_q Phi_q formed once
    per leaf (k_vxc_apass) and contracted from directly loaded LDS with the
    diagonal blocks folded before the scatter (the default), and the
    one-kernel form that forms A per block pair (OCCXC_VXC_APASS=0).  Both
    round A to FP32 the same way and sum 32 points at a time in FP32 but in a
    different order, so they differ by FP32 rounding of the block sums: of
    the size of the mode's own Vxc error against FP64 (4.3e-8 on (H2O)128;
    1.8e-8 measured between the forms here), bounded at 5e-8.  Both meet the
    oracle's tolerances, and Exc and N_elec, which the Vxc stage does not
    touch, agree to rounding."""
    ids = IDS if deriv_ids == "gga" else functional_ids("svwn5")
    atoms = molecules.water_cluster(6, seed=4)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = _grid(atoms)
    D = random_density(basis.nbf, 30, 41)
    K, exc, nel = oracle.xc_build(oracle.OracleBasis.from_flat(basis.flat()), pts, wts, ids, D)
    out = []
    for apass in ("1", "0"):
        monkeypatch.setenv("OCCXC_VXC_APASS", apass)
        for tile in (32, 64, 128):
            ctx = XCContext(basis, pts, wts, ids, precision="fp32", tile_ao=tile)
            V, e, n = ctx.build(D)
            ctx.close()
            assert abs(e - exc) <= 1e-6 and abs(n - nel) <= 1e-8 and np.abs(V - K).max() <= 1e-7
            assert np.abs(V - V.T).max() == 0.0
            out.append((apass, tile, V, e, n))
    for apass, tile, V, e, n in out[3:]:
        _, _, V0, e0, n0 = out[(32, 64, 128).index(tile)]
        assert np.abs(V - V0).max() <= 5e-8, tile
        assert abs(e - e0) <= 1e-12 and abs(n - n0) <= 1e-12, tile


def test_fp32_a_plane_is_accounted(monkeypatch):
    """The FP32 mode's A plane (one FP32 value per leaf point and AO, beside
    the four AO planes of a GGA) is counted in the context's device bytes:
    with it, device_bytes grows by exactly ao_bytes / 4 over the one-kernel
    form (OCCXC_VXC_APASS=0); FP64 and FP32-cache contexts hold no A plane."""
    atoms = molecules.water_cluster(3, seed=2)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = _grid(atoms)
    info = {}
    for precision in ("fp32", "fp32cache", "fp64"):
        for apass in ("1", "0"):
            monkeypatch.setenv("OCCXC_VXC_APASS", apass)
            ctx = XCContext(basis, pts, wts, IDS, precision=precision)
            info[precision, apass] = ctx.info()
            ctx.close()
    with_a, without = info["fp32", "1"], info["fp32", "0"]
    assert with_a["device_bytes"] - without["device_bytes"] == with_a["ao_bytes"] / 4
    for precision in ("fp32cache", "fp64"):
        assert info[precision, "1"]["device_bytes"] == info[precision, "0"]["device_bytes"]
