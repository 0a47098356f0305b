"""GPU parity: every stage of the HIP path against the CPU oracle, through
the C ABI.  Tolerances (stated per result): |dExc| <= 1e-6 Ha,
max|dVxc| <= 1e-7, |dN_elec| <= 1e-8 for the FP64 path."""
import numpy as np
import pytest

import grid_ref
import oracle
from conftest import random_density
from occ_amd import AOBasis, molecules
from occ_amd.dft import DFT, DensityFunctional, XCContext, evaluate_basis, evaluate_decay_cutoff
from occ_amd.grid import GridSettings, MolecularGrid

pytestmark = pytest.mark.gpu

TOL_E, TOL_V, TOL_N = 1e-6, 1e-7, 1e-8


def _grid(atoms, nrad=31, nang=110):
    g = MolecularGrid(atoms, GridSettings(radial_points=nrad, angular_points=nang))
    return g.build(device=0)


def _xc_case(atoms, basis_name, method, D_seed, pure, nrad=31, nang=110, precision="fp64",
             screen=True):
    basis = AOBasis.load(atoms, basis_name, pure=pure)
    pts, wts = _grid(atoms, nrad, nang)
    nocc = max(1, molecules.nuclear_charge(atoms) // 2)
    D = random_density(basis.nbf, min(nocc, basis.nbf), D_seed)
    print(f"case {basis_name} {method} nbf={basis.nbf} npts={len(wts)} {precision}", flush=True)
    dft = DFT(method, basis, pts, wts, precision=precision, screen=screen)
    V = dft.compute_vxc(D)
    ob = oracle.OracleBasis.from_flat(basis.flat())
    K, exc, nel = oracle.xc_build(ob, pts, wts, dft.ids, D, screened=screen)
    return (abs(dft.exchange_correlation_energy() - exc), float(np.abs(V - K).max()),
            abs(dft.total_density() - nel), exc, nel)


def test_evaluate_basis_matches_oracle(rng):
    atoms = molecules.water_cluster(2, seed=3)
    for pure in (True, False):
        basis = AOBasis.load(atoms, "def2-svp", pure=pure)
        pts = rng.normal(size=(500, 3)) * 3.0
        out = evaluate_basis(basis, pts, deriv=1)
        ob = oracle.OracleBasis.from_flat(basis.flat())
        fb = basis.first_bf()
        for s, sh in enumerate(basis.shells):
            ref = oracle.eval_shell(sh.l, sh.pure, sh.exponents, sh.coefficients, sh.origin, pts)
            got = out[:, fb[s]:fb[s] + sh.size(), :]
            np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14)
        assert ob.nbf == basis.nbf


def test_evaluate_basis_l3_l4(rng):
    atoms = molecules.water()
    basis = AOBasis.load(atoms, "def2-tzvp", pure=True)  # O has f
    pts = rng.normal(size=(300, 3)) * 2.0
    out = evaluate_basis(basis, pts, deriv=1)
    fb = basis.first_bf()
    for s, sh in enumerate(basis.shells):
        ref = oracle.eval_shell(sh.l, sh.pure, sh.exponents, sh.coefficients, sh.origin, pts)
        np.testing.assert_allclose(out[:, fb[s]:fb[s] + sh.size(), :], ref, rtol=1e-12,
                                   atol=1e-14)


def test_decay_cutoff_matches_oracle():
    atoms = molecules.water_cluster(2, seed=5)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    ob = oracle.OracleBasis.from_flat(basis.flat())
    np.testing.assert_array_equal(evaluate_decay_cutoff(basis, 1e-10),
                                  oracle.decay_cutoff(ob, 1e-10))


@pytest.mark.parametrize("name,fid", [("lda_x", 1), ("lda_c_vwn", 7), ("lda_c_pw_mod", 13),
                                      ("gga_x_pbe", 101), ("gga_x_b88", 106),
                                      ("gga_c_pbe", 130), ("gga_c_lyp", 131)])
def test_functional_matches_oracle(rng, name, fid):
    rho = 10 ** rng.uniform(-12, 3, 4000)
    sigma = 10 ** rng.uniform(-14, 2, 4000) * rho ** (8 / 3)
    exc, vr, vs = DensityFunctional(name).evaluate(rho, sigma)
    e0, v0, s0 = oracle.functional_eval(fid, rho, sigma)
    np.testing.assert_allclose(exc, e0, rtol=1e-11, atol=1e-200)
    np.testing.assert_allclose(vr, v0, rtol=1e-10, atol=1e-200)
    if fid > 100:
        np.testing.assert_allclose(vs, s0, rtol=1e-10, atol=1e-200)


def test_becke_partition_matches_numpy():
    # compared as partition fractions P_own / sum_k P_k: near cell boundaries
    # the cell function 0.5 (1 - v) cancels, so weights there agree in
    # absolute fraction, not in relative value
    atoms = molecules.water_cluster(2, seed=7)
    g = MolecularGrid(atoms, GridSettings(radial_points=21, angular_points=50))
    pts, w_gpu = g.build(device=0)
    _, w_cpu = grid_ref.molecular_grid(g)
    _, w0, _ = g.unpartitioned()
    assert np.all(np.abs(w_gpu - w_cpu) <= 1e-10 * w0 + 1e-300)
    assert np.abs(w_gpu - w_cpu).max() / w0.max() < 1e-14


@pytest.mark.parametrize("method", ["svwn5", "pbe", "blyp"])
def test_xc_water_631g(method):
    de, dv, dn, exc, nel = _xc_case(molecules.water(), "6-31g", method, 11, pure=False)
    print(f"water/6-31G {method}: |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}")
    assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N


@pytest.mark.parametrize("method", ["pbe", "blyp", "svwn5"])
def test_xc_water_dimer_def2svp(method):
    de, dv, dn, exc, nel = _xc_case(molecules.water_cluster(2, seed=1), "def2-svp", method, 12,
                                    pure=True)
    print(f"(H2O)2/def2-SVP {method}: |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}")
    assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N


def test_xc_dense_path():
    de, dv, dn, exc, nel = _xc_case(molecules.water(), "6-31g", "pbe", 13, pure=False,
                                    screen=False)
    assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N


def test_xc_alkane_cartesian_d():
    de, dv, dn, exc, nel = _xc_case(molecules.alkane(3), "def2-svp", "pbe", 14, pure=False,
                                    nrad=21, nang=50)
    assert de <= TOL_E and dv <= TOL_V and dn <= TOL_N


def test_mfma_matches_reference_kernels(monkeypatch):
    """The MFMA contractions against the FP64 reference kernels of kernels.hip."""
    atoms = molecules.water_cluster(2, seed=4)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = _grid(atoms)
    D = random_density(basis.nbf, 10, 21)
    out = {}
    for ref in (False, True):
        if ref:
            monkeypatch.setenv("OCCXC_REFERENCE_KERNELS", "1")
        dft = DFT("blyp", basis, pts, wts)
        out[ref] = (dft.compute_vxc(D), dft.exchange_correlation_energy(), dft.total_density())
    assert abs(out[True][1] - out[False][1]) <= 1e-10
    assert np.abs(out[True][0] - out[False][0]).max() <= 1e-11
    assert abs(out[True][2] - out[False][2]) <= 1e-11


@pytest.mark.parametrize("method", ["pbe", "blyp"])
def test_xc_fp32_cache_accuracy(method):
    """FP32 AO cache with the FP64 density matrix and FP64 arithmetic: the
    AO rounding errors are independent per (point, AO) and largely average
    out of Exc and Vxc.  N_elec is dominated by the few points that carry each
    tight core function, so its error (~2e-8 here) is reported, not gated."""
    de, dv, dn, exc, nel = _xc_case(molecules.water_cluster(2, seed=1), "def2-svp", method, 12,
                                    pure=True, precision="fp32cache")
    print(f"fp32cache {method}: |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}")
    assert de <= TOL_E and dv <= TOL_V and dn <= 1e-6


@pytest.mark.parametrize("variant", ["0", "1", "2", "reference"])
def test_fp32_cache_kernels_agree(monkeypatch, variant):
    """Both MFMA density kernels and the FP64 reference kernels read the same
    FP32 AO cache and agree to FP64 rounding."""
    atoms = molecules.water_cluster(2, seed=4)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = _grid(atoms)
    D = random_density(basis.nbf, 10, 22)
    base = DFT("pbe", basis, pts, wts, precision="fp32cache")
    V0 = base.compute_vxc(D)
    if variant == "reference":
        monkeypatch.setenv("OCCXC_REFERENCE_KERNELS", "1")
    else:
        monkeypatch.setenv("OCCXC_DENSITY_VARIANT", variant)
    dft = DFT("pbe", basis, pts, wts, precision="fp32cache")
    V = dft.compute_vxc(D)
    assert abs(dft.exchange_correlation_energy() - base.exchange_correlation_energy()) <= 1e-10
    assert abs(dft.total_density() - base.total_density()) <= 1e-11
    assert np.abs(V - V0).max() <= 1e-11


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_xc_mixed_precision_reported(precision):
    de, dv, dn, exc, nel = _xc_case(molecules.water_cluster(2, seed=1), "def2-svp", "pbe", 12,
                                    pure=True, precision=precision)
    print(f"{precision}: |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}")
    assert de <= TOL_E


@pytest.mark.parametrize("precision", ["fp32", "fp32cache"])
def test_fp32_storage_n_elec_corrected(precision):
    """FP32 AO storage: N_elec carries the first-order correction sum D2 dS
    (dS = sum_p w_p (r Phi^T + Phi r^T) from the rounding residuals r), so it
    matches the FP64 oracle to well inside 1e-8, while Exc and Vxc keep the
    FP32 storage error (within their stated tolerances).  (The split-BF16
    mode's N_elec error comes from its density contraction, not the storage.)"""
    de, dv, dn, exc, nel = _xc_case(molecules.water_cluster(3, seed=2), "def2-svp", "blyp", 17,
                                    pure=True, precision=precision)
    print(f"{precision}: |dExc|={de:.2e} max|dVxc|={dv:.2e} |dN|={dn:.2e}")
    assert dn <= 1e-10
    assert de <= TOL_E
    assert dv <= TOL_V


def test_cpp_driver_drop_in(tmp_path):
    """A C++ caller through include/occxc_dft.hpp gives the Python result and
    repeats it on the second build of a resident context."""
    import subprocess
    atoms = molecules.water()
    basis = AOBasis.load(atoms, "6-31g", pure=False)
    pts, wts = _grid(atoms)
    D = random_density(basis.nbf, 5, 31)
    f = basis.flat()
    ids = np.array([101, 130], dtype=np.int32)
    np.array([len(f["l"]), len(f["exponents"]), len(wts), basis.nbf, len(ids)],
             dtype=np.int64).tofile(tmp_path / "header.bin")
    for k in ("l", "pure", "nprim"):
        f[k].astype(np.int32).tofile(tmp_path / f"{k}.bin")
    for k in ("origin", "exponents", "coefficients"):
        f[k].astype(np.float64).tofile(tmp_path / f"{k}.bin")
    ids.tofile(tmp_path / "ids.bin")
    np.ascontiguousarray(pts).tofile(tmp_path / "points.bin")
    wts.tofile(tmp_path / "weights.bin")
    np.asfortranarray(D).ravel(order="F").tofile(tmp_path / "density.bin")
    exe = __import__("pathlib").Path(__file__).parent / "native" / "xc_driver"
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = [l.split() for l in out.stdout.strip().splitlines()]
    # Exc and N_elec repeat bit for bit (fixed-order leaf reduction); Vxc is
    # summed with FP64 atomics, whose order varies between builds
    assert lines[0][2:6] == lines[1][2:6]
    assert float(lines[0][7]) == pytest.approx(float(lines[1][7]), rel=1e-13)
    dft = DFT("pbe", basis, pts, wts)
    V = dft.compute_vxc(D)
    assert float(lines[0][3]) == pytest.approx(dft.exchange_correlation_energy(), abs=1e-12)
    assert float(lines[0][5]) == pytest.approx(dft.total_density(), abs=1e-12)
    cs = float(np.sum(V.ravel(order="F") * ((np.arange(V.size) % 7) + 1)))
    assert float(lines[0][7]) == pytest.approx(cs, rel=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_matches_single_and_oracle(precision):
    """Molecules built together in one batch context give their solo results
    (same kernels, one launch) and the oracle's within the stated tolerances
    (in FP32 mode through the A pass and the folded scatter, per molecule).
    FP32 mode's Vxc carries FP32 rounding of A and of the stored AO values,
    relative to the entry: on this coarse grid (21 x 50) a few points near the
    nucleus carry the core entries, so V[0, 0] of (H2O)2 (-1.94) is off by
    1.2e-7 (6e-8 relative).  Its bound is 1e-7 times the largest |V| (at
    least 1); Exc and N_elec keep the absolute tolerances."""
    from occ_amd.dft import XCBatch
    mols = [molecules.water(), molecules.alkane(2), molecules.water_cluster(2, seed=9)]
    bases = [AOBasis.load(m, "6-31g", pure=False) for m in mols]
    grids = [_grid(m, 21, 50) for m in mols]
    Ds = [random_density(b.nbf, 4, 40 + i) for i, b in enumerate(bases)]
    batch = XCBatch(bases, grids, [101, 130], precision=precision)
    Vs, excs, nels = batch.build(Ds)
    for i, (b, (pts, w)) in enumerate(zip(bases, grids)):
        dft = DFT("pbe", b, pts, w, precision=precision)
        V = dft.compute_vxc(Ds[i])
        assert abs(excs[i] - dft.exchange_correlation_energy()) <= 1e-11
        assert abs(nels[i] - dft.total_density()) <= 1e-11
        assert np.abs(Vs[i] - V).max() <= 1e-11
        K, exc, nel = oracle.xc_build(oracle.OracleBasis.from_flat(b.flat()), pts, w,
                                      [101, 130], Ds[i])
        assert abs(excs[i] - exc) <= TOL_E and abs(nels[i] - nel) <= TOL_N
        vscale = 1.0 if precision == "fp64" else max(1.0, float(np.abs(K).max()))
        assert np.abs(Vs[i] - K).max() <= TOL_V * vscale


def test_density_variants_agree(monkeypatch):
    atoms = molecules.water_cluster(2, seed=6)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = _grid(atoms)
    D = random_density(basis.nbf, 10, 51)
    # the per-leaf kernels (OCCXC_TILED=0) in each density variant, then the
    # tiled kernels (the default)
    variants = ("0", "1", "2", "4", "tiled")
    out = []
    for v in variants:
        if v == "tiled":
            monkeypatch.delenv("OCCXC_TILED", raising=False)
        else:
            monkeypatch.setenv("OCCXC_TILED", "0")
            monkeypatch.setenv("OCCXC_DENSITY_VARIANT", v)
        dft = DFT("pbe", basis, pts, wts)
        out.append((dft.compute_vxc(D), dft.exchange_correlation_energy(), dft.total_density()))
    for k in range(1, len(variants)):
        assert abs(out[0][1] - out[k][1]) <= 1e-11 and abs(out[0][2] - out[k][2]) <= 1e-11
        assert np.abs(out[0][0] - out[k][0]).max() <= 1e-12


def test_density_kernels_split_by_leaf_width(monkeypatch):
    """Leaves wider than the per-wavefront kernels' LDS AO map (2048 AOs) go
    to the LDS-shared kernel while the rest stay on the per-wavefront one:
    same results as the LDS-shared kernel throughout.  (H2O)90 / def2-SVP
    (2160 AOs) on a coarse grid has leaves of every width."""
    atoms = molecules.water_cluster(90, seed=3)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = _grid(atoms, 11, 26)
    D = random_density(basis.nbf, 40, 53)
    out = []
    monkeypatch.setenv("OCCXC_TILED", "0")
    for v in ("2", "1"):
        monkeypatch.setenv("OCCXC_DENSITY_VARIANT", v)
        dft = DFT("pbe", basis, pts, wts)
        out.append((dft.compute_vxc(D), dft.exchange_correlation_energy(), dft.total_density()))
        print(f"density variant {v}: nbf {basis.nbf}, {len(wts)} points, "
              f"mean m_b {dft.ctx.info()['mean_local_nbf']:.0f}", flush=True)
    assert abs(out[0][1] - out[1][1]) <= 1e-10 and abs(out[0][2] - out[1][2]) <= 1e-10
    assert np.abs(out[0][0] - out[1][0]).max() <= 1e-12


def test_device_setup_matches_host_setup(monkeypatch):
    """The leaf layout built on the device (Morton sort, leaf spheres,
    screening, leaf tables) is the host restatement's, bit for bit: same
    leaves and work, identical builds.  Includes grids of <= 256 points (one
    leaf) and a batch of molecules."""
    from occ_amd.dft import XCBatch, XCContext, functional_ids

    def both(make):
        out = []
        for host in (True, False):
            if host:
                monkeypatch.setenv("OCCXC_HOST_SETUP", "1")
            else:
                monkeypatch.delenv("OCCXC_HOST_SETUP", raising=False)
            out.append(make())
        return out

    ids = functional_ids("pbe")
    cases = [(molecules.water_cluster(4, seed=2), 31, 110), (molecules.water(), 3, 26),
             (molecules.water(), 11, 26)]
    for atoms, nrad, nang in cases:
        basis = AOBasis.load(atoms, "def2-svp", pure=True)
        pts, wts = _grid(atoms, nrad, nang)
        D = random_density(basis.nbf, 5, 70)
        print(f"set-up case: {len(atoms)} atoms, {len(wts)} points", flush=True)

        def make():
            ctx = XCContext(basis, pts, wts, ids)
            V, e, n = ctx.build(D)
            return ctx.info(), V, e, n
        (i0, V0, e0, n0), (i1, V1, e1, n1) = both(make)
        for k in ("nleaves", "mean_local_nbf", "contraction_flops", "ao_bytes", "ao_write_bytes"):
            assert i0[k] == i1[k], k
        assert e0 == e1 and n0 == n1 and np.abs(V0 - V1).max() <= 1e-14
    mols = [molecules.water_cluster(k, seed=k) for k in (1, 2, 3)]
    bases = [AOBasis.load(a, "6-31g", pure=False) for a in mols]
    grids = [_grid(a, 21, 50) for a in mols]
    Ds = [random_density(b.nbf, 5, 71 + k) for k, b in enumerate(bases)]

    def make_batch():
        bt = XCBatch(bases, grids, ids)
        return bt.info(), bt.build(Ds)
    (i0, r0), (i1, r1) = both(make_batch)
    assert i0["nleaves"] == i1["nleaves"]
    assert np.array_equal(r0[1], r1[1]) and np.array_equal(r0[2], r1[2])
    for a, b in zip(r0[0], r1[0]):
        assert np.abs(a - b).max() <= 1e-14


def test_graph_replay_matches_direct_builds():
    """Builds replayed from the captured graph give the direct builds' results,
    on every replay, with eager work and synchronisation between replays, and
    re-capture when the output pointers change."""
    from occ_amd.dft import XCContext
    from occ_amd.device import DeviceArray, synchronize
    atoms = molecules.water_cluster(2, seed=3)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = _grid(atoms)
    n = basis.nbf
    Ds = [random_density(n, 10, 60 + k) for k in range(3)]
    ctx = XCContext(basis, pts, wts, [101, 130])
    dD, dV, dr = DeviceArray(n * n), DeviceArray(n * n), DeviceArray(2)
    dV2, dr2 = DeviceArray(n * n), DeviceArray(2)
    direct = []
    for D in Ds:
        dD.upload(D)
        ctx.build_device(dD.ptr, dV.ptr, dr.ptr)
        ctx.synchronize()
        direct.append((dV.download((n, n)), dr.download()))
    ctx.set_graph(True)
    for rep in range(2):
        for k, D in enumerate(Ds):
            dD.upload(D)
            outV, outr = (dV, dr) if rep == 0 else (dV2, dr2)
            ctx.build_device(dD.ptr, outV.ptr, outr.ptr)
            ctx.synchronize()
            synchronize()
            V, r = outV.download((n, n)), outr.download()
            assert np.abs(V - direct[k][0]).max() <= 1e-13
            assert abs(r[0] - direct[k][1][0]) <= 1e-12 and abs(r[1] - direct[k][1][1]) <= 1e-12
    ctx.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_vxc_variants_agree(monkeypatch, precision):
    """The symmetric tile-pair Vxc kernel (upper tiles only), the general one
    and the two-pass form (stored blocks + tile-pair reduction) agree, and the
    FP64 result matches the oracle."""
    atoms = molecules.water_cluster(3, seed=8)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = _grid(atoms)
    D = random_density(basis.nbf, 12, 71)
    out = []
    for v in ("0", "1", "2", "tiled", "ws"):
        # per-leaf kernels (OCCXC_TILED=0) in each Vxc variant, then the tiled
        # kernels one item per workgroup and in the persistent forms
        monkeypatch.setenv("OCCXC_TILED", "1" if v in ("tiled", "ws") else "0")
        monkeypatch.setenv("OCCXC_VXC_VARIANT", v if v in ("0", "1", "2") else "0")
        dft = DFT("pbe", basis, pts, wts, precision=precision)
        if v in ("tiled", "ws"):
            dft.ctx.close()
            dft.ctx = XCContext(basis, pts, wts, dft.ids, precision=precision, tile_ao=128,
                                persistent=3 if v == "ws" else 0)
        out.append((dft.compute_vxc(D), dft.exchange_correlation_energy(), dft.total_density()))
        if v == "2":
            # the two-pass form sums in a fixed order: builds repeat bit for bit
            assert np.array_equal(dft.compute_vxc(D), out[-1][0])
    tol = 1e-12 if precision == "fp64" else 2e-6
    for k in range(1, len(out)):
        assert np.abs(out[0][0] - out[k][0]).max() <= tol
    assert np.abs(out[2][0] - out[2][0].T).max() == 0.0
    assert np.abs(out[0][0] - out[0][0].T).max() == 0.0
    if precision == "fp64":
        K, exc, nel = oracle.xc_build(oracle.OracleBasis.from_flat(basis.flat()), pts, wts,
                                      [101, 130], D)
        assert np.abs(out[0][0] - K).max() <= TOL_V
        assert abs(out[0][1] - exc) <= TOL_E and abs(out[0][2] - nel) <= TOL_N


def test_two_pass_vxc_batch_and_large_leaves(monkeypatch):
    """Two-pass Vxc on a batch context (several molecules, one reduction
    launch) and on a system whose leaves exceed the symmetric kernel's size
    (general-kernel blocks in the same reduction)."""
    from occ_amd.dft import XCBatch
    monkeypatch.setenv("OCCXC_VXC_VARIANT", "2")
    mols = [molecules.water(), molecules.water_cluster(10, seed=2)]
    bases = [AOBasis.load(m, "def2-svp", pure=True) for m in mols]
    grids = [_grid(m, 21, 50) for m in mols]
    Ds = [random_density(b.nbf, 6, 80 + i) for i, b in enumerate(bases)]
    batch = XCBatch(bases, grids, [101, 130])
    Vs, excs, nels = batch.build(Ds)
    for i, (b, (pts, w)) in enumerate(zip(bases, grids)):
        K, exc, nel = oracle.xc_build(oracle.OracleBasis.from_flat(b.flat()), pts, w,
                                      [101, 130], Ds[i])
        assert abs(excs[i] - exc) <= TOL_E and abs(nels[i] - nel) <= TOL_N
        assert np.abs(Vs[i] - K).max() <= TOL_V


def test_batch_estimate_bytes_tracks_context_footprint():
    """Per-molecule footprints from occxc_batch_estimate_bytes (no large
    allocations) against the device bytes the contexts really hold."""
    from occ_amd.dft import XCBatch, XCContext, estimate_batch_bytes
    mols = [molecules.water(), molecules.alkane(4), molecules.water_cluster(3, seed=5)]
    bases = [AOBasis.load(m, "def2-svp", pure=True) for m in mols]
    grids = [_grid(m, 21, 50) for m in mols]
    est = estimate_batch_bytes(bases, grids, [101, 130])
    assert est.shape == (3,) and np.all(est > 0)
    for b, (pts, w), e in zip(bases, grids, est):
        ctx = XCContext(b, pts, w, [101, 130])
        real = ctx.info()["device_bytes"]
        ctx.close()
        assert 0.5 * real <= e <= 1.5 * real, (e, real)
    batch = XCBatch(bases, grids, [101, 130])
    real = batch.info()["device_bytes"]
    batch.close()
    assert 0.5 * real <= est.sum() <= 1.5 * real


def test_xc_on_the_occ_default_grid():
    """A build on OCC's default molecular grid (LMG radial rule, NWChem
    pruning, reduced H grid, Becke partition on the GPU) against the oracle."""
    from occ_amd.grid import OccGridSettings
    atoms = molecules.water_cluster(2, seed=1)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, wts = MolecularGrid(atoms, OccGridSettings(), basis=basis).build(device=0)
    D = random_density(basis.nbf, 10, 31)
    print(f"occ default grid: {len(wts)} points", flush=True)
    dft = DFT("pbe", basis, pts, wts)
    V = dft.compute_vxc(D)
    K, exc, nel = oracle.xc_build(oracle.OracleBasis.from_flat(basis.flat()), pts, wts,
                                  dft.ids, D)
    assert abs(dft.exchange_correlation_energy() - exc) <= TOL_E
    assert np.abs(V - K).max() <= TOL_V and abs(dft.total_density() - nel) <= TOL_N
