"""CPU: the seeded workload generators the benchmark configurations use."""
import numpy as np

from occ_amd import AOBasis, molecules


def test_batch_set_is_seeded_and_mixed():
    a = molecules.batch_set(64, seed=3)
    b = molecules.batch_set(64, seed=3)
    assert [x[0] for x in a] == [x[0] for x in b]
    assert all(np.allclose(np.array(x[1]), np.array(y[1])) for x, y in zip(a, b))
    waters = [x for x in a if x[0].startswith("(H2O)")]
    alkanes = [x for x in a if x[0].startswith("C")]
    assert len(waters) == 32 and len(alkanes) == 32
    assert {x[2] for x in a} == {"6-31g", "def2-svp"}
    for name, atoms, _ in waters:
        n = int(name[5:])
        assert 1 <= n <= 16 and len(atoms) == 3 * n
    for name, atoms, _ in alkanes:
        n = int(name[1:name.index("H")])
        assert 1 <= n <= 24 and len(atoms) == 3 * n + 2


def test_alkane_geometry():
    atoms = np.array([a[1:] for a in molecules.alkane(5)]) * molecules.BOHR_TO_ANGSTROM
    zs = [a[0] for a in molecules.alkane(5)]
    c = atoms[[i for i, z in enumerate(zs) if z == 6]]
    h = atoms[[i for i, z in enumerate(zs) if z == 1]]
    assert len(c) == 5 and len(h) == 12
    assert np.allclose(np.linalg.norm(np.diff(c, axis=0), axis=1), 1.54)
    dch = np.linalg.norm(h[:, None, :] - c[None, :, :], axis=2).min(axis=1)
    assert np.allclose(dch, 1.09)


def test_sized_molecule_sizes():
    rng = np.random.default_rng(0)
    for natoms in (3, 10, 50, 300):
        name, atoms = molecules.sized_molecule(natoms, rng)
        assert abs(len(atoms) - natoms) <= 3


def test_water_cluster_basis_sizes():
    atoms = molecules.water_cluster(32, seed=0)
    assert AOBasis.load(atoms, "def2-svp", pure=True).nbf == 768
    assert AOBasis.load(molecules.water(), "6-31g", pure=False).nbf == 13
