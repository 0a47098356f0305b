"""GPU: the native serving loop (occxc_server_*) -- continuous batching of
builds from several SCFs, admission by device-memory footprint, and results
that are the same alone and in the mix (Vxc to 1e-12: the atomics that sum
Vxc have no fixed order; Exc and N_elec to 1e-12).  Every served build is
also checked against the CPU oracle (|dExc| <= 1e-6, max|dVxc| <= 1e-7,
|dN| <= 1e-8)."""
import numpy as np
import pytest

import oracle
from conftest import random_density
from occ_amd import AOBasis, molecules
from occ_amd.dft import DFT, XCContext
from occ_amd.grid import GridSettings, MolecularGrid
from occ_amd.serving import XCServer

pytestmark = pytest.mark.gpu

FUNCS = [101, 130]


def _systems():
    mols = [("water", molecules.water(), "6-31g", False),
            ("dimer", molecules.water_cluster(2, seed=5), "def2-svp", True),
            ("propane", molecules.alkane(3), "6-31g", False)]
    out = []
    for name, atoms, bname, pure in mols:
        basis = AOBasis.load(atoms, bname, pure=pure)
        pts, w = MolecularGrid(atoms, GridSettings(radial_points=31,
                                                   angular_points=110)).build(device=0)
        Ds = [random_density(basis.nbf, 4, 100 * len(out) + k) for k in range(3)]
        out.append((name, basis, pts, w, Ds))
    return out


def _oracle(basis, pts, w, D):
    return oracle.xc_build(oracle.OracleBasis.from_flat(basis.flat()), pts, w, FUNCS, D)


def _matches_oracle(out, ref):
    V, e, n = out[:3]
    K, exc, nel = ref
    return np.abs(V - K).max() <= 1e-7 and abs(e - exc) <= 1e-6 and abs(n - nel) <= 1e-8


def _close(a, b):
    Va, ea, na = a
    Vb, eb, nb = b
    return np.abs(Va - Vb).max() <= 1e-12 and abs(ea - eb) <= 1e-12 and abs(na - nb) <= 1e-12


def test_server_results_same_alone_and_in_mix():
    systems = _systems()
    alone = {}
    for name, basis, pts, w, Ds in systems:
        with XCServer(slots=1) as srv:
            scf = srv.open(basis, pts, w, FUNCS)
            for k, D in enumerate(Ds):
                V, e, n, _ = srv.wait(srv.submit(scf, D))
                alone[name, k] = (V, e, n)
                assert _matches_oracle(alone[name, k], _oracle(basis, pts, w, D)), (name, k)
        # the serving path gives the plain context's result
        dft = DFT("pbe", basis, pts, w)
        V = dft.compute_vxc(Ds[0])
        assert _close(alone[name, 0], (V, dft.exchange_correlation_energy(),
                                       dft.total_density()))
    with XCServer(slots=4) as srv:
        scfs = {name: srv.open(basis, pts, w, FUNCS) for name, basis, pts, w, _ in systems}
        tickets = []
        for k in range(3):           # every SCF's iteration k in flight together
            for name, basis, pts, w, Ds in systems:
                tickets.append((name, k, srv.submit(scfs[name], Ds[k])))
        for name, k, t in tickets:
            V, e, n, lat = srv.wait(t)
            assert lat > 0.0
            assert _close(alone[name, k], (V, e, n)), (name, k)
        st = srv.stats()
        assert st["builds_completed"] == 9 and st["builds_failed"] == 0
        assert st["scfs_admitted"] == 3


def test_server_admits_by_memory_budget():
    """With room for one SCF at a time, every build still completes with the
    right result: later SCFs wait for memory and idle ones are evicted."""
    systems = _systems()
    need = []
    for name, basis, pts, w, Ds in systems:
        ctx = XCContext(basis, pts, w, FUNCS)
        need.append(ctx.info()["device_bytes"])
        ctx.close()
    budget = 1.05 * max(need)
    ref = {}
    for name, basis, pts, w, Ds in systems:
        dft = DFT("pbe", basis, pts, w)
        for k in range(2):
            V = dft.compute_vxc(Ds[k])
            ref[name, k] = (V, dft.exchange_correlation_energy(), dft.total_density())
    sysmap = {name: (basis, pts, w, Ds) for name, basis, pts, w, Ds in systems}
    with XCServer(slots=2, memory_budget=budget) as srv:
        scfs = {name: srv.open(basis, pts, w, FUNCS) for name, basis, pts, w, _ in systems}
        for k in range(2):
            tickets = [(name, srv.submit(scfs[name], Ds[k])) for name, _, _, _, Ds in systems]
            for name, t in tickets:
                out = srv.wait(t)
                assert _close(ref[name, k], out[:3]), (name, k)
                basis, pts, w, Ds = sysmap[name]
                assert _matches_oracle(out, _oracle(basis, pts, w, Ds[k])), (name, k)
            st = srv.stats()
            assert st["resident_bytes"] <= budget
        st = srv.stats()
        assert st["builds_completed"] == 6
        assert st["evictions"] + st["admissions_deferred"] > 0
        for s in scfs.values():
            srv.close_scf(s)


def test_server_rejects_scf_larger_than_budget():
    name, basis, pts, w, Ds = _systems()[1]
    with XCServer(slots=1, memory_budget=1 << 20) as srv:
        scf = srv.open(basis, pts, w, FUNCS)
        t = srv.submit(scf, Ds[0])
        from occ_amd._lib import OccxcError
        with pytest.raises(OccxcError) as e:
            srv.wait(t)
        assert e.value.status == 4


def test_server_cancels_builds_that_have_not_started():
    """occxc_server_cancel completes a queued build at once with
    OCCXC_ERR_CANCELLED; a build already running finishes normally.  Closing
    an SCF cancels its queued builds."""
    from occ_amd._lib import OccxcError
    name, basis, pts, w, Ds = _systems()[1]
    with XCServer(slots=1) as srv:
        scf = srv.open(basis, pts, w, FUNCS)
        tickets = [srv.submit(scf, Ds[k % 3]) for k in range(6)]
        cancelled = {t: srv.cancel(t) for t in tickets[3:]}
        assert any(cancelled.values())
        for k, t in enumerate(tickets):
            if cancelled.get(t):
                with pytest.raises(OccxcError) as e:
                    srv.wait(t)
                assert e.value.status == 6
            else:
                out = srv.wait(t)
                assert _matches_oracle(out, _oracle(basis, pts, w, Ds[k % 3])), k
        st = srv.stats()
        assert st["builds_cancelled"] == sum(cancelled.values())
        assert st["builds_completed"] == 6 - sum(cancelled.values())
        more = [srv.submit(scf, Ds[0]) for _ in range(4)]
        srv.close_scf(scf)
        n_cancelled = 0
        for t in more:
            try:
                out = srv.wait(t)
                assert _matches_oracle(out, _oracle(basis, pts, w, Ds[0]))
            except OccxcError as err:
                assert err.status == 6
                n_cancelled += 1
        assert n_cancelled >= 1
        assert srv.stats()["builds_cancelled"] == sum(cancelled.values()) + n_cancelled
