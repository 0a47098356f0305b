"""The allocation-free Becke partition (occxc_becke_partition_noalloc:
caller-owned device buffers, its inter-atomic distance table in the
library's fixed device scratch array) against occxc_becke_partition, through
the C ABI.  Both run the same partition kernel on the same distance table
(computed in the same operation order), so the weights must be identical.
Two calls in a row, the second on a smaller molecule and grid, show that the
scratch array's previous contents do not leak into a result."""
import ctypes

import numpy as np
import pytest

from occ_amd import _lib, molecules
from occ_amd.device import DeviceArray, hip, synchronize
from occ_amd.grid import GridSettings, MolecularGrid

pytestmark = pytest.mark.gpu


def _inputs(atoms, nrad, nang):
    g = MolecularGrid(atoms, GridSettings(radial_points=nrad, angular_points=nang))
    pts, w, own = g.unpartitioned()
    xyz, radii = g.atom_arrays()
    return (np.ascontiguousarray(xyz), np.ascontiguousarray(radii), np.ascontiguousarray(pts),
            np.ascontiguousarray(w), np.ascontiguousarray(own, dtype=np.int32))


def _host(xyz, radii, pts, w, own, ta):
    w = w.copy()
    _lib.check(_lib.lib().occxc_becke_partition(
        len(radii), _lib.dptr(xyz), _lib.dptr(radii), ta, len(w), _lib.dptr(pts),
        _lib.iptr(own), _lib.dptr(w), 0))
    return w


def _device_int32(a):
    buf = DeviceArray((a.size + 1) // 2)
    err = hip().hipMemcpy(ctypes.c_void_p(buf.ptr), a.ctypes.data_as(ctypes.c_void_p),
                          4 * a.size, 1)
    assert err == 0
    return buf


def _noalloc(xyz, radii, pts, w, own, ta):
    d_xyz, d_r, d_p, d_w = (DeviceArray(xyz.size), DeviceArray(radii.size), DeviceArray(pts.size),
                            DeviceArray(w.size))
    d_xyz.upload(xyz.ravel())
    d_r.upload(radii)
    d_p.upload(pts.ravel())
    d_w.upload(w)
    d_own = _device_int32(own)
    _lib.check(_lib.lib().occxc_becke_partition_noalloc(
        len(radii), d_xyz.ptr, d_r.ptr, ta, len(w), d_p.ptr, d_own.ptr, d_w.ptr, None))
    synchronize()
    out = d_w.download()
    for b in (d_xyz, d_r, d_p, d_w, d_own):
        b.free()
    return out


def test_noalloc_partition_matches_allocating_entry_point():
    cases = [(molecules.water_cluster(8, seed=4), 21, 50), (molecules.water(), 11, 26)]
    for atoms, nrad, nang in cases:   # the second call is smaller
        args = _inputs(atoms, nrad, nang)
        for ta in (1, 0):
            print(f"noalloc Becke: {len(atoms)} atoms, {len(args[3])} points, ta={ta}", flush=True)
            w0 = _host(*args, ta)
            w1 = _noalloc(*args, ta)
            assert np.all(np.isfinite(w1))
            assert np.array_equal(w0, w1)


def test_noalloc_partition_rejects_more_atoms_than_its_table():
    d = DeviceArray(8)
    st = _lib.lib().occxc_becke_partition_noalloc(1025, d.ptr, d.ptr, 1, 1, d.ptr, d.ptr, d.ptr,
                                                   None)
    assert st == 1   # OCCXC_ERR_INVALID_ARGUMENT
    assert b"1024" in _lib.lib().occxc_last_error_message()
    d.free()
