"""Serving: XC builds from many concurrent SCFs on one GPU (occxc_server_*).

OCC runs one SCF per process (scf_impl.h:539 calls DFT::compute_vxc each
iteration).  XCServer wraps the library's native serving loop: SCFs are
opened with their basis, grid and functional, admitted by device-memory
footprint, kept resident between builds, and their builds run concurrently
on several HIP streams.  The C++ loop owns the threads; Python only submits
and waits (the ctypes calls release the GIL).
"""
import ctypes

import numpy as np

from . import _lib
from .dft import PRECISIONS, default_options


class XCServer:
    def __init__(self, slots=4, memory_budget=0.0, device=0):
        L = _lib.lib()
        o = _lib.CServerOptions()
        L.occxc_server_default_options(ctypes.byref(o))
        o.device, o.slots, o.memory_budget = device, slots, float(memory_budget)
        h = ctypes.c_void_p()
        _lib.check(L.occxc_server_create(ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._nbf = {}
        self._out = {}

    def open(self, basis, points, weights, funcs, precision="fp64", screen=True):
        """Registers an SCF; returns its id.  Inputs are copied."""
        st, keep = basis.c_struct()
        pts = _lib.f64(points).reshape(-1)
        w = _lib.f64(weights)
        grid = _lib.CGrid(len(w), _lib.dptr(pts), _lib.dptr(w))
        ids = _lib.i32(funcs)
        func = _lib.CFunctional(len(ids), _lib.iptr(ids))
        o = default_options()
        o.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        o.screen = int(bool(screen))
        scf = ctypes.c_longlong()
        _lib.check(_lib.lib().occxc_server_open(self._h, ctypes.byref(st), ctypes.byref(grid),
                                                ctypes.byref(func), ctypes.byref(o),
                                                ctypes.byref(scf)))
        self._nbf[scf.value] = basis.nbf
        return scf.value

    def submit(self, scf, D):
        """Queues one build of SCF scf with density D; returns a ticket."""
        n = self._nbf[scf]
        D = np.asfortranarray(D, dtype=np.float64)
        if D.shape != (n, n):
            raise ValueError(f"D must be {n}x{n}")
        V = np.zeros((n, n), order="F")
        r = np.zeros(2)
        t = ctypes.c_longlong()
        _lib.check(_lib.lib().occxc_server_submit(
            self._h, scf, D.ctypes.data_as(_lib._dp), V.ctypes.data_as(_lib._dp),
            _lib.dptr(r[0:1]), _lib.dptr(r[1:2]), ctypes.byref(t)))
        self._out[t.value] = (V, r)
        return t.value

    def wait(self, ticket):
        """(Vxc, Exc, N_elec, latency in ms) of a submitted build."""
        lat = ctypes.c_double()
        V, r = self._out.pop(ticket)
        _lib.check(_lib.lib().occxc_server_wait(self._h, ticket, ctypes.byref(lat)))
        return V, float(r[0]), float(r[1]), lat.value

    def poll(self, ticket):
        d = ctypes.c_int()
        _lib.check(_lib.lib().occxc_server_poll(self._h, ticket, ctypes.byref(d)))
        return bool(d.value)

    def cancel(self, ticket):
        """Cancels a build that has not started; True if it was cancelled
        (its wait() then raises with status CANCELLED)."""
        c = ctypes.c_int()
        _lib.check(_lib.lib().occxc_server_cancel(self._h, ticket, ctypes.byref(c)))
        return bool(c.value)

    def close_scf(self, scf):
        _lib.check(_lib.lib().occxc_server_close(self._h, scf))

    def stats(self):
        s = _lib.CServerStats()
        _lib.check(_lib.lib().occxc_server_get_stats(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.check(_lib.lib().occxc_server_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
