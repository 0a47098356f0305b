"""ctypes binding of liboccxc.so (the C ABI declared in include/occxc.h).

The library is the only compute path: if it is not built, importing the
parts of occ_amd that need it raises immediately.
"""
import ctypes
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboccxc.so"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_llp = ctypes.POINTER(ctypes.c_longlong)

OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "UNSUPPORTED", 3: "DEVICE",
                4: "OUT_OF_MEMORY", 5: "ADMISSION", 6: "CANCELLED", 7: "INTERNAL"}


class OccxcError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class CBasis(ctypes.Structure):
    _fields_ = [("nshell", ctypes.c_int), ("l", _ip), ("pure", _ip), ("nprim", _ip),
                ("origin", _dp), ("exponents", _dp), ("coefficients", _dp)]


class CGrid(ctypes.Structure):
    _fields_ = [("npts", ctypes.c_longlong), ("points", _dp), ("weights", _dp)]


class CFunctional(ctypes.Structure):
    _fields_ = [("nfunc", ctypes.c_int), ("ids", _ip)]


class COptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("density_threshold", ctypes.c_double),
                ("screening_threshold", ctypes.c_double), ("screen", ctypes.c_int),
                ("device", ctypes.c_int), ("tile_ao", ctypes.c_int), ("persistent", ctypes.c_int)]


class CContextInfo(ctypes.Structure):
    _fields_ = [("npts", ctypes.c_longlong), ("nleaves", ctypes.c_longlong),
                ("nbf", ctypes.c_int), ("nshell", ctypes.c_int),
                ("mean_local_nbf", ctypes.c_double), ("contraction_flops", ctypes.c_double),
                ("ao_bytes", ctypes.c_double), ("ao_write_bytes", ctypes.c_double),
                ("device_bytes", ctypes.c_double), ("max_local_nbf", ctypes.c_int),
                ("tile_ao", ctypes.c_int), ("persistent", ctypes.c_int),
                ("tuning_row", ctypes.c_int)]


class CServerOptions(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("slots", ctypes.c_int),
                ("memory_budget", ctypes.c_double)]


class CServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_longlong) for k in (
        "scfs_opened", "scfs_admitted", "scfs_closed", "scfs_resident", "scfs_waiting",
        "admissions_deferred", "evictions", "builds_submitted", "builds_completed",
        "builds_failed")] + [("resident_bytes", ctypes.c_double),
                             ("budget_bytes", ctypes.c_double),
                             ("builds_cancelled", ctypes.c_longlong)]


# every symbol declared in include/occxc.h: name -> (restype, argtypes)
SIGNATURES = {
    "occxc_default_options": (None, [ctypes.POINTER(COptions)]),
    "occxc_context_create": (ctypes.c_int, [ctypes.POINTER(CBasis), ctypes.POINTER(CGrid),
                                            ctypes.POINTER(CFunctional),
                                            ctypes.POINTER(COptions),
                                            ctypes.POINTER(ctypes.c_void_p)]),
    "occxc_context_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "occxc_context_get_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CContextInfo)]),
    "occxc_build": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp]),
    "occxc_xc_gradient": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int, _ip, _dp]),
    "occxc_build_unrestricted": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp]),
    "occxc_xc_gradient_unrestricted": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int, _ip,
                                                      _dp]),
    "occxc_build_unrestricted_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                       ctypes.c_void_p, ctypes.c_void_p,
                                                       ctypes.c_void_p]),
    "occxc_functional_evaluate_polarized": (ctypes.c_int, [ctypes.c_int, ctypes.c_longlong, _dp,
                                                           _dp, _dp, _dp, _dp, ctypes.c_int]),
    "occxc_functional_evaluate_mgga": (ctypes.c_int, [ctypes.c_int, ctypes.c_longlong, _dp, _dp,
                                                      _dp, _dp, _dp, _dp, _dp, ctypes.c_int]),
    "occxc_functional_evaluate_mgga_polarized": (ctypes.c_int, [ctypes.c_int, ctypes.c_longlong,
                                                                _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                                                ctypes.c_int]),
    "occxc_lmg_radial_grid": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp,
                                             ctypes.c_double, ctypes.c_int, _dp, _dp, _ip]),
    "occxc_nlc_vv10": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double, _dp,
                                      _dp]),
    "occxc_nlc_vv10_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "occxc_xc_gradient_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                _ip, ctypes.c_void_p, ctypes.c_void_p]),
    "occxc_xc_gradient_unrestricted_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                             ctypes.c_int, _ip, ctypes.c_void_p,
                                                             ctypes.c_void_p]),
    "occxc_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(CBasis),
                                          ctypes.POINTER(CGrid), ctypes.POINTER(CFunctional),
                                          ctypes.POINTER(COptions),
                                          ctypes.POINTER(ctypes.c_void_p)]),
    "occxc_batch_build": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_dp),
                                         ctypes.POINTER(_dp), _dp, _dp]),
    "occxc_batch_layout": (ctypes.c_int, [ctypes.c_void_p, _ip, _ip, _llp]),
    "occxc_batch_estimate_bytes": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(CBasis),
                                                  ctypes.POINTER(CGrid),
                                                  ctypes.POINTER(CFunctional),
                                                  ctypes.POINTER(COptions), _dp]),
    "occxc_build_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p]),
    "occxc_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "occxc_context_evaluate_ao": (ctypes.c_int, [ctypes.c_void_p]),
    "occxc_context_set_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "occxc_context_set_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "occxc_context_get_timing": (ctypes.c_int, [ctypes.c_void_p, _dp, _llp]),
    "occxc_evaluate_basis": (ctypes.c_int, [ctypes.POINTER(CBasis), ctypes.c_longlong, _dp,
                                            ctypes.c_int, _dp, ctypes.c_int]),
    "occxc_evaluate_decay_cutoff": (ctypes.c_int, [ctypes.POINTER(CBasis), ctypes.c_double,
                                                   _dp, ctypes.c_int]),
    "occxc_functional_evaluate": (ctypes.c_int, [ctypes.c_int, ctypes.c_longlong, _dp, _dp,
                                                 _dp, _dp, _dp, ctypes.c_int]),
    "occxc_grid_hierarchy": (ctypes.c_int, [ctypes.c_longlong, _dp, _dp, _dp, _dp, _llp, _llp,
                                            _llp, _dp, _dp, _llp]),
    "occxc_radial_grid": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp,
                                         _ip]),
    "occxc_becke_partition": (ctypes.c_int, [ctypes.c_int, _dp, _dp, ctypes.c_int,
                                             ctypes.c_longlong, _dp, _ip, _dp, ctypes.c_int]),
    "occxc_becke_partition_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p,
                                                     ctypes.c_void_p, ctypes.c_int,
                                                     ctypes.c_longlong, ctypes.c_void_p,
                                                     ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_void_p]),
    "occxc_server_default_options": (None, [ctypes.POINTER(CServerOptions)]),
    "occxc_server_create": (ctypes.c_int, [ctypes.POINTER(CServerOptions),
                                           ctypes.POINTER(ctypes.c_void_p)]),
    "occxc_server_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "occxc_server_open": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CBasis),
                                         ctypes.POINTER(CGrid), ctypes.POINTER(CFunctional),
                                         ctypes.POINTER(COptions), _llp]),
    "occxc_server_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _dp, _dp, _dp,
                                           _dp, _llp]),
    "occxc_server_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _dp]),
    "occxc_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _ip]),
    "occxc_server_cancel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _ip]),
    "occxc_server_close": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong]),
    "occxc_server_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CServerStats)]),
    "occxc_last_error_message": (ctypes.c_char_p, []),
    "occxc_version": (ctypes.c_char_p, []),
}

_lib = None


def lib():
    """Load liboccxc.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != OK:
        msg = lib().occxc_last_error_message()
        raise OccxcError(status, msg.decode() if msg else "")


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def llptr(a):
    return a.ctypes.data_as(_llp)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
