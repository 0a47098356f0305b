"""The C ABI library loads, exports every symbol include/occxc.h declares,
and its host-side argument checks return status codes (no GPU needed: every
call here is rejected before any device work)."""
import ctypes
import pathlib
import re

import numpy as np
import pytest

from occ_amd import _lib

HEADER = pathlib.Path(__file__).resolve().parent.parent / "include" / "occxc.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(occxc_[a-z_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, f"{n} missing from the ctypes signature table"


def test_version_and_empty_error():
    L = _lib.lib()
    assert L.occxc_version().startswith(b"occxc")
    assert L.occxc_last_error_message() is not None


def test_default_options():
    o = _lib.COptions()
    _lib.lib().occxc_default_options(ctypes.byref(o))
    assert o.precision == 0 and o.screen == 1
    assert o.density_threshold == 1e-10 and o.screening_threshold == 1e-10


def test_context_create_rejects_bad_input_without_gpu():
    L = _lib.lib()
    h = ctypes.c_void_p()
    # null basis / grid / functional
    st = L.occxc_context_create(None, None, None, None, ctypes.byref(h))
    assert st == 1
    ids = np.array([9999], dtype=np.int32)
    f = _lib.CFunctional(1, _lib.iptr(ids))
    pts = np.zeros(3)
    w = np.ones(1)
    g = _lib.CGrid(1, _lib.dptr(pts), _lib.dptr(w))
    from occ_amd import AOBasis, molecules
    b, keep = AOBasis.load(molecules.water(), "sto-3g").c_struct()
    st = L.occxc_context_create(ctypes.byref(b), ctypes.byref(g), ctypes.byref(f), None,
                                ctypes.byref(h))
    assert st == 2  # unknown functional id -> UNSUPPORTED
    assert b"not found" in L.occxc_last_error_message()
    o = _lib.COptions()
    L.occxc_default_options(ctypes.byref(o))
    o.precision = 7
    st = L.occxc_context_create(ctypes.byref(b), ctypes.byref(g), ctypes.byref(f),
                                ctypes.byref(o), ctypes.byref(h))
    assert st == 1
    st = L.occxc_batch_create(0, ctypes.byref(b), ctypes.byref(g), ctypes.byref(f), None,
                              ctypes.byref(h))
    assert st == 1


def test_unsupported_angular_momentum_rejected():
    from occ_amd.basis import Shell, AOBasis
    from occ_amd.dft import evaluate_decay_cutoff
    b = AOBasis([(8, 0, 0, 0)], [Shell(5, [1.0], [1.0], (0, 0, 0), pure=True)])
    with pytest.raises(_lib.OccxcError) as e:
        evaluate_decay_cutoff(b)
    assert e.value.status == 2


def test_functional_rejects_missing_sigma():
    L = _lib.lib()
    r = np.ones(4)
    out = np.zeros(4)
    st = L.occxc_functional_evaluate(101, 4, _lib.dptr(r), None, _lib.dptr(out),
                                     _lib.dptr(out), None, 0)
    assert st == 1


def test_radial_grid_rejects_unknown_method():
    r = np.zeros(4)
    n = np.zeros(1, dtype=np.int32)
    st = _lib.lib().occxc_radial_grid(42, 4, 1.0, _lib.dptr(r), _lib.dptr(r), _lib.iptr(n))
    assert st == 2


def test_server_rejects_bad_options_without_gpu():
    L = _lib.lib()
    o = _lib.CServerOptions()
    L.occxc_server_default_options(ctypes.byref(o))
    assert o.slots == 4 and o.memory_budget == 0.0
    o.slots = 0
    h = ctypes.c_void_p()
    assert L.occxc_server_create(ctypes.byref(o), ctypes.byref(h)) == 1
    assert L.occxc_server_submit(None, 1, None, None, None, None, None) == 1
    assert L.occxc_server_destroy(None) == 0


def test_xc_gradient_rejects_null_context_without_gpu():
    L = _lib.lib()
    G = np.zeros(3)
    sa = np.zeros(1, dtype=np.int32)
    st = L.occxc_xc_gradient(None, None, 1, _lib.iptr(sa), G.ctypes.data_as(_lib._dp))
    assert st == 1
    assert b"null" in L.occxc_last_error_message()


def test_unrestricted_and_polarised_reject_without_gpu():
    L = _lib.lib()
    st = L.occxc_build_unrestricted(None, None, None, None, None)
    assert st == 1
    x = np.zeros(6)
    st = L.occxc_functional_evaluate_polarized(9999, 1, _lib.dptr(x), _lib.dptr(x), _lib.dptr(x),
                                               _lib.dptr(x), _lib.dptr(x), 0)
    assert st == 2


def test_launch_wrappers_refuse_short_coefficient_buffers(tmp_path):
    """The density and Vxc launch wrappers check that the per-point
    coefficient buffers hold every plane their kernels access, before any
    launch (tests/native/launch_guard.cpp calls them with a one-plane buffer
    for a GGA and with no tau buffer for a meta-GGA: OCCXC_ERR_INTERNAL)."""
    import subprocess
    root = pathlib.Path(__file__).resolve().parent.parent
    exe = tmp_path / "launch_guard"
    subprocess.run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{root / 'occ_amd' / 'csrc'}", f"-I{root / 'include'}",
                    str(root / "tests" / "native" / "launch_guard.cpp"),
                    f"-L{root / 'occ_amd'}", "-loccxc", "-L/opt/rocm/lib", "-lamdhip64",
                    f"-Wl,-rpath,{root / 'occ_amd'}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count("status 7") == 3


@pytest.mark.parametrize("field,value", [("tile_ao", 48), ("tile_ao", 256), ("persistent", 4),
                                         ("persistent", -2)])
def test_context_create_rejects_bad_launch_settings(field, value):
    from occ_amd.dft import default_options
    o = default_options()
    assert o.tile_ao == 0 and o.persistent == -1
    setattr(o, field, value)
    shells = _lib.i32([0])
    b = _lib.CBasis(1, _lib.iptr(shells), _lib.iptr(shells), _lib.iptr(_lib.i32([1])),
                    _lib.dptr(np.zeros(3)), _lib.dptr(np.ones(1)), _lib.dptr(np.ones(1)))
    pts = np.zeros(3)
    w = np.ones(1)
    g = _lib.CGrid(1, _lib.dptr(pts), _lib.dptr(w))
    ids = _lib.i32([1])
    f = _lib.CFunctional(1, _lib.iptr(ids))
    h = ctypes.c_void_p()
    st = _lib.lib().occxc_context_create(ctypes.byref(b), ctypes.byref(g), ctypes.byref(f),
                                         ctypes.byref(o), ctypes.byref(h))
    assert st == 1
    assert field.encode() in _lib.lib().occxc_last_error_message()
