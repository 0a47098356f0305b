"""ctypes access to the CPU oracle (liboccxc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  The product package
(occ_amd) never imports this module.
"""
import ctypes
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboccxc_oracle.so"
REF_GG_PATH = HERE / "_ref" / "libgau2grid_ref.so"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


class _Basis(ctypes.Structure):
    _fields_ = [("nshell", ctypes.c_int), ("l", _ip), ("pure", _ip),
                ("nprim", _ip), ("prim_offset", _ip), ("origin", _dp),
                ("exps", _dp), ("coefs", _dp)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        L = ctypes.CDLL(str(LIB_PATH))
        L.oracle_normalize_shell.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.oracle_common_fac.restype = ctypes.c_double
        L.oracle_common_fac.argtypes = [ctypes.c_int, ctypes.c_int]
        L.oracle_cart_to_sph.argtypes = [ctypes.c_int, _dp]
        L.oracle_eval_shell.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                        ctypes.c_long, _dp, ctypes.c_int, _dp, _dp, _dp, _dp,
                                        ctypes.c_long]
        L.oracle_decay_cutoff.argtypes = [ctypes.POINTER(_Basis), ctypes.c_double, _dp]
        L.oracle_build_hierarchy.restype = ctypes.c_long
        L.oracle_build_hierarchy.argtypes = [ctypes.c_long, _dp, _dp, _dp, _dp, _lp, _lp, _lp,
                                             _dp, _dp]
        L.oracle_functional_family.argtypes = [ctypes.c_int]
        L.oracle_functional_eval.argtypes = [ctypes.c_int, ctypes.c_long, _dp, _dp, _dp, _dp, _dp]
        L.oracle_functional_energy_density.restype = ctypes.c_double
        L.oracle_functional_energy_density.argtypes = [ctypes.c_int, ctypes.c_double,
                                                       ctypes.c_double]
        L.oracle_xc_build.argtypes = [ctypes.POINTER(_Basis), ctypes.c_long, _dp, _dp,
                                      ctypes.c_int, _ip, _dp, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.oracle_eval_shell2.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                         _dp, ctypes.c_long, _dp, _dp, ctypes.c_long]
        L.oracle_xc_gradient.argtypes = [ctypes.POINTER(_Basis), _ip, ctypes.c_int,
                                         ctypes.c_long, _dp, _dp, ctypes.c_int, _ip, _dp,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_int, _dp]
        L.oracle_functional_eval_pol.argtypes = [ctypes.c_int, ctypes.c_long, _dp, _dp, _dp, _dp,
                                                 _dp]
        L.oracle_functional_energy_density_pol.restype = ctypes.c_double
        L.oracle_functional_energy_density_pol.argtypes = [ctypes.c_int] + [ctypes.c_double] * 5
        L.oracle_xc_build_unrestricted.argtypes = [ctypes.POINTER(_Basis), ctypes.c_long, _dp, _dp,
                                                   ctypes.c_int, _ip, _dp, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                   _dp, _dp, _dp]
        L.oracle_xc_gradient_unrestricted.argtypes = L.oracle_xc_gradient.argtypes
        L.oracle_functional_eval_mgga.argtypes = [ctypes.c_int, ctypes.c_long] + [_dp] * 7
        L.oracle_functional_eval_mgga_pol.argtypes = [ctypes.c_int, ctypes.c_long] + [_dp] * 7
        L.oracle_functional_energy_density_mgga_pol.argtypes = [ctypes.c_int, _dp]
        L.oracle_functional_energy_density_mgga_pol.restype = ctypes.c_double
        L.oracle_vv10.argtypes = [ctypes.POINTER(_Basis), ctypes.c_long, _dp, _dp, _dp,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.oracle_vv10_kernel.argtypes = [ctypes.c_long, _dp, _dp, _dp, _dp, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, _dp, _dp, _dp]
        L.oracle_functional_energy_density_mgga.restype = ctypes.c_double
        L.oracle_functional_energy_density_mgga.argtypes = [ctypes.c_int] + [ctypes.c_double] * 3
        L.oracle_screened_work.restype = ctypes.c_long
        L.oracle_screened_work.argtypes = [ctypes.POINTER(_Basis), ctypes.c_long, _dp, _dp,
                                           ctypes.c_double, ctypes.c_long, _lp, _lp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _l(a):
    return a.ctypes.data_as(_lp)


class OracleBasis:
    """Flat arrays of a basis (shell order, normalised coefficients)."""

    def __init__(self, l, pure, nprim, origin, exps, coefs):
        self.l = np.ascontiguousarray(l, dtype=np.int32)
        self.pure = np.ascontiguousarray(pure, dtype=np.int32)
        self.nprim = np.ascontiguousarray(nprim, dtype=np.int32)
        self.prim_offset = np.zeros(len(self.l), dtype=np.int32)
        self.prim_offset[1:] = np.cumsum(self.nprim)[:-1]
        self.origin = np.ascontiguousarray(origin, dtype=np.float64).reshape(-1)
        self.exps = np.ascontiguousarray(exps, dtype=np.float64)
        self.coefs = np.ascontiguousarray(coefs, dtype=np.float64)
        self._s = _Basis(len(self.l), _i(self.l), _i(self.pure), _i(self.nprim),
                         _i(self.prim_offset), _d(self.origin), _d(self.exps), _d(self.coefs))

    @classmethod
    def from_flat(cls, flat):
        """From occ_amd.basis.AOBasis.flat() style dict."""
        return cls(flat["l"], flat["pure"], flat["nprim"], flat["origin"],
                   flat["exponents"], flat["coefficients"])

    @property
    def nbf(self):
        return int(sum((2 * l + 1) if p else (l + 1) * (l + 2) // 2
                       for l, p in zip(self.l, self.pure)))

    @property
    def ptr(self):
        return ctypes.byref(self._s)


def normalize_shell(l, exps, raw):
    exps = np.ascontiguousarray(exps, dtype=np.float64)
    raw = np.ascontiguousarray(raw, dtype=np.float64)
    out = np.zeros_like(raw)
    lib().oracle_normalize_shell(l, len(exps), _d(exps), _d(raw), _d(out))
    return out


def common_fac(l, pure):
    return lib().oracle_common_fac(l, int(pure))


def cart_to_sph(l):
    ncart = (l + 1) * (l + 2) // 2
    out = np.zeros((2 * l + 1) * ncart)
    lib().oracle_cart_to_sph(l, _d(out))
    return out.reshape(2 * l + 1, ncart)


def eval_shell(l, pure, exps, coefs, center, xyz, deriv=1):
    exps = np.ascontiguousarray(exps, dtype=np.float64)
    coefs = np.ascontiguousarray(coefs, dtype=np.float64)
    center = np.ascontiguousarray(center, dtype=np.float64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    nc = (2 * l + 1) if pure else (l + 1) * (l + 2) // 2
    out = np.zeros((4, nc, n))
    lib().oracle_eval_shell(l, int(pure), len(exps), _d(exps), _d(coefs), _d(center), n,
                            _d(xyz), deriv, _d(out[0]), _d(out[1]), _d(out[2]), _d(out[3]), n)
    return out if deriv else out[:1]


def eval_shell2(l, pure, exps, coefs, center, xyz):
    """(10, ncomp, npts): value, gradient, Hessian xx,xy,xz,yy,yz,zz."""
    exps = np.ascontiguousarray(exps, dtype=np.float64)
    coefs = np.ascontiguousarray(coefs, dtype=np.float64)
    center = np.ascontiguousarray(center, dtype=np.float64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    nc = (2 * l + 1) if pure else (l + 1) * (l + 2) // 2
    out = np.zeros((10, nc, n))
    lib().oracle_eval_shell2(l, int(pure), len(exps), _d(exps), _d(coefs), _d(center), n,
                             _d(xyz), _d(out), n)
    return out


def decay_cutoff(basis, threshold=1e-10):
    ext = np.zeros(len(basis.l))
    lib().oracle_decay_cutoff(basis.ptr, threshold, _d(ext))
    return ext


def build_hierarchy(xyz, w):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    w = np.ascontiguousarray(w, dtype=np.float64)
    n = len(w)
    cap = n // 32 + 2
    sxyz = np.zeros((n, 3))
    sw = np.zeros(n)
    perm = np.zeros(n, dtype=np.int64)
    off = np.zeros(cap, dtype=np.int64)
    cnt = np.zeros(cap, dtype=np.int64)
    cen = np.zeros((cap, 3))
    rad = np.zeros(cap)
    nl = lib().oracle_build_hierarchy(n, _d(xyz), _d(w), _d(sxyz), _d(sw), _l(perm), _l(off),
                                      _l(cnt), _d(cen), _d(rad))
    return dict(points=sxyz, weights=sw, perm=perm, offset=off[:nl], count=cnt[:nl],
                center=cen[:nl], radius=rad[:nl])


def functional_eval(fid, rho, sigma=None):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    n = len(rho)
    fam = lib().oracle_functional_family(fid)
    if fam < 0:
        raise ValueError(f"unknown functional id {fid}")
    sig = np.ascontiguousarray(sigma if sigma is not None else np.zeros(n), dtype=np.float64)
    exc, vrho, vsig = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().oracle_functional_eval(fid, n, _d(rho), _d(sig), _d(exc), _d(vrho), _d(vsig))
    return exc, vrho, vsig


def energy_density(fid, rho, sigma=0.0):
    return lib().oracle_functional_energy_density(fid, float(rho), float(sigma))


def functional_eval_mgga(fid, rho, sigma, tau):
    """Meta-GGA evaluation (total rho, sigma, tau per point) -> exc, vrho,
    vsigma, vtau."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    n = len(rho)
    if lib().oracle_functional_family(fid) < 0:
        raise ValueError(f"unknown functional id {fid}")
    sig = np.ascontiguousarray(sigma, dtype=np.float64)
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    exc, vrho, vsig, vtau = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    lib().oracle_functional_eval_mgga(fid, n, _d(rho), _d(sig), _d(tau), _d(exc), _d(vrho),
                                      _d(vsig), _d(vtau))
    return exc, vrho, vsig, vtau


def energy_density_mgga(fid, rho, sigma, tau):
    return lib().oracle_functional_energy_density_mgga(fid, float(rho), float(sigma), float(tau))


def xc_build(basis, xyz, w, func_ids, D, density_threshold=1e-10,
             screening_threshold=1e-10, screened=True, nthreads=0):
    """compute_K_dft restated on the CPU: returns (K, Exc, N_elec)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    w = np.ascontiguousarray(w, dtype=np.float64)
    ids = np.ascontiguousarray(func_ids, dtype=np.int32)
    D = np.asfortranarray(D, dtype=np.float64)
    nbf = basis.nbf
    assert D.shape == (nbf, nbf)
    K = np.zeros((nbf, nbf), order="F")
    exc = ctypes.c_double()
    nel = ctypes.c_double()
    rc = lib().oracle_xc_build(basis.ptr, len(w), _d(xyz), _d(w), len(ids), _i(ids),
                               D.ctypes.data_as(_dp), density_threshold, screening_threshold,
                               int(screened), int(nthreads), K.ctypes.data_as(_dp),
                               ctypes.byref(exc), ctypes.byref(nel))
    if rc != 0:
        raise ValueError("oracle_xc_build rejected its input")
    return K, exc.value, nel.value


def functional_eval_pol(fid, rho, sigma=None):
    """Polarised evaluation: rho (n, 2), sigma (n, 3) -> exc (n), vrho (n, 2),
    vsigma (n, 3)."""
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1, 2)
    n = rho.shape[0]
    if lib().oracle_functional_family(fid) < 0:
        raise ValueError(f"unknown functional id {fid}")
    sig = np.ascontiguousarray(sigma if sigma is not None else np.zeros((n, 3)),
                               dtype=np.float64).reshape(-1, 3)
    exc, vrho, vsig = np.zeros(n), np.zeros((n, 2)), np.zeros((n, 3))
    lib().oracle_functional_eval_pol(fid, n, _d(rho), _d(sig), _d(exc), _d(vrho), _d(vsig))
    return exc, vrho, vsig


def energy_density_pol(fid, ra, rb, saa=0.0, sab=0.0, sbb=0.0):
    return lib().oracle_functional_energy_density_pol(fid, float(ra), float(rb), float(saa),
                                                      float(sab), float(sbb))


def functional_eval_mgga_pol(fid, rho, sigma, tau):
    """Polarised meta-GGA: rho (n, 2), sigma (n, 3), tau (n, 2) -> exc (n),
    vrho (n, 2), vsigma (n, 3), vtau (n, 2)."""
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1, 2)
    n = rho.shape[0]
    if lib().oracle_functional_family(fid) < 0:
        raise ValueError(f"unknown functional id {fid}")
    sig = np.ascontiguousarray(sigma, dtype=np.float64).reshape(-1, 3)
    tau = np.ascontiguousarray(tau, dtype=np.float64).reshape(-1, 2)
    exc, vrho, vsig, vtau = np.zeros(n), np.zeros((n, 2)), np.zeros((n, 3)), np.zeros((n, 2))
    lib().oracle_functional_eval_mgga_pol(fid, n, _d(rho), _d(sig), _d(tau), _d(exc), _d(vrho),
                                          _d(vsig), _d(vtau))
    return exc, vrho, vsig, vtau


def energy_density_mgga_pol(fid, ra, rb, saa, sab, sbb, ta, tb):
    x = np.array([ra, rb, saa, sab, sbb, ta, tb], dtype=np.float64)
    return lib().oracle_functional_energy_density_mgga_pol(fid, _d(x))


def xc_build_unrestricted(basis, xyz, w, func_ids, D, density_threshold=1e-10,
                          screening_threshold=1e-10, screened=True, nthreads=0):
    """compute_K_dft<d, Unrestricted> on the CPU: D is (2 nbf, nbf) (alpha
    block over beta block); returns (K (2 nbf, nbf), Exc, (N_alpha, N_beta))."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    w = np.ascontiguousarray(w, dtype=np.float64)
    ids = np.ascontiguousarray(func_ids, dtype=np.int32)
    D = np.asfortranarray(D, dtype=np.float64)
    nbf = basis.nbf
    assert D.shape == (2 * nbf, nbf)
    K = np.zeros((2 * nbf, nbf), order="F")
    exc = ctypes.c_double()
    nel = np.zeros(2)
    rc = lib().oracle_xc_build_unrestricted(basis.ptr, len(w), _d(xyz), _d(w), len(ids), _i(ids),
                                            D.ctypes.data_as(_dp), density_threshold,
                                            screening_threshold, int(screened), int(nthreads),
                                            K.ctypes.data_as(_dp), ctypes.byref(exc), _d(nel))
    if rc != 0:
        raise ValueError("oracle_xc_build_unrestricted rejected its input")
    return K, exc.value, (float(nel[0]), float(nel[1]))


def xc_gradient(basis, shell_atom, natoms, xyz, w, func_ids, D, density_threshold=1e-10,
                screening_threshold=1e-10, screened=True, nthreads=0):
    """compute_xc_gradient restated on the CPU: returns the 3 x natoms XC
    contribution to the nuclear gradient."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    w = np.ascontiguousarray(w, dtype=np.float64)
    ids = np.ascontiguousarray(func_ids, dtype=np.int32)
    sa = np.ascontiguousarray(shell_atom, dtype=np.int32)
    D = np.asfortranarray(D, dtype=np.float64)
    nbf = basis.nbf
    assert D.shape == (nbf, nbf) and len(sa) == len(basis.l)
    G = np.zeros((3, natoms), order="F")
    rc = lib().oracle_xc_gradient(basis.ptr, _i(sa), int(natoms), len(w), _d(xyz), _d(w),
                                  len(ids), _i(ids), D.ctypes.data_as(_dp), density_threshold,
                                  screening_threshold, int(screened), int(nthreads),
                                  G.ctypes.data_as(_dp))
    if rc != 0:
        raise ValueError("oracle_xc_gradient rejected its input")
    return G


def xc_gradient_unrestricted(basis, shell_atom, natoms, xyz, w, func_ids, D,
                             density_threshold=1e-10, screening_threshold=1e-10, screened=True,
                             nthreads=0):
    """compute_xc_gradient for an unrestricted D ((2 nbf, nbf), alpha block
    over beta block): the 3 x natoms XC gradient."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    w = np.ascontiguousarray(w, dtype=np.float64)
    ids = np.ascontiguousarray(func_ids, dtype=np.int32)
    sa = np.ascontiguousarray(shell_atom, dtype=np.int32)
    D = np.asfortranarray(D, dtype=np.float64)
    nbf = basis.nbf
    assert D.shape == (2 * nbf, nbf) and len(sa) == len(basis.l)
    G = np.zeros((3, natoms), order="F")
    rc = lib().oracle_xc_gradient_unrestricted(basis.ptr, _i(sa), int(natoms), len(w), _d(xyz),
                                               _d(w), len(ids), _i(ids), D.ctypes.data_as(_dp),
                                               density_threshold, screening_threshold,
                                               int(screened), int(nthreads), G.ctypes.data_as(_dp))
    if rc != 0:
        raise ValueError("oracle_xc_gradient_unrestricted rejected its input")
    return G


def screened_work(basis, xyz, w, screening_threshold=1e-10):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    w = np.ascontiguousarray(w, dtype=np.float64)
    cap = len(w) // 32 + 2
    ln = np.zeros(cap, dtype=np.int64)
    lm = np.zeros(cap, dtype=np.int64)
    k = lib().oracle_screened_work(basis.ptr, len(w), _d(xyz), _d(w), screening_threshold, cap,
                                   _l(ln), _l(lm))
    return ln[:k], lm[:k]


# ---- the reference's own gau2grid, compiled from its sources into _ref/ ----

GG_SPHERICAL_CCA = 300
GG_CARTESIAN_CCA = 400
_gg = None


def gau2grid_available():
    return REF_GG_PATH.exists()


def gau2grid():
    global _gg
    if _gg is None:
        L = ctypes.CDLL(str(REF_GG_PATH))
        L.gg_collocation_deriv1.argtypes = [ctypes.c_int, ctypes.c_ulong, _dp, ctypes.c_ulong,
                                            ctypes.c_int, _dp, _dp, _dp, ctypes.c_int,
                                            _dp, _dp, _dp, _dp]
        L.gg_collocation.argtypes = [ctypes.c_int, ctypes.c_ulong, _dp, ctypes.c_ulong,
                                     ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp]
        L.gg_collocation_deriv2.argtypes = [ctypes.c_int, ctypes.c_ulong, _dp, ctypes.c_ulong,
                                            ctypes.c_int, _dp, _dp, _dp, ctypes.c_int] + [_dp] * 10
        _gg = L
    return _gg


def gau2grid_deriv1(l, pure, exps, coefs, center, xyz):
    """gg_collocation_deriv1 exactly as gto.cpp calls it (xyz stride 3), WITHOUT
    the common_fac scaling that evaluate_basis applies afterwards."""
    exps = np.ascontiguousarray(exps, dtype=np.float64)
    coefs = np.ascontiguousarray(coefs, dtype=np.float64)
    center = np.ascontiguousarray(center, dtype=np.float64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    nc = (2 * l + 1) if pure else (l + 1) * (l + 2) // 2
    out = np.zeros((4, nc, n))
    order = GG_SPHERICAL_CCA if pure else GG_CARTESIAN_CCA
    gau2grid().gg_collocation_deriv1(l, n, _d(xyz), 3, len(exps), _d(coefs), _d(exps),
                                     _d(center), order, _d(out[0]), _d(out[1]), _d(out[2]),
                                     _d(out[3]))
    return out


os.environ.setdefault("OMP_PROC_BIND", "false")


def gau2grid_deriv2(l, pure, exps, coefs, center, xyz):
    """gg_collocation_deriv2 as gto.cpp:267-270 calls it: (10, ncomp, npts) =
    value, gradient, Hessian xx,xy,xz,yy,yz,zz, WITHOUT common_fac."""
    exps = np.ascontiguousarray(exps, dtype=np.float64)
    coefs = np.ascontiguousarray(coefs, dtype=np.float64)
    center = np.ascontiguousarray(center, dtype=np.float64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    nc = (2 * l + 1) if pure else (l + 1) * (l + 2) // 2
    out = np.zeros((10, nc, n))
    order = GG_SPHERICAL_CCA if pure else GG_CARTESIAN_CCA
    gau2grid().gg_collocation_deriv2(l, n, _d(xyz), 3, len(exps), _d(coefs), _d(exps),
                                     _d(center), order, *[_d(out[q]) for q in range(10)])
    return out


def vv10(basis, xyz, w, D, b=6.0, C=0.01, nthreads=0):
    """NonLocalCorrelationFunctional::vv10 on the CPU: (energy, Vxc) of the
    restricted density D on the grid in the given point order."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    w = np.ascontiguousarray(w, dtype=np.float64)
    D = np.asfortranarray(D, dtype=np.float64)
    K = np.zeros((basis.nbf, basis.nbf), order="F")
    e = ctypes.c_double()
    rc = lib().oracle_vv10(basis.ptr, len(w), _d(xyz), _d(w), D.ctypes.data_as(_dp), b, C,
                           int(nthreads), ctypes.byref(e), K.ctypes.data_as(_dp))
    if rc != 0:
        raise ValueError("oracle_vv10 rejected its input")
    return e.value, K


def vv10_kernel(rho, sigma, xyz, w, b=6.0, C=0.01, nthreads=0):
    """vv10_kernel: weighted exc, vrho, vsigma per point."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    w = np.ascontiguousarray(w, dtype=np.float64)
    n = len(rho)
    exc, vr, vs = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().oracle_vv10_kernel(n, _d(rho), _d(sigma), _d(xyz), _d(w), b, C, int(nthreads),
                             _d(exc), _d(vr), _d(vs))
    return exc, vr, vs
