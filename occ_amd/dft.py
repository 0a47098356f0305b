"""Host-side mirror of occ::dft::DFT's XC interface over the occxc C ABI.

    dft = DFT("pbe", basis, points, weights)        # DFT::DFT (src/dft/dft.cpp:19-35)
    K = dft.compute_vxc(D)                          # DFT::compute_vxc (dft.h:395-429)
    dft.exchange_correlation_energy()               # dft.h:83
    dft.total_density()                             # N_elec, logged only in OCC (dft.h:384-385)

Method names resolve to libxc components as get_dft_method does
(src/dft/dft_method.cpp:161-227, share/methods/dft_methods.json).
"""
import ctypes

import numpy as np

from . import _lib

LIBXC_IDS = {"lda_x": 1, "lda_c_vwn": 7, "lda_c_pw": 12, "lda_c_pw_mod": 13,
             "gga_x_pbe": 101, "gga_x_b88": 106, "gga_c_pbe": 130, "gga_c_lyp": 131,
             "mgga_x_tpss": 202, "mgga_c_tpss": 231}
MGGA_IDS = {202, 231}

# method -> libxc components (entries of share/methods/dft_methods.json whose
# components this build implements)
METHODS = {"svwn5": ["lda_x", "lda_c_vwn"], "vwn5": ["lda_c_vwn"],
           "pbe": ["gga_x_pbe", "gga_c_pbe"], "pbepbe": ["gga_x_pbe", "gga_c_pbe"],
           "blyp": ["gga_x_b88", "gga_c_lyp"], "bpbe": ["gga_x_b88", "gga_c_pbe"],
           "slater": ["lda_x"], "tpss": ["mgga_x_tpss", "mgga_c_tpss"]}

PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2, "fp32cache": 3}


def functional_ids(method):
    """libxc ids of a method name or of a list of libxc component names."""
    if isinstance(method, str):
        key = method.lower()
        if key in METHODS:
            comps = METHODS[key]
        elif key.startswith("xc_") and key[3:] in LIBXC_IDS:
            comps = [key[3:]]
        elif key in LIBXC_IDS:
            comps = [key]
        else:
            raise ValueError(f"Unknown functional name {method}")
    else:
        comps = list(method)
    return [LIBXC_IDS[c] if isinstance(c, str) else int(c) for c in comps]


def default_options():
    o = _lib.COptions()
    _lib.lib().occxc_default_options(ctypes.byref(o))
    return o


class XCContext:
    """One occxc context: basis + grid + functional resident on one GPU."""

    def __init__(self, basis, points, weights, funcs, precision="fp64", screen=True,
                 density_threshold=1e-10, screening_threshold=1e-10, device=0, tile_ao=0,
                 persistent=-1):
        L = _lib.lib()
        self.nbf = basis.nbf
        self._basis, self._keep = basis.c_struct()
        self._pts = _lib.f64(points).reshape(-1)
        self._wts = _lib.f64(weights)
        if len(self._pts) != 3 * len(self._wts):
            raise ValueError("points must be [npts, 3] matching weights")
        self._grid = _lib.CGrid(len(self._wts), _lib.dptr(self._pts), _lib.dptr(self._wts))
        self._ids = _lib.i32(funcs)
        self._func = _lib.CFunctional(len(self._ids), _lib.iptr(self._ids))
        o = default_options()
        o.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        o.screen = int(bool(screen))
        o.density_threshold = density_threshold
        o.screening_threshold = screening_threshold
        o.device = device
        o.tile_ao = tile_ao
        o.persistent = persistent
        self.options = o
        h = ctypes.c_void_p()
        _lib.check(L.occxc_context_create(ctypes.byref(self._basis), ctypes.byref(self._grid),
                                          ctypes.byref(self._func), ctypes.byref(o),
                                          ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def info(self):
        i = _lib.CContextInfo()
        _lib.check(_lib.lib().occxc_context_get_info(self._h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in i._fields_}

    def build(self, D):
        D = np.asfortranarray(D, dtype=np.float64)
        if D.shape != (self.nbf, self.nbf):
            raise ValueError(f"D must be {self.nbf}x{self.nbf}")
        V = np.zeros((self.nbf, self.nbf), order="F")
        exc = ctypes.c_double()
        nel = ctypes.c_double()
        _lib.check(_lib.lib().occxc_build(self._h, D.ctypes.data_as(_lib._dp),
                                          V.ctypes.data_as(_lib._dp), ctypes.byref(exc),
                                          ctypes.byref(nel)))
        return V, exc.value, nel.value

    def build_device(self, d_D, d_V, d_result, stream=None):
        _lib.check(_lib.lib().occxc_build_device(self._h, ctypes.c_void_p(d_D),
                                                 ctypes.c_void_p(d_V),
                                                 ctypes.c_void_p(d_result),
                                                 ctypes.c_void_p(stream or 0)))

    def nlc_vv10(self, D, b=6.0, C=0.01):
        """VV10 energy and Vxc of the restricted density D on this context's
        grid (NonLocalCorrelationFunctional::vv10)."""
        D = np.asfortranarray(D, dtype=np.float64)
        if D.shape != (self.nbf, self.nbf):
            raise ValueError(f"D must be {self.nbf}x{self.nbf}")
        V = np.zeros((self.nbf, self.nbf), order="F")
        e = ctypes.c_double()
        _lib.check(_lib.lib().occxc_nlc_vv10(self._h, D.ctypes.data_as(_lib._dp), float(b),
                                             float(C), ctypes.byref(e),
                                             V.ctypes.data_as(_lib._dp)))
        return e.value, V

    def nlc_vv10_device(self, d_D, d_V, d_result, b=6.0, C=0.01, stream=None):
        _lib.check(_lib.lib().occxc_nlc_vv10_device(self._h, ctypes.c_void_p(d_D), float(b),
                                                    float(C), ctypes.c_void_p(d_V),
                                                    ctypes.c_void_p(d_result),
                                                    ctypes.c_void_p(stream or 0)))

    def build_unrestricted(self, D):
        """Unrestricted build (occxc_build_unrestricted): D is (2 nbf, nbf),
        alpha block over beta block; returns (Vxc (2 nbf, nbf), Exc,
        (N_alpha, N_beta))."""
        D = np.asfortranarray(D, dtype=np.float64)
        if D.shape != (2 * self.nbf, self.nbf):
            raise ValueError(f"D must be {2 * self.nbf}x{self.nbf}")
        V = np.zeros((2 * self.nbf, self.nbf), order="F")
        exc = ctypes.c_double()
        nel = np.zeros(2)
        _lib.check(_lib.lib().occxc_build_unrestricted(self._h, D.ctypes.data_as(_lib._dp),
                                                       V.ctypes.data_as(_lib._dp),
                                                       ctypes.byref(exc), _lib.dptr(nel)))
        return V, exc.value, (float(nel[0]), float(nel[1]))

    def xc_gradient_unrestricted(self, D, shell_atom, natoms=None):
        """Unrestricted XC gradient (occxc_xc_gradient_unrestricted): D is
        (2 nbf, nbf); returns 3 x natoms."""
        D = np.asfortranarray(D, dtype=np.float64)
        if D.shape != (2 * self.nbf, self.nbf):
            raise ValueError(f"D must be {2 * self.nbf}x{self.nbf}")
        sa = _lib.i32(shell_atom)
        natoms = int(natoms if natoms is not None else (sa.max() + 1 if len(sa) else 0))
        G = np.zeros((3, max(natoms, 1)), order="F")
        _lib.check(_lib.lib().occxc_xc_gradient_unrestricted(
            self._h, D.ctypes.data_as(_lib._dp), natoms, _lib.iptr(sa),
            G.ctypes.data_as(_lib._dp)))
        return G

    def build_unrestricted_device(self, d_D, d_V, d_result, stream=None):
        """occxc_build_unrestricted_device: device pointers to D and Vxc
        (2 nbf x nbf each) and {Exc, N_alpha, N_beta}."""
        _lib.check(_lib.lib().occxc_build_unrestricted_device(
            self._h, ctypes.c_void_p(d_D), ctypes.c_void_p(d_V), ctypes.c_void_p(d_result),
            ctypes.c_void_p(stream or 0)))

    def xc_gradient(self, D, shell_atom, natoms=None):
        """XC contribution to the nuclear gradient (occxc_xc_gradient):
        3 x natoms, column per atom."""
        D = np.asfortranarray(D, dtype=np.float64)
        if D.shape != (self.nbf, self.nbf):
            raise ValueError(f"D must be {self.nbf}x{self.nbf}")
        sa = _lib.i32(shell_atom)
        natoms = int(natoms if natoms is not None else (sa.max() + 1 if len(sa) else 0))
        G = np.zeros((3, max(natoms, 1)), order="F")
        _lib.check(_lib.lib().occxc_xc_gradient(self._h, D.ctypes.data_as(_lib._dp), natoms,
                                                _lib.iptr(sa), G.ctypes.data_as(_lib._dp)))
        return G

    def xc_gradient_device(self, d_D, shell_atom, natoms, d_grad, stream=None):
        """occxc_xc_gradient_device: D and the 3 x natoms gradient in device
        memory (pointers), enqueued without host synchronisation."""
        self._sa = _lib.i32(shell_atom)
        _lib.check(_lib.lib().occxc_xc_gradient_device(self._h, ctypes.c_void_p(d_D),
                                                       int(natoms), _lib.iptr(self._sa),
                                                       ctypes.c_void_p(d_grad),
                                                       ctypes.c_void_p(stream or 0)))

    def xc_gradient_unrestricted_device(self, d_D, shell_atom, natoms, d_grad, stream=None):
        """occxc_xc_gradient_unrestricted_device: D (2 nbf x nbf) and the
        3 x natoms gradient in device memory, enqueued without host
        synchronisation."""
        self._sa = _lib.i32(shell_atom)
        _lib.check(_lib.lib().occxc_xc_gradient_unrestricted_device(
            self._h, ctypes.c_void_p(d_D), int(natoms), _lib.iptr(self._sa),
            ctypes.c_void_p(d_grad), ctypes.c_void_p(stream or 0)))

    STAGES = ("prepare", "density", "vxc", "finalize", "ao", "gradient", "nlc")

    def set_timing(self, enable=True):
        _lib.check(_lib.lib().occxc_context_set_timing(self._h, int(bool(enable))))

    def set_graph(self, enable=True):
        """Replay builds from a captured HIP graph (see occxc_context_set_graph)."""
        _lib.check(_lib.lib().occxc_context_set_graph(self._h, int(bool(enable))))

    def timing(self):
        """{stage: (total ms, launches)} of the timed builds since set_timing."""
        ms = np.zeros(len(self.STAGES))
        n = np.zeros(len(self.STAGES), dtype=np.int64)
        _lib.check(_lib.lib().occxc_context_get_timing(self._h, _lib.dptr(ms), _lib.llptr(n)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.STAGES)}

    def evaluate_ao(self):
        _lib.check(_lib.lib().occxc_context_evaluate_ao(self._h))

    def synchronize(self):
        _lib.check(_lib.lib().occxc_synchronize(self._h))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.check(_lib.lib().occxc_context_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def estimate_batch_bytes(bases, grids, funcs, precision="fp64", screen=True, device=0):
    """Device bytes each molecule adds to an XCBatch (occxc_batch_estimate_bytes):
    hierarchy and screening only.  Returns an array, one entry per molecule."""
    n = len(bases)
    keep = []
    cb = (_lib.CBasis * n)()
    cg = (_lib.CGrid * n)()
    for i, (b, (pts, w)) in enumerate(zip(bases, grids)):
        st, k = b.c_struct()
        pts = _lib.f64(pts).reshape(-1)
        w = _lib.f64(w)
        keep += [k, pts, w]
        cb[i] = st
        cg[i] = _lib.CGrid(len(w), _lib.dptr(pts), _lib.dptr(w))
    ids = _lib.i32(funcs)
    func = _lib.CFunctional(len(ids), _lib.iptr(ids))
    o = default_options()
    o.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    o.screen = int(bool(screen))
    o.device = device
    out = np.zeros(n)
    _lib.check(_lib.lib().occxc_batch_estimate_bytes(n, cb, cg, ctypes.byref(func),
                                                     ctypes.byref(o), _lib.dptr(out)))
    return out


class XCBatch:
    """Many molecules in one occxc context: one launch per stage for all."""

    def __init__(self, bases, grids, funcs, precision="fp64", screen=True, device=0, tile_ao=0,
                 persistent=-1):
        L = _lib.lib()
        n = len(bases)
        if n != len(grids) or n == 0:
            raise ValueError("need one (points, weights) grid per basis")
        self._keep = []
        cb = (_lib.CBasis * n)()
        cg = (_lib.CGrid * n)()
        for i, (b, (pts, w)) in enumerate(zip(bases, grids)):
            st, keep = b.c_struct()
            pts = _lib.f64(pts).reshape(-1)
            w = _lib.f64(w)
            self._keep += [keep, pts, w]
            cb[i] = st
            cg[i] = _lib.CGrid(len(w), _lib.dptr(pts), _lib.dptr(w))
        self._cb, self._cg = cb, cg
        self.nbf = [b.nbf for b in bases]
        self._ids = _lib.i32(funcs)
        self._func = _lib.CFunctional(len(self._ids), _lib.iptr(self._ids))
        o = default_options()
        o.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        o.screen = int(bool(screen))
        o.device = device
        o.tile_ao = tile_ao
        o.persistent = persistent
        h = ctypes.c_void_p()
        _lib.check(L.occxc_batch_create(n, cb, cg, ctypes.byref(self._func), ctypes.byref(o),
                                        ctypes.byref(h)))
        self._h = h
        nm = np.zeros(1, dtype=np.int32)
        nbf = np.zeros(n, dtype=np.int32)
        off = np.zeros(n, dtype=np.int64)
        _lib.check(L.occxc_batch_layout(h, _lib.iptr(nm), _lib.iptr(nbf), _lib.llptr(off)))
        self.offsets = off

    @property
    def handle(self):
        return self._h

    def build(self, Ds):
        n = len(self.nbf)
        Ds = [np.asfortranarray(D, dtype=np.float64) for D in Ds]
        Vs = [np.zeros((k, k), order="F") for k in self.nbf]
        dp = (_lib._dp * n)(*[D.ctypes.data_as(_lib._dp) for D in Ds])
        vp = (_lib._dp * n)(*[V.ctypes.data_as(_lib._dp) for V in Vs])
        exc = np.zeros(n)
        nel = np.zeros(n)
        _lib.check(_lib.lib().occxc_batch_build(self._h, dp, vp, _lib.dptr(exc),
                                                _lib.dptr(nel)))
        return Vs, exc, nel

    def build_device(self, d_D, d_V, d_result, stream=None):
        _lib.check(_lib.lib().occxc_build_device(self._h, ctypes.c_void_p(d_D),
                                                 ctypes.c_void_p(d_V),
                                                 ctypes.c_void_p(d_result),
                                                 ctypes.c_void_p(stream or 0)))

    def info(self):
        i = _lib.CContextInfo()
        _lib.check(_lib.lib().occxc_context_get_info(self._h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in i._fields_}

    set_timing = XCContext.set_timing
    set_graph = XCContext.set_graph
    evaluate_ao = XCContext.evaluate_ao
    timing = XCContext.timing
    STAGES = XCContext.STAGES
    synchronize = XCContext.synchronize
    close = XCContext.close
    __del__ = XCContext.__del__


class DFT:
    """compute_vxc / exchange_correlation_energy for a restricted density."""

    def __init__(self, method, basis, points, weights, precision="fp64", screen=True,
                 device=0):
        self.method = method
        self.basis = basis
        self.ids = functional_ids(method)
        self.ctx = XCContext(basis, points, weights, self.ids, precision=precision,
                             screen=screen, device=device)
        self._exc = 0.0
        self._nel = 0.0

    def compute_vxc(self, D):
        """Returns the matrix OCC's compute_vxc returns (Fock adds it, dft.h:435).
        A (2 nbf, nbf) D is unrestricted (alpha block over beta block, as
        compute_vxc dispatches on mo.kind, dft.h:398-424); total_density() is
        then (N_alpha, N_beta)."""
        D = np.asarray(D)
        if D.shape == (2 * self.basis.nbf, self.basis.nbf):
            V, self._exc, self._nel = self.ctx.build_unrestricted(D)
            return V
        V, self._exc, self._nel = self.ctx.build(D)
        return V

    def exchange_correlation_energy(self):
        return self._exc

    def compute_xc_gradient(self, D):
        """The 3 x natoms matrix OCC's compute_xc_gradient returns (dft.h:531);
        a (2 nbf, nbf) D is unrestricted."""
        D = np.asarray(D)
        if D.shape == (2 * self.basis.nbf, self.basis.nbf):
            return self.ctx.xc_gradient_unrestricted(D, self.basis.shell_atom,
                                                     len(self.basis.atoms))
        return self.ctx.xc_gradient(D, self.basis.shell_atom, len(self.basis.atoms))

    def total_density(self):
        return self._nel


def evaluate_basis(basis, points, deriv=1, device=0):
    """gto::evaluate_basis on the GPU: array [ncomp, nbf, npts]."""
    st, keep = basis.c_struct()
    pts = _lib.f64(points).reshape(-1)
    n = len(pts) // 3
    nc = 4 if deriv else 1
    out = np.zeros((nc, basis.nbf, n))
    _lib.check(_lib.lib().occxc_evaluate_basis(ctypes.byref(st), n, _lib.dptr(pts), int(deriv),
                                               _lib.dptr(out), device))
    return out


def evaluate_decay_cutoff(basis, threshold=1e-10, device=0):
    st, keep = basis.c_struct()
    ext = np.zeros(basis.nsh)
    _lib.check(_lib.lib().occxc_evaluate_decay_cutoff(ctypes.byref(st), threshold,
                                                      _lib.dptr(ext), device))
    return ext


class DensityFunctional:
    """DensityFunctional::evaluate for an unpolarised density on the GPU."""

    def __init__(self, name):
        ids = functional_ids(name)
        if len(ids) != 1:
            raise ValueError("DensityFunctional takes one libxc component")
        self.id = ids[0]

    def evaluate(self, rho, sigma=None, device=0):
        rho = _lib.f64(rho)
        n = len(rho)
        exc, vrho = np.zeros(n), np.zeros(n)
        gga = self.id > 100
        sig = _lib.f64(sigma) if gga else None
        vsig = np.zeros(n) if gga else None
        _lib.check(_lib.lib().occxc_functional_evaluate(
            self.id, n, _lib.dptr(rho), _lib.dptr(sig) if gga else None, _lib.dptr(exc),
            _lib.dptr(vrho), _lib.dptr(vsig) if gga else None, device))
        return exc, vrho, vsig

    def evaluate_mgga(self, rho, sigma, tau, device=0):
        """Family::MGGA evaluation (functional.cpp:49-57): total rho, sigma and
        tau = (1/2) sum_i |grad psi_i|^2 per point -> exc, vrho, vsigma, vtau."""
        rho = _lib.f64(rho)
        n = len(rho)
        sig, t = _lib.f64(sigma), _lib.f64(tau)
        exc, vrho, vsig, vtau = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        _lib.check(_lib.lib().occxc_functional_evaluate_mgga(
            self.id, n, _lib.dptr(rho), _lib.dptr(sig), _lib.dptr(t), _lib.dptr(exc),
            _lib.dptr(vrho), _lib.dptr(vsig), _lib.dptr(vtau), device))
        return exc, vrho, vsig, vtau

    def evaluate_mgga_polarized(self, rho, sigma, tau, device=0):
        """Family::MGGA on a spin-polarised density: rho (n, 2), sigma (n, 3),
        tau (n, 2) -> exc (n), vrho (n, 2), vsigma (n, 3), vtau (n, 2).  For an
        LDA id sigma and tau may be None (vsigma and vtau are then None), for
        a GGA id tau may be None (vtau is then None)."""
        rho = _lib.f64(rho).reshape(-1, 2)
        n = rho.shape[0]
        gga = self.id > 100
        mgga = self.id in MGGA_IDS
        sig = _lib.f64(sigma).reshape(-1, 3) if gga else None
        t = _lib.f64(tau).reshape(-1, 2) if mgga else None
        exc, vrho = np.zeros(n), np.zeros((n, 2))
        vsig = np.zeros((n, 3)) if gga else None
        vtau = np.zeros((n, 2)) if mgga else None

        def ptr(a):
            return _lib.dptr(a) if a is not None else None

        _lib.check(_lib.lib().occxc_functional_evaluate_mgga_polarized(
            self.id, n, _lib.dptr(rho), ptr(sig), ptr(t), _lib.dptr(exc), _lib.dptr(vrho),
            ptr(vsig), ptr(vtau), device))
        return exc, vrho, vsig, vtau

    def evaluate_polarized(self, rho, sigma=None, device=0):
        """Spin-polarised: rho (n, 2), sigma (n, 3) -> exc (n), vrho (n, 2),
        vsigma (n, 3) (None for LDA)."""
        rho = _lib.f64(rho).reshape(-1, 2)
        n = rho.shape[0]
        exc, vrho = np.zeros(n), np.zeros((n, 2))
        gga = self.id > 100
        sig = _lib.f64(sigma).reshape(-1, 3) if gga else None
        vsig = np.zeros((n, 3)) if gga else None
        _lib.check(_lib.lib().occxc_functional_evaluate_polarized(
            self.id, n, _lib.dptr(rho), _lib.dptr(sig) if gga else None, _lib.dptr(exc),
            _lib.dptr(vrho), _lib.dptr(vsig) if gga else None, device))
        return exc, vrho, vsig


class NonLocalCorrelationFunctional:
    """occ::dft::NonLocalCorrelationFunctional (include/occ/dft/
    nonlocal_correlation.h, src/dft/nonlocal_correlation.cpp): VV10 on the GPU.

        nlc = NonLocalCorrelationFunctional()
        nlc.set_parameters(b=6.0, C=0.01)
        nlc.set_integration_grid(basis, points, weights)   # the NLC grid
        energy, Vxc = nlc(D)
    """

    def __init__(self, b=6.0, C=0.01):
        self.b, self.C = float(b), float(C)
        self.ctx = None

    def set_parameters(self, b=6.0, C=0.01):
        self.b, self.C = float(b), float(C)

    def set_integration_grid(self, basis, points, weights, precision="fp64", screen=True,
                             device=0):
        if self.ctx is not None:
            self.ctx.close()
        # the functional id only selects a GGA AO cache (phi and grad phi)
        self.ctx = XCContext(basis, points, weights, [LIBXC_IDS["gga_x_pbe"]],
                             precision=precision, screen=screen, device=device)

    def __call__(self, D):
        if self.ctx is None:
            raise RuntimeError("set_integration_grid first")
        return self.ctx.nlc_vv10(D, self.b, self.C)

    def close(self):
        if self.ctx is not None:
            self.ctx.close()
            self.ctx = None


def grid_hierarchy(points, weights):
    pts = _lib.f64(points).reshape(-1)
    w = _lib.f64(weights)
    n = len(w)
    cap = n // 32 + 2
    sp = np.zeros(3 * n)
    sw = np.zeros(n)
    perm = np.zeros(n, dtype=np.int64)
    off = np.zeros(cap, dtype=np.int64)
    cnt = np.zeros(cap, dtype=np.int64)
    cen = np.zeros(3 * cap)
    rad = np.zeros(cap)
    nl = np.zeros(1, dtype=np.int64)
    _lib.check(_lib.lib().occxc_grid_hierarchy(n, _lib.dptr(pts), _lib.dptr(w), _lib.dptr(sp),
                                               _lib.dptr(sw), _lib.llptr(perm), _lib.llptr(off),
                                               _lib.llptr(cnt), _lib.dptr(cen), _lib.dptr(rad),
                                               _lib.llptr(nl)))
    k = int(nl[0])
    return dict(points=sp.reshape(n, 3), weights=sw, perm=perm, offset=off[:k], count=cnt[:k],
                center=cen[:3 * k].reshape(k, 3), radius=rad[:k])
