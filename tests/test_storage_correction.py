"""CPU: the N_elec correction of the FP32 AO storage modes.

The GPU stores AO values as FP32 (s) and keeps, per context, dS = sum_p w_p
(r Phi^T + Phi r^T) with r = v - s the value plane's rounding residuals
(api.cpp / contract.hip run_storage_overlap).  Each build then adds
sum D2 (.) dS to N_elec = sum_p w_p s^T D2 s, D2 being the doubled
symmetrised density the contractions use (rho = Phi^T D2 Phi).  Here the same
quantities are formed in numpy from the oracle's AO values, and the corrected
N_elec is checked against the exact one and the oracle's; a flipped sign of r,
a missing factor of the symmetric form or D in place of D2 fails these.
"""
import numpy as np

import grid_ref
import oracle
from conftest import random_density
from occ_amd import AOBasis, molecules
from occ_amd.grid import GridSettings, MolecularGrid


def _setup():
    atoms = molecules.water_cluster(2, seed=7)
    basis = AOBasis.load(atoms, "def2-svp", pure=True)
    pts, w = grid_ref.molecular_grid(MolecularGrid(atoms, GridSettings(radial_points=31,
                                                                       angular_points=110)))
    phi = np.concatenate([oracle.eval_shell(sh.l, sh.pure, sh.exponents, sh.coefficients,
                                            sh.origin, pts, deriv=0)[0]
                          for sh in basis.shells], axis=0)  # (nbf, npts)
    D = random_density(basis.nbf, 10, 5)
    return basis, pts, w, phi, D


def test_first_order_storage_correction_recovers_n_elec():
    basis, pts, w, phi, D = _setup()
    D2 = D + D.T  # doubled, symmetrised (k_prepare_density)
    n_of = lambda p: float(np.einsum("p,mp,mn,np->", w, p, D2, p))
    n_exact = n_of(phi)
    # the oracle's N_elec for the same D (unscreened): rho = Phi^T D2 Phi
    ob = oracle.OracleBasis.from_flat(basis.flat())
    _, _, nel = oracle.xc_build(ob, pts, w, [1, 7], D, screened=False)
    assert abs(n_exact - nel) <= 1e-9 * abs(nel)
    s = phi.astype(np.float32).astype(np.float64)
    r = phi - s
    n32 = n_of(s)
    dS = (s * w) @ r.T
    dS = dS + dS.T
    corrected = n32 + float((D2 * dS).sum())
    err32, errc = abs(n32 - n_exact), abs(corrected - n_exact)
    assert err32 > 1e-9            # the FP32 values alone miss N_elec
    assert errc <= 1e-3 * err32    # first order: the rest is r^T D2 r
    assert errc <= 1e-12 * abs(n_exact)
    # a residual of the wrong sign doubles the error instead
    assert abs(n32 - float((D2 * dS).sum()) - n_exact) > 1.5 * err32
