"""Writes gau2grid_collocation.npz: phi, phi_x, phi_y, phi_z (and, under
hess<k>, phi_xx, phi_xy, phi_xz, phi_yy, phi_yz, phi_zz from
gg_collocation_deriv2) from the
reference's gau2grid (src/3rdparty/gau2grid, compiled from its own sources
into oracle/_ref by `make -C oracle ref`) for L = 0..4, spherical CCA and
Cartesian, nprim 1/3/6, points at the centre, mid-range and in the tails.
The values are WITHOUT the common_fac(l) scaling evaluate_basis applies
afterwards (src/gto/gto.cpp:245-256).
"""
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent.parent / "oracle"))

import oracle  # noqa: E402
from test_oracle_basis import _collocation_cases  # noqa: E402


def main():
    out = {}
    cases = _collocation_cases()
    for k, (l, pure, exps, coefs, cen, xyz) in enumerate(cases):
        out[f"l{k}"] = l
        out[f"pure{k}"] = pure
        out[f"exps{k}"] = exps
        out[f"coefs{k}"] = coefs
        out[f"center{k}"] = cen
        out[f"xyz{k}"] = xyz
        out[f"phi{k}"] = oracle.gau2grid_deriv1(l, pure, exps, coefs, cen, xyz)
        out[f"hess{k}"] = oracle.gau2grid_deriv2(l, pure, exps, coefs, cen, xyz)[4:]
    out["ncase"] = len(cases)
    np.savez_compressed(HERE / "gau2grid_collocation.npz", **out)
    print(len(cases), "cases")


if __name__ == "__main__":
    main()
