"""Gaussian basis sets: BSE JSON loading and shell normalisation.

Mirrors occ::gto::AOBasis::load (src/gto/shell.cpp:499-571) and
Shell::incorporate_shell_norm (shell.cpp:180-186): general contractions are
split into one shell per contraction, primitives are scaled by
gto_norm(l, alpha) = 1/sqrt(gint(2l+2, 2 alpha)) and the contraction by
1/sqrt(c^T E c) with E_ij = gint(2l+2, alpha_i + alpha_j) (shell.cpp:20-64).
Purity is a basis-wide setting, as AOBasis::set_pure.
"""
import json
import math
import pathlib

import numpy as np

from . import _lib

DATA = pathlib.Path(__file__).resolve().parent / "data" / "basis"


def _gint(n, alpha):
    """Gamma((n+1)/2) / (2 alpha^((n+1)/2)) (shell.cpp:20-23)."""
    n1_2 = 0.5 * (n + 1)
    return math.gamma(n1_2) / (2.0 * np.power(alpha, n1_2))


def normalize_contraction(l, exponents, coefficients):
    exps = np.asarray(exponents, dtype=np.float64)
    c = np.asarray(coefficients, dtype=np.float64) / np.sqrt(_gint(2 * l + 2, 2.0 * exps))
    ee = _gint(2 * l + 2, exps[:, None] + exps[None, :])
    return c / math.sqrt(float(c @ ee @ c))


class Shell:
    def __init__(self, l, exponents, coefficients, origin, pure=False, normalize=True):
        self.l = int(l)
        self.exponents = np.asarray(exponents, dtype=np.float64)
        c = np.asarray(coefficients, dtype=np.float64)
        self.coefficients = normalize_contraction(self.l, self.exponents, c) if normalize else c
        self.origin = np.asarray(origin, dtype=np.float64)
        self.pure = bool(pure)

    def size(self):
        return 2 * self.l + 1 if self.pure else (self.l + 1) * (self.l + 2) // 2

    def num_primitives(self):
        return len(self.exponents)


class AOBasis:
    def __init__(self, atoms, shells, name=""):
        self.atoms = list(atoms)   # [(Z, x, y, z)] in Bohr
        self.shells = list(shells)
        self.name = name
        self.shell_atom = []

    @classmethod
    def load(cls, atoms, name, pure=False):
        """atoms: iterable of (Z, x, y, z) in Bohr."""
        path = DATA / f"{name.lower()}.json"
        if not path.exists():
            raise FileNotFoundError(f"basis {name} not shipped ({path})")
        elements = json.loads(path.read_text())["elements"]
        shells, shell_atom = [], []
        for ia, (Z, x, y, z) in enumerate(atoms):
            key = str(int(Z))
            if key not in elements:
                raise ValueError(f"element {Z} not found in basis {name}")
            for s in elements[key]["electron_shells"]:
                exps = [float(e) for e in s["exponents"]]
                am = s["angular_momentum"]
                for i, coefs in enumerate(s["coefficients"]):
                    shells.append(Shell(am[i % len(am)], exps, [float(c) for c in coefs],
                                        (x, y, z), pure))
                    shell_atom.append(ia)
        b = cls(atoms, shells, name)
        b.shell_atom = shell_atom
        return b

    def set_pure(self, pure):
        for s in self.shells:
            s.pure = bool(pure)

    @property
    def nbf(self):
        return sum(s.size() for s in self.shells)

    @property
    def nsh(self):
        return len(self.shells)

    def first_bf(self):
        out, n = [], 0
        for s in self.shells:
            out.append(n)
            n += s.size()
        return out

    def l_max(self):
        return max(s.l for s in self.shells)

    def flat(self):
        """Flat arrays in the layout of occxc_basis."""
        return dict(
            l=np.array([s.l for s in self.shells], dtype=np.int32),
            pure=np.array([int(s.pure) for s in self.shells], dtype=np.int32),
            nprim=np.array([s.num_primitives() for s in self.shells], dtype=np.int32),
            origin=np.concatenate([s.origin for s in self.shells]).astype(np.float64),
            exponents=np.concatenate([s.exponents for s in self.shells]).astype(np.float64),
            coefficients=np.concatenate([s.coefficients for s in self.shells]).astype(np.float64))

    def c_struct(self):
        """(CBasis, keep-alive arrays) for the C ABI."""
        f = self.flat()
        st = _lib.CBasis(len(self.shells), _lib.iptr(f["l"]), _lib.iptr(f["pure"]),
                         _lib.iptr(f["nprim"]), _lib.dptr(f["origin"]),
                         _lib.dptr(f["exponents"]), _lib.dptr(f["coefficients"]))
        return st, f
