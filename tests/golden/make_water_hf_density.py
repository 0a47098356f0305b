"""Writes tests/golden/water_hf_321g.npz: the converged RHF density of water in
the 3-21G basis (spherical AOs), the density the reference's whole-path XC
known answers are quoted on (tests/isolated_xc_tests.cpp:31-48: O at the
origin, H at (0, -+0.757, 0.587) Angstrom, 3-21G, set_pure(true), HF SCF).

The SCF is written from scratch here: one- and two-electron integrals over
contracted s and p Gaussians by the McMurchie-Davidson scheme (Helgaker,
Jorgensen, Olsen, Molecular Electronic-Structure Theory, ch. 9: Hermite
expansion coefficients E, Hermite Coulomb integrals R with the Boys function),
closed-shell Roothaan-Hall iterations with DIIS on the orbital gradient
FDS - SDF until it is below 1e-12.  Integrals are formed over Cartesian
primitives and carried into the AO functions the XC build evaluates (the
package's normalised contractions, common_fac(l), the CCA order of p
functions), so the density matrix is in the XC build's AO order and
convention (D = C_occ C_occ^T, no factor 2).

Checks made before writing: the overlap matrix equals tests/gaussian_integrals
(independent code path, 1e-14), tr(2 D S) = 10, the virial ratio -V/T is
near 2, and the energy is stationary (orbital gradient below 1e-12).

Run in the build container: python tests/golden/make_water_hf_density.py
"""
import math
import pathlib
import sys

import numpy as np
from scipy.special import gammainc, gamma

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT / "oracle"))

from occ_amd import AOBasis  # noqa: E402
from occ_amd.molecules import ANGSTROM_TO_BOHR  # noqa: E402

OUT = pathlib.Path(__file__).resolve().parent / "water_hf_321g.npz"

# tests/isolated_xc_tests.cpp:31-38 (Angstrom)
WATER_ANGSTROM = [(8, 0.0, 0.0, 0.0), (1, 0.0, -0.757, 0.587), (1, 0.0, 0.757, 0.587)]


def boys(nmax, T):
    """F_n(T) for n = 0..nmax: Gamma(n+1/2) P(n+1/2, T) / (2 T^(n+1/2)), with
    the Taylor series sum_k (-T)^k / (k! (2n+2k+1)) for small T."""
    out = np.zeros(nmax + 1)
    for n in range(nmax + 1):
        if T < 1e-3:
            s, term, k = 0.0, 1.0, 0
            while True:
                add = term / (2 * n + 2 * k + 1)
                s += add
                k += 1
                term *= -T / k
                if abs(add) < 1e-18:
                    break
            out[n] = s
        else:
            a = n + 0.5
            out[n] = gamma(a) * gammainc(a, T) / (2.0 * T ** a)
    return out


def hermite_E(la, lb, Xab, a, b):
    """E[i][j][t] for i <= la, j <= lb (Helgaker eqs. 9.5.6-9.5.7)."""
    p = a + b
    mu = a * b / p
    XPA = -b / p * Xab
    XPB = a / p * Xab
    E = np.zeros((la + 1, lb + 1, la + lb + 2))
    E[0, 0, 0] = math.exp(-mu * Xab * Xab)
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = XPA * E[i - 1, j, t]
                    if t > 0:
                        v += E[i - 1, j, t - 1] / (2 * p)
                    v += (t + 1) * E[i - 1, j, t + 1]
                else:
                    v = XPB * E[i, j - 1, t]
                    if t > 0:
                        v += E[i, j - 1, t - 1] / (2 * p)
                    v += (t + 1) * E[i, j - 1, t + 1]
                E[i, j, t] = v
    return E


def hermite_R(tmax, p, PC):
    """R[t][u][v] = R^0_tuv for t+u+v <= tmax (Helgaker eqs. 9.9.18-9.9.20)."""
    T = p * float(PC @ PC)
    F = boys(tmax, T)
    n = tmax + 1
    Rn = np.zeros((n + 1, n, n, n))
    for m in range(tmax + 1):
        Rn[m, 0, 0, 0] = (-2.0 * p) ** m * F[m]
    X, Y, Z = PC
    for s in range(1, tmax + 1):
        for m in range(tmax - s + 1):
            for t in range(s + 1):
                for u in range(s - t + 1):
                    v = s - t - u
                    if t > 0:
                        val = X * Rn[m + 1, t - 1, u, v]
                        if t > 1:
                            val += (t - 1) * Rn[m + 1, t - 2, u, v]
                    elif u > 0:
                        val = Y * Rn[m + 1, t, u - 1, v]
                        if u > 1:
                            val += (u - 1) * Rn[m + 1, t, u - 2, v]
                    else:
                        val = Z * Rn[m + 1, t, u, v - 1]
                        if v > 1:
                            val += (v - 1) * Rn[m + 1, t, u, v - 2]
                    Rn[m, t, u, v] = val
    return Rn[0]


def cart_powers(l):
    return [(a, b, l - a - b) for a in range(l, -1, -1) for b in range(l - a, -1, -1)]


def shell_transform(sh):
    """Rows: the XC build's AO functions of the shell; columns: Cartesian
    monomials x^a y^b z^c times the contracted radial part (normalised
    coefficients), with common_fac(l) applied (src/gto/gto.cpp:7-16)."""
    if sh.l == 0:
        M = np.eye(1)
    elif sh.l == 1:
        # CCA spherical p order is (y, z, x); Cartesian order is (x, y, z)
        M = np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]], float) if sh.pure else np.eye(3)
    else:
        raise ValueError("only s and p shells are needed for 3-21G")
    fac = {0: 0.282094791773878143, 1: 0.488602511902919921}[sh.l]
    return fac * M


def one_electron(basis, nuclei):
    shells = basis.shells
    fb = basis.first_bf()
    nbf = basis.nbf
    S = np.zeros((nbf, nbf))
    Tk = np.zeros((nbf, nbf))
    V = np.zeros((nbf, nbf))
    for I, A in enumerate(shells):
        for J, B in enumerate(shells):
            ca, cb = cart_powers(A.l), cart_powers(B.l)
            s_c = np.zeros((len(ca), len(cb)))
            t_c = np.zeros_like(s_c)
            v_c = np.zeros_like(s_c)
            for a, da in zip(A.exponents, A.coefficients):
                for b, db in zip(B.exponents, B.coefficients):
                    p = a + b
                    P = (a * A.origin + b * B.origin) / p
                    Ex = [hermite_E(A.l, B.l + 2, A.origin[k] - B.origin[k], a, b)
                          for k in range(3)]
                    pref = da * db
                    Rs = [hermite_R(A.l + B.l, p, P - np.array(C[1:])) for C in nuclei]
                    for u, pa in enumerate(ca):
                        for v, pb in enumerate(cb):
                            s1 = [Ex[k][pa[k], pb[k], 0] * math.sqrt(math.pi / p)
                                  for k in range(3)]
                            s_c[u, v] += pref * s1[0] * s1[1] * s1[2]
                            # kinetic: -1/2 <a| d^2/dk^2 |b> per direction
                            tk = 0.0
                            for k in range(3):
                                j = pb[k]
                                d2 = 4 * b * b * Ex[k][pa[k], j + 2, 0]
                                d2 -= 2 * b * (2 * j + 1) * Ex[k][pa[k], j, 0]
                                if j >= 2:
                                    d2 += j * (j - 1) * Ex[k][pa[k], j - 2, 0]
                                d2 *= math.sqrt(math.pi / p)
                                other = [s1[m] for m in range(3) if m != k]
                                tk += -0.5 * d2 * other[0] * other[1]
                            t_c[u, v] += pref * tk
                            vv = 0.0
                            for (Z, *_), R in zip(nuclei, Rs):
                                acc = 0.0
                                for t in range(pa[0] + pb[0] + 1):
                                    for uu in range(pa[1] + pb[1] + 1):
                                        for w in range(pa[2] + pb[2] + 1):
                                            acc += (Ex[0][pa[0], pb[0], t] * Ex[1][pa[1], pb[1], uu]
                                                    * Ex[2][pa[2], pb[2], w] * R[t, uu, w])
                                vv += -Z * acc
                            v_c[u, v] += pref * 2 * math.pi / p * vv
            MA, MB = shell_transform(A), shell_transform(B)
            sl = slice(fb[I], fb[I] + A.size())
            sm = slice(fb[J], fb[J] + B.size())
            S[sl, sm] = MA @ s_c @ MB.T
            Tk[sl, sm] = MA @ t_c @ MB.T
            V[sl, sm] = MA @ v_c @ MB.T
    return S, Tk, V


def two_electron(basis):
    shells = basis.shells
    fb = basis.first_bf()
    nbf = basis.nbf
    eri = np.zeros((nbf, nbf, nbf, nbf))
    ns = len(shells)
    done = {}
    for I in range(ns):
        for J in range(I + 1):
            for K in range(ns):
                for L in range(K + 1):
                    if (K, L) > (I, J):
                        continue
                    blk = _eri_shell_quartet(shells[I], shells[J], shells[K], shells[L])
                    done[(I, J, K, L)] = blk
    for (I, J, K, L), blk in done.items():
        si = slice(fb[I], fb[I] + shells[I].size())
        sj = slice(fb[J], fb[J] + shells[J].size())
        sk = slice(fb[K], fb[K] + shells[K].size())
        sl = slice(fb[L], fb[L] + shells[L].size())
        for (a, b, c, d, x) in [(si, sj, sk, sl, blk),
                                (sj, si, sk, sl, blk.transpose(1, 0, 2, 3)),
                                (si, sj, sl, sk, blk.transpose(0, 1, 3, 2)),
                                (sj, si, sl, sk, blk.transpose(1, 0, 3, 2))]:
            eri[a, b, c, d] = x
            eri[c, d, a, b] = x.transpose(2, 3, 0, 1)
    return eri


def _eri_shell_quartet(A, B, C, D):
    ca, cb, cc, cd = (cart_powers(s.l) for s in (A, B, C, D))
    out = np.zeros((len(ca), len(cb), len(cc), len(cd)))
    for a, da in zip(A.exponents, A.coefficients):
        for b, db in zip(B.exponents, B.coefficients):
            p = a + b
            P = (a * A.origin + b * B.origin) / p
            Eab = [hermite_E(A.l, B.l, A.origin[k] - B.origin[k], a, b) for k in range(3)]
            for c, dc in zip(C.exponents, C.coefficients):
                for d, dd in zip(D.exponents, D.coefficients):
                    q = c + d
                    Q = (c * C.origin + d * D.origin) / q
                    Ecd = [hermite_E(C.l, D.l, C.origin[k] - D.origin[k], c, d)
                           for k in range(3)]
                    alpha = p * q / (p + q)
                    R = hermite_R(A.l + B.l + C.l + D.l, alpha, P - Q)
                    pref = (da * db * dc * dd * 2 * math.pi ** 2.5
                            / (p * q * math.sqrt(p + q)))
                    for i, pa in enumerate(ca):
                        for j, pb in enumerate(cb):
                            for k, pc in enumerate(cc):
                                for m, pd in enumerate(cd):
                                    acc = 0.0
                                    for t in range(pa[0] + pb[0] + 1):
                                        e1 = Eab[0][pa[0], pb[0], t]
                                        for u in range(pa[1] + pb[1] + 1):
                                            e2 = e1 * Eab[1][pa[1], pb[1], u]
                                            for v in range(pa[2] + pb[2] + 1):
                                                e3 = e2 * Eab[2][pa[2], pb[2], v]
                                                for tau in range(pc[0] + pd[0] + 1):
                                                    f1 = Ecd[0][pc[0], pd[0], tau]
                                                    for nu in range(pc[1] + pd[1] + 1):
                                                        f2 = f1 * Ecd[1][pc[1], pd[1], nu]
                                                        for phi in range(pc[2] + pd[2] + 1):
                                                            f3 = f2 * Ecd[2][pc[2], pd[2], phi]
                                                            sgn = -1.0 if (tau + nu + phi) % 2 else 1.0
                                                            acc += (e3 * f3 * sgn
                                                                    * R[t + tau, u + nu, v + phi])
                                    out[i, j, k, m] += pref * acc
    MA, MB, MC, MD = (shell_transform(s) for s in (A, B, C, D))
    return np.einsum("ai,bj,ck,dl,ijkl->abcd", MA, MB, MC, MD, out)


def nuclear_repulsion(nuclei):
    e = 0.0
    for i in range(len(nuclei)):
        for j in range(i):
            r = np.linalg.norm(np.array(nuclei[i][1:]) - np.array(nuclei[j][1:]))
            e += nuclei[i][0] * nuclei[j][0] / r
    return e


def rhf(S, H, eri, nocc, tol=1e-12, maxiter=200):
    s, U = np.linalg.eigh(S)
    X = U @ np.diag(s ** -0.5) @ U.T
    e, Cp = np.linalg.eigh(X.T @ H @ X)
    C = X @ Cp
    D = C[:, :nocc] @ C[:, :nocc].T
    focks, errs = [], []
    for it in range(maxiter):
        J = np.einsum("ijkl,kl->ij", eri, D)
        K = np.einsum("ikjl,kl->ij", eri, D)
        F = H + 2 * J - K
        err = F @ D @ S - S @ D @ F
        en = float(np.sum(D * (H + F)))
        gnorm = float(np.abs(err).max())
        if gnorm < tol:
            return D, en, F, it, gnorm
        focks.append(F)
        errs.append(X.T @ err @ X)
        focks, errs = focks[-8:], errs[-8:]
        n = len(focks)
        Bm = -np.ones((n + 1, n + 1))
        Bm[n, n] = 0.0
        for i in range(n):
            for j in range(n):
                Bm[i, j] = float(np.sum(errs[i] * errs[j]))
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        coef = np.linalg.lstsq(Bm, rhs, rcond=None)[0][:n]
        Fd = sum(c * f for c, f in zip(coef, focks))
        e, Cp = np.linalg.eigh(X.T @ Fd @ X)
        C = X @ Cp
        D = C[:, :nocc] @ C[:, :nocc].T
    raise RuntimeError("SCF did not converge")


def water_atoms():
    return [(z, x * ANGSTROM_TO_BOHR, y * ANGSTROM_TO_BOHR, w * ANGSTROM_TO_BOHR)
            for z, x, y, w in WATER_ANGSTROM]


def main():
    atoms = water_atoms()
    basis = AOBasis.load(atoms, "3-21g", pure=True)
    S, T, V = one_electron(basis, atoms)
    import gaussian_integrals
    S_check = gaussian_integrals.overlap_matrix(basis)
    assert np.abs(S - S_check).max() < 1e-13, np.abs(S - S_check).max()
    eri = two_electron(basis)
    H = T + V
    D, e_elec, F, iters, gnorm = rhf(S, H, eri, nocc=5)
    e_nuc = nuclear_repulsion(atoms)
    ntr = float(np.sum(2 * D * S))
    ekin = float(np.sum(2 * D * T))
    etot = e_elec + e_nuc
    virial = -(etot - ekin) / ekin
    print(f"nbf={basis.nbf} iterations={iters} |FDS-SDF|={gnorm:.2e}")
    print(f"E(RHF)={etot:.12f}  E_nuc={e_nuc:.12f}  tr(2DS)={ntr:.14f}  -V/T={virial:.6f}")
    assert abs(ntr - 10.0) < 1e-12
    assert abs(virial - 2.0) < 0.02
    np.savez(OUT, D=D, S=S, T=T, V=V, energy=etot, atoms=np.array(atoms, dtype=np.float64))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
