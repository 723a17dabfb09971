"""Restricted Hartree-Fock: ``rhf`` for s, p, d Gaussian bases (integrals from
integrals.py) and ``rhf_s``, the closed-form s-only version kept as an
independent check.

Original notes of the s-only solver:

pyscf (used by the reference's calculator, qmctorch/scf/calculator/pyscf.py:78-101)
is unavailable offline.  For s-only bases (e.g. H2/STO-3G) the one- and
two-electron integrals have the closed forms of Szabo & Ostlund, "Modern
Quantum Chemistry", Appendix A (A.9, A.11, A.33, A.41) and this module solves
the Roothaan equations with them.  The orbitals are expressed on the contracted
functions of the QMCTorch AO layer, i.e. sum_p c_p (2 a_p / pi)^(3/4) exp(-a_p r^2),
and are column-normalised like calculator_base.py:36-46.
"""
from math import erf, pi, sqrt

import numpy as np


def _f0(t: float) -> float:
    if t < 1e-12:
        return 1.0 - t / 3.0
    return 0.5 * sqrt(pi / t) * erf(sqrt(t))


def _contracted(basis, atom_coords):
    """[(center, exps, coeffs_with_norm)] per contracted AO."""
    index_ctr = np.asarray(basis.index_ctr)
    coords = np.asarray(atom_coords, dtype=np.float64)
    bas_atom = np.repeat(np.arange(len(basis.nshells)), basis.nshells)
    out = []
    for mu in range(basis.nao):
        sel = np.where(index_ctr == mu)[0]
        a = np.asarray(basis.bas_exp)[sel]
        c = np.asarray(basis.bas_coeffs)[sel] * (2 * a / pi) ** 0.75
        out.append((coords[bas_atom[sel[0]]], a, c))
    return out


def s_integrals(basis, atom_coords, charges):
    if np.any(np.asarray(basis.bas_kx) + np.asarray(basis.bas_ky) + np.asarray(basis.bas_kz)):
        raise ValueError("s-only RHF called on a basis with l>0 functions")
    fns = _contracted(basis, atom_coords)
    n = len(fns)
    S = np.zeros((n, n))
    T = np.zeros((n, n))
    V = np.zeros((n, n))
    coords = np.asarray(atom_coords, dtype=np.float64)
    for i, (A, ai, ci) in enumerate(fns):
        for j, (B, aj, cj) in enumerate(fns):
            ab2 = float(((A - B) ** 2).sum())
            for a, ca in zip(ai, ci):
                for b, cb in zip(aj, cj):
                    p = a + b
                    k = np.exp(-a * b / p * ab2)
                    s = (pi / p) ** 1.5 * k
                    S[i, j] += ca * cb * s
                    T[i, j] += ca * cb * a * b / p * (3.0 - 2.0 * a * b / p * ab2) * s
                    P = (a * A + b * B) / p
                    for C, Z in zip(coords, charges):
                        pc2 = float(((P - C) ** 2).sum())
                        V[i, j] += ca * cb * (-2.0 * pi / p * Z * k * _f0(p * pc2))
    eri = np.zeros((n, n, n, n))
    for i, (A, ai, ci) in enumerate(fns):
        for j, (B, aj, cj) in enumerate(fns):
            ab2 = float(((A - B) ** 2).sum())
            for k_, (C, ak, ck) in enumerate(fns):
                for l_, (D, al, cl) in enumerate(fns):
                    cd2 = float(((C - D) ** 2).sum())
                    val = 0.0
                    for a, ca in zip(ai, ci):
                        for b, cb in zip(aj, cj):
                            p = a + b
                            P = (a * A + b * B) / p
                            kab = np.exp(-a * b / p * ab2)
                            for c, cc in zip(ak, ck):
                                for d, cd in zip(al, cl):
                                    q = c + d
                                    Q = (c * C + d * D) / q
                                    kcd = np.exp(-c * d / q * cd2)
                                    pq2 = float(((P - Q) ** 2).sum())
                                    val += (
                                        ca * cb * cc * cd * 2.0 * pi ** 2.5
                                        / (p * q * sqrt(p + q)) * kab * kcd
                                        * _f0(p * q / (p + q) * pq2)
                                    )
                    eri[i, j, k_, l_] = val
    return S, T, V, eri


def rhf_s(basis, atom_coords, charges, nelec: int, maxiter: int = 100, tol: float = 1e-12):
    """Closed-shell SCF; returns (E_total, mo_coeff (nao x nao), column-normalised)."""
    S, T, V, eri = s_integrals(basis, atom_coords, charges)
    H = T + V
    w, U = np.linalg.eigh(S)
    X = U @ np.diag(w ** -0.5) @ U.T
    nocc = nelec // 2
    P = np.zeros_like(S)
    e_old = 0.0
    coords = np.asarray(atom_coords, dtype=np.float64)
    vnn = 0.0
    for a in range(len(charges)):
        for b in range(a + 1, len(charges)):
            vnn += charges[a] * charges[b] / np.sqrt(((coords[a] - coords[b]) ** 2).sum())
    for _ in range(maxiter):
        J = np.einsum("ijkl,kl->ij", eri, P)
        K = np.einsum("ikjl,kl->ij", eri, P)
        F = H + J - 0.5 * K
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        C = X @ Cp
        P = 2.0 * C[:, :nocc] @ C[:, :nocc].T
        e_elec = 0.5 * float((P * (H + F)).sum())
        if abs(e_elec - e_old) < tol:
            break
        e_old = e_elec
    C = C * np.where(C[0:1, :] < 0, -1.0, 1.0)
    return e_elec + vnn, C


def nuclear_repulsion(atom_coords, charges) -> float:
    coords = np.asarray(atom_coords, dtype=np.float64)
    vnn = 0.0
    for a in range(len(charges)):
        for b in range(a + 1, len(charges)):
            vnn += charges[a] * charges[b] / np.sqrt(((coords[a] - coords[b]) ** 2).sum())
    return float(vnn)


def rhf(basis, atom_coords, charges, nelec: int, maxiter: int = 200, tol: float = 1e-11, ao_transform=None):
    """Closed-shell Roothaan-Hall SCF with DIIS (Pulay, Chem. Phys. Lett. 73, 393
    (1980)) in the AO basis of the AO layer (s, p, d Cartesian Gaussians; the
    integrals of integrals.py).  ``ao_transform`` (nao x m) restricts the
    variational space to its columns (e.g. the five pure d combinations).
    Returns (E_total, C (nao x nmo)); columns are the canonical orbitals in
    increasing orbital energy."""
    from .integrals import ao_integrals

    if nelec % 2:
        raise ValueError("closed-shell RHF needs an even electron count")
    S, T, V, eri = ao_integrals(basis, atom_coords, charges)
    H = T + V
    X0 = np.eye(S.shape[0]) if ao_transform is None else np.asarray(ao_transform, dtype=np.float64)
    Sx = X0.T @ S @ X0
    w, U = np.linalg.eigh(Sx)
    keep = w > 1e-10 * w.max()
    X = X0 @ (U[:, keep] / np.sqrt(w[keep])[None, :])
    nocc = nelec // 2
    e, Cp = np.linalg.eigh(X.T @ H @ X)
    C = X @ Cp
    P = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    e_old = None
    focks, errs = [], []
    for _ in range(maxiter):
        J = np.einsum("ijkl,kl->ij", eri, P, optimize=True)
        K = np.einsum("ikjl,kl->ij", eri, P, optimize=True)
        F = H + J - 0.5 * K
        e_elec = 0.5 * float((P * (H + F)).sum())
        err = X.T @ (F @ P @ S - S @ P @ F) @ X
        focks.append(F)
        errs.append(err)
        if len(focks) > 8:
            focks.pop(0)
            errs.pop(0)
        if e_old is not None and abs(e_elec - e_old) < tol and np.abs(err).max() < 1e-8:
            break
        e_old = e_elec
        if len(focks) >= 2:
            m = len(focks)
            B = -np.ones((m + 1, m + 1))
            B[m, m] = 0.0
            for i in range(m):
                for j in range(m):
                    B[i, j] = float((errs[i] * errs[j]).sum())
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            try:
                c = np.linalg.solve(B, rhs)[:m]
                F = sum(ci * Fi for ci, Fi in zip(c, focks))
            except np.linalg.LinAlgError:
                pass
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        C = X @ Cp
        P = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    else:
        raise RuntimeError("RHF did not converge in %d iterations" % maxiter)
    C = C * np.where(C[np.argmax(np.abs(C), axis=0), np.arange(C.shape[1])] < 0, -1.0, 1.0)[None, :]
    return e_elec + nuclear_repulsion(atom_coords, charges), C


def pure_d_transform(basis) -> np.ndarray:
    """Columns spanning the AO space with each Cartesian d shell reduced to its
    five pure (l = 2) combinations xy, yz, xz, xx - yy, 2zz - xx - yy (the
    variational space of an SCF run with spherical d functions, pyscf's default
    for the bases used here); other AOs pass through.  nao x nmo."""
    nao = int(basis.nao)
    cols = []
    d_aos = set()
    for (_, l, _, a0) in getattr(basis, "shell_info", []):
        if l == 2:
            xx, xy, xz, yy, yz, zz = range(a0, a0 + 6)
            d_aos.update(range(a0, a0 + 6))
            for v in ({xy: 1.0}, {yz: 1.0}, {xz: 1.0}, {xx: 1.0, yy: -1.0}, {zz: 2.0, xx: -1.0, yy: -1.0}):
                c = np.zeros(nao)
                for k, w in v.items():
                    c[k] = w
                cols.append(c)
    for mu in range(nao):
        if mu not in d_aos:
            c = np.zeros(nao)
            c[mu] = 1.0
            cols.append(c)
    return np.stack(cols, 1)
