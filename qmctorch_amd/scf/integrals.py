"""One- and two-electron integrals over the Cartesian Gaussian AOs of the basis.

McMurchie-Davidson scheme (L. E. McMurchie, E. R. Davidson, J. Comput. Phys.
26, 218 (1978); recurrences as written in Helgaker, Jorgensen, Olsen,
"Molecular Electronic-Structure Theory", ch. 9): Hermite expansion
coefficients E_t^{ij} of Gaussian products, Hermite Coulomb integrals
R_{tuv}^n from the Boys function, and the kinetic energy from shifted
overlaps.  Everything is evaluated on primitive Cartesian shells, vectorised
over all shell pairs / quartets of one angular-momentum class, and then
contracted with the AO layer's own coefficients
(AO_mu = sum_p norm_cst_p * bas_coeffs_p * x^kx y^ky z^kz exp(-a_p r^2),
atomic_orbitals.py:236-249, 654-669), so the SCF runs in exactly the AO basis
the wavefunction evaluates.
"""
from math import pi
from typing import Dict, List, Tuple

import numpy as np
from scipy.special import gamma as _gamma
from scipy.special import gammainc

CART = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)],
}


def boys(nmax: int, T) -> np.ndarray:
    """F_n(T) = int_0^1 t^(2n) exp(-T t^2) dt for n = 0..nmax; shape (nmax+1,)+T.shape.
    Series sum_k (-T)^k / (k! (2n+2k+1)) for T < 1, otherwise
    Gamma(n+1/2) P(n+1/2, T) / (2 T^(n+1/2)) with the regularised incomplete gamma."""
    T = np.asarray(T, dtype=np.float64)
    out = np.empty((nmax + 1,) + T.shape)
    small = T < 1.0
    Ts = T[small]
    Tl = T[~small]
    for n in range(nmax + 1):
        if Ts.size:
            term = np.ones_like(Ts)
            acc = term / (2 * n + 1)
            for k in range(1, 30):
                term = term * (-Ts) / k
                acc = acc + term / (2 * n + 2 * k + 1)
            out[n][small] = acc
        if Tl.size:
            a = n + 0.5
            out[n][~small] = _gamma(a) * gammainc(a, Tl) / (2.0 * Tl ** a)
    return out


def hermite_E(la: int, lb: int, a, b, qx) -> List[List[List[np.ndarray]]]:
    """E[i][j][t] (i <= la, j <= lb, t <= i+j) for one Cartesian axis.
    a, b exponents, qx = A_x - B_x (broadcasting arrays)."""
    p = a + b
    mu = a * b / p
    xpa = -b * qx / p
    xpb = a * qx / p
    zero = np.zeros(np.broadcast(a, b, qx).shape)
    E = [[[zero] * (la + lb + 2) for _ in range(lb + 1)] for _ in range(la + 1)]
    E[0][0] = [zero] * (la + lb + 2)
    E[0][0][0] = np.exp(-mu * qx * qx) + zero
    inv2p = 0.5 / p

    def get(lst, t):
        return lst[t] if 0 <= t < len(lst) else zero

    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            if i > 0:
                prev = E[i - 1][j]
                new = [inv2p * get(prev, t - 1) + xpa * get(prev, t) + (t + 1) * get(prev, t + 1)
                       for t in range(la + lb + 2)]
            else:
                prev = E[i][j - 1]
                new = [inv2p * get(prev, t - 1) + xpb * get(prev, t) + (t + 1) * get(prev, t + 1)
                       for t in range(la + lb + 2)]
            E[i][j] = new
    return E


def hermite_R(L: int, alpha, X, Y, Z) -> Dict[Tuple[int, int, int], np.ndarray]:
    """R_{tuv}^0 for t+u+v <= L (Hermite Coulomb integrals)."""
    R2 = X * X + Y * Y + Z * Z
    F = boys(L, alpha * R2)
    memo = {}

    def r(t, u, v, n):
        key = (t, u, v, n)
        if key in memo:
            return memo[key]
        if t < 0 or u < 0 or v < 0:
            val = 0.0
        elif t == 0 and u == 0 and v == 0:
            val = (-2.0 * alpha) ** n * F[n]
        elif t > 0:
            val = X * r(t - 1, u, v, n + 1)
            if t > 1:
                val = val + (t - 1) * r(t - 2, u, v, n + 1)
        elif u > 0:
            val = Y * r(t, u - 1, v, n + 1)
            if u > 1:
                val = val + (u - 1) * r(t, u - 2, v, n + 1)
        else:
            val = Z * r(t, u, v - 1, n + 1)
            if v > 1:
                val = val + (v - 1) * r(t, u, v - 2, n + 1)
        memo[key] = val
        return val

    return {(t, u, v): r(t, u, v, 0) for t in range(L + 1) for u in range(L + 1 - t) for v in range(L + 1 - t - u)}


class PrimitiveShells:
    """Primitive Cartesian shells of a basis namespace and the matrix that
    contracts them into the AO layer's AOs."""

    def __init__(self, basis, atom_coords, norm_cst=None):
        from ..wavefunction.plan import atomic_orbital_norm

        coords = np.asarray(atom_coords, dtype=np.float64)
        nsh = np.asarray(basis.nshells)
        atom_of = np.repeat(np.arange(len(nsh)), nsh)
        kx, ky, kz = (np.asarray(v, dtype=np.int64) for v in (basis.bas_kx, basis.bas_ky, basis.bas_kz))
        lval = kx + ky + kz
        if lval.max(initial=0) > 2:
            raise ValueError("integrals: s, p and d functions only")
        if getattr(basis, "radial_type", "gto_pure") != "gto_pure":
            raise ValueError("integrals: Gaussian (gto_pure) bases only")
        ex = np.asarray(basis.bas_exp, dtype=np.float64)
        norm = np.asarray(atomic_orbital_norm(basis) if norm_cst is None else norm_cst, dtype=np.float64)
        K = norm * np.asarray(basis.bas_coeffs, dtype=np.float64)
        idx = np.asarray(basis.index_ctr)
        self.shells: List[Tuple[int, float, int]] = []  # (atom, exponent, l)
        key_to_shell: Dict[Tuple[int, float, int], int] = {}
        for e in range(len(ex)):
            key = (int(atom_of[e]), float(ex[e]), int(lval[e]))
            if key not in key_to_shell:
                key_to_shell[key] = len(self.shells)
                self.shells.append(key)
        self.offset = []
        n = 0
        for (_, _, l) in self.shells:
            self.offset.append(n)
            n += len(CART[l])
        self.nprim_cart = n
        self.centers = coords
        self.C = np.zeros((n, int(basis.nao)))
        for e in range(len(ex)):
            s = key_to_shell[(int(atom_of[e]), float(ex[e]), int(lval[e]))]
            comp = CART[int(lval[e])].index((int(kx[e]), int(ky[e]), int(kz[e])))
            self.C[self.offset[s] + comp, idx[e]] += K[e]

    def by_l(self, l: int) -> List[int]:
        return [i for i, s in enumerate(self.shells) if s[2] == l]

    def _pairs(self, la: int, lb: int, lpad: int = 0):
        ia, ib = self.by_l(la), self.by_l(lb)
        if not ia or not ib:
            return None
        I, J = np.meshgrid(ia, ib, indexing="ij")
        I, J = I.ravel(), J.ravel()
        a = np.array([self.shells[i][1] for i in I])
        b = np.array([self.shells[j][1] for j in J])
        A = self.centers[[self.shells[i][0] for i in I]]
        B = self.centers[[self.shells[j][0] for j in J]]
        p = a + b
        P = (a[:, None] * A + b[:, None] * B) / p[:, None]
        E = [hermite_E(la, lb + lpad, a, b, A[:, d] - B[:, d]) for d in range(3)]
        return I, J, a, b, p, P, E


def one_electron(ps: PrimitiveShells, charges) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Overlap, kinetic and nuclear-attraction matrices over primitive Cartesian functions."""
    n = ps.nprim_cart
    S = np.zeros((n, n))
    T = np.zeros((n, n))
    V = np.zeros((n, n))
    charges = np.asarray(charges, dtype=np.float64)
    for la in range(3):
        for lb in range(3):
            pr = ps._pairs(la, lb, lpad=2)
            if pr is None:
                continue
            I, J, a, b, p, P, E = pr
            sq = np.sqrt(pi / p)

            def s1(d, i, j):
                return E[d][i][j][0] * sq if j >= 0 else 0.0

            L = la + lb
            Rs = [hermite_R(L, p, P[:, 0] - C[0], P[:, 1] - C[1], P[:, 2] - C[2]) for C in ps.centers]
            for ca, (ax, ay, az) in enumerate(CART[la]):
                for cb, (bx, by, bz) in enumerate(CART[lb]):
                    ij = ((ax, bx), (ay, by), (az, bz))
                    ov = [s1(d, i, j) for d, (i, j) in enumerate(ij)]
                    sval = ov[0] * ov[1] * ov[2]
                    kin = 0.0
                    for d, (i, j) in enumerate(ij):
                        d2 = (j * (j - 1) * s1(d, i, j - 2) if j >= 2 else 0.0) - 2.0 * b * (2 * j + 1) * s1(d, i, j) \
                            + 4.0 * b * b * s1(d, i, j + 2)
                        others = np.prod([ov[e] for e in range(3) if e != d], axis=0)
                        kin = kin - 0.5 * d2 * others
                    vv = 0.0
                    for Z, R in zip(charges, Rs):
                        acc = 0.0
                        for t in range(ax + bx + 1):
                            for u in range(ay + by + 1):
                                for v in range(az + bz + 1):
                                    acc = acc + E[0][ax][bx][t] * E[1][ay][by][u] * E[2][az][bz][v] * R[(t, u, v)]
                        vv = vv - Z * 2.0 * pi / p * acc
                    r = np.array([ps.offset[i] + ca for i in I])
                    c = np.array([ps.offset[j] + cb for j in J])
                    S[r, c] = sval
                    T[r, c] = kin
                    V[r, c] = vv
    return S, T, V


def _hermite_table(E, la, lb):
    """Per Cartesian component pair: list of (t, u, v, coefficient array)."""
    out = []
    for (ax, ay, az) in CART[la]:
        for (bx, by, bz) in CART[lb]:
            terms = []
            for t in range(ax + bx + 1):
                for u in range(ay + by + 1):
                    for v in range(az + bz + 1):
                        terms.append((t, u, v, E[0][ax][bx][t] * E[1][ay][by][u] * E[2][az][bz][v]))
            out.append(terms)
    return out


def two_electron(ps: PrimitiveShells, chunk: int = 200000) -> np.ndarray:
    """(ab|cd) over primitive Cartesian functions, chemists' order, full tensor."""
    n = ps.nprim_cart
    eri = np.zeros((n, n, n, n))
    pairs = {}
    for la in range(3):
        for lb in range(3):
            pr = ps._pairs(la, lb)
            if pr is not None:
                I, J, a, b, p, P, E = pr
                pairs[(la, lb)] = (I, J, p, P, _hermite_table(E, la, lb))
    for (la, lb), (I1, J1, p, P, H1) in pairs.items():
        for (lc, ld), (I2, J2, q, Q, H2) in pairs.items():
            L = la + lb + lc + ld
            n1, n2 = len(p), len(q)
            step = max(1, chunk // max(n2, 1))
            for s0 in range(0, n1, step):
                sl = slice(s0, min(n1, s0 + step))
                pp = p[sl][:, None]
                qq = q[None, :]
                alpha = pp * qq / (pp + qq)
                PQ = P[sl][:, None, :] - Q[None, :, :]
                R = hermite_R(L, alpha, PQ[..., 0], PQ[..., 1], PQ[..., 2])
                pref = 2.0 * pi ** 2.5 / (pp * qq * np.sqrt(pp + qq))
                for kab, t1 in enumerate(H1):
                    ca, cb = divmod(kab, len(CART[lb]))
                    rows_a = np.array([ps.offset[i] + ca for i in I1[sl]])
                    rows_b = np.array([ps.offset[j] + cb for j in J1[sl]])
                    for kcd, t2 in enumerate(H2):
                        cc, cd = divmod(kcd, len(CART[ld]))
                        acc = 0.0
                        for (t, u, v, e1) in t1:
                            inner = 0.0
                            for (tau, nu, phi, e2) in t2:
                                sgn = -1.0 if (tau + nu + phi) & 1 else 1.0
                                inner = inner + sgn * e2[None, :] * R[(t + tau, u + nu, v + phi)]
                            acc = acc + e1[sl][:, None] * inner
                        val = pref * acc
                        rows_c = np.array([ps.offset[i] + cc for i in I2])
                        rows_d = np.array([ps.offset[j] + cd for j in J2])
                        eri[rows_a[:, None], rows_b[:, None], rows_c[None, :], rows_d[None, :]] = val
    return eri


def ao_integrals(basis, atom_coords, charges, norm_cst=None):
    """S, T, V (nao x nao) and the ERI tensor (nao^4) in the AO layer's basis."""
    ps = PrimitiveShells(basis, atom_coords, norm_cst)
    S, T, V = one_electron(ps, charges)
    eri = two_electron(ps)
    C = ps.C
    S, T, V = (C.T @ M @ C for M in (S, T, V))
    eri = np.einsum("pqrs,pi->iqrs", eri, C, optimize=True)
    eri = np.einsum("iqrs,qj->ijrs", eri, C, optimize=True)
    eri = np.einsum("ijrs,rk->ijks", eri, C, optimize=True)
    eri = np.einsum("ijks,sl->ijkl", eri, C, optimize=True)
    return S, T, V, eri
