"""The trace-trick identities of the excitation kernels, pinned against
explicit determinants (slater_pooling.py:147-260, 411-825).

With A the ground-state block (columns 0..n-1), V the virtual columns, B the
kinetic rows, M = A^-1 V, X = A^-1 B and Mk = X_virt - X_occ M, a configuration
that replaces columns h_i by p_i in place has D/D0 = det(M[h, p]) and
tr(A'^-1 B') = tr(A^-1 B_occ) + tr(M[h, p]^-1 Mk[h, p]); listing the columns in
another order multiplies D by the permutation's sign.  The table rule
(holes ascending, particles ascending, sign from the inversions of the listed
order against the in-place order) is the one qmc_plan_create applies.  A wrong
hole/particle pairing, a dropped X_occ M term or a lost permutation sign each
fail here.
"""
import itertools

import numpy as np


def excitation_entry(conf, n):
    ground = list(range(n))
    holes = sorted(set(ground) - set(conf))
    parts = sorted(set(conf) - set(ground))
    inplace = ground.copy()
    for h, p in zip(holes, parts):
        inplace[h] = p
    perm = [inplace.index(c) for c in conf]
    inv = sum(1 for i in range(n) for j in range(i + 1, n) if perm[i] > perm[j])
    return holes, [p - n for p in parts], (-1.0) ** inv


def _ratio_and_trace(M, Mk, t0, holes, parts):
    if not holes:
        return 1.0, t0
    if len(holes) == 1:
        m = M[holes[0], parts[0]]
        return m, t0 + Mk[holes[0], parts[0]] / m
    T = M[np.ix_(holes, parts)]
    K = Mk[np.ix_(holes, parts)]
    return np.linalg.det(T), t0 + np.trace(np.linalg.solve(T, K))


def _configs(n, nmo, rng):
    out = [list(range(n))]
    for h in range(n):
        for p in range(n, nmo):
            c = [i for i in range(n) if i != h] + [p]
            out.append(list(rng.permutation(c)))
    for h1, h2 in itertools.combinations(range(n), 2):
        for p1, p2 in itertools.combinations(range(n, nmo), 2):
            c = [i for i in range(n) if i not in (h1, h2)] + [p2, p1]
            out.append(list(rng.permutation(c)))
    return out


def test_trace_trick_matches_explicit_determinants():
    rng = np.random.default_rng(5)
    for n, nmo in ((1, 5), (2, 6), (5, 9)):
        A_full = rng.standard_normal((n, nmo))
        B_full = rng.standard_normal((n, nmo))
        A = A_full[:, :n]
        d0 = np.linalg.det(A)
        M = np.linalg.solve(A, A_full[:, n:])
        X = np.linalg.solve(A, B_full)
        Mk = X[:, n:] - X[:, :n] @ M
        t0 = np.trace(X[:, :n])
        for conf in _configs(n, nmo, rng):
            holes, parts, sign = excitation_entry(conf, n)
            R, tr = _ratio_and_trace(M, Mk, t0, holes, parts)
            Ac, Bc = A_full[:, conf], B_full[:, conf]
            d = np.linalg.det(Ac)
            assert abs(sign * d0 * R - d) <= 1e-12 * abs(d) + 1e-14
            assert abs(tr - np.trace(np.linalg.solve(Ac, Bc))) <= 1e-11 * max(1.0, abs(tr))
