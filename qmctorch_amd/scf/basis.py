"""Gaussian basis sets in the layout the QMCTorch AO layer consumes.

The reference builds its basis namespace from pyscf
(qmctorch/scf/calculator/pyscf.py:104-252).  pyscf is not available, so this
module reads published basis sets (NWChem format, ``basis_data/*.nw``) and
reproduces the conventions of that function:

* shells of one atom ordered by angular momentum, primitives by decreasing
  exponent (pyscf's basis parser / ``make_bas_env``);
* contraction coefficients normalised so that each contracted radial function
  is normalised with radially normalised primitives (pyscf
  ``_nomalize_contracted_ao`` followed by ``bas_ctr_coeff``);
* one entry per (shell, contraction column, Cartesian component, primitive),
  Cartesian component order xx, xy, xz, yy, yz, zz (pyscf.py:114-116);
* contraction column ``ictr`` scaled by ``1/(ictr+1)`` (pyscf.py:169-173);
* ``radial_type='gto_pure'``, ``harmonics_type='cart'`` (pyscf.py:125-126),
  ``nao`` = number of Cartesian functions, ``nmo`` = number of spherical ones.
"""
import os
from types import SimpleNamespace
from typing import Dict, List, Tuple

import numpy as np

_DATA = os.path.join(os.path.dirname(__file__), "basis_data")

# Cartesian exponents per angular momentum, pyscf ordering (pyscf.py:114-116)
CART_KX = {0: [0], 1: [1, 0, 0], 2: [2, 1, 1, 0, 0, 0]}
CART_KY = {0: [0], 1: [0, 1, 0], 2: [0, 1, 0, 2, 1, 0]}
CART_KZ = {0: [0], 1: [0, 0, 1], 2: [0, 0, 1, 0, 1, 2]}
SPH_M = {0: [0], 1: [-1, 0, 1], 2: [-2, -1, 0, 1, 2]}

_L = {"S": 0, "P": 1, "D": 2}

ALIASES = {
    "sto-3g": "sto-3g",
    "sto3g": "sto-3g",
    "6-31g": "6-31g",
    "631g": "6-31g",
    "cc-pvdz": "cc-pvdz",
    "ccpvdz": "cc-pvdz",
}


def available_basis() -> List[str]:
    return sorted(set(ALIASES.values()))


def _parse_nwchem(path: str) -> Dict[str, List[Tuple[int, np.ndarray, np.ndarray]]]:
    """Parse an NWChem-format file into {element: [(l, exps, coeffs[nprim, nctr])]}."""
    out: Dict[str, List] = {}
    cur = None
    with open(path) as f:
        for raw in f:
            line = raw.split("#")[0].strip()
            if not line or line.upper() == "END":
                continue
            tok = line.split()
            if tok[1].upper() in ("S", "P", "D", "SP"):
                cur = [tok[0], tok[1].upper(), []]
                out.setdefault(tok[0], []).append(cur)
            else:
                cur[2].append([float(t) for t in tok])
    shells: Dict[str, List] = {}
    for el, lst in out.items():
        sh = []
        for _, lab, rows in lst:
            rows = np.asarray(rows, dtype=np.float64)
            if lab == "SP":
                sh.append((0, rows[:, 0].copy(), rows[:, 1:2].copy()))
                sh.append((1, rows[:, 0].copy(), rows[:, 2:3].copy()))
            else:
                sh.append((_L[lab], rows[:, 0].copy(), rows[:, 1:].copy()))
        shells[el] = sh
    return shells


def gaussian_int(n: int, alpha) -> np.ndarray:
    """int_0^inf r^n exp(-alpha r^2) dr = Gamma((n+1)/2) / (2 alpha^((n+1)/2))."""
    from math import gamma

    return gamma((n + 1) * 0.5) / (2.0 * np.asarray(alpha) ** ((n + 1) * 0.5))


def gto_norm(l: int, alpha) -> np.ndarray:
    """Radial normalisation of r^l exp(-alpha r^2)."""
    return 1.0 / np.sqrt(gaussian_int(2 * l + 2, 2 * np.asarray(alpha)))


def normalized_contraction(l: int, exps: np.ndarray, coeffs: np.ndarray) -> np.ndarray:
    """Contraction coefficients on radially normalised primitives, with every
    contracted column normalised to one (pyscf make_bas_env + bas_ctr_coeff)."""
    n = gto_norm(l, exps)
    cs = coeffs * n[:, None]
    ee = gaussian_int(2 * l + 2, exps[:, None] + exps[None, :])
    s1 = 1.0 / np.sqrt(np.einsum("pi,pq,qi->i", cs, ee, cs))
    return coeffs * s1[None, :]


def load_shells(basis_name: str, element: str):
    key = ALIASES.get(basis_name.lower().replace("_", "-"))
    if key is None:
        raise ValueError(
            "basis %s not available offline; available: %s" % (basis_name, available_basis())
        )
    shells = _parse_nwchem(os.path.join(_DATA, key + ".nw"))
    if element not in shells:
        raise ValueError("element %s not in basis %s" % (element, key))
    sh = shells[element]
    # pyscf orders the shells of an atom by angular momentum (stable)
    sh = sorted(sh, key=lambda s: s[0])
    res = []
    for l, e, c in sh:
        order = np.argsort(-e, kind="stable")
        e, c = e[order], c[order]
        res.append((l, e, normalized_contraction(l, e, c)))
    return res


def build_basis(atoms: List[str], atom_coords, basis_name: str) -> SimpleNamespace:
    """Assemble the basis namespace of pyscf.py:104-252 (without the MOs)."""
    bas_coeff, bas_exp, index_ctr = [], [], []
    bas_n, bas_l, bas_m, bas_kx, bas_ky, bas_kz = [], [], [], [], [], []
    bas_norm = []
    nshells = [0] * len(atoms)
    nmo = 0
    iao = 0
    shell_info = []  # (atom, l, ictr, first_ao) per contracted shell column
    for iat, el in enumerate(atoms):
        for l, exps, coeffs in load_shells(basis_name, el):
            nprim, nctr = coeffs.shape
            ncart = len(CART_KX[l])
            for ictr in range(nctr):
                c = coeffs[:, ictr] / (ictr + 1)
                bas_coeff += c.tolist() * ncart
                bas_exp += exps.tolist() * ncart
                bas_n += [l + 1] * nprim * ncart
                bas_l += [l] * nprim * ncart
                bas_norm += gto_norm(l, exps).tolist() * ncart
                nshells[iat] += nprim * ncart
                shell_info.append((iat, l, ictr, iao))
                for _ in range(ncart):
                    index_ctr += [iao] * nprim
                    iao += 1
                for m in SPH_M[l]:
                    bas_m += [m] * nprim
                for k in CART_KX[l]:
                    bas_kx += [k] * nprim
                for k in CART_KY[l]:
                    bas_ky += [k] * nprim
                for k in CART_KZ[l]:
                    bas_kz += [k] * nprim
                nmo += 2 * l + 1

    b = SimpleNamespace()
    b.radial_type = "gto_pure"
    b.harmonics_type = "cart"
    b.nao = iao
    b.nmo = nmo
    b.nshells = np.asarray(nshells, dtype=np.int64)
    b.index_ctr = np.asarray(index_ctr, dtype=np.int64)
    intervals = np.concatenate(([0], np.cumsum(nshells)))
    b.nao_per_atom = np.asarray(
        [len(np.unique(b.index_ctr[s:e])) for s, e in zip(intervals[:-1], intervals[1:])],
        dtype=np.int64,
    )
    _, counts = np.unique(b.index_ctr, return_counts=True)
    b.nctr_per_ao = counts.astype(np.int64)
    b.bas_coeffs = np.asarray(bas_coeff, dtype=np.float64)
    b.bas_exp = np.asarray(bas_exp, dtype=np.float64)
    b.bas_norm = np.asarray(bas_norm, dtype=np.float64)
    b.bas_n = bas_n
    b.bas_l = bas_l
    b.bas_m = bas_m
    b.bas_kr = np.zeros_like(b.bas_exp)
    b.bas_kx = np.asarray(bas_kx, dtype=np.int64)
    b.bas_ky = np.asarray(bas_ky, dtype=np.int64)
    b.bas_kz = np.asarray(bas_kz, dtype=np.int64)
    b.atom_coords_internal = np.asarray(atom_coords, dtype=np.float64)
    b.shell_info = shell_info
    b.TotalEnergy = 0.0
    return b


def random_orthonormal_mos(nao: int, nmo: int, seed: int = 0) -> np.ndarray:
    """Seeded random orthonormal (QR of N(0,1)) MO coefficients, column-normalised."""
    rng = np.random.default_rng(seed)
    q, r = np.linalg.qr(rng.standard_normal((nao, nmo)))
    q = q * np.sign(np.diag(r))[None, :]
    return normalize_columns(q)


def normalize_columns(mat: np.ndarray) -> np.ndarray:
    """Column normalisation of calculator_base.py:36-46."""
    return mat / np.sqrt((mat ** 2).sum(0))
