"""CPU FP64 restatement of the QMCTorch VMC hot path -- TEST INFRASTRUCTURE ONLY.

This module is the oracle against which the HIP kernels of ``qmctorch_amd`` are
checked.  It is imported only by ``tests/``, ``__graft_entry__.smoke()`` and the
baseline legs of ``bench.py``; the product never calls it.

It re-expresses, operation for operation on torch CPU tensors (float64), the
reference functions cited in each docstring (paths relative to the reference
repository root).  Parity status of every function is recorded in the table at
the bottom of this file (``PARITY``): each entry names the independent checks
(closed forms, autograd/finite differences, reference-held fixtures) that pin it
in ``tests/test_oracle_*.py``.
"""
import math
from itertools import combinations, product
from types import SimpleNamespace
from typing import List, Optional, Tuple

import numpy as np
import torch

DT = torch.float64


# --------------------------------------------------------------------------- #
# utils/algebra_utils.py, utils/torch_utils.py
# --------------------------------------------------------------------------- #
def double_factorial(n):
    """(n)!! with the reference's convention that non-positive arguments give 1
    (qmctorch/utils/algebra_utils.py:45-55 maps scipy's 0 for n<0 to 1)."""
    out = []
    for v in np.atleast_1d(np.asarray(n, dtype=np.int64)):
        r = 1
        k = int(v)
        while k > 1:
            r *= k
            k -= 2
        out.append(r)
    return np.asarray(out, dtype=np.float64)


def fast_power(x: torch.Tensor, k: torch.Tensor) -> torch.Tensor:
    """x**k for k in {0,1,2} by masks (qmctorch/utils/torch_utils.py:24-59)."""
    if int(k.max()) < 3:
        out = x.clone()
        out = out.masked_fill(k == 0, 1.0)
        if int(k.max()) > 1:
            m2 = (k == 2).expand_as(out)
            out = torch.where(m2, out * out, out)
        return out
    return x ** k


def btrace(M: torch.Tensor) -> torch.Tensor:
    """Batched trace (qmctorch/utils/algebra_utils.py:7-16)."""
    return torch.diagonal(M, dim1=-2, dim2=-1).sum(-1)


# --------------------------------------------------------------------------- #
# wavefunction/orbitals/norm_orbital.py
# --------------------------------------------------------------------------- #
def norm_gaussian_cartesian(a, b, c, exp):
    """(2a/pi)^(3/4) prod_d (4 alpha)^(k_d/2)/sqrt((2k_d-1)!!)  (norm_orbital.py:136-161)."""
    exp = np.asarray(exp, dtype=np.float64)
    pref = (2 * exp / np.pi) ** 0.75
    out = pref.copy()
    for k in (a, b, c):
        k = np.asarray(k)
        out = out * (4 * exp) ** (k / 2) / np.sqrt(double_factorial(2 * k - 1))
    return torch.as_tensor(out, dtype=DT)


def norm_slater_cartesian(a, b, c, n, exp):
    """STO Cartesian norm, l = a+b+c+n+1 (norm_orbital.py:95-133)."""
    a, b, c, n = [np.asarray(v) for v in (a, b, c, n)]
    exp = np.asarray(exp, dtype=np.float64)
    lvals = a + b + c + n + 1.0
    lfact = np.asarray([math.factorial(int(2 * i)) for i in lvals], dtype=np.float64)
    prefact = 4 * np.pi * lfact / ((2 * exp) ** (2 * lvals + 1))
    num = (
        double_factorial(2 * a.astype(int) - 1)
        * double_factorial(2 * b.astype(int) - 1)
        * double_factorial(2 * c.astype(int) - 1)
    )
    denom = double_factorial((2 * a + 2 * b + 2 * c + 1).astype(int))
    return torch.as_tensor(np.sqrt(1.0 / (prefact * num / denom)), dtype=DT)


def atomic_orbital_norm(basis) -> torch.Tensor:
    """Dispatch of norm_orbital.py:8-46 (Cartesian harmonics only)."""
    if basis.harmonics_type != "cart":
        raise ValueError("oracle covers Cartesian harmonics only")
    if basis.radial_type.startswith("sto"):
        return norm_slater_cartesian(basis.bas_kx, basis.bas_ky, basis.bas_kz, basis.bas_kr, basis.bas_exp)
    return norm_gaussian_cartesian(basis.bas_kx, basis.bas_ky, basis.bas_kz, basis.bas_exp)


# --------------------------------------------------------------------------- #
# wavefunction/orbitals/radial_functions.py and spherical_harmonics.py
# --------------------------------------------------------------------------- #
def radial_gaussian_pure(R, bas_exp, xyz, derivative):
    """exp(-a r^2) with gradient (per axis) and Laplacian (radial_functions.py:241-320).

    Returns a list with one entry per requested order: value (..,Np); gradient
    (..,Np,3); sum of the diagonal Hessian (..,Np)."""
    R2 = R * R
    er = torch.exp(-bas_exp * R2)
    out = []
    if any(d in derivative for d in (1, 2)):
        bexp_er = bas_exp * er
        nabla_er = -2 * bexp_er.unsqueeze(-1) * xyz
    for d in derivative:
        if d == 0:
            out.append(er)
        elif d == 1:
            out.append(nabla_er)
        elif d == 2:
            out.append(bas_exp * er * (4 * bas_exp * R2 - 6))
    return out


def radial_slater(R, bas_n, bas_exp, xyz, derivative):
    """r^n exp(-a r) with gradient and Laplacian (radial_functions.py:6-123)."""
    rn = fast_power(R, bas_n)
    er = torch.exp(-bas_exp * R)
    out = []
    if any(d in derivative for d in (1, 2)):
        Rnm2 = R ** (bas_n - 2)
        nRnm2 = bas_n * Rnm2
        bexp_er = bas_exp * er
        nabla_rn = nRnm2.unsqueeze(-1) * xyz
        nabla_er = -bexp_er.unsqueeze(-1) * xyz / R.unsqueeze(-1)
    for d in derivative:
        if d == 0:
            out.append(rn * er)
        elif d == 1:
            out.append(nabla_rn * er.unsqueeze(-1) + rn.unsqueeze(-1) * nabla_er)
        elif d == 2:
            lap_rn = nRnm2 * (bas_n + 1)
            lap_er = bexp_er * (bas_exp - 2.0 / R)
            out.append(lap_rn * er + 2 * (nabla_rn * nabla_er).sum(-1) + rn * lap_er)
    return out


def cartesian_harmonics(xyz, k, derivative):
    """x^kx y^ky z^kz with per-axis gradient and Laplacian
    (spherical_harmonics.py:102-199).  ``k`` is (Np,3) int."""
    xyz_k = fast_power(xyz, k)
    out = []
    kx, ky, kz = k[:, 0], k[:, 1], k[:, 2]
    for d in derivative:
        if d == 0:
            out.append(xyz_k.prod(-1))
        elif d == 1:
            km1 = (k - 1).clamp(min=0)
            p1 = fast_power(xyz, km1)
            dx = kx * p1[..., 0] * xyz_k[..., 1] * xyz_k[..., 2]
            dy = ky * xyz_k[..., 0] * p1[..., 1] * xyz_k[..., 2]
            dz = kz * xyz_k[..., 0] * xyz_k[..., 1] * p1[..., 2]
            out.append(torch.stack((dx, dy, dz), dim=-1))
        elif d == 2:
            km2 = (k - 2).clamp(min=0)
            p2 = fast_power(xyz, km2)
            d2x = kx * (kx - 1) * p2[..., 0] * xyz_k[..., 1] * xyz_k[..., 2]
            d2y = ky * (ky - 1) * xyz_k[..., 0] * p2[..., 1] * xyz_k[..., 2]
            d2z = kz * (kz - 1) * xyz_k[..., 0] * xyz_k[..., 1] * p2[..., 2]
            out.append(d2x + d2y + d2z)
    return out


# --------------------------------------------------------------------------- #
# wavefunction/pooling/orbital_configurations.py, orbital_projector.py
# --------------------------------------------------------------------------- #
def get_configs(configs, nup: int, ndown: int, nmo: int):
    """ground_state / cas(n,m) / explicit tuple (orbital_configurations.py:14-63,79-89,180-238)."""
    if isinstance(configs, tuple):
        return configs
    configs = configs.lower()
    if configs == "ground_state":
        return torch.LongTensor([list(range(nup))]), torch.LongTensor([list(range(ndown))])
    if configs.startswith("cas("):
        nel, norb = [int(v) for v in configs[4:-1].split(",")]
        if nel > nup + ndown or norb > nmo:
            raise ValueError("CAS larger than the system")
        nocc = (nel // 2, nel // 2) if nel % 2 == 0 else (nel // 2 + 1, nel // 2)
        nvirt = (norb - nocc[0], norb - nocc[1])
        if nup != ndown:
            raise ValueError("CAS active space not possible with spin polarized calculation")
        lo, hi = nup - nocc[0], nup + nvirt[0]
        frz = list(range(lo))
        cup = [frz + list(c) for c in combinations(range(lo, hi), nel // 2)]
        cdn = [frz + list(c) for c in combinations(range(lo, hi), nel // 2)]
        pairs = list(product(cup, cdn))
        return torch.LongTensor([p[0] for p in pairs]), torch.LongTensor([p[1] for p in pairs])
    raise ValueError("oracle configs: ground_state, cas(n,m) or explicit tuple")


def unique_configs(configs):
    """torch.unique over rows with inverse indices (orbital_projector.py:29-54)."""
    def _u(c):
        if c.numel() == 0:  # a spin channel without electrons (one-electron test systems)
            return c[:1], torch.zeros(c.shape[0], dtype=torch.long)
        return torch.unique(c, dim=0, return_inverse=True)

    (cu, iu), (cd, idn) = _u(configs[0]), _u(configs[1])
    return (cu, cd), (iu, idn)


def split_orbitals(mat, nup, ucfg):
    """Gather the Slater matrices of the unique configurations
    (orbital_projector.py:56-94; zip semantics kept)."""
    ups, dns = [], []
    for cu, cd in zip(ucfg[0], ucfg[1]):
        ups.append(mat[..., :nup, cu].unsqueeze(0))
        dns.append(mat[..., nup:, cd].unsqueeze(0))
    return torch.cat(ups, 0), torch.cat(dns, 0)


# --------------------------------------------------------------------------- #
# Jastrow distances and Pade kernels
# --------------------------------------------------------------------------- #
def ee_distance(pos, nelec, derivative=0):
    """Quadratic-expansion e-e distance with diagonal eps
    (jastrows/distance/electron_electron_distance.py:47-190)."""
    eps = 1e-16
    x = pos.view(-1, nelec, 3)
    norm = (x ** 2).sum(-1).unsqueeze(-1)
    dist = norm + norm.transpose(1, 2) - 2.0 * torch.bmm(x, x.transpose(1, 2))
    eye = torch.diag(dist.new_ones(nelec)).expand_as(dist)
    diag = torch.diag_embed(torch.diagonal(dist, dim1=-1, dim2=-2))
    dist = torch.sqrt(dist - diag + eps * eye)
    if derivative == 0:
        return dist
    diff = x.transpose(1, 2).unsqueeze(3)
    diff = diff - diff.transpose(2, 3)
    if derivative == 1:
        return diff * (1.0 / (dist + eps * eye)).unsqueeze(1)
    d2 = (diff ** 2)[:, [[1, 2], [2, 0], [0, 1]], ...].sum(2)
    return d2 * (1.0 / (dist ** 3 + eps * eye)).unsqueeze(1)


def en_distance(pos, atoms, nelec, derivative=0):
    """Quadratic-expansion e-n distance (jastrows/distance/electron_nuclei_distance.py:53-162)."""
    eps = 1e-16
    x = pos.view(-1, nelec, 3)
    norm = (x ** 2).sum(-1).unsqueeze(-1)
    norm_at = (atoms ** 2).sum(-1).unsqueeze(-1).T
    dist = torch.sqrt(norm + norm_at - 2.0 * x @ atoms.T)
    if derivative == 0:
        return dist
    if derivative == 1:
        invr = (1.0 / (dist + eps)).unsqueeze(-1)
        diff = (x.unsqueeze(-1) - atoms.T).transpose(2, 3)
        return (diff * invr).permute(0, 3, 1, 2)
    invr3 = (1.0 / (dist ** 3 + eps)).unsqueeze(1)
    diff = x.transpose(1, 2).unsqueeze(3)
    diff = (diff - atoms.T.unsqueeze(1)) ** 2
    diff = diff[:, [[1, 2], [2, 0], [0, 1]], ...].sum(2)
    return diff * invr3


def pade(static_weight, w, r):
    """w0 r / (1 + w r) (elec_elec/kernels/pade_jastrow_kernel.py:89-104)."""
    return static_weight * r / (1.0 + w * r)


def pade_d1(static_weight, w, r, dr):
    """(elec_elec/kernels/pade_jastrow_kernel.py:106-131)."""
    r_ = r.unsqueeze(1)
    den = 1.0 / (1.0 + w * r_)
    return static_weight * dr * den - static_weight * w * r_ * dr * den ** 2


def pade_d2(static_weight, w, r, dr, d2r):
    """(elec_elec/kernels/pade_jastrow_kernel.py:133-153)."""
    r_ = r.unsqueeze(1)
    den = 1.0 / (1.0 + w * r_)
    den2 = den ** 2
    dr2 = dr * dr
    a = static_weight * d2r * den
    b = -2 * static_weight * w * dr2 * den2
    c = -static_weight * w * r_ * d2r * den2
    d = 2 * static_weight * w ** 2 * r_ * dr2 * den ** 3
    return a + b + c + d


def ee_static_weight(nup, ndown):
    """0.25 same spin, 0.5 opposite spin, upper triangle (pade_jastrow_kernel.py:34-66)."""
    n = nup + ndown
    sw = torch.full((n, n), 0.5, dtype=DT)
    sw[:nup, :nup] = 0.25
    sw[nup:, nup:] = 0.25
    mask = torch.triu(torch.ones(n, n), diagonal=1).bool()
    return sw.masked_select(mask)


class JastrowEE:
    """Electron-electron Pade Jastrow factor
    (elec_elec/jastrow_factor_electron_electron.py:11-260)."""

    def __init__(self, nup, ndown, w: torch.Tensor):
        self.nup, self.ndown, self.nelec = nup, ndown, nup + ndown
        self.w = w
        self.sw = ee_static_weight(nup, ndown)
        n = self.nelec
        self.mask = torch.triu(torch.ones(n, n), diagonal=1).bool()
        rows, cols = [], []
        for i in range(n - 1):
            for j in range(i + 1, n):
                rows.append(i)
                cols.append(j)
        self.index_row = torch.LongTensor(rows)
        self.index_col = torch.LongTensor(cols)

    def _tri(self, m):
        nb = m.shape[0]
        return m.masked_select(self.mask).view(nb, -1)

    def _edist(self, pos, d):
        if d == 0:
            return self._tri(ee_distance(pos, self.nelec, 0))
        nb = pos.shape[0]
        return self._tri(ee_distance(pos, self.nelec, d)).view(nb, 3, -1)

    def _scatter(self, v, sign_col):
        out = torch.zeros(list(v.shape[:-1]) + [self.nelec], dtype=v.dtype, device=v.device)
        out = out.index_add(-1, self.index_row, v)
        return out.index_add(-1, self.index_col, sign_col * v)

    def __call__(self, pos, derivative=0):
        r = self._edist(pos, 0)
        jast = torch.exp(pade(self.sw, self.w, r).sum(-1)).unsqueeze(-1)
        if derivative == 0:
            return jast
        dr = self._edist(pos, 1)
        djast = pade_d1(self.sw, self.w, r, dr) * jast.unsqueeze(-1)
        grad = self._scatter(djast, -1.0)  # (Nw,3,Ne)  (:188-214)
        d2r = self._edist(pos, 2)
        d2 = pade_d2(self.sw, self.w, r, dr, d2r).sum(-2)
        hess = self._scatter(d2, 1.0)
        part = (self._scatter(pade_d1(self.sw, self.w, r, dr), -1.0) ** 2).sum(-2)
        lap = (hess + part) * jast  # (:216-260)
        return jast, grad, lap


class JastrowEN:
    """Electron-nucleus Pade Jastrow factor, static weight +1
    (elec_nuclei/jastrow_factor_electron_nuclei.py:9-161, kernels/pade_jastrow_kernel.py:8-116)."""

    def __init__(self, nelec, atoms: torch.Tensor, w: torch.Tensor):
        self.nelec = nelec
        self.atoms = atoms
        self.w = w
        self.sw = torch.as_tensor([1.0], dtype=DT)

    def __call__(self, pos, derivative=0):
        r = en_distance(pos, self.atoms, self.nelec, 0)
        jast = torch.exp(pade(self.sw, self.w, r).sum([-1, -2])).unsqueeze(-1)
        if derivative == 0:
            return jast
        dr = en_distance(pos, self.atoms, self.nelec, 1)
        d2r = en_distance(pos, self.atoms, self.nelec, 2)
        grad = pade_d1(self.sw, self.w, r, dr).sum(3) * jast.unsqueeze(-1)
        d2 = pade_d2(self.sw, self.w, r, dr, d2r).sum((1, 3))
        dj = ((pade_d1(self.sw, self.w, r, dr).sum(3)) ** 2).sum(1)
        return jast, grad, (d2 + dj) * jast


def combine_jastrow(terms, pos):
    """Product of Jastrow factors with product-rule derivatives
    (jastrows/combine_jastrow.py:33-195, derivative=[0,1,2] branch)."""
    vals = [t(pos, 2) for t in terms]
    if len(vals) == 1:
        return vals[0]
    j = [v[0] for v in vals]
    dj = [v[1] for v in vals]
    d2j = [v[2] for v in vals]
    out_j = j[0]
    for t in j[1:]:
        out_j = out_j * t
    out_d2 = 0.0
    for i in range(len(j)):
        tmp = list(j)
        tmp[i] = d2j[i]
        p = tmp[0]
        for t in tmp[1:]:
            p = p * t
        out_d2 = out_d2 + p
    for i in range(len(j) - 1):
        for k in range(i + 1, len(j)):
            tmp = [t.unsqueeze(-1) for t in j]
            tmp[i] = dj[i]
            tmp[k] = dj[k]
            p = tmp[0]
            for t in tmp[1:]:
                p = p * t
            out_d2 = out_d2 + (2.0 * p).sum(1)
    ju = [t.unsqueeze(-1) for t in j]
    out_d = 0.0
    for i in range(len(j)):
        tmp = list(ju)
        tmp[i] = dj[i]
        p = tmp[0]
        for t in tmp[1:]:
            p = p * t
        out_d = out_d + p
    return out_j, out_d, out_d2


# --------------------------------------------------------------------------- #
# The Slater-Jastrow wavefunction
# --------------------------------------------------------------------------- #
class OracleSlaterJastrow(torch.nn.Module):
    """Restatement of SlaterJastrow (wavefunction/slater_jastrow.py:30-482) and
    WaveFunction (wavefunction/wf_base.py:7-277) for ground_state / cas / explicit
    configurations, Cartesian GTO (gto_pure) or STO (sto) radial parts, and
    ee / en Pade Jastrow factors."""

    def __init__(
        self,
        basis,
        atom_coords,
        atomic_number,
        nup: int,
        ndown: int,
        configs="ground_state",
        mo_coeffs=None,
        ci_weights=None,
        jastrow_ee: Optional[float] = 1.0,
        jastrow_en: Optional[float] = None,
        include_all_mo: bool = True,
    ):
        super().__init__()
        self.basis = basis
        self.nup, self.ndown = nup, ndown
        self.nelec = nup + ndown
        self.atom_coords = torch.nn.Parameter(torch.as_tensor(np.asarray(atom_coords), dtype=DT))
        self.atoms_static = torch.as_tensor(np.asarray(atom_coords), dtype=DT)
        self.atomic_number = list(atomic_number)
        self.natom = len(self.atomic_number)
        self.nshells = torch.as_tensor(np.asarray(basis.nshells))
        self.index_ctr = torch.as_tensor(np.asarray(basis.index_ctr))
        self.contract = len(torch.unique(self.index_ctr)) != len(self.index_ctr)
        self.bas_coeffs = torch.as_tensor(np.asarray(basis.bas_coeffs), dtype=DT)
        self.bas_exp = torch.nn.Parameter(torch.as_tensor(np.asarray(basis.bas_exp), dtype=DT))
        self.bas_n = torch.as_tensor(np.asarray(basis.bas_kr), dtype=DT)
        self.bas_k = torch.stack(
            [torch.as_tensor(np.asarray(v)) for v in (basis.bas_kx, basis.bas_ky, basis.bas_kz)]
        ).T.contiguous()
        self.radial_type = basis.radial_type
        self.norm_cst = atomic_orbital_norm(basis)
        self.norb = int(basis.nao)
        self.configs = get_configs(configs, nup, ndown, int(basis.nmo))
        self.config_method = configs if isinstance(configs, str) else "explicit"
        self.nci = len(self.configs[0])
        self.highest_occ_mo = int(max(int(c.max()) if c.numel() else -1 for c in self.configs)) + 1
        mos = np.asarray(basis.mos if mo_coeffs is None else mo_coeffs, dtype=np.float64)
        if not include_all_mo:
            mos = mos[:, : self.highest_occ_mo]
        self.mo_scf = torch.as_tensor(mos, dtype=DT)
        self.mo_modifier = torch.nn.Parameter(torch.ones_like(self.mo_scf))
        w = torch.zeros(1, self.nci, dtype=DT)
        w[0, 0] = 1.0
        if ci_weights is not None:
            w = torch.as_tensor(np.asarray(ci_weights), dtype=DT).reshape(1, -1).clone()
        self.fc_weight = torch.nn.Parameter(w)
        self.ucfg, self.iucfg = unique_configs(self.configs)
        self.jterms = []
        self.w_ee = self.w_en = None
        if jastrow_ee is not None:
            self.w_ee = torch.nn.Parameter(torch.as_tensor([float(jastrow_ee)], dtype=DT))
            self.jterms.append(JastrowEE(nup, ndown, self.w_ee))
        if jastrow_en is not None:
            self.w_en = torch.nn.Parameter(torch.as_tensor([float(jastrow_en)], dtype=DT))
            self.jterms.append(JastrowEN(self.nelec, self.atoms_static, self.w_en))

    # ----------------------------------------------------------- orbitals
    def _xyz_r(self, pos):
        """Electron-atom vectors, repeated to the primitives (atomic_orbitals.py:611-652)."""
        xyz = pos.view(-1, self.nelec, 1, 3) - self.atom_coords[None, ...]
        r = torch.sqrt((xyz * xyz).sum(3))
        return xyz.repeat_interleave(self.nshells, dim=2), r.repeat_interleave(self.nshells, dim=2)

    def _radial(self, r, xyz, derivative):
        if self.radial_type == "gto_pure":
            return radial_gaussian_pure(r, self.bas_exp, xyz, derivative)
        if self.radial_type == "sto":
            return radial_slater(r, self.bas_n, self.bas_exp, xyz, derivative)
        raise ValueError("oracle radial types: gto_pure, sto")

    def _contract(self, bas, coeff=True):
        """index_add_ over index_ctr (atomic_orbitals.py:654-669)."""
        if coeff:
            bas = self.bas_coeffs * bas if bas.ndim == 3 else self.bas_coeffs.unsqueeze(-1) * bas
        shape = list(bas.shape)
        shape[2] = self.norb
        out = torch.zeros(shape, dtype=bas.dtype, device=bas.device)
        return out.index_add(2, self.index_ctr, bas)

    def ao(self, pos, derivative=0):
        """AO values; with derivative='all' returns (ao, grad (Nw,Ne,Nao,3), lap)
        (atomic_orbitals.py:220-249, 330-356, 397-423, 578-609)."""
        xyz, r = self._xyz_r(pos)
        k = self.bas_k
        if derivative == 0:
            R = self._radial(r, xyz, [0])[0]
            Y = cartesian_harmonics(xyz, k, [0])[0]
            ao = self.norm_cst * R * Y
            return self._contract(ao) if self.contract else ao
        R, dR, d2R = self._radial(r, xyz, [0, 1, 2])
        Y, dY, d2Y = cartesian_harmonics(xyz, k, [0, 1, 2])
        ao = self.norm_cst * R * Y
        ao = self._contract(ao) if self.contract else ao
        g = dR * Y.unsqueeze(-1) + R.unsqueeze(-1) * dY
        g = self.norm_cst.unsqueeze(-1) * self.bas_coeffs.unsqueeze(-1) * g
        g = self._contract(g, coeff=False) if self.contract else g
        lap = self.norm_cst * (d2R * Y + 2.0 * (dR * dY).sum(3) + R * d2Y)
        lap = self._contract(lap) if self.contract else lap
        return ao, g, lap

    def mo(self, ao):
        """ao @ (mo_scf * mo_modifier) (molecular_orbitals.py:79-95)."""
        w = self.mo_scf * self.mo_modifier
        return ao @ w.reshape(1, *w.shape)

    # ------------------------------------------------------------- pooling
    def pool(self, mo):
        """Products of up/down determinants per configuration (slater_pooling.py:64-201)."""
        if self.config_method == "ground_state":
            du = torch.det(mo[:, : self.nup, : self.nup])
            dd = torch.det(mo[:, self.nup :, : self.ndown])
            return (du * dd).unsqueeze(-1)
        up, dn = split_orbitals(mo, self.nup, self.ucfg)
        du, dd = torch.det(up), torch.det(dn)
        return (du[self.iucfg[0], ...] * dd[self.iucfg[1], ...]).transpose(0, 1)

    def operator(self, mo, bop):
        """tr(A^-1 B) per spin, summed over spins (slater_pooling.py:262-409)."""
        if self.config_method == "ground_state":
            iu = torch.inverse(mo[:, : self.nup, : self.nup])
            idn = torch.inverse(mo[:, self.nup :, : self.ndown])
            ou = btrace(iu @ bop[..., : self.nup, : self.nup]).unsqueeze(-1)
            od = btrace(idn @ bop[..., self.nup :, : self.ndown]).unsqueeze(-1)
            return ou + od
        Au, Ad = split_orbitals(mo, self.nup, self.ucfg)
        Bu, Bd = split_orbitals(bop, self.nup, self.ucfg)
        ou = btrace(torch.inverse(Au) @ Bu).transpose(0, 1)
        od = btrace(torch.inverse(Ad) @ Bd).transpose(0, 1)
        return ou[..., self.iucfg[0]] + od[..., self.iucfg[1]]

    def fc(self, x):
        return x @ self.fc_weight.T

    def to_device(self, device):
        """Move every tensor of the restatement (parameters and plain tensors) to
        ``device`` -- used to time the same op sequence on torch-ROCm."""
        self.to(device)
        for k, v in list(vars(self).items()):
            if isinstance(v, torch.Tensor) and not isinstance(v, torch.nn.Parameter):
                setattr(self, k, v.to(device))
        self.ucfg = tuple(c.to(device) for c in self.ucfg)
        self.iucfg = tuple(c.to(device) for c in self.iucfg)
        for t in self.jterms:  # their weights are this module's Parameters, moved in place above
            for k, v in list(vars(t).items()):
                if isinstance(v, torch.Tensor) and not isinstance(v, torch.nn.Parameter):
                    setattr(t, k, v.to(device))
        return self

    # ------------------------------------------------------------ jastrow
    def jastrow(self, pos, derivative=0):
        if not self.jterms:
            return None
        if derivative == 0:
            j = None
            for t in self.jterms:
                j = t(pos, 0) if j is None else j * t(pos, 0)
            return j
        return combine_jastrow(self.jterms, pos)

    # ------------------------------------------------------------ wavefunction
    def forward(self, pos):
        """J * fc(dets) (slater_jastrow.py:243-286)."""
        x = self.fc(self.pool(self.mo(self.ao(pos))))
        J = self.jastrow(pos, 0)
        return x if J is None else J * x

    def pdf(self, pos):
        """psi^2 (wf_base.py:241-247)."""
        return (self.forward(pos) ** 2).reshape(-1)

    def kinetic_operator(self, pos, ao, dao, d2ao, mo):
        """B = -1/2 [d2MO + 2 dJ/J . dMO + d2J/J MO] (slater_jastrow.py:449-482)."""
        bkin = self.mo(d2ao)
        if self.jterms:
            jast, djast, d2jast = self.jastrow(pos, 2)
            djast = djast.transpose(1, 2) / jast.unsqueeze(-1)
            d2jast = d2jast / jast
            dmo = self.mo(dao.transpose(2, 3)).transpose(2, 3)
            bkin = bkin + 2 * (djast.unsqueeze(2) * dmo).sum(-1) + d2jast.unsqueeze(-1) * mo
        return -0.5 * bkin

    def kinetic_energy_jacobi(self, pos):
        """fc(tr(A^-1 B) D) / fc(D) (slater_jastrow.py:312-344)."""
        ao, dao, d2ao = self.ao(pos, "all")
        mo = self.mo(ao)
        bkin = self.kinetic_operator(pos, ao, dao, d2ao, mo)
        kin = self.operator(mo, bkin)
        psi = self.pool(mo)
        return self.fc(kin * psi) / self.fc(psi)

    def kinetic_energy_autograd(self, pos):
        """-1/2 Laplacian(psi)/psi by double autograd (wf_base.py:134-182)."""
        out = self.forward(pos)
        jac = torch.autograd.grad(out, pos, torch.ones_like(out), create_graph=True)[0]
        hess = torch.zeros(jac.shape[0], dtype=DT, device=pos.device)
        for i in range(jac.shape[1]):
            t = torch.autograd.grad(jac[:, i], pos, torch.ones(jac.shape[0], dtype=DT, device=pos.device),
                                    retain_graph=True)[0]
            hess = hess + t[:, i]
        return -0.5 * hess.view(-1, 1) / out

    def electronic_potential(self, pos):
        """sum_{i<j} 1/r_ij with direct differences (wf_base.py:49-70)."""
        pot = torch.zeros(pos.shape[0], dtype=DT, device=pos.device)
        for i in range(self.nelec - 1):
            pi = pos[:, 3 * i : 3 * i + 3]
            for j in range(i + 1, self.nelec):
                pj = pos[:, 3 * j : 3 * j + 3]
                pot = pot + 1.0 / torch.sqrt(((pi - pj) ** 2).sum(1))
        return pot.view(-1, 1)

    def nuclear_potential(self, pos):
        """-sum Z/r_iA (wf_base.py:72-95)."""
        p = torch.zeros(pos.shape[0], dtype=DT, device=pos.device)
        for i in range(self.nelec):
            pe = pos[:, 3 * i : 3 * i + 3]
            for a in range(self.natom):
                r = torch.sqrt(((pe - self.atom_coords[a, :]) ** 2).sum(1))
                p = p - self.atomic_number[a] / r
        return p.view(-1, 1)

    def nuclear_repulsion(self):
        """sum_{A<B} Z_A Z_B / R_AB (wf_base.py:97-116)."""
        v = 0.0
        for a in range(self.natom - 1):
            for b in range(a + 1, self.natom):
                rab = torch.sqrt(((self.atom_coords[a] - self.atom_coords[b]) ** 2).sum())
                v = v + self.atomic_number[a] * self.atomic_number[b] / rab
        return v

    def local_energy(self, pos, kinetic="jacobi"):
        """T + V_en + V_ee + V_nn (wf_base.py:184-215)."""
        ke = self.kinetic_energy_jacobi(pos) if kinetic == "jacobi" else self.kinetic_energy_autograd(pos)
        return ke + self.nuclear_potential(pos) + self.electronic_potential(pos) + self.nuclear_repulsion()

    def gradients_jacobi(self, pos, pdf=False):
        """grad psi by the Jacobi formula (slater_jastrow.py:346-447)."""
        mo = self.mo(self.ao(pos))
        _, gao, _ = self.ao(pos, "all")
        dmo = self.mo(gao.transpose(2, 3)).transpose(2, 3).permute(3, 0, 1, 2)
        eye = torch.eye(self.nelec, dtype=DT, device=pos.device)
        dmo = (dmo.unsqueeze(2) * eye.unsqueeze(-1)).permute(2, 0, 1, 3, 4)
        dmo = dmo.reshape(-1, *dmo.shape[2:])
        gdets = self._operator_multi(mo, dmo)
        dets = self.pool(mo)
        out = self.fc(gdets * dets).transpose(0, 1).squeeze()
        if self.jterms:
            nb = pos.shape[0]
            jast, djast, _ = self.jastrow(pos, 2)
            out = jast * out + (djast.permute(0, 2, 1) * self.fc(dets).unsqueeze(-1)).reshape(nb, -1)
        if pdf:
            out = 2 * out * self.fc(dets)
            if self.jterms:
                out = out * jast
        return out

    def _operator_multi(self, mo, bop):
        """operator() with a leading operator axis (slater_pooling.py:308-409, 4-D case)."""
        if self.config_method == "ground_state":
            iu = torch.inverse(mo[:, : self.nup, : self.nup])
            idn = torch.inverse(mo[:, self.nup :, : self.ndown])
            ou = btrace(iu @ bop[..., : self.nup, : self.nup]).unsqueeze(-1)
            od = btrace(idn @ bop[..., self.nup :, : self.ndown]).unsqueeze(-1)
            return ou + od
        Au, Ad = split_orbitals(mo, self.nup, self.ucfg)
        Bu, Bd = split_orbitals(bop, self.nup, self.ucfg)
        ou = btrace(torch.inverse(Au).unsqueeze(1) @ Bu).permute(1, 2, 0)
        od = btrace(torch.inverse(Ad).unsqueeze(1) @ Bd).permute(1, 2, 0)
        return ou[..., self.iucfg[0]] + od[..., self.iucfg[1]]


# --------------------------------------------------------------------------- #
# sampler/metropolis.py, sampler/walkers.py
# --------------------------------------------------------------------------- #
def init_walkers_normal(nwalkers, nelec, mean, sigma, gen=None):
    """MultivariateNormal(mean, sigma) per electron (walkers.py:98-116)."""
    mvn = torch.distributions.MultivariateNormal(torch.as_tensor(mean, dtype=DT), torch.as_tensor(sigma, dtype=DT))
    return mvn.sample((nwalkers, nelec)).view(nwalkers, nelec * 3)


def metropolis(pdf, pos, nstep, step_size=0.2, ntherm=-1, ndecor=1, proba="normal", seed=None, device="cpu"):
    """All-electron Metropolis (metropolis.py:85-177, 179-225, 256-298).

    Proposal: MultivariateNormal(0, sigma I) with sigma = step/(2 sqrt(2 ln 2))
    used as the *covariance* (metropolis.py:206-212), sampled on the host and
    copied to the device every step (:272-274); acceptance min(1, f'/f) - U >= 0
    (:279-298); f==0 replaced by eps (:103-107,124,146,160).  Returns
    (positions kept, acceptance rate)."""
    if seed is not None:
        torch.manual_seed(seed)
    eps = 1e-16
    if ntherm >= nstep:
        raise ValueError("Thermalisation longer than trajectory")
    if ntherm < 0:
        ntherm = nstep + ntherm
    nw, ndim = pos.shape
    sig = step_size / (2 * torch.sqrt(2 * torch.log(torch.as_tensor(2.0))))
    mvn = torch.distributions.MultivariateNormal(torch.zeros(3), sig * torch.eye(3))
    with torch.no_grad():
        fx = pdf(pos)
        fx[fx == 0] = eps
        kept, rate, idecor = [], 0.0, 0
        for istep in range(nstep):
            if proba == "normal":
                disp = mvn.sample((nw, ndim // 3)).to(device).view(nw, ndim).to(pos.dtype)
            else:
                d = torch.rand((nw, ndim // 3, 3), device=device).view(nw, ndim)
                disp = step_size * (2.0 * d - 1.0)
            xn = pos + disp
            fxn = pdf(xn)
            fxn[fxn == 0.0] = eps
            df = fxn / fx
            df[df > 1] = 1.0
            idx = (df - torch.rand_like(df) >= 0).reshape(-1)
            rate += idx.byte().sum().float().to("cpu") / nw
            pos[idx, :] = xn[idx, :]
            fx[idx] = fxn[idx]
            fx[fx == 0] = eps
            if istep >= ntherm:
                if idecor % ndecor == 0:
                    kept.append(pos.to("cpu").clone())
                idecor += 1
    return torch.cat(kept), float(rate) / nstep


# --------------------------------------------------------------------------- #
# sampler/hamiltonian.py
# --------------------------------------------------------------------------- #
def grad_potential(pdf, q):
    """grad of U = -log pdf by autograd (hamiltonian.py:44-77)."""
    with torch.enable_grad():
        x = q.detach().clone().requires_grad_(True)
        val = -torch.log(pdf(x))
        val.backward(torch.ones_like(val))
    return x.grad.detach()


def hmc_step(pdf, q_init, p, eps, L, u):
    """One step of Hamiltonian._step (hamiltonian.py:143-201) with the momenta p
    (Nw, 3Ne) and the uniforms u (Nw,) given: E = U + |p|^2/2 with U = -log
    pdf; half momentum step, L-1 full (position, momentum) steps, a last
    position step and a closing half momentum step; the move is rejected iff
    exp(E_init - E_new) < u.  Returns (positions, accepted mask)."""
    with torch.no_grad():
        q = q_init.clone()
        e_init = -torch.log(pdf(q)) + 0.5 * (p * p).sum(1)
        p = p - 0.5 * eps * grad_potential(pdf, q)
        for _ in range(L - 1):
            q = q + eps * p
            p = p - eps * grad_potential(pdf, q)
        q = q + eps * p
        p = p - 0.5 * eps * grad_potential(pdf, q)
        p = -p
        e_new = -torch.log(pdf(q)) + 0.5 * (p * p).sum(1)
        rejected = torch.exp(e_init - e_new) < u
        q[rejected] = q_init[rejected]
    return q, ~rejected


def hamiltonian(pdf, pos, nstep, step_size=0.2, L=10, ntherm=-1, ndecor=1, seed=None):
    """Hamiltonian.__call__ (hamiltonian.py:80-141): momenta from torch.randn,
    uniforms from torch.rand, in that order every step, drawn on the device of
    ``pos`` (the reference draws them on the host, which confines it to CPU
    tensors).  Returns (positions kept, acceptance rate)."""
    if seed is not None:
        torch.manual_seed(seed)
    if ntherm < 0:
        ntherm = nstep + ntherm
    q = pos.detach().clone()
    kept, rate, idecor = [], 0.0, 0
    for istep in range(nstep):
        p = torch.randn(q.shape, dtype=q.dtype, device=q.device)
        u = torch.rand(q.shape[0], dtype=q.dtype, device=q.device)
        q, acc = hmc_step(pdf, q, p, step_size, L, u)
        rate += float(acc.double().mean())
        if istep >= ntherm:
            if idecor % ndecor == 0:
                kept.append(q.clone())
            idecor += 1
    return torch.cat(kept), rate / nstep


def single_point(wf, pos):
    """Local energy batch and its statistics (solver_base.py:316-387)."""
    eloc = wf.local_energy(pos)
    e, v = torch.mean(eloc), torch.var(eloc)
    err = torch.sqrt(eloc.var() / eloc.shape[0])
    return SimpleNamespace(pos=pos, local_energy=eloc, energy=e, variance=v, error=err)


def manual_gradient(wf, pos):
    """Gradient estimate of solver.py:372-431: psi.backward(2/N (E_L - <E_L>)/psi)."""
    with torch.no_grad():
        eloc = wf.local_energy(pos)
    psi = wf(pos)
    weight = (eloc - eloc.mean()) / psi.detach() * (2.0 / len(psi))
    params = [p for p in wf.parameters() if p.requires_grad]
    grads = torch.autograd.grad(psi, params, grad_outputs=weight, allow_unused=True)
    return eloc, dict(zip([n for n, p in wf.named_parameters() if p.requires_grad], grads))


PARITY = {
    "double_factorial": "pinned: exact integers (tests/test_oracle_basis.py)",
    "norm_gaussian_cartesian": "pinned: quadrature of chi^2 = 1 (tests/test_oracle_basis.py)",
    "norm_slater_cartesian": "pinned: quadrature of chi^2 = 1 and the stored ADF norms (tests/test_oracle_basis.py)",
    "ao": "pinned: closed-form s/p/d values, autograd gradient/Laplacian, ADF cube grids",
    "mo": "pinned: ADF MO cube grids",
    "pool/operator": "pinned: LAPACK det/inverse vs Leibniz on 2x2/3x3, Jacobi vs autograd",
    "JastrowEE/JastrowEN/combine_jastrow": "pinned: autograd gradient/Laplacian, cusp slope, permutation invariance",
    "kinetic_energy_jacobi/local_energy": "pinned: autograd Laplacian, H-atom and single-Gaussian closed forms, LiH golden fixture",
    "metropolis": "pinned statistically (tests/test_oracle_sampler_gradient.py): hydrogen 1s E_L = -1/2 on every "
                  "kept sample and <r> = 3/2; Gaussian <x^2> = 1/(4 alpha) for normal and uniform moves; first-step "
                  "acceptance (1 + sigma)^(-3/2) from the maximum of exp(-|x|^2/2); flat-density proposal variance "
                  "sigma and step^2/3; kept counts of ntherm / ndecor",
    "hmc_step/hamiltonian": "pinned: leapfrog of a Gaussian target against the closed-form linear map, exact "
                            "energy change, Gaussian moments (tests/test_oracle_hmc.py)",
    "grad_potential": "pinned: -2 grad ln|psi| against gradients_jacobi / psi and central differences",
    "manual_gradient": "pinned (tests/test_oracle_sampler_gradient.py): Richardson-extrapolated central "
                       "differences (h = 1e-5) of the reweighted energy sum r E_L,0 / sum r, r = psi^2/psi_0^2, on "
                       "fixed positions, for mo_modifier, both Pade weights and CI weights (H2O, LiH CAS(2,2))",
}
