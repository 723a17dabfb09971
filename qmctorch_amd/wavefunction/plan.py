"""Host-side construction of the device plan (QmcSystemDesc -> QmcPlan).

The reference evaluates every (primitive, Cartesian component, contraction
column) entry separately (atomic_orbitals.py:611-669: Np entries, each with its
own exp).  Here entries of one atom with identical exponent lists are grouped,
so each exponential is evaluated once per electron and shared by all AOs of the
group (for H2O/cc-pVDZ: 24 exponentials per electron instead of 79).  The
grouping is recomputed from the current values of ``bas_exp`` whenever that
parameter changes, so optimising exponents independently per entry (as the
reference allows) stays exact.
"""
import ctypes
from typing import List, Optional, Tuple

import numpy as np
import torch

from .. import _native as N


def _norm_gaussian_cartesian(kx, ky, kz, a):
    """Per-entry normalisation of Cartesian GTOs (norm_orbital.py:136-161)."""

    def dfact(n):
        out = np.ones_like(n, dtype=np.float64)
        for i, v in enumerate(np.asarray(n).ravel()):
            r, k = 1, int(v)
            while k > 1:
                r *= k
                k -= 2
            out.ravel()[i] = r
        return out

    a = np.asarray(a, dtype=np.float64)
    out = (2 * a / np.pi) ** 0.75
    for k in (kx, ky, kz):
        k = np.asarray(k)
        out = out * (4 * a) ** (k / 2) / np.sqrt(dfact(2 * k - 1))
    return out


def _norm_slater_cartesian(kx, ky, kz, n, a):
    """Per-entry normalisation of Cartesian STOs (norm_orbital.py:95-133)."""
    from math import factorial

    def dfact(v):
        r, k = 1, int(v)
        while k > 1:
            r *= k
            k -= 2
        return r

    kx, ky, kz, n = [np.asarray(v).astype(int) for v in (kx, ky, kz, n)]
    a = np.asarray(a, dtype=np.float64)
    out = np.zeros_like(a)
    for i in range(len(a)):
        l = kx[i] + ky[i] + kz[i] + n[i] + 1
        pref = 4 * np.pi * factorial(2 * l) / (2 * a[i]) ** (2 * l + 1)
        num = dfact(2 * kx[i] - 1) * dfact(2 * ky[i] - 1) * dfact(2 * kz[i] - 1)
        den = dfact(2 * kx[i] + 2 * ky[i] + 2 * kz[i] + 1)
        out[i] = np.sqrt(1.0 / (pref * num / den))
    return out


def atomic_orbital_norm(basis) -> np.ndarray:
    """norm_cst of the AO layer, computed once from the basis (atomic_orbitals.py:90-94)."""
    if basis.harmonics_type != "cart":
        raise NotImplementedError("only Cartesian harmonics are supported (pyscf/ADF both emit 'cart')")
    if basis.radial_type.startswith("sto"):
        return _norm_slater_cartesian(basis.bas_kx, basis.bas_ky, basis.bas_kz, basis.bas_kr, basis.bas_exp)
    return _norm_gaussian_cartesian(basis.bas_kx, basis.bas_ky, basis.bas_kz, basis.bas_exp)


class SystemTables:
    """Numpy tables of one wavefunction; owns the arrays QmcSystemDesc points to."""

    # groups whose largest exponent reaches this value are evaluated in FP64 by the
    # mixed-precision local energy (core primitives: O 1s 11720, Li 1s 642, ...)
    TIGHT_EXPONENT = 50.0

    def __init__(self, basis, atom_coords, atomic_number, nup, ndown, configs, nmo_used,
                 use_ee, use_en, norm_cst, bas_exp_values, vnn_coords=None):
        self.basis = basis
        self.nup, self.ndown = int(nup), int(ndown)
        self.nelec = self.nup + self.ndown
        radial = basis.radial_type
        if radial == "gto_pure":
            self.radial = N.QMC_RADIAL_GTO_PURE
        elif radial == "sto":
            self.radial = N.QMC_RADIAL_STO
        else:
            raise NotImplementedError("radial type %s is not supported by the HIP kernels" % radial)
        nshells = np.asarray(basis.nshells, dtype=np.int64)
        self.natom = len(nshells)
        prim_atom = np.repeat(np.arange(self.natom), nshells)
        index_ctr = np.asarray(basis.index_ctr, dtype=np.int64)
        nao = int(basis.nao)
        coeffs = np.asarray(basis.bas_coeffs, dtype=np.float64)
        contract = len(np.unique(index_ctr)) != len(index_ctr)
        if not contract and np.any(coeffs != 1.0):
            # atomic_orbitals.py applies bas_coeffs to gradients but not to values
            # when no contraction takes place; only unit coefficients keep the two consistent
            raise NotImplementedError("uncontracted basis with non-unit coefficients")
        kval = np.asarray(norm_cst, dtype=np.float64) * (coeffs if contract else 1.0)
        kx = np.asarray(basis.bas_kx).astype(int)
        ky = np.asarray(basis.bas_ky).astype(int)
        kz = np.asarray(basis.bas_kz).astype(int)
        kr = np.asarray(basis.bas_kr).astype(np.float64)
        if max(kx.max(initial=0), ky.max(initial=0), kz.max(initial=0), kr.max(initial=0)) > 3:
            raise NotImplementedError("Cartesian / radial powers above 3")
        ex = np.asarray(bas_exp_values, dtype=np.float64)

        # entries of each AO, in reference order
        entries = [np.where(index_ctr == mu)[0] for mu in range(nao)]
        groups = {}
        order = []
        for mu, ent in enumerate(entries):
            at = int(prim_atom[ent[0]])
            if np.any(prim_atom[ent] != at):
                raise ValueError("AO %d spans several atoms" % mu)
            key = (at, tuple(ex[ent].tolist()))
            if key not in groups:
                groups[key] = []
                order.append(key)
            groups[key].append(mu)
        order.sort(key=lambda k: k[0])  # groups of one atom are contiguous

        g_atom, g_nprim, g_prim0, g_src, g_ao0, g_nao, g_tight = [], [], [], [], [], [], []
        e_mu, e_k, e_K, e_n, e_src = [], [], [], [], []
        P = N.QMC_MAXP
        for key in order:
            at, exps = key
            aos = groups[key]
            rep = entries[aos[0]]
            for c0 in range(0, len(exps), P):
                c1 = min(c0 + P, len(exps))
                g_atom.append(at)
                g_nprim.append(c1 - c0)
                g_prim0.append(len(g_src))
                g_src.extend(rep[c0:c1].tolist())
                g_ao0.append(len(e_mu))
                g_nao.append(len(aos))
                g_tight.append(1 if max(exps) >= self.TIGHT_EXPONENT else 0)
                for mu in aos:
                    ent = entries[mu][c0:c1]
                    k0 = entries[mu][0]
                    e_mu.append(mu)
                    e_k.append(int(kx[k0]) | (int(ky[k0]) << 4) | (int(kz[k0]) << 8))
                    Kp = np.zeros(P)
                    Kp[: len(ent)] = kval[ent]
                    Np = np.zeros(P)
                    Np[: len(ent)] = kr[ent]
                    e_K.append(Kp)
                    e_n.append(Np)
                    Sp = np.zeros(P, dtype=np.int64)
                    Sp[: len(ent)] = ent
                    e_src.append(Sp)
        self.group_key = [(k[0], len(k[1])) for k in order]
        self.grp_atom = np.asarray(g_atom, dtype=np.int32)
        self.grp_nprim = np.asarray(g_nprim, dtype=np.int32)
        self.grp_prim0 = np.asarray(g_prim0, dtype=np.int32)
        self.grp_prim_src = np.asarray(g_src, dtype=np.int32)
        self.grp_ao0 = np.asarray(g_ao0, dtype=np.int32)
        self.grp_nao = np.asarray(g_nao, dtype=np.int32)
        self.grp_tight = np.asarray(g_tight, dtype=np.int32)
        self.gao_mu = np.asarray(e_mu, dtype=np.int32)
        self.gao_k = np.asarray(e_k, dtype=np.int32)
        self.gao_K = np.ascontiguousarray(np.stack(e_K)).astype(np.float64)
        self.gao_n = np.ascontiguousarray(np.stack(e_n)).astype(np.float64)
        self.gao_src = np.ascontiguousarray(np.stack(e_src)).astype(np.int32)
        self.nexp_per_electron = int(self.grp_nprim.sum())
        # grouping signature: which entries must keep equal exponents
        self.members = []
        for key in order:
            aos = groups[key]
            self.members.append(np.stack([entries[mu] for mu in aos]))

        coords = np.asarray(atom_coords, dtype=np.float64).reshape(-1, 3)
        self.atom_z = np.asarray(atomic_number, dtype=np.float64)
        # e-n Jastrow centres: the molecule's coordinates at construction
        # (jastrow_factor_electron_nuclei.py:40); V_nn from the AO layer's
        # current coordinates when given (wf_base.py:97-116)
        self.jastrow_atoms = np.ascontiguousarray(coords).astype(np.float64)
        self.vnn = nuclear_repulsion(coords if vnn_coords is None else vnn_coords, self.atom_z)

        cup, cdn = [np.asarray(c, dtype=np.int64) for c in configs]
        self.nci = cup.shape[0]
        uu, iu = _unique_rows(cup)
        ud, idn = _unique_rows(cdn)
        # beyond QMC_MAXCONF unique configurations per spin the library uses its
        # excitation (trace-trick) kernels, for single and double excitations of
        # the ground state (qmc_plan_create checks)
        self.excitations = len(uu) > N.QMC_MAXCONF or len(ud) > N.QMC_MAXCONF
        self.conf_up = np.ascontiguousarray(uu.reshape(-1)).astype(np.int32)
        self.conf_dn = np.ascontiguousarray(ud.reshape(-1)).astype(np.int32)
        if self.ndown == 0:
            self.conf_dn = np.zeros(1, dtype=np.int32)
        self.nconf_up, self.nconf_dn = len(uu), max(len(ud), 1)
        self.ci_up = iu.astype(np.int32)
        self.ci_dn = idn.astype(np.int32)
        self.nmo = int(nmo_used)
        self.nao = nao
        self.nprim_ref = len(ex)
        self.use_ee, self.use_en = int(bool(use_ee)), int(bool(use_en))

    def grouping_valid(self, bas_exp_values: np.ndarray) -> bool:
        """True while every group's members still share their exponents."""
        for m in self.members:
            v = bas_exp_values[m]
            if not np.all(v == v[0:1, :]):
                return False
        return True

    def desc(self) -> N.QmcSystemDesc:
        def ip(a):
            return a.ctypes.data_as(N.c_i32p)

        def fp(a):
            return a.ctypes.data_as(N.c_f64p)

        d = N.QmcSystemDesc()
        d.nelec, d.nup, d.ndown, d.natom = self.nelec, self.nup, self.ndown, self.natom
        d.nao = self.nao
        d.nprim_ref = self.nprim_ref
        d.radial_type = self.radial
        d.ngroup = len(self.grp_atom)
        d.grp_atom, d.grp_nprim, d.grp_prim0 = ip(self.grp_atom), ip(self.grp_nprim), ip(self.grp_prim0)
        d.grp_prim_src, d.grp_ao0, d.grp_nao = ip(self.grp_prim_src), ip(self.grp_ao0), ip(self.grp_nao)
        d.grp_tight = ip(self.grp_tight)
        d.nentry = len(self.gao_mu)
        d.gao_mu, d.gao_k, d.gao_K, d.gao_n = ip(self.gao_mu), ip(self.gao_k), fp(self.gao_K), fp(self.gao_n)
        d.gao_src = ip(self.gao_src)
        d.atom_z, d.jastrow_atoms = fp(self.atom_z), fp(self.jastrow_atoms)
        d.nmo, d.nci = self.nmo, self.nci
        d.nconf_up, d.nconf_dn = self.nconf_up, self.nconf_dn
        d.conf_up, d.conf_dn = ip(self.conf_up), ip(self.conf_dn)
        d.ci_up, d.ci_dn = ip(self.ci_up), ip(self.ci_dn)
        d.use_jastrow_ee, d.use_jastrow_en = self.use_ee, self.use_en
        d.vnn = self.vnn
        return d


def nuclear_repulsion(coords, charges) -> float:
    """sum_{A<B} Z_A Z_B / |R_A - R_B| (wf_base.py:97-116)."""
    c = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    z = np.asarray(charges, dtype=np.float64)
    vnn = 0.0
    for a in range(len(z)):
        for b in range(a + 1, len(z)):
            vnn += z[a] * z[b] / np.linalg.norm(c[a] - c[b])
    return float(vnn)


def component_plan(mol, configs=None, nmo_used=None, use_ee=False, use_en=False) -> "DevicePlan":
    """Plan of a module used on its own (atomic orbitals, Jastrow factors,
    Slater pooling): the molecule's basis tables, the given configurations
    (ground state by default) and Jastrow flags."""
    if configs is None:
        configs = (np.arange(mol.nup)[None], np.arange(mol.ndown)[None])
    cup, cdn = [np.asarray(c, dtype=np.int64) for c in configs]
    if nmo_used is None:
        nmo_used = max(int(cup.max()) + 1 if cup.size else 1, int(cdn.max()) + 1 if cdn.size else 1,
                       mol.nup, max(mol.ndown, 1))
    b = mol.basis
    tab = SystemTables(b, mol.atom_coords, mol.atomic_number, mol.nup, mol.ndown, (cup, cdn), nmo_used, use_ee,
                       use_en, atomic_orbital_norm(b), np.asarray(b.bas_exp, dtype=np.float64))
    return DevicePlan(tab)


def _unique_rows(c: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Sorted unique rows and inverse index (torch.unique(dim=0) semantics,
    orbital_projector.py:29-54)."""
    if c.size == 0:
        return np.zeros((1, 0), dtype=np.int64), np.zeros(c.shape[0], dtype=np.int64)
    t = torch.as_tensor(c)
    u, inv = torch.unique(t, dim=0, return_inverse=True)
    return u.numpy(), inv.numpy()


class DevicePlan:
    """Owns a QmcPlan handle for one device."""

    def __init__(self, tables: SystemTables):
        self.tables = tables
        lib = N.load()
        h = ctypes.c_void_p()
        desc = tables.desc()
        N.check(lib.qmc_plan_create(ctypes.byref(desc), ctypes.byref(h)))
        self.handle = h
        self._lib = lib

    def __del__(self):
        try:
            if getattr(self, "handle", None) and self.handle.value:
                self._lib.qmc_plan_destroy(self.handle)
                self.handle = ctypes.c_void_p()
        except Exception:
            pass

    def sizes(self) -> N.QmcPlanSizes:
        if getattr(self, "_sizes", None) is None:
            sz = N.QmcPlanSizes()
            N.check(self._lib.qmc_plan_sizes(self.handle, ctypes.byref(sz)))
            self._sizes = sz
        return self._sizes

    @property
    def bucket(self) -> int:
        """Column count of the kernels' MO rows (row stride of the entry-ordered W)."""
        return int(self._lib.qmc_plan_bucket(self.handle))

    @property
    def specialization(self):
        """Name of the basis layout whose generated kernels this plan uses, or None."""
        name = self._lib.qmc_plan_specialization(self.handle)
        return name.decode() if name else None

    def set_specialized(self, enable: bool) -> None:
        """Use the generated kernels of the plan's layout (default) or only the
        table-driven ones."""
        N.check(self._lib.qmc_plan_set_specialized(self.handle, 1 if enable else 0))
