"""Atomic-orbital layer (atomic_orbitals.py:15-695 of the reference).

Same parameters and attributes (``atom_coords``, ``bas_exp`` trainable,
``norm_cst`` computed once at construction like atomic_orbitals.py:90-94);
``forward`` evaluates on the device through ``qmc_ao_eval``.
"""
from typing import List, Optional

import numpy as np
import torch
from torch import nn

from ... import _native as N
from ..plan import DevicePlan, SystemTables, atomic_orbital_norm


class AtomicOrbitals(nn.Module):
    def __init__(self, mol, cuda: Optional[bool] = False) -> None:
        super().__init__()
        dtype = torch.get_default_dtype()
        self.mol = mol
        self.nelec = mol.nelec
        self.norb = int(mol.basis.nao)
        self.ndim = 3
        self.atom_coords = nn.Parameter(torch.as_tensor(np.asarray(mol.basis.atom_coords_internal)).type(dtype))
        self.natoms = len(self.atom_coords)
        self.atomic_number = mol.atomic_number
        self.nshells = torch.as_tensor(np.asarray(mol.basis.nshells))
        self.nao_per_atom = torch.as_tensor(np.asarray(getattr(mol.basis, "nao_per_atom", [])))
        self.index_ctr = torch.as_tensor(np.asarray(mol.basis.index_ctr))
        self.nbas = len(self.index_ctr)
        self.contract = len(torch.unique(self.index_ctr)) != len(self.index_ctr)
        self.bas_coeffs = torch.as_tensor(np.asarray(mol.basis.bas_coeffs)).type(dtype)
        self.bas_exp = nn.Parameter(torch.as_tensor(np.asarray(mol.basis.bas_exp)).type(dtype))
        self.harmonics_type = mol.basis.harmonics_type
        self.radial_type = mol.basis.radial_type
        self.bas_n = torch.as_tensor(np.asarray(mol.basis.bas_kr)).type(dtype)
        self.norm_cst = torch.as_tensor(atomic_orbital_norm(mol.basis)).type(dtype)
        self.backflow_trans = None
        self.cuda = cuda
        self.device = torch.device("cuda" if cuda else "cpu")
        self._plan = None
        self._plan_key = None
        if cuda:
            self.to(self.device)

    def __repr__(self) -> str:
        return "AtomicOrbitals(%s, %s, %d -> (%d,%d) )" % (
            self.radial_type, self.harmonics_type, self.nelec * self.ndim, self.nelec, self.norb)

    # plan shared with the owning wavefunction when there is one
    def device_plan(self, owner_plan=None):
        if owner_plan is not None:
            return owner_plan
        key = self.bas_exp._version
        if self._plan is None or self._plan_key != key:
            ex = self.bas_exp.detach().cpu().numpy()
            tab = SystemTables(self.mol.basis, self.mol.atom_coords, self.mol.atomic_number, self.mol.nup,
                               self.mol.ndown, (np.arange(self.mol.nup)[None], np.arange(self.mol.ndown)[None]),
                               max(self.mol.nup, 1), False, False, self.norm_cst.cpu().numpy(), ex)
            self._plan = DevicePlan(tab)
            self._plan_key = key
        return self._plan

    def forward(self, pos: torch.Tensor, derivative: Optional[List[int]] = [0], sum_grad: Optional[bool] = True,
                sum_hess: Optional[bool] = True, one_elec: Optional[bool] = False, plan=None):
        """AO values (Nw, Ne, Nao) or derivatives, atomic_orbitals.py:131-218."""
        if not isinstance(derivative, list):
            derivative = [derivative]
        if one_elec:
            raise NotImplementedError("one_elec AO updates are not provided by the HIP path")
        N.require_device(pos, "positions")
        dp = self.device_plan(plan)
        nw = pos.shape[0]
        x = pos.detach().to(torch.float64).contiguous()
        need_d = any(d in derivative for d in (1, 2))
        ao = torch.empty(nw, self.nelec, self.norb, dtype=torch.float64, device=pos.device)
        dao = torch.empty(nw, self.nelec, self.norb, 3, dtype=torch.float64, device=pos.device) if need_d else None
        d2ao = torch.empty(nw, self.nelec, self.norb, dtype=torch.float64, device=pos.device) if need_d else None
        prm = N.QmcParams()
        be = self.bas_exp.detach().to(torch.float64).contiguous()
        ac = self.atom_coords.detach().to(torch.float64).contiguous()
        src = torch.as_tensor(dp.tables.grp_prim_src.astype(np.int64), device=be.device)
        gexp = be.index_select(0, src)
        gexpf = gexp.to(torch.float32)
        prm.bas_exp, prm.atom_coords = be.data_ptr(), ac.data_ptr()
        prm.grp_exp, prm.grp_expf = gexp.data_ptr(), gexpf.data_ptr()
        N.check(N.load().qmc_ao_eval(dp.handle, prm, x.data_ptr(), nw, ao.data_ptr(), N.ptr(dao), N.ptr(d2ao),
                                     N.current_stream()))
        dt = torch.get_default_dtype()
        out = []
        for d in derivative:
            if d == 0:
                out.append(ao.to(dt))
            elif d == 1:
                out.append((dao.sum(-1) if sum_grad else dao).to(dt))
            elif d == 2:
                if not sum_hess:
                    raise NotImplementedError("per-axis Hessian diagonal is not provided by the HIP path")
                out.append(d2ao.to(dt))
            else:
                raise NotImplementedError("derivative %s" % d)
        return out[0] if len(out) == 1 else tuple(out)
