"""Atomic-orbital layer (atomic_orbitals.py:15-695 of the reference).

Same parameters and attributes (``atom_coords``, ``bas_exp`` trainable,
``norm_cst`` computed once at construction like atomic_orbitals.py:90-94);
``forward`` evaluates on the device through ``qmc_ao_eval_full`` for every
derivative option of the reference: values, gradients (summed or per axis),
Laplacians or Hessian diagonals, mixed second derivatives, all three at once,
and one-electron updates.
"""
from typing import List, Optional

import numpy as np
import torch
from torch import nn

from ... import _native as N
from ..param_views import f64
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

    def device_plan(self, owner_plan=None):
        """The owning wavefunction's plan, or one of this layer's (rebuilt when
        the exponents stop sharing their groups)."""
        if owner_plan is not None:
            return owner_plan
        key = self.bas_exp._version
        if self._plan is None or self._plan_key != key:
            ex = self.bas_exp.detach().cpu().numpy().astype(np.float64)
            if self._plan is None or not self._plan.tables.grouping_valid(ex):
                tab = SystemTables(self.mol.basis, self.mol.atom_coords, self.mol.atomic_number, self.mol.nup,
                                   self.mol.ndown, (np.arange(self.mol.nup)[None], np.arange(self.mol.ndown)[None]),
                                   max(self.mol.nup, 1), False, False, self.norm_cst.cpu().numpy(), ex)
                self._plan = DevicePlan(tab)
            self._plan_key = key
        return self._plan

    def _evaluate(self, pos, nelec, want, plan=None):
        N.require_device(pos, "positions")
        dp = self.device_plan(plan)
        x = f64(pos)
        nw = x.shape[0]
        dev = x.device
        shape = (nw, nelec, self.norb)
        out = {k: torch.empty(*shape, *( (3,) if k in ("grad", "hess_diag", "hess_mixed") else ()),
                              dtype=torch.float64, device=dev) for k in want}
        if nw:
            be = f64(self.bas_exp)
            ac = f64(self.atom_coords)
            src = getattr(dp, "_grp_src_dev", None)
            if src is None or src.device != dev:
                src = dp._grp_src_dev = torch.as_tensor(dp.tables.grp_prim_src.astype(np.int64), device=dev)
            gexp = be.index_select(0, src)
            prm = N.QmcParams()
            prm.bas_exp, prm.atom_coords, prm.grp_exp = be.data_ptr(), ac.data_ptr(), gexp.data_ptr()
            o = N.QmcAoOut()
            for k, v in out.items():
                setattr(o, k, v.data_ptr())
            N.check(N.load().qmc_ao_eval_full(dp.handle, prm, x.data_ptr(), nw, nelec, o, N.current_stream()))
        return out

    def forward(self, pos: torch.Tensor, derivative: Optional[List[int]] = [0], sum_grad: Optional[bool] = True,
                sum_hess: Optional[bool] = True, one_elec: Optional[bool] = False, plan=None):
        """AO values (Nbatch, Nelec, Nao) or derivatives (atomic_orbitals.py:131-218):
        [1] summed gradient (Nbatch, Nelec, Nao) or per axis (..., 3) with
        sum_grad=False; [2] Laplacian, or the Hessian diagonal (..., 3) with
        sum_hess=False; [3] mixed second derivatives xy, xz, yz (..., 3);
        [0, 1, 2] (values, per-axis gradients, Laplacians).  one_elec=True takes
        the positions of one electron (Nbatch, 3)."""
        if not isinstance(derivative, list):
            derivative = [derivative]
        if not sum_grad:
            assert 1 in derivative
        if not sum_hess:
            assert 2 in derivative
        nelec = 1 if one_elec else self.nelec
        if pos.dim() != 2 or pos.shape[1] != 3 * nelec:
            raise ValueError("positions must have shape (Nbatch, %d)" % (3 * nelec))
        dt = torch.get_default_dtype()
        if derivative == [0]:
            return self._evaluate(pos, nelec, ["ao"], plan)["ao"].to(dt)
        if derivative == [1]:
            g = self._evaluate(pos, nelec, ["grad"], plan)["grad"]
            return (g.sum(-1) if sum_grad else g).to(dt)
        if derivative == [2]:
            if sum_hess:
                return self._evaluate(pos, nelec, ["lap"], plan)["lap"].to(dt)
            return self._evaluate(pos, nelec, ["hess_diag"], plan)["hess_diag"].to(dt)
        if derivative == [3]:
            return self._evaluate(pos, nelec, ["hess_mixed"], plan)["hess_mixed"].to(dt)
        if derivative == [0, 1, 2]:
            o = self._evaluate(pos, nelec, ["ao", "grad", "lap"], plan)
            return o["ao"].to(dt), o["grad"].to(dt), o["lap"].to(dt)
        raise ValueError("derivative must be 0, 1, 2, 3 or [0, 1, 2], got ", derivative)

    def update(self, ao: torch.Tensor, pos: torch.Tensor, idelec: int) -> torch.Tensor:
        """AO matrix with electron idelec's row recomputed at its new position
        (atomic_orbitals.py:671-695)."""
        ao_new = ao.clone()
        ids, ide = idelec * 3, (idelec + 1) * 3
        ao_new[:, idelec, :] = self.forward(pos[:, ids:ide], one_elec=True).squeeze(1)
        return ao_new
