"""MO layer (molecular_orbitals.py:7-95): W = mo_scf * mo_modifier, MO = AO @ W.

``mo_modifier`` is the trainable parameter (state_dict key ``mo.mo_modifier``).
In the fused kernels the contraction happens in registers; ``forward`` here
serves direct calls on AO tensors and is a plain GEMM.
"""
import numpy as np
import torch
from torch import nn


class MolecularOrbitals(nn.Module):
    def __init__(self, mol, include_all_mo: bool, highest_occ_mo: int, mix_mo: bool, orthogonalize_mo: bool,
                 cuda: bool):
        super().__init__()
        if mix_mo:
            raise NotImplementedError("mix_mo is not supported (the reference path is broken as well)")
        if orthogonalize_mo:
            raise Warning("orthogonalize_mo=True has no effect as mix_mo=False")
        dtype = torch.get_default_dtype()
        self.mol = mol
        self.mix_mo = mix_mo
        self.orthogonalize_mo = orthogonalize_mo
        self.cuda = cuda
        self.device = torch.device("cuda" if cuda else "cpu")
        self.include_all_mo = include_all_mo
        self.highest_occ_mo = highest_occ_mo
        self.nmo_opt = int(mol.basis.nmo) if include_all_mo else highest_occ_mo
        self.mo_scf = self.get_mo_coeffs()
        self.mo_modifier = nn.Parameter(torch.ones_like(self.mo_scf, requires_grad=True)).type(dtype)
        self.mo_mixer = None
        if cuda:
            self.mo_scf = self.mo_scf.to(self.device)
            self.to(self.device)

    def get_mo_coeffs(self) -> torch.Tensor:
        mo = torch.as_tensor(np.asarray(self.mol.basis.mos)).type(torch.get_default_dtype())
        if not self.include_all_mo:
            mo = mo[:, : self.highest_occ_mo]
        return mo.requires_grad_(False)

    def weight(self) -> torch.Tensor:
        return self.mo_scf.to(self.mo_modifier.device) * self.mo_modifier

    def forward(self, ao: torch.Tensor) -> torch.Tensor:
        w = self.weight()
        return ao @ w.reshape(1, *w.shape)
