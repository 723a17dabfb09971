"""Electron-nucleus Jastrow factor (jastrow_factor_electron_nuclei.py:9-161).

The e-n centres are ``mol.atom_coords`` (:40), not the AO layer's trainable
coordinates.  Inside a wavefunction the factor is evaluated by the fused
wavefunction kernels; ``forward`` evaluates it on its own (``qmc_jastrow_eval``)
with the reference's output shapes.
"""
import numpy as np
import torch
from torch import nn


class JastrowFactorElectronNuclei(nn.Module):
    kind = "en"

    def __init__(self, mol, jastrow_kernel, kernel_kwargs={}, cuda: bool = False) -> None:
        super().__init__()
        from .kernels.pade_jastrow_kernel import PadeJastrowKernel

        if jastrow_kernel is not PadeJastrowKernel:
            raise NotImplementedError("only the Pade e-n kernel is implemented on the device")
        self.nup, self.ndown = mol.nup, mol.ndown
        self.nelec = self.nup + self.ndown
        self.cuda = cuda
        self.device = torch.device("cuda" if cuda else "cpu")
        atomic_pos = torch.as_tensor(np.asarray(mol.atom_coords), dtype=torch.float64)
        self.atoms = atomic_pos
        self.natoms = self.atoms.shape[0]
        self.ndim = 3
        self.jastrow_kernel = jastrow_kernel(mol.nup, mol.ndown, atomic_pos, cuda, **kernel_kwargs)
        self.requires_autograd = self.jastrow_kernel.requires_autograd
        self.mol = mol

    def __repr__(self) -> str:
        return "en -> " + self.jastrow_kernel.__class__.__name__

    def forward(self, pos, derivative=0, sum_grad=True):
        """J (Nbatch, 1); derivative=1: (Nbatch, Nelec) summed over x, y, z or
        (Nbatch, 3, Nelec) with sum_grad=False; derivative=2: Laplacian of J per
        electron (Nbatch, Nelec); [0, 1, 2]: the three (jastrow_factor_electron_nuclei.py:60-161)."""
        from .._device import evaluate, shaped

        need = derivative if isinstance(derivative, list) else [derivative]
        j, dj, d2j = evaluate(self, self.mol, None, self, pos, 1 in need, 2 in need)
        return shaped(j, dj, d2j, derivative, sum_grad)
