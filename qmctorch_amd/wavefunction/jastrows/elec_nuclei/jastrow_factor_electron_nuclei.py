"""Electron-nucleus Jastrow factor (jastrow_factor_electron_nuclei.py:9-161).

The e-n centres are ``mol.atom_coords`` (:40), not the AO layer's trainable
coordinates; the factor is evaluated by the fused wavefunction kernels.
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

    def __repr__(self) -> str:
        return "en -> " + self.jastrow_kernel.__class__.__name__

    def forward(self, pos, derivative=0, sum_grad=True):
        raise NotImplementedError(
            "the e-n Jastrow is evaluated inside the fused wavefunction kernels "
            "(SlaterJastrow.forward / pdf / local_energy)"
        )
