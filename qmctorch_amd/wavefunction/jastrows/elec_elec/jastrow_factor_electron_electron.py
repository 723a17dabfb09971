"""Electron-electron Jastrow factor (jastrow_factor_electron_electron.py:11-260).

Holds the kernel module (state_dict key ``jastrow.jastrow_kernel.weight``) and
the metadata the fused kernels need; J = exp(sum_{i<j} B(r_ij)) and its
derivatives are evaluated on the device by the wavefunction kernels.
"""
import torch
from torch import nn


class JastrowFactorElectronElectron(nn.Module):
    kind = "ee"

    def __init__(self, mol, jastrow_kernel, kernel_kwargs={}, scale: bool = False, scale_factor: float = 0.6,
                 cuda: bool = False) -> None:
        super().__init__()
        if scale:
            raise NotImplementedError("scaled e-e distances are not supported by the HIP kernels")
        from .kernels.pade_jastrow_kernel import PadeJastrowKernel

        if jastrow_kernel is not PadeJastrowKernel:
            raise NotImplementedError("only the Pade e-e kernel is implemented on the device")
        self.nup, self.ndown = mol.nup, mol.ndown
        self.nelec = self.nup + self.ndown
        self.ndim = 3
        self.cuda = cuda
        self.device = torch.device("cuda" if cuda else "cpu")
        self.jastrow_kernel = jastrow_kernel(mol.nup, mol.ndown, cuda, **kernel_kwargs)
        self.requires_autograd = self.jastrow_kernel.requires_autograd

    def __repr__(self) -> str:
        return "ee -> " + self.jastrow_kernel.__class__.__name__

    def forward(self, pos, derivative=0, sum_grad=True):
        raise NotImplementedError(
            "the e-e Jastrow is evaluated inside the fused wavefunction kernels "
            "(SlaterJastrow.forward / pdf / local_energy)"
        )
