"""Electron-electron Jastrow factor (jastrow_factor_electron_electron.py:11-260).

Holds the kernel module (state_dict key ``jastrow.jastrow_kernel.weight``);
inside a wavefunction J = exp(sum_{i<j} B(r_ij)) and its derivatives are
evaluated by the fused wavefunction kernels, and ``forward`` evaluates them on
their own (``qmc_jastrow_eval``) with the reference's output shapes.
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
        self.mol = mol

    def __repr__(self) -> str:
        return "ee -> " + self.jastrow_kernel.__class__.__name__

    def forward(self, pos, derivative=0, sum_grad=True):
        """J (Nbatch, 1); derivative=1: dJ summed over x, y, z (Nbatch, Nelec) or
        per axis (Nbatch, 3, Nelec) with sum_grad=False; derivative=2: the
        Laplacian of J per electron (Nbatch, Nelec); [0, 1, 2]: the three
        (jastrow_factor_electron_electron.py:124-175)."""
        from .._device import evaluate, shaped

        need = derivative if isinstance(derivative, list) else [derivative]
        j, dj, d2j = evaluate(self, self.mol, self, None, pos, 1 in need, 2 in need)
        return shaped(j, dj, d2j, derivative, sum_grad)
