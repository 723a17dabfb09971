"""Electron-nucleus Pade kernel parameters (elec_nuclei/kernels/pade_jastrow_kernel.py:8-116).

B_iA = r_iA / (1 + w r_iA): the static weight is +1 for every nucleus (:33),
one trainable scalar ``weight``.
"""
import torch
from torch import nn


class PadeJastrowKernel(nn.Module):
    def __init__(self, nup: int, ndown: int, atomic_pos: torch.Tensor, cuda: bool, w: float = 1.0) -> None:
        super().__init__()
        self.nup, self.ndown = nup, ndown
        self.atoms = torch.as_tensor(atomic_pos)
        self.cuda = cuda
        self.device = torch.device("cuda" if cuda else "cpu")
        self.weight = nn.Parameter(torch.as_tensor([w], dtype=torch.get_default_dtype()), requires_grad=True)
        self.static_weight = torch.as_tensor([1.0])
        self.requires_autograd = True

    def forward(self, r: torch.Tensor) -> torch.Tensor:
        return self.static_weight.to(r) * r / (1.0 + self.weight * r)
