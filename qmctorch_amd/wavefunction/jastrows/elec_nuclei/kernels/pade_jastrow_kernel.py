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

    def compute_derivative(self, r: torch.Tensor, dr: torch.Tensor) -> torch.Tensor:
        """d B / d x_k = w0 (dr/dx_k) / (1 + w r)^2 per axis, shape of dr
        (Nbatch, 3, ...) (pade_jastrow_kernel.py:51-79): the
        reference's two terms w0 dr/(1+wr) - w0 w r dr/(1+wr)^2 combined."""
        q = 1.0 / (1.0 + self.weight * r.unsqueeze(1))
        return self.static_weight.to(r) * dr * q * q

    def compute_second_derivative(self, r: torch.Tensor, dr: torch.Tensor, d2r: torch.Tensor) -> torch.Tensor:
        """d^2 B / d x_k^2 = w0 [d2r / (1 + w r)^2 - 2 w dr^2 / (1 + w r)^3]
        (pade_jastrow_kernel.py:81-116, its four terms combined)."""
        q = 1.0 / (1.0 + self.weight * r.unsqueeze(1))
        q2 = q * q
        return self.static_weight.to(r) * (d2r * q2 - 2.0 * self.weight * dr * dr * q2 * q)
