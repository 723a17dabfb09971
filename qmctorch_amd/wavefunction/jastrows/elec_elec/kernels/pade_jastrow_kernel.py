"""Electron-electron Pade kernel parameters (elec_elec/kernels/pade_jastrow_kernel.py:8-153).

B_ij = w0 r_ij / (1 + w r_ij) with w0 = 0.25 for parallel and 0.5 for
antiparallel spins (static weights, :34-66) and one trainable scalar ``weight``.
The kernel itself is evaluated inside the fused HIP wavefunction kernels.
"""
import torch
from torch import nn


class PadeJastrowKernel(nn.Module):
    def __init__(self, nup: int, ndown: int, cuda: bool, w: float = 1.0) -> None:
        super().__init__()
        self.nup, self.ndown = nup, ndown
        self.nelec = nup + ndown
        self.cuda = cuda
        self.device = torch.device("cuda" if cuda else "cpu")
        self.weight = nn.Parameter(torch.as_tensor([w], dtype=torch.get_default_dtype()), requires_grad=True)
        self.static_weight = self.get_static_weight()
        self.requires_autograd = False

    def get_static_weight(self) -> torch.Tensor:
        n = self.nelec
        sw = torch.full((n, n), 0.5)
        sw[: self.nup, : self.nup] = 0.25
        sw[self.nup :, self.nup :] = 0.25
        return sw.masked_select(torch.triu(torch.ones(n, n), diagonal=1).bool())

    def forward(self, r: torch.Tensor) -> torch.Tensor:
        """Kernel values on a (Nbatch, Npair) tensor of distances."""
        return self.static_weight.to(r) * r / (1.0 + self.weight * r)

    def compute_derivative(self, r: torch.Tensor, dr: torch.Tensor) -> torch.Tensor:
        """d B / d x_k = w0 (dr/dx_k) / (1 + w r)^2 per axis, shape of dr
        (Nbatch, 3, ...) (pade_jastrow_kernel.py:84-117): the
        reference's two terms w0 dr/(1+wr) - w0 w r dr/(1+wr)^2 combined."""
        q = 1.0 / (1.0 + self.weight * r.unsqueeze(1))
        return self.static_weight.to(r) * dr * q * q

    def compute_second_derivative(self, r: torch.Tensor, dr: torch.Tensor, d2r: torch.Tensor) -> torch.Tensor:
        """d^2 B / d x_k^2 = w0 [d2r / (1 + w r)^2 - 2 w dr^2 / (1 + w r)^3]
        (pade_jastrow_kernel.py:119-153, its four terms combined)."""
        q = 1.0 / (1.0 + self.weight * r.unsqueeze(1))
        q2 = q * q
        return self.static_weight.to(r) * (d2r * q2 - 2.0 * self.weight * dr * dr * q2 * q)
