"""Slater determinant pooling metadata (pooling/slater_pooling.py:11-849).

The determinants, inverses and traces tr(A^-1 B) are computed by the fused
HIP kernels (row-per-lane Gauss-Jordan in FP64); this module keeps the
configuration bookkeeping of the reference (``configs``, ``nconfs``,
unique configurations per spin).
"""
import torch
from torch import nn


class SlaterPooling(nn.Module):
    def __init__(self, config_method, configs, mol, cuda: bool = False) -> None:
        super().__init__()
        self.config_method = config_method
        self.configs = configs
        self.nconfs = len(configs[0])
        self.nmo = int(mol.basis.nmo)
        self.nup, self.ndown = mol.nup, mol.ndown
        self.nelec = self.nup + self.ndown
        self.index_max_orb_up = int(configs[0].max()) + 1 if configs[0].numel() else 0
        self.index_max_orb_down = int(configs[1].max()) + 1 if configs[1].numel() else 0
        self.device = torch.device("cuda" if cuda else "cpu")

    def forward(self, mo):
        raise NotImplementedError("Slater determinants are evaluated inside the fused wavefunction kernels")
