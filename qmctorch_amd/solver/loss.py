"""Loss of the wavefunction optimisation (solver/loss.py:7-145 of the reference)."""
from typing import Tuple

import torch
from torch import nn


class Loss(nn.Module):
    def __init__(self, wf, method: str = "energy", clip: bool = False, clip_threshold: int = 5):
        super().__init__()
        self.wf = wf
        self.method = method
        self.clip = clip
        self.use_weight = True
        self.clip_num_std = clip_threshold
        self.loss_fn = {"energy": torch.mean, "variance": torch.var}[method]
        self.weight = {"psi": None, "psi0": None}

    def forward(self, pos: torch.Tensor, no_grad: bool = False, deactivate_weight: bool = False):
        with (torch.no_grad() if no_grad else torch.enable_grad()):
            local_energies = self.wf.local_energy_autodiff(pos)
            mask = self.get_clipping_mask(local_energies)
            weight = self.get_sampling_weights(pos, deactivate_weight)
            loss = self.loss_fn((weight * local_energies)[mask])
        return loss, local_energies

    def get_clipping_mask(self, local_energies: torch.Tensor) -> torch.Tensor:
        if self.clip:
            median = torch.median(local_energies)
            std = torch.std(local_energies)
            return torch.abs((local_energies - median) / std) < self.clip_num_std
        return torch.ones_like(local_energies).type(torch.bool)

    def get_sampling_weights(self, pos: torch.Tensor, deactivate_weight: bool) -> torch.Tensor:
        if self.use_weight and not deactivate_weight:
            self.weight["psi"] = self.wf(pos)
            if self.weight["psi0"] is None:
                self.weight["psi0"] = self.weight["psi"].detach().clone()
                return torch.ones_like(self.weight["psi"])
            w = (self.weight["psi"] / self.weight["psi0"]) ** 2
            return w / w.sum()
        return torch.tensor(1.0)
