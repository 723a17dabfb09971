"""WaveFunction base class (wavefunction/wf_base.py:7-277 of the reference).

Statistics helpers and the ``load`` of trained parameters keep the reference
behaviour; the potentials are computed by the fused HIP kernels as part of
``local_energy`` (``electronic_potential`` / ``nuclear_potential`` are exposed
through the same kernel's output).
"""
from typing import Optional

import numpy as np
import torch


class WaveFunction(torch.nn.Module):
    def __init__(self, nelec: int, ndim: int, kinetic: str = "jacobi", cuda: bool = False):
        super().__init__()
        self.ndim = ndim
        self.nelec = nelec
        self.ndim_tot = nelec * ndim
        self.kinetic = kinetic
        self.cuda = cuda
        self.device = torch.device("cuda" if cuda else "cpu")

    def forward(self, x: torch.Tensor):
        raise NotImplementedError()

    def energy(self, pos: torch.Tensor) -> torch.Tensor:
        """Mean local energy (wf_base.py:217-219)."""
        return torch.mean(self.local_energy(pos))

    def variance(self, pos: torch.Tensor) -> torch.Tensor:
        """Variance of the local energy (wf_base.py:221-223)."""
        return torch.var(self.local_energy(pos))

    def sampling_error(self, eloc: torch.Tensor) -> torch.Tensor:
        """sqrt(var/N) assuming uncorrelated samples (wf_base.py:225-229)."""
        return torch.sqrt(eloc.var() / eloc.shape[0])

    def _energy_variance(self, pos: torch.Tensor):
        el = self.local_energy(pos)
        return torch.mean(el), torch.var(el)

    def _energy_variance_error(self, pos: torch.Tensor):
        el = self.local_energy(pos)
        return torch.mean(el), torch.var(el), self.sampling_error(el)

    def pdf(self, pos: torch.Tensor, return_grad: Optional[bool] = False) -> torch.Tensor:
        """psi^2 (wf_base.py:241-247)."""
        if return_grad:
            return self.gradients(pos, pdf=True)
        return (self.forward(pos) ** 2).reshape(-1)

    def get_number_parameters(self) -> int:
        return sum(p.numel() for p in self.parameters() if p.requires_grad)

    def load(self, filename: str, group: Optional[str] = "wf_opt", model: Optional[str] = "best"):
        """Load trained parameters (wf_base.py:257-277): group/models/<model>
        of the solver's HDF5 output (read with h5py or, where it is absent,
        this package's reader), or an .npz/.pt state dict."""
        if filename.endswith(".npz"):
            d = np.load(filename, allow_pickle=False)
            pre = "%s/models/%s/" % (group, model)
            data = {k[len(pre):]: torch.as_tensor(d[k]) for k in d.files if k.startswith(pre)}
        elif filename.endswith(".pt") or filename.endswith(".pth"):
            data = torch.load(filename, weights_only=True)
        else:
            from ..utils.hdf5_utils import open_file

            with open_file(filename, "r") as f5:
                grp = f5[group]["models"][model]
                data = {name: torch.as_tensor(np.asarray(val[()])) for name, val in grp.items()}
        self.load_state_dict(data)
