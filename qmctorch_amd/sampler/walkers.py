"""Walker initialisation (sampler/walkers.py:8-150 of the reference).

Initial positions are drawn on the host with torch's / numpy's global
generators exactly like the reference (so ``torch.manual_seed`` reproduces
them) and then moved to the device once; the Metropolis chain itself stays on
the device.
"""
from typing import Dict, Optional, Union

import numpy as np
import torch

from .. import log


class Walkers:
    def __init__(self, nwalkers: int = 100, nelec: int = 1, ndim: int = 3, init: Optional[Dict] = None,
                 cuda: bool = False) -> None:
        self.nwalkers = nwalkers
        self.ndim = ndim
        self.nelec = nelec
        self.init_domain = init
        self.pos = None
        self.status = None
        self.cuda = cuda
        self.device = torch.device("cuda" if cuda else "cpu")

    def initialize(self, pos: Union[None, torch.Tensor] = None) -> None:
        if self.cuda:
            self.device = torch.device("cuda")
        if pos is not None:
            if len(pos) > self.nwalkers:
                pos = pos[-self.nwalkers :, :]
            self.pos = pos
            return
        log.debug("  Initialize walkers")
        d = self.init_domain
        if "center" in d:
            self.pos = self._init_center()
        elif "min" in d:
            self.pos = self._init_uniform()
        elif "mean" in d:
            self.pos = self._init_multivar()
        elif "atom_coords" in d:
            self.pos = self._init_atomic()
        else:
            raise ValueError("Init walkers not recognized")

    def _init_center(self) -> torch.Tensor:
        eps = 1e-3
        pos = -eps + 2 * eps * torch.rand(self.nwalkers, self.nelec * self.ndim)
        return pos.type(torch.get_default_dtype()).to(device=self.device)

    def _init_uniform(self) -> torch.Tensor:
        pos = torch.rand(self.nwalkers, self.nelec * self.ndim)
        pos *= self.init_domain["max"] - self.init_domain["min"]
        pos += self.init_domain["min"]
        return pos.type(torch.get_default_dtype()).to(device=self.device)

    def _init_multivar(self) -> torch.Tensor:
        multi = torch.distributions.MultivariateNormal(torch.as_tensor(self.init_domain["mean"]),
                                                       torch.as_tensor(self.init_domain["sigma"]))
        pos = multi.sample((self.nwalkers, self.nelec)).type(torch.get_default_dtype())
        return pos.view(self.nwalkers, self.nelec * self.ndim).to(device=self.device)

    def _init_atomic(self) -> torch.Tensor:
        pos = torch.zeros(self.nwalkers, self.nelec * self.ndim)
        idx_ref, nelec_tot, natom = [], 0, 0
        for iat, ne in enumerate(self.init_domain["atom_nelec"]):
            idx_ref += [iat] * ne
            nelec_tot += ne
            natom += 1
        for iw in range(self.nwalkers):
            placed = [0] * natom
            idx = torch.as_tensor(idx_ref)[torch.randperm(nelec_tot)]
            xyz = torch.as_tensor(self.init_domain["atom_coords"])[idx, :]
            for ie in range(nelec_tot):
                a = idx[ie]
                z = self.init_domain["atom_num"][a]
                if placed[a] == 0:
                    s = 1.0 / z
                elif placed[a] < 5:
                    s = 2.0 / (z - 2)
                else:
                    s = 3.0 / (z - 3)
                xyz[ie, :] += np.random.normal(scale=s, size=(1, 3))
                placed[a] += 1
            pos[iw, :] = xyz.view(-1)
        return pos.to(device=self.device)
