"""Sampler base class (sampler/sampler_base.py:7-91 of the reference)."""
from typing import Callable, Dict

import torch

from .. import log
from .walkers import Walkers


class SamplerBase:
    def __init__(self, nwalkers: int, nstep: int, step_size: float, ntherm: int, ndecor: int, nelec: int,
                 ndim: int, init: Dict, cuda: bool) -> None:
        self.nelec = nelec
        self.ndim = ndim
        self.nstep = nstep
        self.step_size = step_size
        self.ntherm = ntherm
        self.ndecor = ndecor
        self.cuda = cuda
        self.device = torch.device("cuda" if cuda else "cpu")
        self.walkers = Walkers(nwalkers=nwalkers, nelec=nelec, ndim=ndim, init=init, cuda=cuda)
        log.info("")
        log.info(" Monte-Carlo Sampler")
        log.info("  Number of walkers   : {0}", self.walkers.nwalkers)
        log.info("  Number of steps     : {0}", self.nstep)
        log.info("  Step size           : {0}", self.step_size)
        log.info("  Thermalization steps: {0}", self.ntherm)
        log.info("  Decorelation steps  : {0}", self.ndecor)
        log.info("  Walkers init pos    : {0}", init["method"])

    def __call__(self, pdf: Callable, *args, **kwargs) -> torch.Tensor:
        raise NotImplementedError("Sampler must have a __call__ method")

    def __repr__(self) -> str:
        return self.__class__.__name__ + " sampler with  %d walkers" % self.walkers.nwalkers

    def get_sampling_size(self) -> int:
        """Number of sampled points (sampler_base.py:84-91)."""
        if self.ntherm == -1:
            return self.walkers.nwalkers
        return self.walkers.nwalkers * int((self.nstep - self.ntherm) / self.ndecor)
