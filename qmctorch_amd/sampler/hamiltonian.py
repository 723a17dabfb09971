"""Hamiltonian Monte Carlo sampler driving the persistent HIP kernel
(sampler/hamiltonian.py:10-201).

Same constructor, ``__call__(pdf, pos=None, with_tqdm=True)`` contract and
semantics as the reference: U = -ln pdf, momenta p ~ N(0, 1) per coordinate,
a leapfrog trajectory of ``L`` position updates of ``step_size`` with half
momentum steps at both ends, acceptance iff exp(E_init - E_new) >= u with
E = U + |p|^2 / 2 and u ~ U[0,1), ``ntherm=-1`` keeps only the last
positions, ``ndecor`` thins.

Differences forced by the device design: grad U = -2 grad ln|psi| comes from
the drift evaluated inside ``qmc_hmc_sweep`` (the lane Gauss-Jordan inverse),
not from autograd through ``-log(pdf)``; the trajectory runs on the device
with the walker state in registers; momenta and uniforms come from
Philox4x32-10 keyed by a seed drawn from torch's global generator (chains agree
with the reference statistically, not sample by sample); a move whose energy
is NaN is rejected (the reference's ``exp(E_init - E_new) < eps`` test keeps
it).
"""
from time import time
from typing import Callable, Dict, Optional

import torch
from tqdm import tqdm

from .. import _native as N
from .. import log
from .sampler_base import SamplerBase

# wavefunction evaluations per kernel launch (steps per launch = this / L)
EVALS_PER_LAUNCH = 500


class Hamiltonian(SamplerBase):
    def __init__(self, nwalkers: int = 100, nstep: int = 100, step_size: float = 0.2, L: int = 10,
                 ntherm: int = -1, ndecor: int = 1, nelec: int = 1, ndim: int = 3,
                 init: Dict = {"min": -5, "max": 5}, cuda: bool = False, seed: Optional[int] = None) -> None:
        if "method" not in init:
            init = dict(init)
            init["method"] = "uniform" if "min" in init else "custom"
        SamplerBase.__init__(self, nwalkers, nstep, step_size, ntherm, ndecor, nelec, ndim, init, cuda)
        if int(L) < 1:
            raise ValueError("the trajectory length L must be >= 1")
        self.traj_length = int(L)
        self.seed = seed
        self.last_acceptance = None
        self.last_walker_steps_per_sec = None

    def _draw_seed(self) -> int:
        if self.seed is not None:
            return int(self.seed)
        return int(torch.randint(0, 2 ** 62, (1,)).item())

    def __call__(self, pdf: Callable, pos: Optional[torch.Tensor] = None, with_tqdm: bool = True) -> torch.Tensor:
        from ..wavefunction.slater_jastrow import SlaterJastrow

        wf = getattr(pdf, "__self__", None)
        if not isinstance(wf, SlaterJastrow) or getattr(pdf, "__name__", "") != "pdf":
            raise NotImplementedError("Hamiltonian samples SlaterJastrow.pdf with the fused HIP kernel; "
                                      "pass wf.pdf of a qmctorch_amd SlaterJastrow")
        if self.ntherm < 0:
            self.ntherm = self.nstep + self.ntherm  # same mutation as hamiltonian.py:99-100

        self.walkers.initialize(pos=pos)
        x = self.walkers.pos.detach()
        if not x.is_cuda:
            x = x.to("cuda")
        x = x.to(torch.float64).contiguous().clone()
        nw = x.shape[0]
        dev = x.device
        logabs = torch.empty(nw, dtype=torch.float64, device=dev)
        sign = torch.empty(nw, dtype=torch.float64, device=dev)
        nacc = torch.zeros(nw, dtype=torch.int32, device=dev)
        nkeep = max(0, (self.nstep - self.ntherm + self.ndecor - 1) // self.ndecor)
        keep = torch.empty(nkeep, nw, x.shape[1], dtype=torch.float64, device=dev)

        plan = wf.device_plan()
        prm, params_keep = wf._params(wf.precision)
        prec = N.QMC_PREC_MIXED if wf.precision == "mixed" else N.QMC_PREC_FP64
        wk = N.QmcWalkers()
        wk.pos, wk.logabs, wk.sign, wk.naccept, wk.walker_id = (x.data_ptr(), logabs.data_ptr(), sign.data_ptr(),
                                                               nacc.data_ptr(), 0)
        hm = N.QmcHmcDesc()
        hm.traj_length = self.traj_length
        hm.step_size = self.step_size
        hm.seed = self._draw_seed()
        hm.keep_from = self.ntherm
        hm.keep_every = self.ndecor
        hm.keep = keep.data_ptr() if nkeep > 0 else 0
        lib = N.load()
        stream = N.current_stream()
        per_launch = max(1, EVALS_PER_LAUNCH // self.traj_length)
        chunks = range(0, self.nstep, per_launch)
        rng = tqdm(chunks, desc="INFO:QMCTorch|  Sampling", disable=not with_tqdm)
        tstart = time()
        for s0 in rng:
            hm.step0 = s0
            hm.nstep = min(per_launch, self.nstep - s0)
            N.check(lib.qmc_hmc_sweep(plan.handle, prm, prec, wk, nw, hm, stream))
        rate = nacc.sum().item() / float(max(1, nw * self.nstep))
        elapsed = time() - tstart
        self.last_acceptance = rate
        self.last_walker_steps_per_sec = nw * self.nstep / max(elapsed, 1e-12)
        self.walkers.pos = x.to(torch.get_default_dtype())
        log.options(style="percent").debug("  Acceptance rate %1.3f %%" % (rate * 100))
        return keep.view(-1, x.shape[1]).to(torch.get_default_dtype()).requires_grad_()
