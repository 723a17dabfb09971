"""Metropolis sampler (sampler/metropolis.py:10-298).

Same constructor, ``__call__(pdf, pos=None, with_tqdm=True)`` contract and
semantics as the reference: all-electron ("all-elec"), one-electron
("one-elec": one uniformly chosen electron per walker and step) and
electron-by-electron ("all-elec-iter": every electron in turn, one
accept/reject each) Gaussian proposals with covariance
sigma = step/(2 sqrt(2 ln 2)) per coordinate (metropolis.py:206-212; the
per-coordinate standard deviation is sqrt(sigma)), uniform-cube proposals
step*(2U-1), acceptance min(1, psi'^2/psi^2) - U >= 0, psi = 0 treated as
psi^2 = eps, ``ntherm=-1`` keeps only the last positions, ``ndecor`` thins.

Two paths:

* ``SlaterJastrow.pdf`` of this package with "all-elec" / "one-elec" moves runs
  inside ``qmc_metropolis_sweep``: the whole chain stays on the device, 50
  steps per persistent launch, Philox4x32-10 proposals;
* any other density callable (and "all-elec-iter" moves) runs the reference's
  step loop on device tensors: torch proposals and uniforms from a generator
  seeded like the sweep, one density evaluation per (sub)step, no host copies
  inside the loop; ``logspace=True`` compares log densities.

Random numbers come from a seed drawn from torch's global generator, so
``torch.manual_seed`` makes runs reproducible, but the streams differ from the
reference's host MultivariateNormal draws (chains agree statistically, not
sample by sample).
"""
import math
from time import time
from typing import Callable, Dict, Union

import torch
from tqdm import tqdm

from .. import _native as N
from .. import log
from .sampler_base import SamplerBase

# steps per kernel launch: long enough to amortise the launch, short enough to
# keep every launch well under a second for the largest systems
STEPS_PER_LAUNCH = 50


class Metropolis(SamplerBase):
    def __init__(self, nwalkers: int = 100, nstep: int = 1000, step_size: float = 0.2, ntherm: int = -1,
                 ndecor: int = 1, nelec: int = 1, ndim: int = 3, init: Dict = {"min": -5, "max": 5},
                 move: Dict = {"type": "all-elec", "proba": "normal"}, logspace: bool = False, symmetry=None,
                 cuda: bool = False, seed=None) -> None:
        if "method" not in init:
            init = dict(init)
            init["method"] = "uniform" if "min" in init else "custom"
        SamplerBase.__init__(self, nwalkers, nstep, step_size, ntherm, ndecor, nelec, ndim, init, cuda)
        self.logspace = logspace
        self.seed = seed
        self.configure_move(move)
        self.symmetry = (lambda x: x) if symmetry is None else symmetry
        self.last_acceptance = None
        self.last_walker_steps_per_sec = None
        log.info("  Move type           : {0}", self.movedict["type"])
        log.info("  Move proba          : {0}", self.movedict["proba"])

    def configure_move(self, move: Dict) -> None:
        """metropolis.py:179-225"""
        self.movedict = dict(move)
        self.movedict.setdefault("type", "one-elec")
        self.movedict.setdefault("proba", "uniform")
        if self.movedict["type"] not in ["one-elec", "all-elec", "all-elec-iter"]:
            raise ValueError(" 'type' in move should be 'one-elec','all-elec', 'all-elec-iter'")
        if self.movedict["proba"] not in ["normal", "uniform"]:
            raise ValueError("'proba' in move should be 'normal' or 'uniform'")
        self.sigma2 = self.step_size / (2 * math.sqrt(2 * math.log(2.0)))
        self._move_per_iter = self.nelec if self.movedict["type"] == "all-elec-iter" else 1

    def _draw_seed(self) -> int:
        if self.seed is not None:
            return int(self.seed)
        return int(torch.randint(0, 2 ** 62, (1,)).item())

    def __call__(self, pdf: Callable, pos: Union[None, torch.Tensor] = None, with_tqdm: bool = True) -> torch.Tensor:
        from ..wavefunction.slater_jastrow import SlaterJastrow

        wf = getattr(pdf, "__self__", None)
        fused = (isinstance(wf, SlaterJastrow) and getattr(pdf, "__name__", "") == "pdf" and not self.logspace
                 and self.movedict["type"] in ("all-elec", "one-elec"))
        if not fused:
            return self._call_generic(pdf, pos, with_tqdm)
        if self.ntherm >= self.nstep:
            raise ValueError("Thermalisation longer than trajectory")
        if self.ntherm < 0:
            self.ntherm = self.nstep + self.ntherm  # same mutation as metropolis.py:113-114

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
        nkeep = (self.nstep - self.ntherm + self.ndecor - 1) // self.ndecor
        keep = torch.empty(nkeep, nw, x.shape[1], dtype=torch.float64, device=dev)

        plan = wf.device_plan()
        prm, params_keep = wf._params(wf.precision)
        prec = N.QMC_PREC_MIXED if wf.precision == "mixed" else N.QMC_PREC_FP64
        wk = N.QmcWalkers()
        wk.pos, wk.logabs, wk.sign, wk.naccept, wk.walker_id = (x.data_ptr(), logabs.data_ptr(), sign.data_ptr(),
                                                               nacc.data_ptr(), 0)
        sw = N.QmcSweepDesc()
        sw.sigma2 = self.sigma2
        sw.seed = self._draw_seed()
        sw.keep_from = self.ntherm
        sw.keep_every = self.ndecor
        sw.uniform_move = 1 if self.movedict["proba"] == "uniform" else 0
        sw.one_electron = 1 if self.movedict["type"] == "one-elec" and self.nelec > 1 else 0
        sw.step_size = self.step_size
        sw.keep = keep.data_ptr()
        lib = N.load()
        stream = N.current_stream()
        chunks = range(0, self.nstep, STEPS_PER_LAUNCH)
        rng = tqdm(chunks, desc="INFO:QMCTorch|  Sampling", disable=not with_tqdm)
        tstart = time()
        for s0 in rng:
            sw.step0 = s0
            sw.nstep = min(STEPS_PER_LAUNCH, self.nstep - s0)
            sw.init_logpsi = 1 if s0 == 0 else 0
            N.check(lib.qmc_metropolis_sweep(plan.handle, prm, prec, wk, nw, sw, stream))
        rate = nacc.sum().item() / float(nw * self.nstep)
        elapsed = time() - tstart
        self.last_acceptance = rate
        self.last_walker_steps_per_sec = nw * self.nstep / max(elapsed, 1e-12)
        self.walkers.pos = x.to(torch.get_default_dtype())
        if with_tqdm:
            log.info("   Acceptance rate     : {:1.2f} %", rate * 100)
            log.info("   Timing statistics   : {:1.2f} steps/sec.", self.nstep / elapsed)
            log.info("   Total Time          : {:1.2f} sec.", elapsed)
        out = keep.view(-1, x.shape[1]).to(torch.get_default_dtype())
        return self.symmetry(out).requires_grad_()

    # ------------------------------------------------------------- generic path
    def _call_generic(self, pdf: Callable, pos, with_tqdm: bool) -> torch.Tensor:
        """The step loop of metropolis.py:85-177 on device tensors for any density."""
        eps = 1e-7 if torch.get_default_dtype() == torch.float32 else 1e-16
        if self.ntherm >= self.nstep:
            raise ValueError("Thermalisation longer than trajectory")
        with torch.no_grad():
            if self.ntherm < 0:
                self.ntherm = self.nstep + self.ntherm
            self.walkers.initialize(pos=pos)
            x = self.walkers.pos.detach().clone()
            if self.cuda and not x.is_cuda:
                x = x.to("cuda")
            dev = x.device
            gen = torch.Generator(device=dev)
            gen.manual_seed(self._draw_seed())
            nw, ne = x.shape[0], self.nelec
            density = (lambda p: torch.log(pdf(p))) if self.logspace else pdf
            fx = density(x).reshape(-1)
            if not self.logspace:
                fx[fx == 0] = eps
            sd = math.sqrt(self.sigma2)
            mtype = self.movedict["type"]
            subs = list(range(ne)) if (mtype == "all-elec-iter" and ne > 1) else [None]
            kept, nacc, idecor = [], 0, 0
            rows = torch.arange(nw, device=dev)
            tstart = time()
            for istep in tqdm(range(self.nstep), desc="INFO:QMCTorch|  Sampling", disable=not with_tqdm):
                for sub in subs:
                    if ne == 1 or mtype == "all-elec":
                        xn = x + self._displacement(nw, ne, dev, gen, sd)
                    else:
                        if sub is None:
                            which = torch.randint(0, ne, (nw,), device=dev, generator=gen)
                        else:
                            which = torch.full((nw,), sub, dtype=torch.long, device=dev)
                        xn = x.clone().view(nw, ne, 3)
                        xn[rows, which, :] += self._displacement(nw, 1, dev, gen, sd)
                        xn = xn.view(nw, 3 * ne)
                    fxn = density(xn).reshape(-1)
                    if self.logspace:
                        tau = torch.log(torch.rand(nw, dtype=fx.dtype, device=dev, generator=gen))
                        acc = (torch.clamp(fxn - fx, max=0.0) - tau) >= 0
                    else:
                        fxn[fxn == 0.0] = eps
                        ratio = torch.clamp(fxn / fx, max=1.0)
                        acc = (ratio - torch.rand(nw, dtype=fx.dtype, device=dev, generator=gen)) >= 0
                    x = torch.where(acc[:, None], xn, x)
                    fx = torch.where(acc, fxn, fx)
                    if not self.logspace:
                        fx[fx == 0] = eps
                    nacc = nacc + acc.sum()
                if istep >= self.ntherm:
                    if idecor % self.ndecor == 0:
                        kept.append(x.clone())
                    idecor += 1
            elapsed = time() - tstart
            rate = float(nacc) / float(nw * self.nstep * len(subs))
            self.last_acceptance = rate
            self.last_walker_steps_per_sec = nw * self.nstep / max(elapsed, 1e-12)
            self.walkers.pos = x
            if with_tqdm:
                log.info("   Acceptance rate     : {:1.2f} %", rate * 100)
                log.info("   Timing statistics   : {:1.2f} steps/sec.", self.nstep / elapsed)
                log.info("   Total Time          : {:1.2f} sec.", elapsed)
        return self.symmetry(torch.cat(kept)).requires_grad_()

    def _displacement(self, nw: int, nel: int, dev, gen, sd: float) -> torch.Tensor:
        """Proposal displacements (metropolis.py:256-277): N(0, sigma I) per
        electron, or step (2U - 1) in a cube."""
        if self.movedict["proba"] == "uniform":
            u = torch.rand(nw, 3 * nel, dtype=torch.get_default_dtype(), device=dev, generator=gen)
            return self.step_size * (2.0 * u - 1.0)
        return sd * torch.randn(nw, 3 * nel, dtype=torch.get_default_dtype(), device=dev, generator=gen)
