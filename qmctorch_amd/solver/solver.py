"""Solver: wavefunction optimisation (solver/solver.py:16-537 of the reference).

``run`` / ``run_epochs`` / resampling / checkpoints / observables follow the
reference.  The manual gradient (default) is psi.backward(weight) exactly as
solver.py:372-431, with psi's backward being the hand-written adjoint kernel
(no autograd graph is recorded); ``grad='auto'`` back-propagates mean(E_L)
through the local-energy adjoint.
"""
from time import time
from typing import Any, Dict, List, Optional, Tuple

import torch
from tqdm import tqdm

from .. import log
from ..utils.hdf5_utils import add_group_attr, dump_to_hdf5
from ..utils.torch_utils import DataLoader
from .loss import Loss
from .solver_base import SolverBase


class Solver(SolverBase):
    def __init__(self, wf=None, sampler=None, optimizer=None, scheduler=None, output: Optional[str] = None,
                 rank: int = 0) -> None:
        SolverBase.__init__(self, wf, sampler, optimizer, scheduler, output, rank)
        self.set_params_requires_grad()
        self.configure(track=["local_energy"], freeze=None, loss="energy", grad="manual", ortho_mo=False,
                       clip_loss=False, resampling={"mode": "update", "resample_every": 1, "nstep_update": 25})

    def configure(self, track: Optional[List[str]] = None, freeze: Optional[List[str]] = None,
                  loss: Optional[str] = None, grad: Optional[str] = None, ortho_mo: Optional[bool] = None,
                  clip_loss: bool = False, clip_threshold: int = 5, resampling: Optional[Dict[str, Any]] = None):
        """solver.py:50-111"""
        self.set_params_requires_grad()
        self.freeze_params_list = freeze
        self.freeze_parameters(freeze)
        if track is not None:
            self.track_observable(track)
        if grad is not None:
            self.grad_method = grad
            self.evaluate_gradient = {"auto": self.evaluate_grad_auto, "manual": self.evaluate_grad_manual}[grad]
        if resampling is not None:
            self.configure_resampling(**resampling)
        if loss is not None:
            self.loss = Loss(self.wf, method=loss, clip=clip_loss, clip_threshold=clip_threshold)
            self.loss.use_weight = self.resampling_options.resample_every > 1
        self.ortho_mo = ortho_mo
        if ortho_mo is True:
            log.warning("Orthogonalization of the MO coeffs via loss penalty is deprecated")

    def set_params_requires_grad(self, wf_params: Optional[bool] = True, geo_params: Optional[bool] = False):
        """solver.py:113-133"""
        self.wf.ao.bas_exp.requires_grad = wf_params
        for p in self.wf.mo.parameters():
            p.requires_grad = wf_params
        self.wf.fc.weight.requires_grad = wf_params
        if getattr(self.wf, "jastrow", None) is not None:
            for p in self.wf.jastrow.parameters():
                p.requires_grad = wf_params
        self.wf.ao.atom_coords.requires_grad = geo_params

    def freeze_parameters(self, freeze: Optional[List[str]]) -> None:
        """solver.py:135-167"""
        if freeze is None:
            return
        if not isinstance(freeze, list):
            freeze = [freeze]
        for name in freeze:
            n = name.lower()
            if n == "ci":
                self.wf.fc.weight.requires_grad = False
            elif n == "mo":
                for p in self.wf.mo.parameters():
                    p.requires_grad = False
            elif n == "ao":
                self.wf.ao.bas_exp.requires_grad = False
            elif n == "jastrow":
                for p in self.wf.jastrow.parameters():
                    p.requires_grad = False
            else:
                raise ValueError("Valid arguments for freeze are :", ["ci", "mo", "ao", "jastrow", "backflow"])

    def save_sampling_parameters(self) -> None:
        self.sampler._nstep_save = self.sampler.nstep
        self.sampler._ntherm_save = self.sampler.ntherm
        if self.resampling_options.mode == "update":
            self.sampler.ntherm = self.resampling_options.ntherm_update
            self.sampler.nstep = self.resampling_options.nstep_update

    def restore_sampling_parameters(self) -> None:
        self.sampler.nstep = self.sampler._nstep_save
        self.sampler.ntherm = self.sampler._ntherm_save

    def run(self, nepoch: int, batchsize: Optional[int] = None, hdf5_group: Optional[str] = "wf_opt",
            chkpt_every: Optional[int] = None, tqdm: Optional[bool] = False):
        """solver.py:186-219"""
        self.prepare_optimization(batchsize, chkpt_every, tqdm)
        self.log_data_opt(nepoch, "wave function optimization")
        self.run_epochs(nepoch)
        self.restore_sampling_parameters()
        self.save_data(hdf5_group)
        return self.observable

    def prepare_optimization(self, batchsize: Optional[int], chkpt_every: Optional[int], tqdm: bool = False):
        log.info("  Initial Sampling    :")
        tstart = time()
        with torch.no_grad():
            pos = self.sampler(self.wf.pdf, with_tqdm=tqdm)
        if batchsize is None:
            batchsize = len(pos)
        self.save_sampling_parameters()
        self.dataloader = DataLoader(pos, batch_size=batchsize, pin_memory=False)
        with torch.no_grad():
            for ibatch, data in enumerate(self.dataloader):
                self.store_observable(data, ibatch=ibatch)
        self.chkpt_every = chkpt_every
        log.info("  done in %1.2f sec." % (time() - tstart))

    def save_data(self, hdf5_group: str):
        self.observable.models.last = dict(self.wf.state_dict())
        grp = dump_to_hdf5(self.observable, self.hdf5file, hdf5_group)
        add_group_attr(self.hdf5file, grp, {"type": "opt"})

    def run_epochs(self, nepoch: int, with_tqdm: Optional[bool] = False, verbose: Optional[bool] = True) -> float:
        """solver.py:266-350 (NaN guard, best model, checkpoints, resampling)."""
        if with_tqdm and verbose:
            raise ValueError("tqdm and verbose are mutually exclusive")
        cumulative_loss, min_loss = 0, 0
        rng = tqdm(range(nepoch), desc="INFO:QMCTorch|  Optimization", disable=not with_tqdm)
        for n in rng:
            tstart = time()
            if verbose:
                log.info("")
                log.info("  epoch %d | %d sampling points" % (n, len(self.dataloader.dataset)))
            cumulative_loss = 0
            self.opt.zero_grad()
            self.wf.zero_grad()
            for ibatch, data in enumerate(self.dataloader):
                lpos = data.to(self.device)
                loss, eloc = self.evaluate_gradient(lpos)
                cumulative_loss += loss.item()
                if torch.isnan(eloc).any():
                    log.info("Error : Nan detected in local energy")
                    return cumulative_loss
                self.store_observable(lpos, local_energy=eloc, ibatch=ibatch)
            self.optimization_step(lpos)
            if n == 0 or cumulative_loss < min_loss:
                min_loss = cumulative_loss
                self.observable.models.best = dict(self.wf.state_dict())
            if self.chkpt_every is not None and n > 0 and n % self.chkpt_every == 0:
                self.save_checkpoint(n, cumulative_loss)
            if verbose:
                self.print_observable(cumulative_loss, verbose=False)
            self.dataloader.dataset = self.resample(n, self.dataloader.dataset)
            if self.scheduler is not None:
                self.scheduler.step()
            if verbose:
                log.info("  epoch done in %1.2f sec." % (time() - tstart))
        return cumulative_loss

    def evaluate_grad_auto(self, lpos: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        """loss = mean(E_L) back-propagated through the local-energy adjoint (solver.py:352-370)."""
        loss, eloc = self.loss(lpos)
        loss.backward()
        return loss, eloc

    def evaluate_grad_manual(self, lpos: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        """dE/dk = 2 <(E_L - <E_L>) dpsi/dk / psi> via psi.backward(weight) (solver.py:372-431)."""
        if self.loss.method not in ["energy", "weighted-energy"]:
            raise ValueError("Manual gradient only for energy minimization")
        with torch.no_grad():
            eloc = self.wf.local_energy(lpos)
        # the sampled positions carry requires_grad (metropolis.py:177); the
        # parameter gradients do not need d psi / d x, so it is not formed
        psi = self.wf(lpos.detach())
        norm = 1.0 / len(psi)
        weight = eloc.clone()
        weight -= torch.mean(eloc)
        weight /= psi.detach().clone()
        weight *= 2.0 * norm
        clip_mask = self.loss.get_clipping_mask(eloc)
        if self.loss.clip:
            weight = weight * clip_mask
        psi.backward(weight)
        return torch.mean(eloc), eloc

    def log_data_opt(self, nepoch, task):
        log.info("")
        log.info("  Optimization")
        log.info("  Task                : {0}", task)
        log.info("  Number Parameters   : {0}", self.wf.get_number_parameters())
        log.info("  Number of epoch     : {0}", nepoch)
        log.info("  Batch size          : {0}", self.sampler.get_sampling_size())
        log.info("  Loss function       : {0}", self.loss.method)
        log.info("  Gradients           : {0}", self.grad_method)
        log.info("  Resampling mode     : {0}", self.resampling_options.mode)
