"""Solver base (solver/solver_base.py:15-546 of the reference).

single_point, resample, observables and checkpoints with the reference's
semantics; sampling and local energies run on the device.
"""
import os
from types import SimpleNamespace
from typing import Any, List, Optional, Tuple, Union

import numpy as np
import torch
from tqdm import tqdm

from .. import log
from ..__version__ import __version__
from ..utils.hdf5_utils import add_group_attr, dump_to_hdf5


class SolverBase:
    def __init__(self, wf=None, sampler=None, optimizer=None, scheduler=None, output: Optional[str] = None,
                 rank: int = 0) -> None:
        self.wf = wf
        self.sampler = sampler
        self.opt = optimizer
        self.scheduler = scheduler
        self.cuda = False
        self.device = torch.device("cpu")
        self.qmctorch_version = __version__
        self.dataloader = None
        self.loss = None
        self.obs_dict = None
        if self.opt is not None and "lpos_needed" not in self.opt.__dict__:
            self.opt.lpos_needed = False
        self.save_model = "model.pth"
        if self.wf.cuda:
            self.device = torch.device("cuda")
            self.cuda = True
            self.sampler.cuda = True
            self.sampler.walkers.cuda = True
        self.hdf5file = output
        if output is None:
            basename = os.path.basename(self.wf.mol.hdf5file).split(".")[0]
            self.hdf5file = basename + "_QMCTorch.hdf5"
        if rank == 0:
            if os.path.isfile(self.hdf5file):
                os.remove(self.hdf5file)
        self.log_data()

    def configure_resampling(self, mode: str = "update", resample_every: int = 1, nstep_update: int = 25,
                             ntherm_update: int = -1, increment: dict = {"every": None, "factor": None}):
        """solver_base.py:76-110"""
        valid = ["never", "full", "update"]
        if mode not in valid:
            raise ValueError(mode, "not a valid update method : ", valid)
        self.resampling_options = SimpleNamespace(mode=mode, resample_every=resample_every,
                                                  ntherm_update=ntherm_update, nstep_update=nstep_update,
                                                  increment=increment)

    def track_observable(self, obs_name: Union[str, List[str]]):
        """solver_base.py:112-173"""
        if not isinstance(obs_name, list):
            obs_name = [obs_name]
        valid = ["energy", "local_energy", "geometry", "parameters", "gradients"]
        for name in obs_name:
            if name not in valid and not hasattr(self.wf, name):
                raise ValueError("Observable not recognized")
        self.observable = SimpleNamespace()
        self.observable.qmctorch_version = self.qmctorch_version
        obs_name = list(obs_name)
        if "energy" not in obs_name:
            obs_name += ["energy"]
        if "geometry" not in obs_name:
            obs_name += ["geometry"]
        for k in obs_name:
            if k == "parameters":
                for key, p in self.wf.named_parameters():
                    if p.requires_grad:
                        setattr(self.observable, key, [])
            elif k == "gradients":
                for key, p in self.wf.named_parameters():
                    if p.requires_grad:
                        setattr(self.observable, key + ".grad", [])
            else:
                setattr(self.observable, k, [])
        self.observable.models = SimpleNamespace()

    def store_observable(self, pos: torch.Tensor, local_energy: Optional[torch.Tensor] = None,
                         ibatch: Optional[int] = None, **kwargs):
        """Running means across minibatches (solver_base.py:175-248)."""
        if self.wf.cuda and pos.device.type == "cpu":
            pos = pos.to(self.device)
        state = None
        for obs in self.observable.__dict__.keys():
            if obs == "energy":
                if local_energy is None:
                    local_energy = self.wf.local_energy(pos)
                data = local_energy.cpu().detach().numpy()
                if ibatch is None or ibatch == 0:
                    self.observable.energy.append(np.mean(data).item())
                else:
                    self.observable.energy[-1] *= ibatch / (ibatch + 1)
                    self.observable.energy[-1] += np.mean(data).item() / (ibatch + 1)
            elif obs == "local_energy" and local_energy is not None:
                data = local_energy.cpu().detach().numpy()
                if ibatch is None or ibatch == 0:
                    self.observable.local_energy.append(data)
                else:
                    self.observable.local_energy[-1] = np.append(self.observable.local_energy[-1], data)
            elif obs == "geometry":
                self.observable.geometry.append(self.wf.geometry(pos))
            else:
                state = state if state is not None else self.wf.state_dict()
                if obs in state:
                    p = state[obs].clone()
                    getattr(self.observable, obs).append(p.data.cpu().numpy())
                    if obs + ".grad" in self.observable.__dict__:
                        param = dict(self.wf.named_parameters()).get(obs)
                        g = param.grad if param is not None else None
                        getattr(self.observable, obs + ".grad").append(
                            g.cpu().numpy() if g is not None else np.zeros_like(p.cpu().numpy()))
                elif obs.endswith(".grad") or obs in ("qmctorch_version", "models"):
                    continue
                elif hasattr(self.wf, obs):
                    data = getattr(self.wf, obs)(pos)
                    if isinstance(data, torch.Tensor):
                        data = data.cpu().detach().numpy()
                    if ibatch is None or ibatch == 0:
                        getattr(self.observable, obs).append(data)
                    else:
                        getattr(self.observable, obs)[-1] = np.append(getattr(self.observable, obs)[-1], data)

    def print_observable(self, cumulative_loss: float, verbose: bool = False):
        for k in self.observable.__dict__.keys():
            if k == "local_energy":
                eloc = self.observable.local_energy[-1]
                e, v = np.mean(eloc), np.var(eloc)
                log.options(style="percent").info("  energy   : %f +/- %f" % (e, np.sqrt(v / len(eloc))))
                log.options(style="percent").info("  variance : %f" % np.sqrt(v))

    def resample(self, n: int, pos: torch.Tensor) -> torch.Tensor:
        """solver_base.py:273-314"""
        if self.resampling_options.mode != "never":
            if n % self.resampling_options.resample_every == 0:
                if self.resampling_options.mode == "update":
                    pos = pos.clone().detach()[: self.sampler.walkers.nwalkers].to(self.device)
                else:
                    pos = None
                inc = self.resampling_options.increment
                if inc["every"] is not None and n % inc["every"] == 0:
                    self.sampler.nstep += inc["factor"] * self.sampler.ndecor
                pos = self.sampler(self.wf.pdf, pos=pos, with_tqdm=False)
                self.dataloader.dataset = pos
            if self.loss.use_weight:
                self.loss.weight["psi0"] = None
        return pos

    def single_point(self, with_tqdm: Optional[bool] = True, batchsize: Optional[int] = None,
                     hdf5_group: str = "single_point"):
        """Sample, then local energies and their statistics (solver_base.py:316-387)."""
        log.info("")
        log.info("  Single Point Calculation : {nw} walkers | {ns} steps", nw=self.sampler.walkers.nwalkers,
                 ns=self.sampler.nstep)
        with torch.no_grad():
            pos = self.sampler(self.wf.pdf, with_tqdm=with_tqdm)
            if self.wf.cuda and pos.device.type == "cpu":
                pos = pos.to(self.device)
            if batchsize is None:
                eloc = self.wf.local_energy(pos)
            else:
                parts = [self.wf.local_energy(pos[i : i + batchsize]) for i in range(0, len(pos), batchsize)]
                eloc = torch.cat(parts)
            e, s, err = torch.mean(eloc), torch.var(eloc), self.wf.sampling_error(eloc)
        log.options(style="percent").info("  Energy   : %f +/- %f" % (e.item(), err.item()))
        log.options(style="percent").info("  Variance : %f" % s.item())
        log.options(style="percent").info("  Size     : %d" % len(eloc))
        obs = SimpleNamespace(pos=pos, local_energy=eloc, energy=e, variance=s, error=err)
        grp = dump_to_hdf5(obs, self.hdf5file, root_name=hdf5_group)
        add_group_attr(self.hdf5file, grp, {"type": "single_point"})
        return obs

    def save_checkpoint(self, epoch: int, loss: float):
        """solver_base.py:389-405 (same keys, including 'optimzier_state_dict')."""
        torch.save({"epoch": epoch, "model_state_dict": self.wf.state_dict(),
                    "optimzier_state_dict": self.opt.state_dict(), "loss": loss},
                   "checkpoint_epoch%d.pth" % epoch)

    def load_checkpoint(self, filename: str) -> Tuple[int, float]:
        data = torch.load(filename, weights_only=True)
        self.wf.load_state_dict(data["model_state_dict"])
        self.opt.load_state_dict(data["optimzier_state_dict"])
        return data["epoch"], data["loss"]

    def optimization_step(self, lpos: torch.Tensor) -> None:
        if self.opt.lpos_needed:
            self.opt.step(lpos)
        else:
            self.opt.step()

    def run(self, nepoch: int, batchsize: Optional[int] = None, loss: str = "variance"):
        raise NotImplementedError()

    def log_data(self) -> None:
        log.info("")
        log.info(" QMC Solver ")
        if self.wf is not None:
            log.info("  WaveFunction        : {0}", self.wf.__class__.__name__)
        if self.sampler is not None:
            log.info("  Sampler             : {0}", self.sampler.__class__.__name__)
        if self.opt is not None:
            log.info("  Optimizer           : {0}", self.opt.__class__.__name__)
        if self.scheduler is not None:
            log.info("  Scheduler           : {0}", self.scheduler.__class__.__name__)
