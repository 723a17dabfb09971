"""Batched device launches over many systems (libqmcmi355x job tables).

``SystemContext`` freezes one wavefunction for serving: its device plan and
the parameter views of the sweep and of the local energy stay resident.
``WalkerSet`` is the device-resident state of one walker ensemble.
``BatchRunner`` turns lists of (system, walkers, steps) into one
``qmc_metropolis_sweep_batch`` / ``qmc_wf_eval_batch`` call: one launch per
register bucket, all systems of the bucket side by side.
"""
import ctypes
import math
from typing import List, Optional, Sequence, Tuple

import torch

from .. import _native as N


class SystemContext:
    """A wavefunction prepared for batched launches (parameters frozen)."""

    def __init__(self, wf, name: Optional[str] = None, step_size: float = 0.2):
        self.wf = wf
        self.name = name or getattr(wf.mol, "name", "system")
        self.nelec = wf.nelec
        self.precision = wf.precision
        self.eloc_precision = wf.eloc_precision
        self.plan = wf.device_plan()
        self.prm_sweep, self._keep_sweep = wf._params(wf.precision)
        self.prm_eval, self._keep_eval = wf._params("fp64")  # the local-energy kernel reads W in FP64
        self.step_size = step_size
        self.sigma2 = step_size / (2 * math.sqrt(2 * math.log(2.0)))  # metropolis.py:206-212

    def walker_bytes(self) -> int:
        """Device bytes held per walker: positions, log|psi|, sign, accept count, E_L."""
        return 8 * 3 * self.nelec + 8 + 8 + 4 + 8


class WalkerSet:
    """Device-resident Metropolis state of ``nw`` walkers of one system."""

    def __init__(self, ctx: SystemContext, pos: torch.Tensor, seed: int):
        N.require_device(pos, "walker positions")
        if pos.dim() != 2 or pos.shape[1] != 3 * ctx.nelec:
            raise ValueError("positions must have shape (nw, %d)" % (3 * ctx.nelec))
        self.ctx = ctx
        self.pos = pos.detach().to(torch.float64).contiguous()
        nw = self.pos.shape[0]
        self.nw = nw
        self.logabs = torch.empty(nw, dtype=torch.float64, device=self.pos.device)
        self.sign = torch.empty(nw, dtype=torch.float64, device=self.pos.device)
        self.naccept = torch.zeros(nw, dtype=torch.int32, device=self.pos.device)
        self.seed = int(seed) & ((1 << 64) - 1)
        self.step = 0  # next chain step
        self.has_logpsi = False


class BatchRunner:
    """Owns a libqmcmi355x batch (job-table ring) for one host thread."""

    def __init__(self, max_jobs: int = 1024):
        self.lib = N.load()
        self.max_jobs = int(max_jobs)
        h = ctypes.c_void_p()
        N.check(self.lib.qmc_batch_create(self.max_jobs, ctypes.byref(h)))
        self.handle = h

    def close(self) -> None:
        if self.handle is not None and self.handle.value:
            N.check(self.lib.qmc_batch_destroy(self.handle))
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _prec(ctxs, attr="precision") -> int:
        precs = {getattr(c, attr) for c in ctxs}
        if len(precs) != 1:
            raise ValueError("one batch runs one precision (got %s)" % sorted(precs))
        return N.QMC_PREC_MIXED if precs.pop() == "mixed" else N.QMC_PREC_FP64

    def sweep(self, jobs: Sequence[Tuple[WalkerSet, int]], stream: Optional[int] = None) -> None:
        """Advance every walker set by its step count (all in one batched call)."""
        jobs = [(w, int(n)) for w, n in jobs if n > 0 and w.nw > 0]
        if not jobs:
            return
        if len(jobs) > self.max_jobs:
            raise ValueError("%d jobs exceed the batch capacity %d" % (len(jobs), self.max_jobs))
        prec = self._prec([w.ctx for w, _ in jobs])
        arr = (N.QmcSweepJob * len(jobs))()
        for k, (w, n) in enumerate(jobs):
            j = arr[k]
            j.plan = w.ctx.plan.handle
            j.params = w.ctx.prm_sweep
            j.walkers.pos, j.walkers.logabs, j.walkers.sign = w.pos.data_ptr(), w.logabs.data_ptr(), w.sign.data_ptr()
            j.walkers.naccept, j.walkers.walker_id = w.naccept.data_ptr(), 0
            j.nw = w.nw
            s = j.sweep
            s.step0, s.nstep, s.init_logpsi = w.step, n, 0 if w.has_logpsi else 1
            s.sigma2, s.seed, s.step_size = w.ctx.sigma2, w.seed, w.ctx.step_size
            s.keep, s.keep_from, s.keep_every, s.uniform_move = 0, 0, 1, 0
        N.check(self.lib.qmc_metropolis_sweep_batch(self.handle, prec, arr, len(jobs),
                                                    N.current_stream() if stream is None else stream))
        for w, n in jobs:
            w.step += n
            w.has_logpsi = True

    def local_energy(self, jobs: Sequence[Tuple[WalkerSet, torch.Tensor]], stream: Optional[int] = None) -> None:
        """E_L of each walker set's current positions into the given (nw,) FP64 tensors."""
        jobs = [(w, e) for w, e in jobs if w.nw > 0]
        if not jobs:
            return
        if len(jobs) > self.max_jobs:
            raise ValueError("%d jobs exceed the batch capacity %d" % (len(jobs), self.max_jobs))
        prec = self._prec([w.ctx for w, _ in jobs], "eloc_precision")
        arr = (N.QmcEvalJob * len(jobs))()
        for k, (w, e) in enumerate(jobs):
            if e.dtype != torch.float64 or e.numel() != w.nw or not e.is_contiguous():
                raise ValueError("E_L output must be a contiguous float64 tensor of %d entries" % w.nw)
            j = arr[k]
            j.plan = w.ctx.plan.handle
            j.params = w.ctx.prm_eval
            j.pos = w.pos.data_ptr()
            j.nw = w.nw
            j.out.eloc = e.data_ptr()
        N.check(self.lib.qmc_wf_eval_batch(self.handle, prec, N.QMC_EVAL_ELOC, arr, len(jobs),
                                           N.current_stream() if stream is None else stream))


def sweep_each(jobs: Sequence[Tuple[WalkerSet, int]], stream: Optional[int] = None) -> None:
    """The same sweeps issued one system at a time (qmc_metropolis_sweep per job)."""
    lib = N.load()
    s_ = N.current_stream() if stream is None else stream
    for w, n in jobs:
        if n <= 0 or w.nw == 0:
            continue
        wk = N.QmcWalkers()
        wk.pos, wk.logabs, wk.sign, wk.naccept, wk.walker_id = (w.pos.data_ptr(), w.logabs.data_ptr(),
                                                               w.sign.data_ptr(), w.naccept.data_ptr(), 0)
        sw = N.QmcSweepDesc()
        sw.step0, sw.nstep, sw.init_logpsi = w.step, int(n), 0 if w.has_logpsi else 1
        sw.sigma2, sw.seed, sw.step_size = w.ctx.sigma2, w.seed, w.ctx.step_size
        sw.keep, sw.keep_from, sw.keep_every, sw.uniform_move = 0, 0, 1, 0
        prec = N.QMC_PREC_MIXED if w.ctx.precision == "mixed" else N.QMC_PREC_FP64
        N.check(lib.qmc_metropolis_sweep(w.ctx.plan.handle, w.ctx.prm_sweep, prec, wk, w.nw, sw, s_))
        w.step += int(n)
        w.has_logpsi = True


__all__: List[str] = ["SystemContext", "WalkerSet", "BatchRunner", "sweep_each"]
