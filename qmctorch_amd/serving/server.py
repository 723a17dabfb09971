"""Continuous-batching VMC request loop on one GPU.

A request asks for ``nsweeps`` all-electron Metropolis steps of ``nwalkers``
walkers of one molecule, started from seeded positions, followed by one
local-energy batch; the answer is the E_L of every walker with its mean,
variance and sampling error (the outputs of SolverBase.single_point,
solver_base.py:316-387; the reference has no request loop).

The loop keeps wavefunction parameters and walker state on the device.  On
every tick it admits the requests that have arrived, as far as the device
memory budget and the job capacity allow, then issues ONE batched sweep over
all admitted requests (each advances by up to ``steps_per_tick`` steps), one
batched local-energy launch for the requests whose sweeps are complete, and
asynchronous copies of their E_L to pinned host memory.  Requests that arrive
while a tick runs join the next tick; a request leaves as soon as its copy has
landed.  The host stays at most ``lookahead`` ticks ahead of the device.

A request's chain is keyed by its own seed, walker index and step number
(Philox), and a walker's arithmetic does not depend on the other jobs of a
launch, so each result is the same alone and in any mix.
"""
import collections
import time
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np
import torch

from . import catalog
from .batch import BatchRunner, SystemContext, WalkerSet


@dataclass
class VMCRequest:
    rid: int
    system: str
    nwalkers: int
    nsweeps: int
    seed: int
    arrival: float = 0.0  # seconds after the start of the run


@dataclass
class VMCResult:
    rid: int
    system: str
    nwalkers: int
    nsweeps: int
    energy: float
    variance: float
    error: float
    acceptance: float
    eloc: np.ndarray
    arrival: float
    admitted: float
    done: float

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class _Active:
    req: VMCRequest
    walkers: WalkerSet
    remaining: int
    nbytes: int
    admitted: float
    eloc_dev: Optional[torch.Tensor] = None
    eloc_host: Optional[torch.Tensor] = None
    nacc_host: Optional[torch.Tensor] = None


def poisson_requests(rate: float, n: int, seed: int = 0, systems: Sequence[str] = catalog.SYSTEM_NAMES,
                     wmin: int = 128, wmax: int = 16384, kmax: int = 20) -> List[VMCRequest]:
    """Poisson arrivals at ``rate`` req/s; system uniform over ``systems``,
    walkers round(exp(U(ln wmin, ln wmax))), sweeps uniform in 1..kmax."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    si = rng.integers(0, len(systems), n)
    w = np.round(np.exp(rng.uniform(np.log(wmin), np.log(wmax), n))).astype(np.int64)
    k = rng.integers(1, kmax + 1, n)
    return [VMCRequest(rid=i, system=systems[si[i]], nwalkers=int(w[i]), nsweeps=int(k[i]),
                       seed=int(seed) * 1_000_003 + i, arrival=float(t[i])) for i in range(n)]


def latency_summary(results: Sequence[VMCResult], wall: float) -> Dict[str, float]:
    lat = np.array([r.latency for r in results]) * 1e3
    ws = float(sum(r.nwalkers * r.nsweeps for r in results))
    return {"requests": len(results), "p50_ms": float(np.percentile(lat, 50)) if len(lat) else float("nan"),
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else float("nan"),
            "mean_ms": float(lat.mean()) if len(lat) else float("nan"),
            "walker_steps_per_s": ws / wall if wall > 0 else float("nan"), "wall_s": wall}


def _stats(e: np.ndarray):
    n = e.shape[0]
    mean = float(e.mean())
    var = float(e.var(ddof=1)) if n > 1 else 0.0  # torch.var default (unbiased), wf_base.py:217-229
    return mean, var, float(np.sqrt(var / n)) if n > 0 else float("nan")


def _tick_summary(ticks) -> Dict[str, float]:
    """Percentiles of the per-tick host submission time, device time and the
    device idle gap before each tick (every event has completed)."""
    host = np.array([h for h, _, _, _ in ticks]) * 1e6
    dev = np.array([a.elapsed_time(b) for _, a, b, _ in ticks]) * 1e3
    gap = np.array([ticks[k - 1][2].elapsed_time(ticks[k][1]) for k in range(1, len(ticks))]) * 1e3
    gap = np.maximum(gap, 0.0)
    pct = lambda x, q: float(np.percentile(x, q)) if len(x) else 0.0
    return {"ticks": len(ticks), "jobs_per_tick_mean": float(np.mean([n for _, _, _, n in ticks])),
            "host_submit_us_p50": pct(host, 50), "host_submit_us_p99": pct(host, 99),
            "device_tick_us_p50": pct(dev, 50), "device_tick_us_p99": pct(dev, 99),
            "idle_gap_us_p50": pct(gap, 50), "idle_gap_us_p99": pct(gap, 99),
            "device_busy_s": float(dev.sum()) * 1e-6, "host_submit_s": float(host.sum()) * 1e-6}


class VMCServer:
    """Continuous-batching request loop (one GPU, one host thread)."""

    def __init__(self, precision: str = "mixed", steps_per_tick: int = 4, max_active: int = 512,
                 hbm_budget_bytes: Optional[int] = None, step_size: float = 0.2, lookahead: int = 1,
                 contexts: Optional[Dict[str, SystemContext]] = None):
        self.precision = precision
        self.steps_per_tick = int(steps_per_tick)
        self.max_active = int(max_active)
        if hbm_budget_bytes is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget_bytes = int(0.5 * free)
        self.hbm_budget = int(hbm_budget_bytes)
        self.step_size = step_size
        self.lookahead = max(0, int(lookahead))
        self.contexts: Dict[str, SystemContext] = dict(contexts or {})
        self.runner = BatchRunner(max_jobs=self.max_active)
        self.ticks = 0
        self.peak_bytes = 0  # largest device-memory reservation seen by the last run
        self.peak_active = 0  # most requests in one tick

    def context(self, system: str) -> SystemContext:
        ctx = self.contexts.get(system)
        if ctx is None:
            ctx = SystemContext(catalog.wavefunction(system, self.precision), system, self.step_size)
            self.contexts[system] = ctx
        elif ctx.precision != self.precision:
            raise ValueError("context %s was built for precision %s" % (system, ctx.precision))
        return ctx

    def request_bytes(self, req: VMCRequest) -> int:
        return req.nwalkers * self.context(req.system).walker_bytes()

    def run(self, requests: Sequence[VMCRequest], realtime: bool = True,
            clock: Callable[[], float] = time.perf_counter, profile_ticks: bool = False) -> List[VMCResult]:
        """Serve ``requests`` (arrival times relative to the start) and return
        their results in completion order.  With realtime=False every request
        is available from the start (arrival times only order the queue).

        profile_ticks: time every tick (timing events around its launches,
        the host clock around its submission) and leave the summary in
        ``self.tick_profile``: host submission time per tick (what a captured
        graph could shorten), device time per tick, and the device idle gap
        between consecutive ticks (zero when the host keeps ahead)."""
        for r in requests:
            if self.request_bytes(r) > self.hbm_budget:
                raise ValueError("request %d needs more device memory than the whole budget" % r.rid)
            self.context(r.system)
        pending = collections.deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        queue: collections.deque = collections.deque()
        active: List[_Active] = []
        inflight: collections.deque = collections.deque()  # (event, [finishing _Active])
        used = 0
        self.peak_bytes = self.peak_active = 0
        results: List[VMCResult] = []
        ticks = [] if profile_ticks else None
        t0 = clock()

        def now() -> float:
            return clock() - t0

        def drain(block: bool) -> None:
            nonlocal used
            while inflight and (block or inflight[0][0].query()):
                ev, fin = inflight.popleft()
                ev.synchronize()
                t = now()
                for a in fin:
                    e = a.eloc_host.numpy().copy()
                    mean, var, err = _stats(e)
                    acc = float(a.nacc_host.numpy().sum()) / float(max(a.req.nwalkers * a.req.nsweeps, 1))
                    results.append(VMCResult(a.req.rid, a.req.system, a.req.nwalkers, a.req.nsweeps, mean, var,
                                             err, acc, e, a.req.arrival, a.admitted, t))
                    used -= a.nbytes
                block = False

        while pending or queue or active or inflight:
            t = now()
            while pending and (not realtime or pending[0].arrival <= t):
                queue.append(pending.popleft())
            # admission: FIFO against the device-memory budget and the job capacity
            while queue and len(active) < self.max_active:
                req = queue[0]
                nb = self.request_bytes(req)
                if used + nb > self.hbm_budget:
                    break
                queue.popleft()
                ctx = self.context(req.system)
                pos = catalog.initial_positions(ctx.wf.mol, req.nwalkers, req.seed)
                active.append(_Active(req, WalkerSet(ctx, pos, req.seed), req.nsweeps, nb, t))
                used += nb
                self.peak_bytes = max(self.peak_bytes, used)
            if active:
                self.peak_active = max(self.peak_active, len(active))
                jobs = [(a.walkers, min(a.remaining, self.steps_per_tick)) for a in active]
                if ticks is not None:
                    ev0 = torch.cuda.Event(enable_timing=True)
                    ev0.record()
                    h0 = time.perf_counter()
                self.runner.sweep(jobs)
                fin = []
                keep = []
                for a, (_, n) in zip(active, jobs):
                    a.remaining -= n
                    (fin if a.remaining == 0 else keep).append(a)
                if fin:
                    for a in fin:
                        a.eloc_dev = torch.empty(a.req.nwalkers, dtype=torch.float64, device="cuda")
                    self.runner.local_energy([(a.walkers, a.eloc_dev) for a in fin])
                    for a in fin:
                        a.eloc_host = torch.empty(a.req.nwalkers, dtype=torch.float64, pin_memory=True)
                        a.eloc_host.copy_(a.eloc_dev, non_blocking=True)
                        a.nacc_host = torch.empty(a.req.nwalkers, dtype=torch.int32, pin_memory=True)
                        a.nacc_host.copy_(a.walkers.naccept, non_blocking=True)
                ev = torch.cuda.Event(enable_timing=ticks is not None)
                ev.record()
                inflight.append((ev, fin))
                if ticks is not None:
                    ticks.append((time.perf_counter() - h0, ev0, ev, len(jobs)))
                active = keep
                self.ticks += 1
                drain(block=len(inflight) > self.lookahead)
            elif inflight:
                drain(block=True)
            elif pending and not queue:
                # idle until the next arrival
                dt = pending[0].arrival - now()
                if dt > 0:
                    time.sleep(min(dt, 1e-3))
            elif queue:
                raise RuntimeError("queued request cannot be admitted (budget or capacity too small)")
        if ticks is not None:
            self.tick_profile = _tick_summary(ticks)
        return results


def serve_per_request(requests: Sequence[VMCRequest], contexts: Dict[str, SystemContext], realtime: bool = True,
                      clock: Callable[[], float] = time.perf_counter) -> List[VMCResult]:
    """Plain path on the same GPU: requests one after the other in arrival
    order, each through the per-request API (Metropolis sampler + local
    energy, one system per launch, a host synchronisation per request)."""
    from ..sampler import Metropolis

    t0 = clock()
    out = []
    for req in sorted(requests, key=lambda r: (r.arrival, r.rid)):
        if realtime:
            while True:
                dt = req.arrival - (clock() - t0)
                if dt <= 0:
                    break
                time.sleep(min(dt, 1e-3))
        start = clock() - t0
        ctx = contexts[req.system]
        wf = ctx.wf
        pos = catalog.initial_positions(wf.mol, req.nwalkers, req.seed)
        sampler = Metropolis(nwalkers=req.nwalkers, nstep=req.nsweeps, step_size=ctx.step_size, nelec=wf.nelec,
                             init=wf.mol.domain("atomic"), move={"type": "all-elec", "proba": "normal"}, cuda=True,
                             seed=req.seed)
        x = sampler(wf.pdf, pos=pos, with_tqdm=False)
        e = wf.local_energy(x.detach()).reshape(-1).cpu().numpy()
        mean, var, err = _stats(e)
        out.append(VMCResult(req.rid, req.system, req.nwalkers, req.nsweeps, mean, var, err,
                             float(sampler.last_acceptance), e, req.arrival, start, clock() - t0))
    return out


__all__ = ["VMCRequest", "VMCResult", "VMCServer", "poisson_requests", "latency_summary", "serve_per_request"]
