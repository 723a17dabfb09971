"""Batched and served VMC on one GPU: many systems per launch
(``BatchRunner``) and a continuous-batching request loop (``VMCServer``).
No counterpart in the reference, whose driver runs one wavefunction per
Solver (solver/solver_base.py:316-387)."""
from . import catalog
from .batch import BatchRunner, SystemContext, WalkerSet, sweep_each
from .server import VMCRequest, VMCResult, VMCServer, latency_summary, poisson_requests, serve_per_request

__all__ = ["catalog", "BatchRunner", "SystemContext", "WalkerSet", "sweep_each", "VMCRequest", "VMCResult",
           "VMCServer", "latency_summary", "poisson_requests", "serve_per_request"]
