"""Kernels generated for the benchmark molecules' basis layouts (MI355X).

Every catalog molecule's plan must find its generated layout (the signature
computed by specialize.py equals the one computed by the library), and the
generated kernels must reproduce the table-driven ones: the FP64 local energy
within 1e-13 relative, the mixed-precision sweep walker for walker, and the
mixed-precision drift of the HMC kernel within FP32 rounding (1e-4 relative
to 1 + |drift|, median below 1e-6).
"""
import math
import sys

import numpy as np
import pytest
import torch

from qmctorch_amd import _native as N
from qmctorch_amd.serving import catalog
from tests import helpers

pytestmark = pytest.mark.gpu

LAYOUTS = [("h2", "ground_state", "h2"), ("lih", "ground_state", "lih"), ("lih", "cas(2,2)", "lih_cas22"),
           ("h2o", "ground_state", "h2o"), ("n2", "ground_state", "n2"), ("c6h6", "ground_state", "c6h6")]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _double():
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


def _wf(system, configs):
    mol = catalog.molecule(system)
    wf, _ = helpers.make_pair(mol, configs=configs, w_ee=1.0, w_en=1.0, perturb=0.05, precision="mixed",
                              ci=None if configs == "ground_state" else [1.0, 0.1, -0.2, 0.05])
    return mol, wf


def _sweep(wf, pos, nstep, seed=5):
    nw = pos.shape[0]
    x = pos.clone().contiguous()
    la = torch.empty(nw, device="cuda")
    sg = torch.empty(nw, device="cuda")
    na = torch.zeros(nw, dtype=torch.int32, device="cuda")
    wk = N.QmcWalkers()
    wk.pos, wk.logabs, wk.sign, wk.naccept, wk.walker_id = x.data_ptr(), la.data_ptr(), sg.data_ptr(), na.data_ptr(), 0
    sw = N.QmcSweepDesc()
    sw.step0, sw.nstep, sw.init_logpsi = 0, nstep, 1
    sw.sigma2, sw.seed, sw.step_size, sw.keep_every = 0.2 / (2 * math.sqrt(2 * math.log(2.0))), seed, 0.2, 1
    prm, keep = wf._params("mixed")
    N.check(N.load().qmc_metropolis_sweep(wf.device_plan().handle, prm, N.QMC_PREC_MIXED, wk, nw, sw,
                                          N.current_stream()))
    torch.cuda.synchronize()
    return x, la, na


@pytest.mark.parametrize("system,configs,name", LAYOUTS)
def test_generated_layout_matches_table_driven(system, configs, name):
    _say("case generated layout", name)
    mol, wf = _wf(system, configs)
    plan = wf.device_plan()
    assert plan.specialization == name
    pos = catalog.initial_positions(mol, 300, seed=7)
    e_spec = wf.local_energy_parts(pos, "fp64")["eloc"]
    xs, ls, ns = _sweep(wf, pos, 20)
    plan.set_specialized(False)
    try:
        assert plan.specialization is None
        e_gen = wf.local_energy_parts(pos, "fp64")["eloc"]
        xg, lg, ng = _sweep(wf, pos, 20)
    finally:
        plan.set_specialized(True)
    rel = ((e_spec - e_gen).abs() / e_gen.abs()).max().item()
    same = (xs == xg).all(dim=1).double().mean().item()
    _say("  E_L max rel diff %.3e, walkers with identical trajectories %.4f, accept %d vs %d"
         % (rel, same, int(ns.sum()), int(ng.sum())))
    assert rel < 1e-13
    assert same >= 0.99


@pytest.mark.parametrize("system,configs,name", LAYOUTS)
def test_generated_hmc_drift_matches_table_driven(system, configs, name):
    _say("case generated hmc", name)
    mol, wf = _wf(system, configs)
    plan = wf.device_plan()
    assert plan.specialization == name
    pos = catalog.initial_positions(mol, 300, seed=8)
    ds, ls, ss = wf.drift(pos, "mixed")
    plan.set_specialized(False)
    try:
        dg, lg, sg = wf.drift(pos, "mixed")
    finally:
        plan.set_specialized(True)
    rel = ((ds - dg).norm(dim=-1) / (1.0 + dg.norm(dim=-1))).view(-1)
    _say("  drift rel diff median %.2e max %.2e, log|psi| max diff %.2e"
         % (float(rel.median()), float(rel.max()), float((ls - lg).abs().max())))
    assert float(rel.median()) < 1e-6 and float(rel.max()) < 1e-4
    assert float((ls - lg).abs().max()) < 1e-4
    assert torch.equal(ss, sg)
