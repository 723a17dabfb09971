"""Edge cases of the device path (needs an MI355X): zero walkers, one walker,
walker counts that leave a wavefront partly empty, argument errors.

Bars: zero walkers is a successful no-op at the C ABI and empty tensors at
the Python surface; one walker gives the oracle's psi and E_L to 1e-10 (the
FP64 bar) and the same chain as walker 0 of a larger batch (bit-exact: a
walker's chain depends on the seed, its id and the step only).
"""
import sys

import pytest
import torch

from qmctorch_amd import _native as N
from tests import helpers, systems

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _double():
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


def test_zero_walkers_at_the_c_abi_and_python_surface():
    _say("case zero walkers")
    mol, kw = helpers.CASES["h2o_ccpvdz_ee_en"]()
    wf, _ = helpers.make_pair(mol, **kw)
    lib, stream = N.load(), N.current_stream()
    plan = wf.device_plan()
    prm, keep = wf._params("fp64")
    ok, err_arg = 0, 1  # QMC_OK, QMC_ERR_ARG
    dummy = torch.zeros(16, dtype=torch.float64, device=DEV)  # valid pointers, never read for 0 walkers
    ptr = dummy.data_ptr()
    o = N.QmcWfOut()
    for k in ("psi", "logabs", "sign", "kinetic", "eloc", "epot"):
        setattr(o, k, ptr)
    assert lib.qmc_wf_eval(plan.handle, prm, N.QMC_PREC_FP64, N.QMC_EVAL_ELOC, ptr, 0, o, stream) == ok
    assert lib.qmc_wf_eval(plan.handle, prm, N.QMC_PREC_FP64, N.QMC_EVAL_ELOC, ptr, -1, o, stream) == err_arg
    assert b"negative" in lib.qmc_last_error()
    wk = N.QmcWalkers()
    wk.pos, wk.logabs, wk.sign, wk.naccept, wk.walker_id = ptr, ptr, ptr, 0, 0
    sw = N.QmcSweepDesc()
    sw.nstep, sw.step_size = 5, 0.2
    assert lib.qmc_metropolis_sweep(plan.handle, prm, N.QMC_PREC_FP64, wk, 0, sw, stream) == ok
    hm = N.QmcHmcDesc()
    hm.nstep, hm.traj_length, hm.step_size = 2, 3, 0.1
    assert lib.qmc_hmc_sweep(plan.handle, prm, N.QMC_PREC_FP64, wk, 0, hm, stream) == ok
    assert lib.qmc_wf_drift(plan.handle, prm, N.QMC_PREC_FP64, ptr, 0, ptr, ptr, ptr, stream) == ok
    assert torch.equal(dummy, torch.zeros_like(dummy))  # nothing written
    empty = torch.empty(0, 3 * mol.nelec, dtype=torch.float64, device=DEV)
    assert wf.local_energy(empty).numel() == 0
    assert wf.pdf(empty).numel() == 0
    dr, la, sg = wf.drift(empty, "fp64")
    assert dr.numel() == 0 and la.numel() == 0
    torch.cuda.synchronize()


@pytest.mark.parametrize("case", ["h2o_ccpvdz_ee_en", "lih_631g_cas22_ee_en", "c6h6_ccpvdz_ee_en"])
def test_one_walker_matches_oracle(case):
    _say("case one walker", case)
    mol, kw = helpers.CASES[case]()
    wf, orc = helpers.make_pair(mol, **kw)
    pos = helpers.positions(mol, 1, seed=8)
    psi = wf(pos.to(DEV)).detach().cpu()
    torch.testing.assert_close(psi, orc(pos).detach(), rtol=1e-10, atol=1e-300)
    el = wf.local_energy_parts(pos.to(DEV))["eloc"].cpu()
    el_o = orc.local_energy(pos).detach().view(-1)
    assert float(((el - el_o).abs() / el_o.abs().clamp(min=1.0)).max()) < 1e-10


@pytest.mark.parametrize("nw", [1, 5, 7])
def test_one_and_few_walkers_chain_is_walker_zero_of_a_batch(nw):
    """Waves that hold fewer walkers than they can (6 for H2O) or a single walker."""
    _say("case few walkers", nw)
    from qmctorch_amd.sampler import Metropolis

    mol = systems.h2o()
    wf, _ = helpers.make_pair(mol, w_ee=1.0, w_en=1.0, precision="mixed")
    p0 = helpers.positions(mol, 64, seed=9).to(DEV)

    def chain(n):
        s = Metropolis(nwalkers=n, nstep=40, step_size=0.2, nelec=mol.nelec, init=mol.domain("normal"),
                       move={"type": "all-elec", "proba": "normal"}, cuda=True, seed=31)
        return s(wf.pdf, pos=p0[:n].clone(), with_tqdm=False).detach()

    a, b = chain(nw), chain(64)
    assert torch.isfinite(a).all()
    assert torch.equal(a, b[:nw])
