"""Solver rows against the oracle (needs an MI355X).

* Solver.single_point (solver_base.py:316-387): energy, variance and error of
  the device chain against the oracle's host-RNG chain, H2/STO-3G and
  LiH/6-31G CAS(2,2), within the combined standard error;
* the gradients Solver.evaluate_grad_manual (solver.py:372-431) and
  Solver.evaluate_grad_auto (solver.py:352-370) hand to the optimiser, on fixed
  positions, against O.manual_gradient and autograd of mean(E_L) through the
  oracle, within 1e-8 relative;
* the fused adjoint reduction against the per-walker row kernels, its
  independence of the workspace size, and empty batches;
* one optimisation epoch of Solver.run against the same epoch on the oracle.
"""
import math
import sys

import numpy as np
import pytest
import torch

from oracle import vmc_oracle as O
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


def _solver(wf, mol, nw, nstep, ntherm=-1, ndecor=1, out="/tmp/qmc_solver_test.hdf5"):
    from qmctorch_amd.sampler import Metropolis
    from qmctorch_amd.solver import Solver

    s = Metropolis(nwalkers=nw, nstep=nstep, step_size=0.2, ntherm=ntherm, ndecor=ndecor, nelec=mol.nelec,
                   init=mol.domain("normal"), move={"type": "all-elec", "proba": "normal"}, cuda=True, seed=17)
    opt = torch.optim.SGD(wf.parameters(), lr=0.01)
    return Solver(wf=wf, sampler=s, optimizer=opt, output=out)


@pytest.mark.parametrize("case", ["h2_sto3g_ee", "lih_631g_cas22_ee_en"])
def test_single_point_matches_oracle_chain(case):
    _say("case single_point", case)
    mol, kw = helpers.CASES[case]()
    wf, orc = helpers.make_pair(mol, **kw)
    nw, nstep = 2000, 600
    solver = _solver(wf, mol, nw, nstep, ntherm=400, ndecor=50)
    torch.manual_seed(0)
    init = helpers.positions(mol, nw, seed=3)
    obs = solver.single_point(with_tqdm=False)  # walkers from mol.domain("normal"), thermalised
    e_d, v_d, err_d = float(obs.energy), float(obs.variance), float(obs.error)
    # the reported statistics are those of the returned local energies
    el = obs.local_energy.view(-1)
    assert e_d == pytest.approx(float(el.mean()), rel=1e-14)
    assert err_d == pytest.approx(math.sqrt(float(el.var()) / len(el)), rel=1e-12)
    pos_o, _ = O.metropolis(orc.pdf, init.clone(), nstep, step_size=0.2, ntherm=400, ndecor=50, seed=2)
    so = O.single_point(orc, pos_o)
    e_o, v_o, err_o = float(so.energy), float(so.variance), float(so.error)
    se = math.sqrt(err_d ** 2 + err_o ** 2)
    _say("  device %.5f +- %.5f (var %.4f)  oracle %.5f +- %.5f (var %.4f)" % (e_d, err_d, v_d, e_o, err_o, v_o))
    assert abs(e_d - e_o) < 5 * se
    # E_L is heavy-tailed (1/r terms near nuclei and nodes), so the variances are
    # compared through their own standard errors from the fourth moments, and the
    # two samples through a two-sample Kolmogorov-Smirnov test (0.1 % level)
    a = el.cpu().numpy()
    b = so.local_energy.detach().view(-1).numpy()

    def var_se(x):
        c = x - x.mean()
        return math.sqrt(max((c ** 4).mean() - (c ** 2).mean() ** 2, 0.0) / len(x))

    sev = math.sqrt(var_se(a) ** 2 + var_se(b) ** 2)
    xa, xb = np.sort(a), np.sort(b)
    grid = np.concatenate([xa, xb])
    ks = np.max(np.abs(np.searchsorted(xa, grid, side="right") / len(xa)
                       - np.searchsorted(xb, grid, side="right") / len(xb)))
    crit = 1.95 * math.sqrt((len(xa) + len(xb)) / (len(xa) * len(xb)))
    _say("  variance difference %.4f (se %.4f); KS %.4f (critical %.4f)" % (abs(v_d - v_o), sev, ks, crit))
    assert abs(v_d - v_o) < 5 * sev
    assert ks < crit


def _fixed_batch(mol, n, seed):
    return helpers.positions(mol, n, seed=seed)


@pytest.mark.parametrize("case", ["h2o_ccpvdz_ee_en", "lih_631g_cas22_ee_en", "lih_adf_sto_sd22",
                                  "lih_631g_sd46_ee_en", "h2o_ccpvdz_sd46_ee"])
def test_solver_manual_gradient_matches_oracle(case):
    """The gradients that Solver.evaluate_grad_manual leaves in .grad equal
    O.manual_gradient (itself pinned by finite differences) on fixed positions.
    The sd46 cases (single_double(4,6): 15 unique configurations per spin, 93
    CI terms) take the excitation (trace-trick) adjoint; the oracle takes
    every determinant explicitly."""
    _say("case solver manual gradient", case)
    mol, kw = helpers.CASES[case]()
    wf, orc = helpers.make_pair(mol, **kw)
    solver = _solver(wf, mol, 64, 1)
    solver.configure(grad="manual", track=["local_energy"])
    pos = _fixed_batch(mol, 96, seed=23)
    wf.zero_grad()
    loss, eloc = solver.evaluate_grad_manual(pos.to(DEV).requires_grad_())
    el_o, g_o = O.manual_gradient(orc, pos.clone().requires_grad_())
    torch.testing.assert_close(eloc.detach().cpu().view(-1), el_o.detach().view(-1), rtol=1e-10, atol=1e-10)
    pairs = [("mo.mo_modifier", "mo_modifier"), ("ao.bas_exp", "bas_exp"), ("fc.weight", "fc_weight"),
             ("jastrow.jastrow_kernel.weight", "w_ee"), ("jastrow.jastrow_terms.0.jastrow_kernel.weight", "w_ee"),
             ("jastrow.jastrow_terms.1.jastrow_kernel.weight", "w_en")]
    names = dict(wf.named_parameters())
    checked = 0
    for kd, ko in pairs:
        if kd in names and g_o.get(ko) is not None and names[kd].requires_grad:
            got = names[kd].grad.detach().cpu().view(-1)
            ref = g_o[ko].detach().view(-1)
            err = float((got - ref).abs().max() / ref.abs().max().clamp(min=1e-300))
            _say("  %-48s rel err %.2e" % (kd, err))
            # zero gradients (fc.weight of a single determinant: sum (E_L - <E_L>) / c = 0)
            # are compared at the round-off of the others
            torch.testing.assert_close(got, ref, rtol=1e-8, atol=1e-12 + 1e-10 * float(ref.abs().max()))
            checked += 1
    assert checked >= 3


@pytest.mark.parametrize("case", ["h2o_ccpvdz_ee_en", "lih_631g_cas22_ee_en", "lih_631g_sd46_ee_en"])
def test_solver_auto_gradient_matches_oracle(case):
    """evaluate_grad_auto back-propagates mean(E_L) (loss.py, use_weight off):
    the device adjoint equals autograd through the oracle's local energy."""
    _say("case solver auto gradient", case)
    mol, kw = helpers.CASES[case]()
    wf, orc = helpers.make_pair(mol, **kw)
    solver = _solver(wf, mol, 64, 1)
    solver.configure(grad="auto", freeze=["ao"], track=["local_energy"],
                     resampling={"mode": "update", "resample_every": 1, "nstep_update": 5})
    assert solver.loss.use_weight is False
    pos = _fixed_batch(mol, 80, seed=29)
    wf.zero_grad()
    solver.evaluate_grad_auto(pos.to(DEV).requires_grad_())
    orc.zero_grad()
    orc.local_energy(pos.clone().requires_grad_()).mean().backward()
    pairs = [("mo.mo_modifier", orc.mo_modifier), ("fc.weight", orc.fc_weight),
             ("jastrow.jastrow_kernel.weight", orc.w_ee), ("jastrow.jastrow_terms.0.jastrow_kernel.weight", orc.w_ee),
             ("jastrow.jastrow_terms.1.jastrow_kernel.weight", orc.w_en)]
    names = dict(wf.named_parameters())
    checked = 0
    for kd, po in pairs:
        if kd in names and po is not None and po.grad is not None and names[kd].requires_grad:
            got = names[kd].grad.detach().cpu().view(-1)
            ref = po.grad.detach().view(-1)
            err = float((got - ref).abs().max() / ref.abs().max().clamp(min=1e-300))
            _say("  %-48s rel err %.2e" % (kd, err))
            torch.testing.assert_close(got, ref, rtol=1e-8, atol=1e-12 + 1e-10 * float(ref.abs().max()))
            checked += 1
    assert checked >= 2


@pytest.mark.parametrize("case", ["h2o_ccpvdz_ee_en", "lih_631g_cas22_ee_en", "c6h6_ccpvdz_ee_en"])
def test_fused_adjoint_equals_row_kernels(case):
    """qmc_logpsi_grad / qmc_eloc_grad (partial sums in slots, no rows in
    memory) against the row kernels contracted by a GEMM, and for two workspace
    sizes (4096 slots or 3 slots: different summation order, same values)."""
    _say("case fused adjoint", case)
    from qmctorch_amd.wavefunction.slater_jastrow import adjoint_reduce, eloc_adjoint_rows, wf_adjoint_rows

    mol, kw = helpers.CASES[case]()
    wf, _ = helpers.make_pair(mol, **kw)
    nw = 48 if mol.nelec > 20 else 700
    x = helpers.positions(mol, nw, seed=31).to(DEV)
    g = torch.linspace(-1.0, 1.5, nw, device=DEV)
    nel = mol.nelec
    gr = g.repeat_interleave(nel)
    rows = wf_adjoint_rows(wf, x, need_exp=True)
    ref_w = rows["ao"].T @ (rows["gamma"] * gr[:, None])
    red = adjoint_reduce(wf, x, g, "logpsi", need_exp=True)
    scale = float(ref_w.abs().max())
    _say("  log|psi|: W %.2e, ci %.2e, jas %.2e, exp %.2e" % (
        float((red["w"] - ref_w).abs().max()) / scale,
        float((red["ci"] - g @ rows["dci"]).abs().max()),
        float((red["jas"] - g @ rows["djas"]).abs().max()),
        float((red["exp"] - gr @ rows["dexp"]).abs().max())))
    torch.testing.assert_close(red["w"], ref_w, rtol=1e-11, atol=1e-12 * scale)
    torch.testing.assert_close(red["ci"], g @ rows["dci"], rtol=1e-11, atol=1e-12)
    torch.testing.assert_close(red["jas"], g @ rows["djas"], rtol=1e-11, atol=1e-12)
    torch.testing.assert_close(red["exp"], gr @ rows["dexp"], rtol=1e-10, atol=1e-11 * float((gr @ rows["dexp"]).abs().max()))
    small = adjoint_reduce(wf, x, g, "logpsi", need_exp=True, workspace_bytes=3 * 8 * (
        wf.device_plan().sizes().nentry * (wf.device_plan().sizes().nmo + 12) + wf.nci + 2))
    torch.testing.assert_close(small["w"], red["w"], rtol=1e-12, atol=1e-13 * scale)
    r = eloc_adjoint_rows(wf, x)
    ref_w = r["ao"].T @ (r["ga"] * gr[:, None]) + r["b"].T @ (r["gb"] * gr[:, None])
    red = adjoint_reduce(wf, x, g, "eloc")
    scale = float(ref_w.abs().max())
    torch.testing.assert_close(red["w"], ref_w, rtol=1e-11, atol=1e-12 * scale)
    torch.testing.assert_close(red["ci"], g @ r["dci"], rtol=1e-11, atol=1e-12)
    torch.testing.assert_close(red["jas"], g @ r["djas"], rtol=1e-11, atol=1e-12)
    torch.testing.assert_close(red["eloc"], r["eloc"], rtol=1e-13, atol=1e-13)


def test_backward_of_empty_batch_gives_zero_gradients():
    _say("case empty backward")
    mol, kw = helpers.CASES["h2o_ccpvdz_ee_en"]()
    wf, _ = helpers.make_pair(mol, **kw)
    x = torch.empty(0, 3 * mol.nelec, device=DEV)
    wf.zero_grad()
    wf(x).sum().backward()
    assert float(wf.mo.mo_modifier.grad.abs().max()) == 0.0
    wf.zero_grad()
    wf.ao.bas_exp.requires_grad_(False)
    wf.local_energy_autodiff(x).sum().backward()
    assert float(wf.mo.mo_modifier.grad.abs().max()) == 0.0


@pytest.mark.parametrize("case", ["lih_631g_cas22_ee_en", "lih_631g_sd46_ee_en"])
def test_solver_run_epoch_matches_oracle_epoch(case):
    """One epoch of Solver.run's manual gradient + SGD step on fixed positions
    (resampling off) moves every parameter as the same step on the oracle
    (sd46: the excitation-plan adjoint)."""
    _say("case solver epoch", case)
    mol, kw = helpers.CASES[case]()
    wf, orc = helpers.make_pair(mol, **kw)
    solver = _solver(wf, mol, 400, 1)
    solver.configure(grad="manual", freeze=["ao"], track=["local_energy"],
                     resampling={"mode": "never", "resample_every": 1, "nstep_update": 1})
    pos = _fixed_batch(mol, 400, seed=37)
    from qmctorch_amd.utils.torch_utils import DataLoader

    solver.dataloader = DataLoader(pos.to(DEV), batch_size=400, pin_memory=False)
    solver.chkpt_every = None
    before = {k: v.detach().cpu().clone() for k, v in wf.named_parameters()}
    solver.run_epochs(1, verbose=False)
    _, g_o = O.manual_gradient(orc, pos.clone().requires_grad_())
    lr = 0.01
    for kd, ko in [("mo.mo_modifier", "mo_modifier"), ("fc.weight", "fc_weight"),
                   ("jastrow.jastrow_terms.0.jastrow_kernel.weight", "w_ee"),
                   ("jastrow.jastrow_terms.1.jastrow_kernel.weight", "w_en")]:
        moved = dict(wf.named_parameters())[kd].detach().cpu() - before[kd]
        torch.testing.assert_close(moved.view(-1), -lr * g_o[ko].detach().view(-1), rtol=1e-8,
                                   atol=1e-12)
    assert torch.equal(dict(wf.named_parameters())["ao.bas_exp"].detach().cpu(), before["ao.bas_exp"])
