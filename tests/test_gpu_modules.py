"""The reference's module-level API on the device (needs an MI355X).

AtomicOrbitals.forward in every derivative mode, the Jastrow factors and their
product, SlaterPooling.forward / operator, the potentials, position gradients
of the wavefunction, the walker initialiser and the Metropolis sampler on any
density, each checked against the oracle or an independent autograd / closed
form.  FP64 throughout; tolerances in each test.
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


def _close(a, b, rtol, atol, what):
    a = a.detach().cpu().double()
    b = b.detach().cpu().double()
    err = float(((a - b).abs() / (b.abs() + atol / rtol)).max()) if a.numel() else 0.0
    _say("  %-28s max rel err %.2e" % (what, err))
    torch.testing.assert_close(a, b, rtol=rtol, atol=atol)


@pytest.mark.parametrize("which", ["h2o", "lih_adf"])
def test_ao_forward_every_derivative_mode(which):
    """Values, gradients (summed / per axis), Laplacians, Hessian diagonals and
    mixed second derivatives against the oracle's AO values differentiated by
    autograd (atomic_orbitals.py:131-218)."""
    _say("case ao modes", which)
    mol = systems.h2o() if which == "h2o" else systems.lih_adf()
    wf, orc = helpers.make_pair(mol, w_ee=None)
    pos = helpers.positions(mol, 5, seed=4)
    x = pos.to(DEV)
    p = pos.clone().requires_grad_()
    ao = orc.ao(p)
    nb, ne, nao = ao.shape
    grad = torch.zeros(nb, ne, nao, 3)
    hd = torch.zeros(nb, ne, nao, 3)
    hm = torch.zeros(nb, ne, nao, 3)
    pairs = [(0, 1), (0, 2), (1, 2)]
    for mu in range(nao):
        g = torch.autograd.grad(ao[:, :, mu].sum(), p, create_graph=True)[0].view(nb, ne, 3)
        grad[:, :, mu] = g.detach()
        for d in range(3):
            h = torch.autograd.grad(g[:, :, d].sum(), p, retain_graph=True)[0].view(nb, ne, 3)
            hd[:, :, mu, d] = h[:, :, d].detach()
            for m, (u, v) in enumerate(pairs):
                if u == d:
                    hm[:, :, mu, m] = h[:, :, v].detach()
    ao = ao.detach()
    _close(wf.ao(x), ao, 1e-11, 1e-13, "values [0]")
    _close(wf.ao(x, derivative=1), grad.sum(-1), 1e-10, 1e-12, "summed gradient [1]")
    _close(wf.ao(x, derivative=1, sum_grad=False), grad, 1e-10, 1e-12, "gradient per axis")
    _close(wf.ao(x, derivative=2), hd.sum(-1), 1e-9, 1e-11, "Laplacian [2]")
    _close(wf.ao(x, derivative=2, sum_hess=False), hd, 1e-9, 1e-11, "Hessian diagonal")
    _close(wf.ao(x, derivative=3), hm, 1e-9, 1e-11, "mixed second derivatives [3]")
    a, g, lap = wf.ao(x, derivative=[0, 1, 2])
    assert g.shape == (nb, ne, nao, 3)
    _close(lap, hd.sum(-1), 1e-9, 1e-11, "[0, 1, 2] Laplacian")
    # one-electron update (atomic_orbitals.py:671-695)
    x2 = x.clone()
    x2[:, 3:6] += 0.3
    upd = wf.ao.update(wf.ao(x), x2, 1)
    _close(upd, wf.ao(x2), 1e-13, 1e-15, "update(ao, pos, 1)")


@pytest.mark.parametrize("kind", ["ee", "en", "combined"])
def test_jastrow_factor_forward(kind):
    """J, dJ per axis / summed, lap J of the e-e factor, the e-n factor and their
    product against the oracle (jastrow_factor_electron_electron.py:124-260,
    jastrow_factor_electron_nuclei.py:60-161, combine_jastrow.py:33-195)."""
    _say("case jastrow", kind)
    from qmctorch_amd.wavefunction.jastrows.combine_jastrow import CombineJastrow
    from qmctorch_amd.wavefunction.jastrows.elec_elec import JastrowFactor as JEE
    from qmctorch_amd.wavefunction.jastrows.elec_elec import PadeJastrowKernel as PEE
    from qmctorch_amd.wavefunction.jastrows.elec_nuclei import JastrowFactor as JEN
    from qmctorch_amd.wavefunction.jastrows.elec_nuclei import PadeJastrowKernel as PEN

    mol = systems.h2o()
    wee, wen = 0.9, 1.3
    ee = JEE(mol, PEE, kernel_kwargs={"w": wee}, cuda=True)
    en = JEN(mol, PEN, kernel_kwargs={"w": wen}, cuda=True)
    atoms = torch.as_tensor(np.asarray(mol.atom_coords))
    oee = O.JastrowEE(mol.nup, mol.ndown, torch.as_tensor([wee]))
    oen = O.JastrowEN(mol.nelec, atoms, torch.as_tensor([wen]))
    if kind == "ee":
        dev_j, terms = ee, [oee]
    elif kind == "en":
        dev_j, terms = en, [oen]
    else:
        dev_j, terms = CombineJastrow([ee, en]), [oee, oen]
    pos = helpers.positions(mol, 7, seed=13)
    j, dj, d2j = O.combine_jastrow(terms, pos)
    x = pos.to(DEV)
    _close(dev_j(x), j, 1e-12, 1e-300, "J")
    _close(dev_j(x, derivative=1, sum_grad=False), dj, 1e-11, 1e-14, "dJ per axis")
    _close(dev_j(x, derivative=1), dj.sum(1), 1e-11, 1e-14, "dJ summed over axes")
    _close(dev_j(x, derivative=2), d2j, 1e-10, 1e-13, "lap J")
    tj, tdj, td2j = dev_j(x, derivative=[0, 1, 2], sum_grad=False)
    _close(tdj, dj, 1e-11, 1e-14, "[0, 1, 2] gradient")


@pytest.mark.parametrize("case", ["lih_631g_cas22_ee_en", "h2o_ccpvdz_ee_en", "lih_adf_sto_sd22"])
def test_slater_pooling_forward_and_operator(case):
    """Determinant products and tr(A^-1 B) of every configuration from the same
    MO and operator matrices as the oracle (slater_pooling.py:64-111, 262-409);
    tr((A^-1 B)^2) and multi-operator stacks against torch.linalg on the CPU."""
    _say("case pooling", case)
    import operator as op

    mol, kw = helpers.CASES[case]()
    wf, orc = helpers.make_pair(mol, **kw)
    pos = helpers.positions(mol, 9, seed=21)
    mo = orc.mo(orc.ao(pos)).detach()
    bop = torch.as_tensor(np.random.default_rng(3).standard_normal(tuple(mo.shape)))
    _close(wf.pool(mo.to(DEV)), orc.pool(mo), 1e-11, 1e-300, "det products")
    _close(wf.pool.operator(mo.to(DEV), bop.to(DEV)), orc.operator(mo, bop), 1e-10, 1e-12, "tr(A^-1 B), op.add")
    up, dn = wf.pool.operator(mo.to(DEV), bop.to(DEV), op=None, op_squared=True)
    cu, cd = wf.configs
    nup = mol.nup
    for n in range(len(cu)):
        A = mo[:, :nup, :][:, :, cu[n]]
        X = torch.linalg.solve(A, bop[:, :nup, :][:, :, cu[n]])
        ref = torch.einsum("bij,bji->b", X, X)
        _close(up[:, n], ref, 1e-9, 1e-11, "tr((A^-1 B)^2) up, term %d" % n)
        A = mo[:, nup:, :][:, :, cd[n]]
        X = torch.linalg.solve(A, bop[:, nup:, :][:, :, cd[n]])
        _close(dn[:, n], torch.einsum("bij,bji->b", X, X), 1e-9, 1e-11, "tr((A^-1 B)^2) dn, term %d" % n)
    stack = torch.stack([bop, 2.0 * bop, -bop])
    multi = wf.pool.operator(mo.to(DEV), stack.to(DEV), op=op.add)
    assert multi.shape == (3, 9, len(cu))
    ref = orc.operator(mo, bop)
    _close(multi[1], 2.0 * ref, 1e-10, 1e-12, "operator stack, entry 1")
    _close(multi[2], -ref, 1e-10, 1e-12, "operator stack, entry 2")


def test_potentials_and_nuclear_repulsion():
    """V_ee, V_en (wf_base.py:49-95) against the oracle; V_nn follows
    ao.atom_coords (wf_base.py:97-116) after a state-dict load."""
    _say("case potentials")
    mol, kw = helpers.CASES["h2o_ccpvdz_ee_en"]()
    wf, orc = helpers.make_pair(mol, **kw)
    pos = helpers.positions(mol, 33, seed=2)
    x = pos.to(DEV)
    _close(wf.electronic_potential(x), orc.electronic_potential(pos), 1e-13, 1e-13, "V_ee")
    _close(wf.nuclear_potential(x), orc.nuclear_potential(pos), 1e-13, 1e-13, "V_en")
    assert abs(float(wf.nuclear_repulsion()) - float(orc.nuclear_repulsion())) < 1e-12
    el0 = wf.local_energy(x).view(-1)
    # move the nuclei of the AO layer: V_nn and V_en follow, E_L changes by their shifts
    sd = wf.state_dict()
    sd["ao.atom_coords"] = sd["ao.atom_coords"] * 1.01
    wf.load_state_dict(sd)
    with torch.no_grad():
        orc.atom_coords.mul_(1.01)
    assert abs(float(wf.nuclear_repulsion()) - float(orc.nuclear_repulsion())) < 1e-12
    el1 = wf.local_energy(x).view(-1).cpu()
    ref = orc.local_energy(pos).view(-1).detach()
    _close(el1, ref, 1e-10, 1e-10, "E_L after moving the AO centres")
    assert float((el1 - el0.cpu()).abs().max()) > 1e-3


@pytest.mark.parametrize("case", ["h2o_ccpvdz_ee_en", "lih_631g_cas22_ee_en", "lih_631g_sd46_ee_en"])
def test_position_gradient_of_psi(case):
    """wf(x).backward() gives d psi / d x (through the drift kernel) equal to
    autograd through the oracle."""
    _say("case position gradient", case)
    mol, kw = helpers.CASES[case]()
    wf, orc = helpers.make_pair(mol, **kw)
    pos = helpers.positions(mol, 11, seed=5)
    p = pos.clone().requires_grad_()
    w = torch.linspace(-1.0, 2.0, 11)
    (orc(p).view(-1) * w).sum().backward()
    x = pos.to(DEV).requires_grad_()
    (wf(x).view(-1) * w.to(DEV)).sum().backward()
    _close(x.grad, p.grad, 1e-9, 1e-12, "d psi / d x")


def test_walker_initialisation_distributions():
    """Device walker initialiser (sampler/walkers.py:8-150 domains).  Atomic
    domain: an electron's width follows its rank on its atom in electron order
    (one oxygen: electron 0 always 1/8, 1-4 1/3, 5-7 3/5); with two distant
    oxygens of two electrons each, every walker puts exactly two electrons on
    each atom, electron 0 is on atom 0 half the time, electron 1 is second on
    its atom with probability 1/3 and electron 3 always second.  Normal domain:
    mean and covariance; uniform box: bounds and variance; the same walker ids
    give the same starts in batches of different sizes."""
    _say("case walker init")
    from qmctorch_amd.sampler.walkers import Walkers, shell_widths

    assert shell_widths(8.0) == [1 / 8, 2 / 6, 3 / 5]
    assert shell_widths(1.0) == [1.0, 1.0, 1.0]
    nw, ne = 200000, 8
    w = Walkers(nwalkers=nw, nelec=ne, cuda=True)
    dom = {"method": "atomic", "atom_coords": [[0.5, -0.25, 1.0]], "atom_num": [8], "atom_nelec": [8]}
    x = w.draw(dom, nw, ne, seed=7).view(nw, ne, 3).cpu() - torch.tensor([0.5, -0.25, 1.0])
    ref = torch.tensor([1 / 64] + [1 / 9] * 4 + [9 / 25] * 3, dtype=torch.float64)
    var_e = x.var(dim=(0, 2))
    _say("  one atom: per-electron variance %s" % var_e.numpy().round(5))
    assert torch.all((var_e - ref).abs() < 0.02 * ref)
    a = w.draw(dom, 64, ne, seed=9)
    b = w.draw(dom, 1000, ne, seed=9)
    assert torch.equal(a, b[:64])
    # two distant atoms: the assignment is a random permutation per walker
    dom2 = {"method": "atomic", "atom_coords": [[0.0, 0.0, 0.0], [100.0, 0.0, 0.0]], "atom_num": [8, 8],
            "atom_nelec": [2, 2]}
    y = Walkers(nwalkers=nw, nelec=4, cuda=True).draw(dom2, nw, 4, seed=5).view(nw, 4, 3).cpu()
    on0 = y[:, :, 0] < 50.0
    assert torch.all(on0.sum(1) == 2)
    frac0 = on0.double().mean(0)
    _say("  two atoms: P(electron on atom 0) %s" % frac0.numpy().round(4))
    assert torch.all((frac0 - 0.5).abs() < 5 * math.sqrt(0.25 / nw))
    centre = torch.where(on0[:, :, None], torch.zeros(3, dtype=y.dtype), torch.tensor([100.0, 0.0, 0.0]))
    d = y - centre
    var_e = d.var(dim=(0, 2))
    ref2 = torch.tensor([1 / 64, (2 / 3) / 64 + (1 / 3) / 9, (1 / 3) / 64 + (2 / 3) / 9, 1 / 9], dtype=torch.float64)
    _say("  two atoms: per-electron variance %s (expected %s)" % (var_e.numpy().round(5), ref2.numpy().round(5)))
    assert torch.all((var_e - ref2).abs() < 0.02 * ref2)
    # normal domain: mean + L z
    mean = np.array([0.1, -0.2, 0.3])
    cov = np.array([[0.5, 0.1, 0.0], [0.1, 0.4, 0.05], [0.0, 0.05, 0.3]])
    z = w.draw({"method": "normal", "mean": mean, "sigma": cov}, nw, ne, seed=3).view(-1, 3).cpu().numpy()
    assert np.abs(z.mean(0) - mean).max() < 5 * math.sqrt(0.5 / len(z))
    assert np.abs(np.cov(z.T) - cov).max() < 0.01
    # uniform box and the centre domain
    u = w.draw({"method": "uniform", "min": -2.0, "max": 3.0}, nw, ne, seed=4).cpu()
    assert float(u.min()) >= -2.0 and float(u.max()) < 3.0
    assert float(u.var()) == pytest.approx(25.0 / 12.0, rel=0.01)
    c = w.draw({"method": "center", "center": [0, 0, 0]}, 1000, ne, seed=4).cpu()
    assert float(c.abs().max()) <= 1e-3


@pytest.mark.parametrize("proba", ["normal", "uniform"])
def test_metropolis_on_any_density(proba):
    """The generic path of Metropolis.__call__ (metropolis.py:85-177) on a
    density that is not a wavefunction: exp(-2 alpha |x|^2) has <x^2> = 1/(4 alpha);
    logspace gives the same statistics."""
    _say("case metropolis generic", proba)
    from qmctorch_amd.sampler import Metropolis

    alpha = 0.5
    pdf = lambda x: torch.exp(-2 * alpha * (x * x).sum(1))  # noqa: E731
    nw = 4096
    for logspace in (False, True):
        s = Metropolis(nwalkers=nw, nstep=300, step_size=1.5, ntherm=100, ndecor=25, nelec=1,
                       init={"min": -3, "max": 3}, move={"type": "all-elec", "proba": proba}, cuda=True,
                       logspace=logspace, seed=5)
        pos = s(pdf, with_tqdm=False).detach()
        x2 = (pos ** 2).mean(1).view(-1, nw).mean(0)
        m, err = float(x2.mean()), float(x2.std()) / math.sqrt(nw)
        _say("  logspace %s: <x^2> %.5f +- %.5f, acceptance %.3f" % (logspace, m, err, s.last_acceptance))
        assert abs(m - 1 / (4 * alpha)) < 6 * err


def test_all_electron_iterative_moves():
    """move type all-elec-iter (metropolis.py:221-254): each electron in turn,
    one accept/reject each.  Exact eigenfunction gives E_L = -1/2 everywhere;
    H2/STO-3G energy agrees with the all-electron device chain."""
    _say("case all-elec-iter")
    from qmctorch_amd.sampler import Metropolis
    from qmctorch_amd.wavefunction import SlaterJastrow

    mol = systems.h2()
    wf, _ = helpers.make_pair(mol, w_ee=1.0, perturb=0.0)
    nw = 2000
    init = helpers.positions(mol, nw, seed=6).to(DEV)
    res = {}
    for mtype in ("all-elec-iter", "all-elec"):
        s = Metropolis(nwalkers=nw, nstep=300, step_size=0.2, ntherm=100, ndecor=40, nelec=mol.nelec,
                       init=mol.domain("normal"), move={"type": mtype, "proba": "normal"}, cuda=True, seed=8)
        pos = s(wf.pdf, pos=init.clone(), with_tqdm=False).detach()
        e = wf.local_energy(pos).view(-1)
        res[mtype] = (float(e.mean()), float(e.var()) / len(e), s.last_acceptance)
    se = math.sqrt(res["all-elec-iter"][1] + res["all-elec"][1])
    _say("  iter %.5f all %.5f se %.5f acc %.3f/%.3f" % (res["all-elec-iter"][0], res["all-elec"][0], se,
                                                        res["all-elec-iter"][2], res["all-elec"][2]))
    assert abs(res["all-elec-iter"][0] - res["all-elec"][0]) < 5 * se
    assert res["all-elec-iter"][2] > res["all-elec"][2]  # single-electron moves are accepted more often
