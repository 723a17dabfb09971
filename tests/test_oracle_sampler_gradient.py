"""Pins for the oracle's Metropolis chain and manual-gradient estimator (CPU).

``O.metropolis`` restates sampler/metropolis.py:85-298; it is pinned here by
what the mathematics fixes, independently of the device sampler:

* an exact eigenfunction (hydrogen 1s STO, zeta = 1) gives E_L = -1/2 on every
  kept sample, and |psi|^2 = e^{-2r}/pi has <r> = 3/2 (a chain that samples
  |psi| instead of |psi|^2, or accepts downhill moves wrongly, misses it);
* the density of a normalised s Gaussian, exp(-2 alpha r^2), has <x^2> = 1/(4 alpha);
* a flat density accepts every move, so one step displaces each coordinate by
  the proposal: variance sigma = step / (2 sqrt(2 ln 2)) (the reference uses
  sigma as the covariance, metropolis.py:206-212) for normal moves and
  step^2/3 for uniform ones;
* the ntherm / ndecor bookkeeping (metropolis.py:113-114, 150-160) keeps
  ceil((nstep - ntherm) / ndecor) configurations per walker.

``O.manual_gradient`` restates Solver.evaluate_grad_manual
(solver/solver.py:372-431): g = 2/N sum_i (E_L,i - <E_L>) d ln psi_i / d theta.
On fixed positions this is exactly the derivative at theta_0 of the reweighted
energy E(theta) = sum_i r_i E_L,0(x_i) / sum_i r_i, r_i = psi_theta(x_i)^2 /
psi_0(x_i)^2, with the local energies held at theta_0.  The test differentiates
E(theta) by central finite differences (h = 1e-5 and 2e-5, Richardson-extrapolated), which evaluates only psi and
never autograd, for mo_modifier entries, both Pade weights and the CI weights.
"""
import math

import numpy as np
import pytest
import torch

from oracle import vmc_oracle as O
from tests import systems


def _wf(mol, configs="ground_state", ee=1.0, en=None, ci=None):
    return O.OracleSlaterJastrow(mol.basis, mol.atom_coords, mol.atomic_number, mol.nup, mol.ndown,
                                 configs=configs, jastrow_ee=ee, jastrow_en=en, ci_weights=ci)


def _gauss(nw, ndim, seed, sigma=1.0):
    g = torch.Generator().manual_seed(seed)
    return sigma * torch.randn(nw, ndim, generator=g, dtype=torch.float64)


def test_metropolis_exact_eigenfunction_and_mean_radius():
    mol = systems.hydrogen_atom("sto", 1.0)
    wf = _wf(mol, ee=None)
    nw = 1024
    pos, rate = O.metropolis(wf.pdf, _gauss(nw, 3, 1), 400, step_size=0.8, ntherm=200, ndecor=50, seed=3)
    el = wf.local_energy(pos).detach()
    assert torch.allclose(el, torch.full_like(el, -0.5), atol=1e-12)
    assert 0.05 < rate < 0.95
    r = pos.norm(dim=1).view(-1, nw)  # (kept, walkers): rows 50 steps apart
    # error bar from the per-walker means (walkers are independent chains)
    per_walker = r.mean(0)
    err = float(per_walker.std()) / math.sqrt(nw)
    assert abs(float(r.mean()) - 1.5) < 5 * err, (float(r.mean()), err)


@pytest.mark.parametrize("proba,step", [("normal", 1.5), ("uniform", 1.5)])
def test_metropolis_gaussian_second_moment(proba, step):
    alpha = 0.5
    pdf = lambda x: torch.exp(-2 * alpha * (x * x).sum(1))  # noqa: E731
    nw = 2048
    pos, _ = O.metropolis(pdf, _gauss(nw, 3, 2, 2.0), 300, step_size=step, ntherm=100, ndecor=25, proba=proba,
                          seed=4)
    x2 = (pos ** 2).mean(1).view(-1, nw).mean(0)
    m, err = float(x2.mean()), float(x2.std()) / math.sqrt(nw)
    assert abs(m - 1 / (4 * alpha)) < 6 * err, (m, err)


def test_metropolis_detailed_balance_rejects_downhill_correctly():
    """Two-level check of the acceptance rule: for the density exp(-x^2/2) in one
    coordinate per electron, starting every walker at x = 0 (the maximum), the
    first-step acceptance of a move d is exp(-d^2/2); its average over the
    proposal N(0, sigma) is 1/sqrt(1 + sigma)."""
    pdf = lambda x: torch.exp(-0.5 * (x * x).sum(1))  # noqa: E731
    nw = 200000
    step = 0.6
    sigma = step / (2 * math.sqrt(2 * math.log(2.0)))
    # three coordinates, density exp(-|x|^2/2): acceptance exp(-|d|^2/2), mean (1+sigma)^(-3/2)
    _, rate = O.metropolis(pdf, torch.zeros(nw, 3, dtype=torch.float64), 1, step_size=step, seed=5)
    exact = (1 + sigma) ** -1.5
    assert abs(rate - exact) < 5 * math.sqrt(exact * (1 - exact) / nw), (rate, exact)


@pytest.mark.parametrize("proba", ["normal", "uniform"])
def test_metropolis_flat_density_proposal(proba):
    nw, step = 100000, 0.2
    pdf = lambda x: torch.ones(x.shape[0], dtype=x.dtype)  # noqa: E731
    x0 = torch.zeros(nw, 6, dtype=torch.float64)
    d, rate = O.metropolis(pdf, x0.clone(), 1, step_size=step, proba=proba, seed=6)
    assert rate == 1.0
    var = step / (2 * math.sqrt(2 * math.log(2.0))) if proba == "normal" else step ** 2 / 3
    v = d.var(0)
    assert torch.all((v - var).abs() < 5 * var * math.sqrt(2.0 / nw)), (v, var)
    assert torch.all(d.mean(0).abs() < 5 * math.sqrt(var / nw))


@pytest.mark.parametrize("nstep,ntherm,ndecor", [(10, -1, 1), (10, 0, 1), (10, 3, 2), (9, 2, 3)])
def test_metropolis_kept_count(nstep, ntherm, ndecor):
    pdf = lambda x: torch.ones(x.shape[0], dtype=x.dtype)  # noqa: E731
    nw = 7
    pos, _ = O.metropolis(pdf, torch.zeros(nw, 3, dtype=torch.float64), nstep, ntherm=ntherm, ndecor=ndecor, seed=0)
    nt = nstep + ntherm if ntherm < 0 else ntherm
    assert pos.shape == (nw * math.ceil((nstep - nt) / ndecor), 3)


def _reweighted_energy(wf, pos, el0, psi0):
    with torch.no_grad():
        r = (wf(pos).view(-1) / psi0) ** 2
    return float((r * el0).sum() / r.sum())


def _central(wf, param, index, pos, el0, psi0, h):
    with torch.no_grad():
        v = float(param.view(-1)[index])
        param.view(-1)[index] = v + h
        ep = _reweighted_energy(wf, pos, el0, psi0)
        param.view(-1)[index] = v - h
        em = _reweighted_energy(wf, pos, el0, psi0)
        param.view(-1)[index] = v
    return (ep - em) / (2 * h)


def _fd(wf, param, index, pos, el0, psi0, h=1e-5):
    """Central differences at h and 2h, Richardson-extrapolated (error O(h^4)):
    near-node walkers give large third derivatives."""
    d1 = _central(wf, param, index, pos, el0, psi0, h)
    d2 = _central(wf, param, index, pos, el0, psi0, 2 * h)
    return (4 * d1 - d2) / 3


@pytest.mark.parametrize("case", ["h2o_ee_en", "lih_cas22_ee_en"])
def test_manual_gradient_equals_fd_of_reweighted_energy(case):
    if case == "h2o_ee_en":
        mol = systems.h2o()
        wf = _wf(mol, ee=0.9, en=1.1)
    else:
        mol = systems.lih()
        wf = _wf(mol, configs="cas(2,2)", ee=0.8, en=1.2, ci=np.array([1.0, 0.1, -0.2, 0.05]))
    rng = np.random.default_rng(0)
    with torch.no_grad():
        wf.mo_modifier.copy_(torch.as_tensor(1.0 + 0.1 * rng.standard_normal(tuple(wf.mo_modifier.shape))))
    pos = torch.as_tensor(systems.gaussian_positions(mol, 48, seed=12))
    eloc, g = O.manual_gradient(wf, pos.clone().requires_grad_())
    el0 = eloc.detach().view(-1)
    with torch.no_grad():
        psi0 = wf(pos).view(-1)
    named = dict(wf.named_parameters())
    checks = []
    nmo = g["mo_modifier"].shape[1]
    # the four mo_modifier entries with the largest gradients among the occupied columns
    gm = g["mo_modifier"][:, : max(mol.nup, mol.ndown) + 1].abs()
    for flat in torch.topk(gm.reshape(-1), 4).indices.tolist():
        i, k = divmod(flat, gm.shape[1])
        checks.append(("mo_modifier", i * nmo + k, float(g["mo_modifier"][i, k])))
    for key in ("w_ee", "w_en"):
        checks.append((key, 0, float(g[key].view(-1)[0])))
    if case == "lih_cas22_ee_en":
        for n in (1, 3):
            checks.append(("fc_weight", n, float(g["fc_weight"].view(-1)[n])))
    for name, i, got in checks:
        fd = _fd(wf, named[name], i, pos, el0, psi0)
        assert abs(fd) > 1e-4 and abs(got - fd) <= 1e-6 * abs(fd) + 1e-8, (name, i, got, fd)
    # and the estimator is not the full derivative of <E_L> (which adds <dE_L/dtheta>)
    p = named["w_ee"]
    with torch.no_grad():
        v = float(p.view(-1)[0])
        p.view(-1)[0] = v + 1e-5
        e_p = float(wf.local_energy(pos).mean())
        p.view(-1)[0] = v - 1e-5
        e_m = float(wf.local_energy(pos).mean())
        p.view(-1)[0] = v
    full = (e_p - e_m) / 2e-5
    assert abs(full - float(g["w_ee"].view(-1)[0])) > 1e-4
