"""Hamiltonian Monte Carlo and the device drift (needs an MI355X).

Parity bars:
  * drift grad_i ln|psi| (FP64 kernels) against the oracle's Jacobi gradient
    divided by psi, 1e-9 relative; log|psi| within 1e-10;
  * the FP64 HMC chain against the oracle's restatement of
    Hamiltonian._step (hamiltonian.py:143-201) driven by the same momenta and
    uniforms (Philox4x32-10 and Box-Muller recomputed on the host):
    identical accept/reject decisions and positions within 1e-8;
  * statistics: zero-variance local energy on an exact eigenfunction,
    <x^2> of a Gaussian density, and the H2 energy of the HMC chain against
    the Metropolis chain within 5 standard errors.
"""
import math
import sys

import numpy as np
import pytest
import torch

from oracle import vmc_oracle as O
from tests import helpers, systems
from tests.test_native_abi import philox_host

pytestmark = pytest.mark.gpu
DEV = "cuda"
TWO32 = 2.0 ** -32


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _double():
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


def _hmc(mol, nw, nstep, L, eps, seed, ntherm=0, ndecor=1, init=None):
    from qmctorch_amd.sampler import Hamiltonian

    return Hamiltonian(nwalkers=nw, nstep=nstep, step_size=eps, L=L, ntherm=ntherm, ndecor=ndecor,
                       nelec=mol.nelec, init=init or mol.domain("normal"), cuda=True, seed=seed)


def _host_noise(seed, nw, nelec, step):
    """Momenta (nw, 3 nelec) and uniforms (nw,) of one HMC step as the kernel draws them."""
    key = (seed & 0xFFFFFFFF, seed >> 32)
    p = np.zeros((nw, 3 * nelec))
    u = np.zeros(nw)
    for w in range(nw):
        for e in range(nelec):
            r = philox_host((w & 0xFFFFFFFF, w >> 32, step, e | 0x40000000), key)
            r1 = math.sqrt(-2.0 * math.log((r[0] + 1.0) * TWO32))
            r2 = math.sqrt(-2.0 * math.log((r[2] + 1.0) * TWO32))
            p[w, 3 * e + 0] = r1 * math.cos(2.0 * math.pi * r[1] * TWO32)
            p[w, 3 * e + 1] = r1 * math.sin(2.0 * math.pi * r[1] * TWO32)
            p[w, 3 * e + 2] = r2 * math.cos(2.0 * math.pi * r[3] * TWO32)
        u[w] = philox_host((w & 0xFFFFFFFF, w >> 32, step, 0x80000000), key)[0] * TWO32
    return torch.as_tensor(p), torch.as_tensor(u)


@pytest.mark.parametrize("case", ["h2_sto3g_ee", "h2o_ccpvdz_ee_en", "lih_631g_cas22_ee_en", "lih_adf_sto_sd22",
                                  "lih_631g_sd46_ee_en", "h2o_ccpvdz_sd46_ee"])
def test_drift_fp64_vs_oracle(case):
    """FP64 drift = Jacobi gradient / psi of the oracle (sd46: 15 unique
    configurations per spin, the excitation path's Gamma rows)."""
    _say("case drift fp64", case)
    mol, kw = helpers.CASES[case]()
    wf, orc = helpers.make_pair(mol, **kw)
    pos = helpers.positions(mol, 45, seed=21)
    dr, la, sg = wf.drift(pos.to(DEV), "fp64")
    psi = orc.forward(pos).detach().view(-1)
    g = orc.gradients_jacobi(pos.clone().requires_grad_()).detach()
    ref = (g / psi[:, None]).view(45, mol.nelec, 3)
    torch.testing.assert_close(dr.cpu(), ref, rtol=1e-9, atol=1e-10)
    torch.testing.assert_close(la.cpu(), torch.log(psi.abs()), rtol=1e-10, atol=1e-10)
    assert torch.equal(sg.cpu(), torch.sign(psi))


@pytest.mark.parametrize("case", ["h2o_ccpvdz_ee_en", "h2o_ccpvdz_sd46_ee"])
def test_drift_mixed_vs_fp64(case):
    _say("case drift mixed", case)
    mol, kw = helpers.CASES[case]()
    wf, _ = helpers.make_pair(mol, precision="mixed", **kw)
    pos = helpers.positions(mol, 512, seed=22).to(DEV)
    d32, la32, _ = wf.drift(pos, "mixed")
    d64, la64, _ = wf.drift(pos, "fp64")
    rel = ((d32 - d64).norm(dim=-1) / (1.0 + d64.norm(dim=-1))).view(-1)
    _say("  drift rel err median %.2e p99 %.2e; log|psi| max %.2e" %
         (float(rel.median()), float(torch.quantile(rel, 0.99)), float((la32 - la64).abs().max())))
    assert float(rel.median()) < 1e-5 and float(torch.quantile(rel, 0.99)) < 1e-3
    assert float((la32 - la64).abs().max()) < 1e-3


@pytest.mark.parametrize("case,eps", [("h2o_ccpvdz_ee_en", 0.1), ("lih_631g_cas22_ee_en", 0.1),
                                      ("lih_631g_cas22_ee_en", 0.8), ("lih_631g_sd46_ee_en", 0.1)])
def test_hmc_fp64_replays_oracle_step(case, eps):
    """Same momenta and uniforms: the device chain is the oracle's chain."""
    _say("case hmc replay", case, eps)
    mol, kw = helpers.CASES[case]()
    wf, orc = helpers.make_pair(mol, **kw)
    nw, nstep, L, seed = 12, 3, 4, 0x1234_5678_9ABC
    init = helpers.positions(mol, nw, seed=23)
    s = _hmc(mol, nw, nstep, L, eps, seed)
    kept = s(wf.pdf, pos=init.to(DEV).clone(), with_tqdm=False).detach().cpu().view(nstep, nw, -1)
    q = init.clone()
    nacc = 0
    for step in range(nstep):
        p, u = _host_noise(seed, nw, mol.nelec, step)
        q, acc = O.hmc_step(orc.pdf, q, p, eps, L, u)
        nacc += int(acc.sum())
        torch.testing.assert_close(kept[step], q, rtol=0, atol=1e-8)
    _say("  accepted %d of %d" % (nacc, nw * nstep))
    assert nacc > 0
    if eps > 0.5 or case.startswith("h2o"):
        assert nacc < nw * nstep  # the rejection branch is exercised too
    assert abs(s.last_acceptance - nacc / (nw * nstep)) < 1e-12


def test_hmc_exact_eigenfunction():
    """H atom, 1s STO zeta=1: every HMC sample has E_L = -1/2."""
    _say("case hmc exact")
    from qmctorch_amd.wavefunction import SlaterJastrow

    mol = systems.hydrogen_atom("sto", 1.0)
    wf = SlaterJastrow(mol, jastrow=None, cuda=True)
    s = _hmc(mol, 512, 40, 5, 0.2, seed=3, ntherm=10, ndecor=5)
    pos = s(wf.pdf, with_tqdm=False)
    assert pos.shape == (512 * 6, 3) and pos.requires_grad
    el = wf.local_energy(pos.detach()).view(-1)
    assert torch.allclose(el, torch.full_like(el, -0.5), atol=1e-12)
    assert 0.3 < s.last_acceptance < 1.0


def test_hmc_gaussian_moments():
    """|psi|^2 of a normalised s Gaussian (alpha = 0.5): <x^2> = 1/(4 alpha)."""
    _say("case hmc moments")
    from qmctorch_amd.wavefunction import SlaterJastrow

    alpha = 0.5
    mol = systems.hydrogen_atom("gto_pure", alpha)
    wf = SlaterJastrow(mol, jastrow=None, cuda=True)
    s = _hmc(mol, 4096, 120, 5, 0.3, seed=4, ntherm=20, ndecor=10)
    pos = s(wf.pdf, with_tqdm=False).detach()
    x2 = (pos ** 2).mean(1)
    m = float(x2.mean())
    err = float(x2.std()) / math.sqrt(4096)
    _say("  <x^2> = %.5f +- %.5f (exact %.5f), acceptance %.3f" % (m, err, 1 / (4 * alpha), s.last_acceptance))
    assert abs(m - 1 / (4 * alpha)) < 6 * err


def test_hmc_batch_invariance():
    """A walker's chain depends on (seed, walker id, step) only."""
    _say("case hmc batch invariance")
    mol = systems.h2o()
    wf, _ = helpers.make_pair(mol, w_ee=1.0, w_en=1.0, precision="mixed")
    p0 = helpers.positions(mol, 96, seed=5).to(DEV)
    a = _hmc(mol, 96, 6, 3, 0.05, seed=9)(wf.pdf, pos=p0.clone(), with_tqdm=False).detach().view(6, 96, -1)
    b = _hmc(mol, 40, 6, 3, 0.05, seed=9)(wf.pdf, pos=p0[:40].clone(), with_tqdm=False).detach().view(6, 40, -1)
    assert torch.equal(a[:, :40], b)


def test_hmc_energy_matches_metropolis():
    """H2/STO-3G with the e-e Jastrow: HMC (mixed) and Metropolis sample the same density."""
    _say("case hmc vs metropolis")
    from qmctorch_amd.sampler import Metropolis

    mol = systems.h2()
    wf, _ = helpers.make_pair(mol, w_ee=1.0, perturb=0.0, precision="mixed")
    nw = 2000
    init = helpers.positions(mol, nw, seed=6).to(DEV)
    ph = _hmc(mol, nw, 60, 5, 0.25, seed=11, ntherm=20, ndecor=10)(wf.pdf, pos=init.clone(), with_tqdm=False)
    pm = Metropolis(nwalkers=nw, nstep=300, step_size=0.2, ntherm=100, ndecor=40, nelec=mol.nelec,
                    init=mol.domain("normal"), move={"type": "all-elec", "proba": "normal"}, cuda=True,
                    seed=12)(wf.pdf, pos=init.clone(), with_tqdm=False)
    eh = wf.local_energy(ph.detach()).view(-1)
    em = wf.local_energy(pm.detach()).view(-1)
    se = math.sqrt(float(eh.var()) / len(eh) + float(em.var()) / len(em))
    _say("  hmc %.5f  metropolis %.5f  se %.5f" % (eh.mean(), em.mean(), se))
    assert abs(float(eh.mean() - em.mean())) < 5 * se


def test_hmc_argument_errors():
    from qmctorch_amd import _native as N

    mol, kw = helpers.CASES["h2_sto3g_ee"]()
    wf, _ = helpers.make_pair(mol, **kw)
    prm, keep = wf._params("fp64")
    x = helpers.positions(mol, 4, seed=1).to(DEV)
    la = torch.empty(4, dtype=torch.float64, device=DEV)
    sg = torch.empty_like(la)
    wk = N.QmcWalkers()
    wk.pos, wk.logabs, wk.sign, wk.naccept, wk.walker_id = x.data_ptr(), la.data_ptr(), sg.data_ptr(), 0, 0
    hm = N.QmcHmcDesc()
    hm.nstep, hm.traj_length, hm.step_size = 2, 0, 0.1
    lib = N.load()
    rc = lib.qmc_hmc_sweep(wf.device_plan().handle, prm, N.QMC_PREC_FP64, wk, 4, hm, N.current_stream())
    assert rc == 1 and b"traj_length" in lib.qmc_last_error()  # QMC_ERR_ARG
