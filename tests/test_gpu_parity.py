"""Parity of the HIP path with the oracle (needs an MI355X).

Tolerances (FP64 path): psi and local energies within 1e-10 relative to the
oracle on the same positions (the bar stated for the FP64 path); AO values
within 1e-12.  The mixed-precision path is measured against the FP64 path.
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


@pytest.mark.parametrize("which", ["h2o", "lih_adf"])
def test_ao_values_and_derivatives(which):
    _say("case ao", which)
    mol = systems.h2o() if which == "h2o" else systems.lih_adf()
    wf, orc = helpers.make_pair(mol, w_ee=None)
    pos = helpers.positions(mol, 37, seed=2)
    ao, dao, d2ao = wf.ao(pos.to(DEV), derivative=[0, 1, 2], sum_grad=False)
    oa, od, ol = orc.ao(pos.clone().requires_grad_(), "all")
    torch.testing.assert_close(ao.cpu(), oa.detach(), rtol=1e-12, atol=1e-13)
    torch.testing.assert_close(dao.cpu(), od.detach(), rtol=1e-11, atol=1e-12)
    torch.testing.assert_close(d2ao.cpu(), ol.detach(), rtol=1e-10, atol=1e-11)


@pytest.mark.parametrize("case", list(helpers.CASES))
def test_psi_and_local_energy_fp64(case):
    _say("case fp64", case)
    mol, kw = helpers.CASES[case]()
    wf, orc = helpers.make_pair(mol, **kw)
    n = 23 if mol.nelec > 20 else 61
    pos = helpers.positions(mol, n, seed=3)
    psi = wf(pos.to(DEV)).detach().cpu()
    psi_o = orc(pos).detach()
    torch.testing.assert_close(psi, psi_o, rtol=1e-10, atol=1e-300)
    parts = wf.local_energy_parts(pos.to(DEV))
    el_o = orc.local_energy(pos).detach().view(-1)
    ke_o = orc.kinetic_energy_jacobi(pos).detach().view(-1)
    el = parts["eloc"].cpu()
    ke = parts["kinetic"].cpu()
    scale = el_o.abs().clamp(min=1.0)
    _say("  max rel dE_L", float(((el - el_o).abs() / scale).max()))
    assert float(((el - el_o).abs() / scale).max()) < 1e-10
    assert float(((ke - ke_o).abs() / ke_o.abs().clamp(min=1.0)).max()) < 1e-10
    pdf = wf.pdf(pos.to(DEV)).cpu()
    torch.testing.assert_close(pdf, psi_o.view(-1) ** 2, rtol=1e-10, atol=1e-300)


def test_lih_golden_local_energy_on_device():
    """E_L of the reference's stored positions with the stored wavefunction state."""
    _say("case golden LiH")
    g = np.load(systems.GOLDEN + "/LiH_golden.npz")
    mol = systems.lih_adf()
    p = lambda k: g["SolverSlaterJastrow/wf/" + k]  # noqa: E731
    from qmctorch_amd.wavefunction import SlaterJastrow

    wf = SlaterJastrow(mol, configs="single_double(2,2)", cuda=True)
    W = p("mo_scf/weight").T @ p("mo/weight").T
    with torch.no_grad():
        wf.mo.mo_scf = torch.as_tensor(W).to(DEV)
        wf.fc.weight.copy_(torch.as_tensor(p("fc/weight")))
        wf.jastrow.jastrow_kernel.weight.copy_(torch.as_tensor(p("jastrow/weight")))
    el = wf.local_energy(torch.as_tensor(g["single_point/pos"]).to(DEV)).cpu().numpy()
    ref = g["single_point/local_energy"]
    _say("  max |dE_L| vs reference golden", np.abs(el - ref).max())
    assert np.abs(el - ref).max() < 1e-9
    assert abs(el.mean() - float(g["single_point/energy"])) < 1e-10


@pytest.mark.parametrize("case", ["lih_631g_cas22_ee_en", "h2o_ccpvdz_ee_en", "c6h6_ccpvdz_ee_en"])
def test_mixed_precision_error(case):
    _say("case mixed", case)
    mol, kw = helpers.CASES[case]()
    wf, _ = helpers.make_pair(mol, **kw)
    pos = helpers.positions(mol, 512, seed=4).to(DEV)
    e64 = wf.local_energy_parts(pos, "fp64")["eloc"]
    e32 = wf.local_energy_parts(pos, "mixed")["eloc"]
    d = (e32 - e64).abs() / mol.nelec
    _say("  mixed-vs-fp64 |dE_L|/Ne  max %.3e  median %.3e  mean %.3e" % (d.max(), d.median(), d.mean()))
    l64, _ = wf.log_psi(pos, "fp64")
    l32, _ = wf.log_psi(pos, "mixed")
    assert float((l64 - l32).abs().median()) < 1e-5
    assert float(d.median()) < 1e-4


def test_philox_device_matches_host_and_kat():
    _say("case philox")
    from qmctorch_amd import _native as N
    from tests.test_native_abi import KAT, philox_host

    rng = np.random.default_rng(0)
    ins = [list(c) + list(k) for c, k, _ in KAT] + rng.integers(0, 2 ** 32, size=(61, 6)).tolist()
    a = torch.tensor(np.asarray(ins, dtype=np.uint32).view(np.int32), device=DEV)
    out = torch.empty(len(ins), 4, dtype=torch.int32, device=DEV)
    N.check(N.load().qmc_philox4x32_10(a.data_ptr(), out.data_ptr(), len(ins), N.current_stream()))
    got = out.cpu().numpy().view(np.uint32)
    for i, row in enumerate(ins):
        assert tuple(got[i]) == tuple(philox_host(row[:4], row[4:6]))
    for i, (_, _, ref) in enumerate(KAT):
        assert tuple(got[i]) == ref


def _sampler(mol, nw, nstep, step=0.2, seed=11, init=None, ntherm=-1, ndecor=1):
    from qmctorch_amd.sampler import Metropolis

    return Metropolis(nwalkers=nw, nstep=nstep, step_size=step, ntherm=ntherm, ndecor=ndecor, nelec=mol.nelec,
                      init=init or mol.domain("normal"), move={"type": "all-elec", "proba": "normal"}, cuda=True,
                      seed=seed)


def test_sweep_exact_eigenfunction():
    """H atom, 1s STO zeta=1: every sampled configuration has E_L = -1/2."""
    _say("case sweep exact")
    from qmctorch_amd.wavefunction import SlaterJastrow

    mol = systems.hydrogen_atom("sto", 1.0)
    wf = SlaterJastrow(mol, jastrow=None, cuda=True)
    s = _sampler(mol, 512, 200, step=0.5)
    pos = s(wf.pdf, with_tqdm=False)
    el = wf.local_energy(pos).view(-1)
    assert torch.allclose(el, torch.full_like(el, -0.5), atol=1e-12)
    assert 0.05 < s.last_acceptance < 0.999


def test_sweep_proposal_distribution():
    """A nearly flat density (s Gaussian with alpha = 1e-9) accepts every move, so one
    step displaces each coordinate by the proposal itself: mean 0 and variance
    sigma2 = step / (2 sqrt(2 ln 2)) per coordinate (metropolis.py:206-212), the
    same in x, y and z, Gaussian kurtosis 3 -- for normal and uniform-cube moves."""
    _say("case sweep proposal")
    from qmctorch_amd.sampler import Metropolis
    from qmctorch_amd.wavefunction import SlaterJastrow

    mol = systems.hydrogen_atom("gto_pure", 1e-9)
    wf = SlaterJastrow(mol, jastrow=None, cuda=True)
    nw = 1 << 17
    for proba, var, kurt in (("normal", 0.2 / (2 * math.sqrt(2 * math.log(2.0))), 3.0), ("uniform", 0.2 ** 2 / 3, 1.8)):
        s = Metropolis(nwalkers=nw, nstep=1, step_size=0.2, nelec=1, init=mol.domain("normal"),
                       move={"type": "all-elec", "proba": proba}, cuda=True, seed=21)
        x0 = torch.zeros(nw, 3, dtype=torch.float64, device=DEV)
        d = s(wf.pdf, pos=x0.clone(), with_tqdm=False).detach()
        assert s.last_acceptance > 0.9999
        m = d.mean(0)
        v = d.var(0)
        k = (d ** 4).mean(0) / v ** 2
        _say("  %s: mean %s var %s (exact %.6f) kurtosis %s" % (proba, m.tolist(), v.tolist(), var, k.tolist()))
        assert torch.all(m.abs() < 5 * math.sqrt(var / nw))
        assert torch.all((v - var).abs() < 5 * var * math.sqrt(2.0 / nw))
        assert torch.all((k - kurt).abs() < 0.05)


def test_sweep_gaussian_moments():
    """|psi|^2 of a normalised s Gaussian (alpha = 0.5) gives <x^2> = 1/(4 alpha)."""
    _say("case sweep moments")
    from qmctorch_amd.wavefunction import SlaterJastrow

    alpha = 0.5
    mol = systems.hydrogen_atom("gto_pure", alpha)
    wf = SlaterJastrow(mol, jastrow=None, cuda=True)
    s = _sampler(mol, 4096, 400, step=1.0, ntherm=100, ndecor=20)
    pos = s(wf.pdf, with_tqdm=False).detach()
    x2 = (pos ** 2).mean(1)  # mean over x, y, z
    m = float(x2.mean())
    # independent walkers x 15 kept steps (decorrelated by 20 steps): conservative error bar
    err = float(x2.std()) / math.sqrt(4096)
    _say("  <x^2> = %.5f +- %.5f (exact %.5f)" % (m, err, 1 / (4 * alpha)))
    assert abs(m - 1 / (4 * alpha)) < 6 * err


@pytest.mark.parametrize("case", ["lih_631g_cas28_ee", "c6h6_ccpvdz_ee_en", "lih_631g_sd46_ee_en",
                                  "h2o_ccpvdz_sd46_ee"])
def test_sweep_logpsi_matches_fp64_evaluation(case):
    """The log|psi| a mixed-precision sweep keeps for its walkers equals an FP64
    evaluation at the final positions (register buckets above 8: the LDS-row
    elimination, with many walkers per wavefront for LiH CAS(2,8))."""
    _say("case sweep logpsi", case)
    from qmctorch_amd import _native as N

    mol, kw = helpers.CASES[case]()
    wf, _ = helpers.make_pair(mol, precision="mixed", **kw)
    nw = 96 if mol.nelec > 20 else 700
    x = helpers.positions(mol, nw, seed=31).to(DEV).contiguous()
    la = torch.empty(nw, dtype=torch.float64, device=DEV)
    sg = torch.empty_like(la)
    na = torch.zeros(nw, dtype=torch.int32, device=DEV)
    wk = N.QmcWalkers()
    wk.pos, wk.logabs, wk.sign, wk.naccept, wk.walker_id = x.data_ptr(), la.data_ptr(), sg.data_ptr(), na.data_ptr(), 0
    sw = N.QmcSweepDesc()
    sw.step0, sw.nstep, sw.init_logpsi = 0, 20, 1
    sw.sigma2, sw.seed, sw.step_size, sw.keep_every = 0.2 / (2 * math.sqrt(2 * math.log(2.0))), 5, 0.2, 1
    prm, keep = wf._params("mixed")
    N.check(N.load().qmc_metropolis_sweep(wf.device_plan().handle, prm, N.QMC_PREC_MIXED, wk, nw, sw,
                                          N.current_stream()))
    l64, s64 = wf.log_psi(x, "fp64")
    d = (la - l64).abs()
    _say("  accepted %d, max |log|psi| - FP64| %.2e" % (int(na.sum()), float(d.max())))
    assert int(na.sum()) > 0
    assert float(d.max()) < 1e-3 * max(1.0, float(l64.abs().max()) / 50)
    assert torch.equal(sg, s64)


def test_sweep_invariant_under_wavefunction_scale():
    """psi^2 is sampled up to its normalisation (metropolis.py replaces only an
    exactly zero density by eps): scaling every MO coefficient by 1e-6 shifts
    log|psi| by a constant and leaves the chain unchanged, also where |psi| is
    far below 1e-8."""
    _say("case sweep scale invariance")
    mol, kw = helpers.CASES["h2o_ccpvdz_ee_en"]()
    wf, _ = helpers.make_pair(mol, precision="fp64", **kw)
    wfs, _ = helpers.make_pair(mol, precision="fp64", **kw)
    with torch.no_grad():
        wfs.mo.mo_modifier.mul_(1e-6)
    init = helpers.positions(mol, 240, seed=41).to(DEV)
    a = _sampler(mol, 240, 30, seed=3)(wf.pdf, pos=init.clone(), with_tqdm=False).detach()
    b = _sampler(mol, 240, 30, seed=3)(wfs.pdf, pos=init.clone(), with_tqdm=False).detach()
    la, _ = wfs.log_psi(b, "fp64")
    same = (a == b).all(dim=1).double().mean().item()
    _say("  log|psi| of the scaled chain in [%.1f, %.1f], identical walkers %.4f"
         % (float(la.min()), float(la.max()), same))
    assert float(la.max()) < -40.0
    assert same >= 0.99


def test_sweep_batch_invariance():
    """A walker's chain depends on (seed, walker id, step) only, not on the batch."""
    _say("case sweep batch invariance")
    mol = systems.h2o()
    wf, _ = helpers.make_pair(mol, w_ee=1.0, w_en=1.0, precision="mixed")
    torch.manual_seed(0)
    p0 = helpers.positions(mol, 96, seed=5).to(DEV)
    a = _sampler(mol, 96, 30)(wf.pdf, pos=p0.clone(), with_tqdm=False).detach()
    b = _sampler(mol, 40, 30)(wf.pdf, pos=p0[:40].clone(), with_tqdm=False).detach()
    assert torch.equal(a[:40], b)


def test_sweep_energy_matches_oracle_chain():
    """H2/STO-3G: the device chain and the oracle's host-RNG chain sample the same
    distribution (mean local energy within the combined standard error)."""
    _say("case sweep vs oracle chain")
    mol = systems.h2()
    wf, orc = helpers.make_pair(mol, w_ee=1.0, perturb=0.0)
    nw, nstep = 2000, 300
    torch.manual_seed(0)
    init = helpers.positions(mol, nw, seed=6)
    pos_d = _sampler(mol, nw, nstep, ntherm=100, ndecor=40)(wf.pdf, pos=init.to(DEV).clone(), with_tqdm=False)
    e_d = wf.local_energy(pos_d.detach()).view(-1).cpu()
    pos_o, _ = O.metropolis(orc.pdf, init.clone(), nstep, step_size=0.2, ntherm=100, ndecor=40, seed=1)
    e_o = orc.local_energy(pos_o).detach().view(-1)
    se = math.sqrt(float(e_d.var()) / len(e_d) + float(e_o.var()) / len(e_o))
    _say("  device %.5f  oracle %.5f  se %.5f" % (e_d.mean(), e_o.mean(), se))
    assert abs(float(e_d.mean() - e_o.mean())) < 5 * se


@pytest.mark.parametrize("case", ["h2o_ccpvdz_ee_en", "lih_631g_cas22_ee_en", "lih_adf_sto_sd22"])
def test_manual_gradient_adjoint_vs_oracle_autograd(case):
    """psi.backward(2/N (E_L - <E_L>)/psi) (solver.py:372-431): adjoint kernel vs autograd."""
    _say("case manual gradient", case)
    mol, kw = helpers.CASES[case]()
    wf, orc = helpers.make_pair(mol, **kw)
    pos = helpers.positions(mol, 64, seed=8)
    eloc_o, g_o = O.manual_gradient(orc, pos.clone().requires_grad_())
    wf.zero_grad()
    x = pos.to(DEV)
    with torch.no_grad():
        eloc = wf.local_energy(x)
    psi = wf(x)
    w = (eloc - eloc.mean()) / psi.detach() * (2.0 / len(psi))
    psi.backward(w)
    torch.testing.assert_close(wf.mo.mo_modifier.grad.cpu(), g_o["mo_modifier"], rtol=1e-8, atol=1e-10)
    torch.testing.assert_close(wf.ao.bas_exp.grad.cpu(), g_o["bas_exp"], rtol=1e-8, atol=1e-10)
    if g_o.get("fc_weight") is not None and wf.nci > 1:
        torch.testing.assert_close(wf.fc.weight.grad.cpu(), g_o["fc_weight"], rtol=1e-8, atol=1e-10)
    names = dict(wf.named_parameters())
    for key_d, key_o in [("jastrow.jastrow_kernel.weight", "w_ee"),
                         ("jastrow.jastrow_terms.0.jastrow_kernel.weight", "w_ee"),
                         ("jastrow.jastrow_terms.1.jastrow_kernel.weight", "w_en")]:
        if key_d in names and g_o.get(key_o) is not None:
            torch.testing.assert_close(names[key_d].grad.cpu(), g_o[key_o], rtol=1e-8, atol=1e-10)


def test_solver_single_point_and_run():
    _say("case solver")
    from qmctorch_amd.solver import Solver

    mol = systems.h2()
    wf, _ = helpers.make_pair(mol, w_ee=1.0, perturb=0.0)
    s = _sampler(mol, 1000, 200)
    opt = torch.optim.Adam(wf.parameters(), lr=0.01)
    solver = Solver(wf=wf, sampler=s, optimizer=opt, output="/tmp/qmc_test_h2.hdf5")
    obs = solver.single_point(with_tqdm=False)
    assert obs.local_energy.shape == (1000, 1)
    assert -1.4 < float(obs.energy) < -0.8
    before = wf.mo.mo_modifier.detach().clone()
    solver.configure(freeze=["ao"], track=["local_energy", "parameters"])
    out = solver.run(3)
    assert len(out.energy) == 4 and np.all(np.isfinite(out.energy))
    assert not torch.equal(before, wf.mo.mo_modifier.detach())
    # the HDF5 output in the reference's layout (solver_base.py:381-385, solver.py:254-264)
    from qmctorch_amd.utils.hdf5_utils import open_file

    with open_file("/tmp/qmc_test_h2.hdf5", "r") as f5:
        sp, opt_grp = f5["single_point"], f5["wf_opt"]
        kind = sp.attrs["type"]
        assert (kind.decode() if isinstance(kind, bytes) else str(kind)) == "single_point"
        np.testing.assert_array_equal(sp["local_energy"][()], obs.local_energy.cpu().numpy())
        np.testing.assert_array_equal(sp["pos"][()], obs.pos.detach().cpu().numpy())
        assert float(sp["energy"][()]) == float(obs.energy)
        np.testing.assert_allclose(opt_grp["energy"][()], np.asarray(out.energy, dtype=np.float64), rtol=0, atol=0)
    last = {k: v.detach().clone() for k, v in wf.state_dict().items()}
    with torch.no_grad():
        wf.mo.mo_modifier.add_(1.0)
    wf.load("/tmp/qmc_test_h2.hdf5", "wf_opt", "last")
    for k, v in wf.state_dict().items():
        assert torch.equal(v.cpu(), last[k].cpu()), k


@pytest.mark.parametrize("case", ["h2o_ccpvdz_ee_en", "lih_631g_cas22_ee_en", "lih_adf_sto_sd22",
                                  "lih_631g_sd46_ee_en", "h2o_ccpvdz_sd46_ee"])
def test_local_energy_adjoint_vs_oracle_autograd(case):
    """d(sum_w g_w E_L,w)/d theta: adjoint kernel vs autograd through the oracle's E_L
    (sd46: the excitation path's adjoint against every determinant taken explicitly)."""
    _say("case eloc adjoint", case)
    mol, kw = helpers.CASES[case]()
    wf, orc = helpers.make_pair(mol, **kw)
    pos = helpers.positions(mol, 48, seed=9)
    gw = torch.as_tensor(np.random.default_rng(1).standard_normal(48))
    orc.bas_exp.requires_grad_(False)
    orc.atom_coords.requires_grad_(False)
    el_o = orc.local_energy(pos.clone()).view(-1)
    params_o = {"mo_modifier": orc.mo_modifier, "fc_weight": orc.fc_weight}
    if orc.w_ee is not None:
        params_o["w_ee"] = orc.w_ee
    if orc.w_en is not None:
        params_o["w_en"] = orc.w_en
    go = dict(zip(params_o, torch.autograd.grad((el_o * gw).sum(), list(params_o.values()))))
    wf.ao.bas_exp.requires_grad_(False)
    wf.zero_grad()
    el = wf.local_energy_autodiff(pos.to(DEV)).view(-1)
    torch.testing.assert_close(el.detach().cpu(), el_o.detach(), rtol=1e-10, atol=1e-10)
    (el * gw.to(DEV)).sum().backward()
    torch.testing.assert_close(wf.mo.mo_modifier.grad.cpu(), go["mo_modifier"], rtol=1e-7, atol=1e-9)
    if wf.nci > 1:
        torch.testing.assert_close(wf.fc.weight.grad.cpu(), go["fc_weight"], rtol=1e-7, atol=1e-9)
    ee, en = wf._jastrow_terms()
    if ee is not None:
        torch.testing.assert_close(ee.jastrow_kernel.weight.grad.cpu(), go["w_ee"], rtol=1e-7, atol=1e-9)
    if en is not None:
        torch.testing.assert_close(en.jastrow_kernel.weight.grad.cpu(), go["w_en"], rtol=1e-7, atol=1e-9)


def test_local_energy_adjoint_vs_central_differences():
    """d<E_L>/d theta on fixed positions vs central finite differences, h = 1e-5."""
    _say("case eloc adjoint vs finite differences")
    mol, kw = helpers.CASES["h2o_ccpvdz_ee_en"]()
    wf, _ = helpers.make_pair(mol, **kw)
    wf.ao.bas_exp.requires_grad_(False)
    x = helpers.positions(mol, 256, seed=10).to(DEV)
    wf.zero_grad()
    wf.local_energy_autodiff(x).mean().backward()
    h = 1e-5
    ee, en = wf._jastrow_terms()
    checks = [(ee.jastrow_kernel.weight, (0,)), (en.jastrow_kernel.weight, (0,)),
              (wf.mo.mo_modifier, (0, 0)), (wf.mo.mo_modifier, (7, 3)), (wf.mo.mo_modifier, (20, 4))]
    for prm, idx in checks:
        with torch.no_grad():
            v0 = prm[idx].item()
            prm[idx] = v0 + h
            ep = wf.local_energy(x).mean().item()
            prm[idx] = v0 - h
            em = wf.local_energy(x).mean().item()
            prm[idx] = v0
        fd = (ep - em) / (2 * h)
        ad = prm.grad[idx].item()
        _say("  fd %.8e adjoint %.8e" % (fd, ad))
        assert abs(fd - ad) <= 1e-5 * max(1.0, abs(ad))


def test_gradients_jacobi_vs_oracle():
    _say("case gradients_jacobi")
    mol, kw = helpers.CASES["lih_631g_cas22_ee_en"]()
    wf, orc = helpers.make_pair(mol, **kw)
    pos = helpers.positions(mol, 33, seed=12)
    g = wf.gradients_jacobi(pos.to(DEV)).cpu()
    g_o = orc.gradients_jacobi(pos.clone().requires_grad_()).detach()
    torch.testing.assert_close(g, g_o, rtol=1e-9, atol=1e-12)
    gp = wf.gradients_jacobi(pos.to(DEV), pdf=True).cpu()
    gp_o = orc.gradients_jacobi(pos.clone().requires_grad_(), pdf=True).detach()
    torch.testing.assert_close(gp, gp_o, rtol=1e-9, atol=1e-12)


def test_solver_run_auto_gradient():
    _say("case solver auto")
    from qmctorch_amd.solver import Solver

    mol = systems.h2o()
    wf, _ = helpers.make_pair(mol, w_ee=1.0, w_en=1.0, perturb=0.0)
    s = _sampler(mol, 512, 50)
    opt = torch.optim.SGD(wf.parameters(), lr=1e-4)
    solver = Solver(wf=wf, sampler=s, optimizer=opt, output="/tmp/qmc_test_auto.hdf5")
    solver.configure(freeze=["ao", "ci"], grad="auto", track=["local_energy"])
    before = wf.mo.mo_modifier.detach().clone()
    out = solver.run(2)
    assert np.all(np.isfinite(out.energy))
    assert not torch.equal(before, wf.mo.mo_modifier.detach())


def test_excitation_plans_refuse_entry_points_without_that_path():
    """Plans with more than 8 unique configurations per spin run psi, pdf, the
    local energy, the sweep, the drift / HMC, both adjoints and the batched
    launches through the excitation kernels; the per-walker adjoint rows say
    that they have no such path."""
    _say("case excitation refusals")
    from qmctorch_amd.wavefunction.slater_jastrow import wf_adjoint_rows

    mol, kw = helpers.CASES["lih_631g_sd46_ee_en"]()
    wf, _ = helpers.make_pair(mol, **kw)
    pos = helpers.positions(mol, 8, seed=2).to(DEV)
    with pytest.raises(RuntimeError, match="excitation"):
        wf_adjoint_rows(wf, pos)