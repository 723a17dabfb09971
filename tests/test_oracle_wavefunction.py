"""Pins for the oracle's wavefunction, Jastrow and local-energy restatement.

Checks: autograd derivatives (the reference's own recipe,
tests/wavefunction/base_test_cases.py:92-121), closed-form energies, the ADF
cube grids (tests/wavefunction/orbitals/test_ao_values_adf.py:90-112,
test_mo_values_adf.py:74-101) and the golden (pos, E_L) pairs of
tests/hdf5/LiH_adf_dz_QMCTorch.hdf5.
"""
import numpy as np
import pytest
import torch

from oracle import vmc_oracle as O
from tests import systems


def _wf(mol, configs="ground_state", ee=1.0, en=None, ci=None, seed=0):
    return O.OracleSlaterJastrow(mol.basis, mol.atom_coords, mol.atomic_number, mol.nup, mol.ndown,
                                 configs=configs, jastrow_ee=ee, jastrow_en=en, ci_weights=ci)


def _pos(mol, n, seed=0, sigma=1.0):
    return torch.as_tensor(systems.gaussian_positions(mol, n, seed, sigma)).requires_grad_()


@pytest.mark.parametrize("which", ["h2o", "lih_adf"])
def test_ao_derivatives_vs_autograd(which):
    mol = systems.h2o() if which == "h2o" else systems.lih_adf()
    wf = _wf(mol, ee=None)
    pos = _pos(mol, 7, seed=3)
    ao, gao, lap = wf.ao(pos, "all")
    nb, ne, nao = ao.shape
    g_auto = torch.zeros(nb, ne, nao, 3, dtype=torch.float64)
    l_auto = torch.zeros(nb, ne, nao, dtype=torch.float64)
    for mu in range(nao):
        s = ao[:, :, mu].sum()
        g = torch.autograd.grad(s, pos, create_graph=True)[0].view(nb, ne, 3)
        g_auto[:, :, mu, :] = g.detach()
        for d in range(3):
            h = torch.autograd.grad(g[:, :, d].sum(), pos, retain_graph=True)[0].view(nb, ne, 3)
            l_auto[:, :, mu] += h[:, :, d].detach()
    assert torch.allclose(gao.detach(), g_auto, rtol=1e-10, atol=1e-12)
    assert torch.allclose(lap.detach(), l_auto, rtol=1e-9, atol=1e-10)


def _cube_points():
    x = torch.linspace(-1, 1, 21, dtype=torch.float64)
    xx, yy = torch.meshgrid((x, x), indexing="ij")
    return torch.cat((xx.reshape(-1, 1), yy.reshape(-1, 1), torch.zeros(441, 1, dtype=torch.float64)), 1)


def test_sto_ao_values_vs_adf_cube():
    mol = systems.c_adf()
    wf = _wf(mol, ee=None)
    pos = torch.zeros(441, mol.nelec * 3, dtype=torch.float64)
    pos[:, :3] = _cube_points()
    ao = wf.ao(pos).detach().numpy()
    cube = np.load(systems.GOLDEN + "/C_cube.npz")
    for iorb in range(mol.basis.nao):
        ref = cube["AO/%d" % iorb].reshape(21, 21)
        delta = np.abs(ref - ao[:, 0, iorb].reshape(21, 21))
        assert delta.mean() < 1e-3  # the reference's criterion


def test_mo_values_vs_adf_cube():
    mol = systems.c_adf()
    wf = _wf(mol, ee=None)
    pos = 10 * torch.ones(441, mol.nelec * 3, dtype=torch.float64)
    pos[:, :3] = _cube_points()
    mo = wf.mo(wf.ao(pos)).detach().numpy()
    cube = np.load(systems.GOLDEN + "/C_cube.npz")
    for iorb in range(mol.basis.nmo):
        ref = cube["MO/%d" % (iorb + 1)].reshape(21, 21) ** 2
        delta = np.sort(np.abs(ref - mo[:, 0, iorb].reshape(21, 21) ** 2).ravel())[:-1]
        assert delta.mean() < 1e-3


def test_hydrogen_atom_exact_eigenfunction():
    # 1s STO with zeta = 1 is the exact ground state: E_L = -1/2 everywhere
    mol = systems.hydrogen_atom("sto", 1.0)
    wf = _wf(mol, ee=None)
    pos = torch.as_tensor(np.random.default_rng(1).standard_normal((50, 3)))
    el = wf.local_energy(pos).detach()
    assert torch.allclose(el, torch.full_like(el, -0.5), atol=1e-12)


@pytest.mark.parametrize("alpha", [0.3, 1.1])
def test_single_gaussian_local_energy_closed_form(alpha):
    # normalised s Gaussian on a proton: E_L(r) = 3a - 2a^2 r^2 - 1/r
    mol = systems.hydrogen_atom("gto_pure", alpha)
    wf = _wf(mol, ee=None)
    pos = torch.as_tensor(np.random.default_rng(2).standard_normal((40, 3)))
    r2 = (pos ** 2).sum(1, keepdim=True)
    ref = 3 * alpha - 2 * alpha ** 2 * r2 - 1 / torch.sqrt(r2)
    assert torch.allclose(wf.local_energy(pos).detach(), ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("case", ["h2", "lih_cas", "h2o", "lih_adf_sd"])
def test_jacobi_kinetic_equals_autograd(case):
    if case == "h2":
        mol, wf = systems.h2(), None
        wf = _wf(mol)
    elif case == "lih_cas":
        mol = systems.lih()
        ci = np.array([1.0, 0.1, -0.2, 0.05])
        wf = _wf(mol, configs="cas(2,2)", en=1.0, ci=ci)
    elif case == "h2o":
        mol = systems.h2o()
        wf = _wf(mol, en=0.8)
    else:
        mol = systems.lih_adf()
        cfg = (torch.LongTensor([[0, 1], [0, 2], [0, 1], [0, 2]]), torch.LongTensor([[0, 1], [0, 1], [0, 2], [0, 2]]))
        wf = _wf(mol, configs=cfg, ee=0.7, ci=np.array([0.9, 0.2, 0.1, 0.4]))
    pos = _pos(mol, 9, seed=5)
    kj = wf.kinetic_energy_jacobi(pos).detach()
    ka = wf.kinetic_energy_autograd(pos).detach()
    assert torch.allclose(kj, ka, rtol=1e-9, atol=1e-9)


def test_gradients_jacobi_vs_autograd():
    mol = systems.lih()
    wf = _wf(mol, configs="cas(2,2)", en=1.0, ci=np.array([1.0, 0.3, -0.1, 0.2]))
    pos = _pos(mol, 6, seed=11)
    g = wf.gradients_jacobi(pos).detach()
    psi = wf(pos)
    ga = torch.autograd.grad(psi.sum(), pos)[0]
    assert torch.allclose(g, ga, rtol=1e-10, atol=1e-12)


def test_jastrow_derivatives_vs_autograd():
    mol = systems.h2o()
    wf = _wf(mol, ee=0.9, en=1.3)
    pos = _pos(mol, 5, seed=7)
    j, dj, d2j = wf.jastrow(pos, 2)
    nb = pos.shape[0]
    g = torch.autograd.grad(j.sum(), pos, create_graph=True)[0]
    assert torch.allclose(dj.detach(), g.detach().view(nb, mol.nelec, 3).transpose(1, 2), rtol=1e-10)
    lap = torch.zeros(nb, mol.nelec, dtype=torch.float64)
    for k in range(3 * mol.nelec):
        h = torch.autograd.grad(g[:, k].sum(), pos, retain_graph=True)[0]
        lap[:, k // 3] += h[:, k].detach()
    assert torch.allclose(d2j.detach(), lap, rtol=1e-9, atol=1e-12)


def test_ee_cusp_and_spin_weights():
    # d(ln J)/dr at r -> 0 equals w0: 1/2 antiparallel, 1/4 parallel (pade_jastrow_kernel.py:34-66)
    from types import SimpleNamespace

    for nup, ndown, w0 in [(1, 1, 0.5), (2, 0, 0.25)]:
        jee = O.JastrowEE(nup, ndown, torch.as_tensor([1.0], dtype=torch.float64))
        h = 1e-6
        p0 = torch.tensor([[0, 0, 0, 0, 0, 1e-9]], dtype=torch.float64)
        p1 = torch.tensor([[0, 0, 0, 0, 0, 1e-9 + h]], dtype=torch.float64)
        slope = (torch.log(jee(p1)) - torch.log(jee(p0))).item() / h
        assert abs(slope - w0) < 1e-5


def test_permutation_and_antisymmetry():
    mol = systems.h2o()
    wf = _wf(mol, en=1.0)
    pos = _pos(mol, 4, seed=9).detach()
    x = pos.view(4, mol.nelec, 3)
    for a, b in [(0, 1), (5, 7)]:
        perm = list(range(mol.nelec))
        perm[a], perm[b] = b, a
        xp = x[:, perm, :].reshape(4, -1)
        assert torch.allclose(wf.jastrow(xp), wf.jastrow(pos), rtol=1e-13)
        assert torch.allclose(wf(xp), -wf(pos), rtol=1e-10)


def test_small_determinants_match_leibniz():
    mol = systems.lih()
    wf = _wf(mol, configs="cas(2,2)", ee=None, ci=np.array([1.0, 0.5, 0.25, 0.125]))
    pos = _pos(mol, 3, seed=4).detach()
    mo = wf.mo(wf.ao(pos)).detach()
    cu, cd = wf.configs
    psi = wf.pool(mo)
    for n in range(4):
        A = mo[:, :2, :][:, :, cu[n]]
        B = mo[:, 2:, :][:, :, cd[n]]
        du = A[:, 0, 0] * A[:, 1, 1] - A[:, 0, 1] * A[:, 1, 0]
        dd = B[:, 0, 0] * B[:, 1, 1] - B[:, 0, 1] * B[:, 1, 0]
        assert torch.allclose(psi[:, n], du * dd, rtol=1e-12)


def test_lih_golden_local_energy():
    """E_L recomputed at the reference's stored positions with the stored state."""
    g = np.load(systems.GOLDEN + "/LiH_golden.npz")
    mol = systems.lih_adf()
    p = lambda k: g["SolverSlaterJastrow/wf/" + k]  # noqa: E731
    cfg = (torch.as_tensor(p("configs_0")), torch.as_tensor(p("configs_1")))
    W = p("mo_scf/weight").T @ p("mo/weight").T
    wf = O.OracleSlaterJastrow(mol.basis, mol.atom_coords, mol.atomic_number, 2, 2, configs=cfg, mo_coeffs=W,
                               ci_weights=p("fc/weight"), jastrow_ee=float(p("jastrow/weight")[0]))
    pos = torch.as_tensor(g["single_point/pos"])
    el = wf.local_energy(pos).detach().numpy()
    ref = g["single_point/local_energy"]
    assert np.abs(el - ref).max() < 1e-9
    assert abs(el.mean() - float(g["single_point/energy"])) < 1e-10
    assert abs(el.var(ddof=1) - float(g["single_point/variance"])) < 1e-8
