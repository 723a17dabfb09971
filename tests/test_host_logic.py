"""CPU checks of the host logic: configurations, norms, plan tables."""
import numpy as np
import pytest
import torch

from oracle import vmc_oracle as O
from qmctorch_amd.wavefunction.plan import SystemTables, atomic_orbital_norm
from qmctorch_amd.wavefunction.pooling.orbital_configurations import OrbitalConfigurations
from tests import systems


def test_single_double_configs_match_reference_fixture():
    # configs_0/configs_1 dumped by the reference for single_double(2,2) on LiH
    g = np.load(systems.GOLDEN + "/LiH_golden.npz")
    cu, cd = OrbitalConfigurations(systems.lih_adf()).get_configs("single_double(2,2)")
    assert np.array_equal(cu.numpy(), g["SolverSlaterJastrow/wf/configs_0"])
    assert np.array_equal(cd.numpy(), g["SolverSlaterJastrow/wf/configs_1"])


def test_cas_configs_match_oracle():
    mol = systems.lih()
    cu, cd = OrbitalConfigurations(mol).get_configs("cas(2,2)")
    ou, od = O.get_configs("cas(2,2)", mol.nup, mol.ndown, mol.basis.nmo)
    assert torch.equal(cu, ou) and torch.equal(cd, od)
    assert cu.tolist() == [[0, 1], [0, 1], [0, 2], [0, 2]] and cd.tolist() == [[0, 1], [0, 2], [0, 1], [0, 2]]


@pytest.mark.parametrize("which", ["h2o", "lih_adf", "c_adf"])
def test_product_norms_equal_oracle_norms(which):
    mol = {"h2o": systems.h2o, "lih_adf": systems.lih_adf, "c_adf": systems.c_adf}[which]()
    assert np.allclose(atomic_orbital_norm(mol.basis), O.atomic_orbital_norm(mol.basis).numpy(), rtol=1e-14)


def _tables(mol, ex=None):
    ex = np.asarray(mol.basis.bas_exp) if ex is None else ex
    return SystemTables(mol.basis, mol.atom_coords, mol.atomic_number, mol.nup, mol.ndown,
                        (np.arange(mol.nup)[None], np.arange(mol.ndown)[None]), mol.nup, True, True,
                        atomic_orbital_norm(mol.basis), ex)


def test_exponent_groups_h2o():
    t = _tables(systems.h2o())
    assert t.nexp_per_electron == 24  # O: 9 s + 4 p + 1 d, H: 4 s + 1 p
    # every reference entry appears exactly once in the term tables
    used = sorted(t.gao_src[t.gao_K != 0].tolist() + t.gao_src[(t.gao_K == 0) & (t.gao_src != 0)].tolist())
    assert set(np.arange(len(systems.h2o().basis.bas_exp))) >= set(used)
    assert t.gao_K.shape[1] == 12


def test_grouping_tracks_diverging_exponents():
    mol = systems.h2o()
    t = _tables(mol)
    ex = np.asarray(mol.basis.bas_exp).copy()
    assert t.grouping_valid(ex)
    # train one Cartesian component's exponent away from its siblings (the reference allows it)
    i = int(np.where(np.asarray(mol.basis.bas_kx) == 1)[0][0])
    ex[i] *= 1.01
    assert not t.grouping_valid(ex)
    t2 = _tables(mol, ex)
    assert t2.nexp_per_electron > t.nexp_per_electron


def test_vnn_h2o():
    t = _tables(systems.h2o())
    c = np.array(systems.h2o().atom_coords)
    z = [8, 1, 1]
    ref = sum(z[a] * z[b] / np.linalg.norm(c[a] - c[b]) for a in range(3) for b in range(a + 1, 3))
    assert abs(t.vnn - ref) < 1e-12
