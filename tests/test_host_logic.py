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


def test_poisson_requests_are_seeded_and_within_ranges():
    from qmctorch_amd.serving import poisson_requests

    a = poisson_requests(500.0, 4000, seed=7)
    b = poisson_requests(500.0, 4000, seed=7)
    assert [(r.system, r.nwalkers, r.nsweeps, r.arrival) for r in a] == \
        [(r.system, r.nwalkers, r.nsweeps, r.arrival) for r in b]
    t = np.array([r.arrival for r in a])
    assert np.all(np.diff(t) > 0)
    assert abs(len(a) / t[-1] - 500.0) < 30.0  # mean rate of the Poisson process
    w = np.array([r.nwalkers for r in a])
    k = np.array([r.nsweeps for r in a])
    assert w.min() >= 128 and w.max() <= 16384 and k.min() == 1 and k.max() == 20
    # log-uniform walker counts: the median of exp(U(ln a, ln b)) is sqrt(a b)
    assert abs(np.log(np.median(w)) - 0.5 * np.log(128 * 16384)) < 0.1
    assert {r.system for r in a} == {"h2", "lih", "h2o", "n2", "c6h6"}


def test_uneven_batch_draw():
    from qmctorch_amd.serving import catalog

    jobs = catalog.uneven_batch(64, seed=0)
    assert len(jobs) == 64 and jobs == catalog.uneven_batch(64, seed=0)
    assert all(256 <= n <= 8192 for _, n in jobs)
    assert {s for s, _ in jobs} <= set(catalog.SYSTEM_NAMES)


def test_latency_summary():
    from qmctorch_amd.serving.server import VMCResult, latency_summary

    res = [VMCResult(i, "h2", 10, 2, 0.0, 0.0, 0.0, 0.5, np.zeros(10), 0.0, 0.0, (i + 1) * 1e-3)
           for i in range(100)]
    s = latency_summary(res, 2.0)
    assert abs(s["p50_ms"] - 50.5) < 1e-9 and abs(s["p99_ms"] - 99.01) < 1e-9
    assert s["walker_steps_per_s"] == 100 * 20 / 2.0


def test_serving_catalog_molecules():
    from qmctorch_amd.serving import catalog

    counts = {n: (catalog.molecule(n).nelec, catalog.molecule(n).basis.nao) for n in catalog.SYSTEM_NAMES}
    assert counts == {"h2": (2, 2), "lih": (4, 11), "h2o": (10, 25), "n2": (14, 30), "c6h6": (42, 120)}
