"""One-electron Metropolis moves (move type "one-elec", metropolis.py:228-250)
in the device sweep (needs an MI355X).

Bars: on a nearly flat density every step moves exactly one electron of each
walker, each electron equally often (within 5 standard errors) and by the
proposal's variance; a walker's chain does not depend on the batch
(bit-exact); the H2 energy of a one-electron chain agrees with the
all-electron chain within 5 standard errors.
"""
import math
import sys
from types import SimpleNamespace

import numpy as np
import pytest
import torch

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


def _two_electrons_one_gaussian(alpha):
    """Helium-like namespace: two electrons (one per spin) in one s Gaussian."""
    mol = SimpleNamespace()
    mol.atoms = np.array(["He"])
    mol.atom_coords = [[0.0, 0.0, 0.0]]
    mol.atomic_number = [2]
    mol.atomic_nelec = [2]
    mol.natom, mol.nelec, mol.nup, mol.ndown, mol.spin = 1, 2, 1, 1, 0
    mol.name = "He"
    mol.hdf5file = "He_atom.hdf5"
    mol.basis = systems.single_function_basis("gto_pure", alpha)
    mol.domain = lambda method: {"method": "normal", "mean": np.zeros(3), "sigma": np.eye(3)}
    return mol


def _metropolis(mol, nw, nstep, move, seed, step=0.2, ntherm=-1, ndecor=1):
    from qmctorch_amd.sampler import Metropolis

    return Metropolis(nwalkers=nw, nstep=nstep, step_size=step, ntherm=ntherm, ndecor=ndecor, nelec=mol.nelec,
                      init=mol.domain("normal"), move={"type": move, "proba": "normal"}, cuda=True, seed=seed)


def test_one_electron_moves_on_a_flat_density():
    _say("case one-elec proposal")
    from qmctorch_amd.wavefunction import SlaterJastrow

    mol = _two_electrons_one_gaussian(1e-9)
    wf = SlaterJastrow(mol, jastrow=None, cuda=True)
    nw = 1 << 16
    s = _metropolis(mol, nw, 1, "one-elec", seed=41)
    x0 = torch.zeros(nw, 6, dtype=torch.float64, device=DEV)
    d = s(wf.pdf, pos=x0.clone(), with_tqdm=False).detach().view(nw, 2, 3)
    assert s.last_acceptance > 0.9999
    moved = (d != 0).any(-1)  # (nw, 2)
    nmoved = moved.sum(-1)
    assert int(nmoved.max()) == 1
    assert float((nmoved == 1).double().mean()) > 0.9999
    f0 = float(moved[:, 0].double().mean())
    _say("  electron 0 moved in %.4f of the walkers" % f0)
    assert abs(f0 - 0.5) < 5 * math.sqrt(0.25 / nw)
    var = 0.2 / (2 * math.sqrt(2 * math.log(2.0)))
    disp = d[moved]  # (n, 3): the displacement of the moved electron
    v = disp.var(0)
    assert torch.all((v - var).abs() < 5 * var * math.sqrt(2.0 / len(disp)))


def test_one_electron_chain_does_not_depend_on_the_batch():
    _say("case one-elec batch invariance")
    mol = systems.h2o()
    wf, _ = helpers.make_pair(mol, w_ee=1.0, w_en=1.0, precision="mixed")
    p0 = helpers.positions(mol, 96, seed=12).to(DEV)
    a = _metropolis(mol, 96, 30, "one-elec", seed=13)(wf.pdf, pos=p0.clone(), with_tqdm=False).detach()
    b = _metropolis(mol, 40, 30, "one-elec", seed=13)(wf.pdf, pos=p0[:40].clone(), with_tqdm=False).detach()
    assert torch.equal(a[:40], b)
    # per step at most one electron of a walker changed position
    c = _metropolis(mol, 96, 1, "one-elec", seed=14)(wf.pdf, pos=p0.clone(), with_tqdm=False).detach()
    changed = (c.view(96, mol.nelec, 3) != p0.view(96, mol.nelec, 3)).any(-1).sum(-1)
    assert int(changed.max()) <= 1


def test_one_electron_energy_matches_all_electron():
    _say("case one-elec energy")
    mol = systems.h2()
    wf, _ = helpers.make_pair(mol, w_ee=1.0, perturb=0.0, precision="mixed")
    nw = 2000
    init = helpers.positions(mol, nw, seed=15).to(DEV)
    p1 = _metropolis(mol, nw, 800, "one-elec", seed=16, ntherm=200, ndecor=60)(wf.pdf, pos=init.clone(),
                                                                                with_tqdm=False)
    pa = _metropolis(mol, nw, 400, "all-elec", seed=17, ntherm=100, ndecor=30)(wf.pdf, pos=init.clone(),
                                                                               with_tqdm=False)
    e1 = wf.local_energy(p1.detach()).view(-1)
    ea = wf.local_energy(pa.detach()).view(-1)
    se = math.sqrt(float(e1.var()) / len(e1) + float(ea.var()) / len(ea))
    _say("  one-elec %.5f  all-elec %.5f  se %.5f" % (e1.mean(), ea.mean(), se))
    assert abs(float(e1.mean() - ea.mean())) < 5 * se

