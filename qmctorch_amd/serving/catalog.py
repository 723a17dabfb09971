"""Molecules of the mixed-size VMC workload and their wavefunctions.

The set is {H2/STO-3G, LiH/6-31G, H2O/cc-pVDZ, N2/cc-pVDZ, C6H6/cc-pVDZ}
(geometries in bohr): H2 0.74 A on z, LiH 1.595 A on z, H2O O-H 0.9572 A and
104.52 deg, N2 1.098 A on z, benzene D6h with C-C 1.39 A and C-H 1.09 A.
Every wavefunction is a single determinant x Pade e-e x Pade e-n Jastrow
(weights 1.0), as SlaterJastrow with the reference's Pade kernels
(slater_jastrow.py:30-117, jastrows/elec_elec/kernels/pade_jastrow_kernel.py,
jastrows/elec_nuclei/kernels/pade_jastrow_kernel.py).
"""
from typing import Dict, List, Tuple

import numpy as np
import torch

from ..scf import Molecule

SYSTEM_NAMES = ("h2", "lih", "h2o", "n2", "c6h6")


def _benzene_atoms() -> str:
    at = []
    for k in range(6):
        t = np.pi / 3 * k
        at.append("C %.9f %.9f 0" % (2.626719 * np.cos(t), 2.626719 * np.sin(t)))
    for k in range(6):
        t = np.pi / 3 * k
        at.append("H %.9f %.9f 0" % (4.686520 * np.cos(t), 4.686520 * np.sin(t)))
    return "; ".join(at)


_GEOMETRY: Dict[str, Tuple[str, str]] = {
    "h2": ("H 0 0 -0.699198616; H 0 0 0.699198616", "sto-3g"),
    "lih": ("Li 0 0 0; H 0 0 3.014112952", "6-31g"),
    "h2o": ("O 0 0 0; H 1.430428706 1.107156964 0; H -1.430428706 1.107156964 0", "cc-pvdz"),
    "n2": ("N 0 0 -1.037459568; N 0 0 1.037459568", "cc-pvdz"),
    "c6h6": (_benzene_atoms(), "cc-pvdz"),
}


def molecule(name: str, **kw) -> Molecule:
    if name not in _GEOMETRY:
        raise KeyError("unknown system %r (known: %s)" % (name, ", ".join(SYSTEM_NAMES)))
    atom, basis = _GEOMETRY[name]
    return Molecule(atom, basis=basis, unit="bohr", name=name, **kw)


def wavefunction(name: str, precision: str = "mixed", w_ee: float = 1.0, w_en: float = 1.0, mol=None):
    """SlaterJastrow (ground state, e-e + e-n Pade) of one system on the device."""
    from ..wavefunction import SlaterJastrow
    from ..wavefunction.jastrows.elec_elec import JastrowFactor as JEE, PadeJastrowKernel as PEE
    from ..wavefunction.jastrows.elec_nuclei import JastrowFactor as JEN, PadeJastrowKernel as PEN

    mol = mol if mol is not None else molecule(name)
    jas = [JEE(mol, PEE, kernel_kwargs={"w": w_ee}, cuda=True), JEN(mol, PEN, kernel_kwargs={"w": w_en}, cuda=True)]
    return SlaterJastrow(mol, jastrow=jas, configs="ground_state", cuda=True, precision=precision)


def uneven_batch(nsys: int = 64, seed: int = 0, wmin: int = 256, wmax: int = 8192,
                 names=SYSTEM_NAMES) -> List[Tuple[str, int]]:
    """Seeded draw of (system, walker count): systems uniform over ``names``,
    walker counts round(exp(U(ln wmin, ln wmax)))."""
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(names), nsys)
    w = np.round(np.exp(rng.uniform(np.log(wmin), np.log(wmax), nsys))).astype(np.int64)
    return [(names[i], int(n)) for i, n in zip(idx, w)]


def _electron_centers(mol, device) -> torch.Tensor:
    """Nucleus (i mod Nat) of every electron i, kept on the device with the
    molecule: a host-to-device copy per request would wait for the stream and
    stall the serving loop behind the running tick."""
    key = str(torch.device(device))
    cache = mol.__dict__.setdefault("_electron_centers", {})
    c = cache.get(key)
    if c is None:
        at = np.asarray(mol.atom_coords, dtype=np.float64).reshape(-1, 3)
        c = torch.as_tensor(at[np.arange(mol.nelec) % len(at)], device=device)
        cache[key] = c
    return c


def initial_positions(mol, nwalkers: int, seed: int, sigma: float = 1.0, device="cuda") -> torch.Tensor:
    """Electron i ~ N(nucleus (i mod Nat), sigma^2 I), drawn on the device from
    a generator seeded with ``seed`` (independent of any other draw)."""
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    centers = _electron_centers(mol, device)
    z = torch.randn(nwalkers, mol.nelec, 3, generator=g, dtype=torch.float64, device=device)
    return (centers[None] + sigma * z).reshape(nwalkers, 3 * mol.nelec).contiguous()
