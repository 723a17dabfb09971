"""Molecule front end: geometry, electron counts, basis namespace, MOs.

Same constructor and attributes as qmctorch/scf/molecule.py:22-162 and the same
``domain()`` contract (:176-214), so the walker initialisation and every caller
of the reference keep working.  Differences, all forced by the offline image:

* no SCF program (pyscf/ADF are absent): the basis comes from the published
  basis sets in ``basis_data`` laid out as pyscf.py:104-252 lays them out, and
  the MO coefficients come from this package's RHF (``rhf.py``, integrals in
  ``integrals.py``) for closed-shell molecules up to RHF_MAX_NAO AOs, in the
  pure-d space as pyscf's spherical SCF, expressed on the AO layer's Cartesian
  functions; otherwise from a seeded random orthonormal matrix
  (``mo_coeffs='random'``), or from the caller (``mo_coeffs=<array>``);
* ``load=`` reads ``.npz`` files converted from the reference's HDF5 molecule
  caches (``tests/golden``), since h5py is not installed for this interpreter.
"""
import os
from types import SimpleNamespace
from typing import Dict, List, Optional, Union

import numpy as np

from .. import log
from ..utils.constants import ANGS2BOHR
from .basis import build_basis, normalize_columns, random_orthonormal_mos

ATOMIC_NUMBER = {
    "H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9,
    "Ne": 10, "Na": 11, "Mg": 12, "Al": 13, "Si": 14, "P": 15, "S": 16, "Cl": 17,
    "Ar": 18,
}


# largest AO count for which Molecule runs its own RHF by default (the integral
# code is numpy; larger systems get seeded random orthonormal orbitals)
RHF_MAX_NAO = 64
_RHF_CACHE = {}


def _cached_rhf(mol, basis):
    """RHF of the molecule in the AO layer's basis (pure d space), cached per
    process by geometry, basis and electron count."""
    from .rhf import pure_d_transform, rhf

    key = (tuple(mol.atoms), np.asarray(mol.atom_coords, dtype=np.float64).round(12).tobytes(),
           mol.basis_name, mol.nelec)
    if key not in _RHF_CACHE:
        log.info("  Running RHF (own integrals, {0} AOs)", basis.nao)
        _RHF_CACHE[key] = rhf(basis, mol.atom_coords, mol.atomic_number, mol.nelec,
                              ao_transform=pure_d_transform(basis))
    e, C = _RHF_CACHE[key]
    return e, C.copy()


class Molecule:
    def __init__(
        self,
        atom: Optional[str] = None,
        calculator: str = "pyscf",
        scf: str = "hf",
        basis: str = "sto-3g",
        unit: str = "bohr",
        charge: int = 0,
        spin: int = 0,
        name: Optional[str] = None,
        load: Optional[str] = None,
        save_scf_file: bool = False,
        redo_scf: bool = False,
        rank: int = 0,
        mpi_size: int = 0,
        mo_coeffs: Union[str, np.ndarray] = "auto",
        mo_seed: int = 0,
    ) -> None:
        self.atom_coords: list = []
        self.atomic_nelec: list = []
        self.atomic_number: list = []
        self.atoms: list = []
        self.atoms_str = atom
        self.hdf5file = None
        self.max_angular = 2
        self.name = name
        self.natom = 0
        self.ndown = 0
        self.nelec = 0
        self.nup = 0
        self.charge = charge
        self.spin = spin
        self.unit = unit
        self.basis = SimpleNamespace()
        self.calculator_name = calculator
        self.basis_name = basis
        self.save_scf_file = save_scf_file
        self.scf_level = scf
        self.mo_source = None

        if load is not None:
            log.info("  Loading data from {file}", file=load)
            if str(load).endswith(".npz"):
                self._load_npz(load)
            else:
                self._load_hdf5(load)
            self.hdf5file = load
            return

        if self.unit not in ["angs", "bohr"]:
            raise ValueError("unit should be angs or bohr")
        self._process_atom_str()
        self.hdf5file = "_".join([self.name, calculator, basis]) + ".hdf5"
        self.basis = build_basis(list(self.atoms), self.atom_coords, basis)
        self._set_mos(mo_coeffs, mo_seed)
        self.log_data()

    # ------------------------------------------------------------------ setup
    def _set_mos(self, mo_coeffs, seed: int) -> None:
        b = self.basis
        if isinstance(mo_coeffs, str):
            mode = mo_coeffs
            if mode == "auto":
                closed = self.nelec % 2 == 0 and self.nup == self.ndown
                mode = "rhf" if closed and b.nao <= RHF_MAX_NAO else "random"
            if mode == "rhf":
                e, C = _cached_rhf(self, b)
                b.TotalEnergy = e
                b.mos = normalize_columns(C)
            elif mode == "random":
                b.mos = random_orthonormal_mos(b.nao, b.nmo, seed)
                b.TotalEnergy = 0.0
            else:
                raise ValueError("mo_coeffs must be 'auto', 'rhf', 'random' or an array")
            self.mo_source = mode
        else:
            mos = np.asarray(mo_coeffs, dtype=np.float64)
            if mos.shape[0] != b.nao:
                raise ValueError("mo_coeffs must have %d rows" % b.nao)
            b.mos = mos
            b.nmo = mos.shape[1]
            self.mo_source = "explicit"

    def _process_atom_str(self) -> None:
        if self.atoms_str.endswith(".xyz"):
            if not os.path.isfile(self.atoms_str):
                raise FileNotFoundError("File %s not  found" % self.atoms_str)
            with open(self.atoms_str) as f:
                data = f.readlines()
            natom = int(data[0])
            atoms = data[2 : 2 + natom]
        else:
            atoms = self.atoms_str.split(";")
        self._get_atomic_properties(atoms)

    def _get_atomic_properties(self, atoms: List[str]) -> None:
        """Atom list, coordinates in bohr, electron counts (molecule.py:229-266)."""
        for a in atoms:
            tok = a.split()
            if not tok:
                continue
            sym = tok[0]
            conv = ANGS2BOHR if self.unit == "angs" else 1.0
            self.atoms.append(sym)
            self.atom_coords.append([float(tok[1]) * conv, float(tok[2]) * conv, float(tok[3]) * conv])
            z = ATOMIC_NUMBER[sym]
            self.atomic_number.append(z)
            self.atomic_nelec.append(z)
            self.nelec += z
        self.nelec -= self.charge
        self.natom = len(self.atoms)
        if (self.nelec - self.spin) % 2 != 0:
            raise ValueError("%d electrons and spin %d doesn't make sense" % (self.nelec, self.spin))
        self.nup = int((self.nelec - self.spin) / 2) + self.spin
        self.ndown = int((self.nelec - self.spin) / 2)
        if self.name is None:
            self.name = self._get_mol_name(self.atoms)
        self.atoms = np.array(self.atoms)

    @staticmethod
    def _get_mol_name(atoms: List[str]) -> str:
        name = ""
        for ua in sorted(set(atoms), key=atoms.index):
            name += ua
            n = atoms.count(ua)
            if n > 1:
                name += str(n)
        return name

    def _load_hdf5(self, filename: str) -> None:
        """Molecule cache in the reference's HDF5 layout (molecule.py:394-424):
        group "molecule" loaded attribute by attribute, then the integer and
        string fields cast back from their stored types."""
        from ..utils.hdf5_utils import load_from_hdf5

        load_from_hdf5(self, filename, "molecule")
        self.atoms = np.asarray(self.atoms).astype("U")
        self.atomic_nelec = [int(z) for z in self.atomic_nelec]
        self.atomic_number = [int(z) for z in np.asarray(self.atomic_number).ravel()]
        self.atom_coords = np.asarray(self.atom_coords, dtype=np.float64).tolist()
        for k in ("nelec", "nup", "ndown", "natom"):
            setattr(self, k, int(getattr(self, k)))
        b = self.basis
        b.nao, b.nmo = int(b.nao), int(b.nmo)
        for k in ("radial_type", "harmonics_type"):
            if isinstance(getattr(b, k, None), (bytes, np.bytes_)):
                setattr(b, k, getattr(b, k).decode("utf-8"))
        self.mo_source = "file"

    def _load_npz(self, filename: str) -> None:
        """Molecule cache converted from the reference's HDF5 layout
        (molecule.py:394-424 reads the same fields)."""
        d = np.load(filename, allow_pickle=False)
        g = lambda k: d["molecule/" + k]  # noqa: E731
        self.atoms = np.array([str(a) for a in g("atoms")])
        self.atom_coords = g("atom_coords").tolist()
        self.atomic_number = [int(z) for z in g("atomic_number")]
        self.atomic_nelec = [int(z) for z in g("atomic_nelec")]
        self.natom = len(self.atoms)
        self.nelec = int(g("nelec"))
        self.nup = int(g("nup"))
        self.ndown = int(g("ndown"))
        self.name = str(g("name"))
        self.basis_name = str(g("basis_name"))
        self.calculator_name = str(g("calculator_name"))
        b = SimpleNamespace()
        for k in d.files:
            if k.startswith("molecule/basis/"):
                v = d[k]
                setattr(b, k[len("molecule/basis/"):], v.item() if v.ndim == 0 else v)
        b.radial_type = str(b.radial_type)
        b.harmonics_type = str(b.harmonics_type)
        b.nao = int(b.nao)
        b.nmo = int(b.nmo)
        self.basis = b
        self.mo_source = "file"

    # ----------------------------------------------------------------- public
    def log_data(self) -> None:
        log.info("  Molecule name       : {0}", self.name)
        log.info("  Number of electrons : {0}", self.nelec)
        log.info("  SCF calculator      : {0}", self.calculator_name)
        log.info("  Basis set           : {0}", self.basis_name)
        log.info("  MO coefficients     : {0}", self.mo_source)
        log.info("  Number of AOs       : {0}", self.basis.nao)
        log.info("  Number of MOs       : {0}", self.basis.nmo)

    def get_total_energy(self) -> float:
        return float(getattr(self.basis, "TotalEnergy", 0.0))

    def print_total_energy(self) -> None:
        log.info("== SCF Energy : {e}", e=self.get_total_energy())

    def domain(self, method: str) -> Dict:
        """Walker initialisation domain (molecule.py:176-214)."""
        domain = {"method": method}
        if method == "center":
            domain["center"] = np.mean(self.atom_coords, 0)
        elif method == "uniform":
            domain["min"] = np.min(self.atom_coords) - 0.5
            domain["max"] = np.max(self.atom_coords) + 0.5
        elif method == "normal":
            domain["mean"] = np.mean(self.atom_coords, 0)
            domain["sigma"] = np.diag(np.std(self.atom_coords, 0) + 0.25)
        elif method == "atomic":
            domain["atom_coords"] = self.atom_coords
            domain["atom_num"] = self.atomic_number
            domain["atom_nelec"] = self.atomic_nelec
        else:
            raise ValueError("Method to initialize the walkers not recognized")
        return domain
