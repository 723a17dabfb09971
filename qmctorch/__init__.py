"""Import alias: ``import qmctorch`` resolves to the MI355X implementation.

Scripts written for the reference (``from qmctorch.scf import Molecule``,
``from qmctorch.wavefunction import SlaterJastrow``, ``from qmctorch.sampler
import Metropolis``, ``from qmctorch.solver import Solver``,
``from qmctorch.utils import set_torch_double_precision``, the Jastrow import
paths) run unchanged against qmctorch_amd.
"""
import importlib
import sys

import qmctorch_amd as _impl

__version__ = _impl.__version__
log = _impl.log

_SUBMODULES = [
    "scf", "scf.molecule", "wavefunction", "wavefunction.wf_base", "wavefunction.slater_jastrow",
    "wavefunction.orbitals", "wavefunction.orbitals.atomic_orbitals", "wavefunction.orbitals.molecular_orbitals",
    "wavefunction.jastrows", "wavefunction.jastrows.elec_elec", "wavefunction.jastrows.elec_elec.kernels",
    "wavefunction.jastrows.elec_nuclei", "wavefunction.jastrows.elec_nuclei.kernels",
    "wavefunction.jastrows.combine_jastrow", "wavefunction.pooling", "wavefunction.pooling.orbital_configurations",
    "wavefunction.pooling.slater_pooling", "sampler", "sampler.metropolis", "sampler.walkers",
    "sampler.sampler_base", "solver", "solver.solver", "solver.solver_base", "solver.loss", "utils",
    "utils.torch_utils", "utils.constants",
]
for _name in _SUBMODULES:
    _mod = importlib.import_module("qmctorch_amd." + _name)
    sys.modules["qmctorch." + _name] = _mod
    _parent, _, _leaf = ("qmctorch." + _name).rpartition(".")
    setattr(sys.modules[_parent], _leaf, _mod)
