from .sampler_base import SamplerBase
from .metropolis import Metropolis
from .hamiltonian import Hamiltonian
from .walkers import Walkers

__all__ = ["SamplerBase", "Metropolis", "Hamiltonian", "Walkers"]
