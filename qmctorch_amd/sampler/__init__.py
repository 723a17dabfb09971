from .sampler_base import SamplerBase
from .metropolis import Metropolis
from .walkers import Walkers

__all__ = ["SamplerBase", "Metropolis", "Walkers"]
