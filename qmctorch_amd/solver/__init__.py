from .solver_base import SolverBase
from .solver import Solver

__all__ = ["SolverBase", "Solver"]
