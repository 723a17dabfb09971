from .molecule import Molecule

__all__ = ["Molecule"]
