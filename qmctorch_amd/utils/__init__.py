from .torch_utils import set_torch_double_precision, set_torch_single_precision, DataLoader
from .constants import ANGS2BOHR, BOHR2ANGS
from .logger import log

__all__ = [
    "set_torch_double_precision",
    "set_torch_single_precision",
    "DataLoader",
    "ANGS2BOHR",
    "BOHR2ANGS",
    "log",
]
